'use strict';
// WASM-shaped facade: lays out ke / edofs / evals / active / p in the
// facade's WebAssembly.Memory the way the reference's worker does before
// calling the matrix-ops exports, calls applyAEbe3D and ebePCG by byte
// offset, and prints the results (argv[2] = {n, evals, p, dens, F, free}).
const path = require('path');
const NODE = path.join(__dirname, '..', '..', 'topologyoptimization_amd', 'node');
const { createWasmFacade, addon } = require(NODE);

const inp = JSON.parse(process.argv[2]);
const n = inp.n;
const nel = n * n * n;
const ndof = 3 * (n + 1) ** 3;
const wasm = createWasmFacade(64);
let top = 1024;
const alloc = (bytes) => { const off = top; top += (bytes + 15) & ~15; return off; };
const kePtr = alloc(576 * 8), edPtr = alloc(nel * 24 * 4), evPtr = alloc(nel * 8), acPtr = alloc(nel * 4);
const pPtr = alloc(ndof * 8), apPtr = alloc(ndof * 8);
addon.k0(0.3, new Float64Array(wasm.memory.buffer, kePtr, 576));
addon.edofs(n, n, n, new Int32Array(wasm.memory.buffer, edPtr, nel * 24));
new Float64Array(wasm.memory.buffer, evPtr, nel).set(inp.evals);
const act = new Int32Array(wasm.memory.buffer, acPtr, nel);
let cnt = 0;
for (let e = 0; e < nel; e++) if (inp.evals[e] > 0) act[cnt++] = e;
new Float64Array(wasm.memory.buffer, pPtr, ndof).set(inp.p);
wasm.applyAEbe3D(kePtr, edPtr, evPtr, acPtr, cnt, pPtr, apPtr, ndof, 0);
const Ap = Array.from(new Float64Array(wasm.memory.buffer, apPtr, ndof));

const dPtr = alloc(nel * 8), fPtr = alloc(ndof * 8), uPtr = alloc(ndof * 8), frPtr = alloc(inp.free.length * 4);
new Float64Array(wasm.memory.buffer, dPtr, nel).set(inp.dens);
new Float64Array(wasm.memory.buffer, fPtr, ndof).set(inp.F);
new Int32Array(wasm.memory.buffer, frPtr, inp.free.length).set(inp.free);
const iters = wasm.ebePCG(dPtr, kePtr, edPtr, fPtr, uPtr, frPtr, nel, 24, ndof, inp.free.length,
    1e-9, 1.0, 3.0, 400, 1e-8, 0);
const U = Array.from(new Float64Array(wasm.memory.buffer, uPtr, ndof));
process.stdout.write(JSON.stringify({ Ap, iters, U, activeCount: cnt }) + '\n');
