'use strict';
// WASM-shaped facade: lays out ke / edofs / evals / active / p in the
// facade's WebAssembly.Memory the way the reference's worker does before
// calling the matrix-ops exports, calls applyAEbe3D, ebePCG and (when
// argv[2].bddc is given) ebeKSP_BDDC by byte offset, and prints the results
// (argv[2] = {n, evals, p, dens, F, free, bddc?}).
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

// ebeKSP_BDDC with the subdomain descriptors laid out like
// _solveWithWasmKSP_BDDC (js/optimizer-worker-3d.js:3939-3962)
let bddc = null;
if (inp.bddc) {
    const B = inp.bddc;
    const ns = B.subs.length;
    const subsPtr = alloc(32 * ns);
    const dv = new DataView(wasm.memory.buffer);
    for (let si = 0; si < ns; si++) {
        const [el, ds, li] = B.subs[si];
        const eo = alloc(el.length * 4), d0 = alloc(ds.length * 4), lo = alloc(ndof * 8);
        new Int32Array(wasm.memory.buffer, eo, el.length).set(el);
        new Int32Array(wasm.memory.buffer, d0, ds.length).set(ds);
        new Float64Array(wasm.memory.buffer, lo, ndof).set(li);
        const o = subsPtr + 32 * si;
        dv.setInt32(o, el.length, true); dv.setInt32(o + 4, ds.length, true);
        dv.setUint32(o + 8, eo, true); dv.setUint32(o + 12, d0, true); dv.setUint32(o + 16, lo, true);
    }
    const cPtr = alloc(B.coarse.length * 4), e2 = alloc(nel * 8), a2 = alloc(B.active.length * 4);
    const fxPtr = alloc(ndof), ivPtr = alloc(ndof * 8), f2 = alloc(ndof * 8), u2 = alloc(ndof * 8);
    new Int32Array(wasm.memory.buffer, cPtr, B.coarse.length).set(B.coarse);
    new Float64Array(wasm.memory.buffer, e2, nel).set(B.evals);
    new Int32Array(wasm.memory.buffer, a2, B.active.length).set(B.active);
    new Uint8Array(wasm.memory.buffer, fxPtr, ndof).set(B.fixed);
    new Float64Array(wasm.memory.buffer, ivPtr, ndof).set(B.invD);
    new Float64Array(wasm.memory.buffer, f2, ndof).set(B.F);
    const it2 = wasm.ebeKSP_BDDC(subsPtr, ns, cPtr, B.coarse.length, kePtr, edPtr, e2, a2, B.active.length,
        fxPtr, ivPtr, f2, u2, ndof, B.maxIter, B.tol, 3, 0);
    bddc = { iters: it2, U: Array.from(new Float64Array(wasm.memory.buffer, u2, ndof)) };
}
// ebePCG on the 2-D quad4 grid (edofSize 8, the use of
// tests/test-wasm-ebePCG.js) when argv[2].twod = {nelx, nely, dens, F, free}
let twod = null;
if (inp.twod) {
    const D = inp.twod;
    const nel2 = D.nelx * D.nely, ndof2 = 2 * (D.nelx + 1) * (D.nely + 1);
    const k2 = alloc(64 * 8), ed2 = alloc(nel2 * 8 * 4), d2 = alloc(nel2 * 8), f2 = alloc(ndof2 * 8);
    const u2 = alloc(ndof2 * 8), fr2 = alloc(D.free.length * 4);
    addon.k0_2d(0.3, new Float64Array(wasm.memory.buffer, k2, 64));
    addon.edofs2d(D.nelx, D.nely, new Int32Array(wasm.memory.buffer, ed2, nel2 * 8));
    new Float64Array(wasm.memory.buffer, d2, nel2).set(D.dens);
    new Float64Array(wasm.memory.buffer, f2, ndof2).set(D.F);
    new Int32Array(wasm.memory.buffer, fr2, D.free.length).set(D.free);
    const it = wasm.ebePCG(d2, k2, ed2, f2, u2, fr2, nel2, 8, ndof2, D.free.length, 1e-9, 1.0, 3.0, 2000, 1e-8, 0);
    twod = { iters: it, U: Array.from(new Float64Array(wasm.memory.buffer, u2, ndof2)) };
}
process.stdout.write(JSON.stringify({ Ap, iters, U, activeCount: cnt, bddc, twod }) + '\n');
