'use strict';
/**
 * WASM-shaped facade of the HIP library: the same object shape as the
 * reference's instantiated matrix-ops module exports -- `memory` (a real
 * WebAssembly.Memory) plus `applyAEbe3D`, `ebePCG`, `ebeMGPCG`,
 * `ebeKSP_BDDC` taking byte offsets into that memory in the argument order
 * of assembly/index.ts (:403-413, :504-521, :1076-1087, :1410-1429). Call
 * sites such as js/optimizer-worker-3d.js:822-825, :3605-3609, :3783-3787
 * and :3972-3979 work unchanged against it; the arrays they wrote into
 * `memory` are handed to the C ABI without copies.
 */
const path = require('path');
const addon = require(path.join(__dirname, 'topo_addon.node'));

// dims of a structured hex8 edofs array (reference numbering)
function gridOf(edofs, ndof) {
    const nnz = edofs[9] / 3;
    const nnyz = edofs[3] / 3;
    const nny = nnyz / nnz;
    const nnx = ndof / 3 / nnyz;
    return { nelx: nnx - 1, nely: nny - 1, nelz: nnz - 1, nel: (nnx - 1) * (nny - 1) * (nnz - 1) };
}

function createWasmFacade(initialPages) {
    const memory = new WebAssembly.Memory({ initial: initialPages || 256, maximum: 65536 });
    const f64 = (off, n) => new Float64Array(memory.buffer, off, n);
    const i32 = (off, n) => new Int32Array(memory.buffer, off, n);
    const u8 = (off, n) => new Uint8Array(memory.buffer, off, n);

    function applyAEbe3D(kePtr, edofsPtr, evalsPtr, activePtr, activeCount, pPtr, apPtr, ndof, scratchPtr) {
        const head = i32(edofsPtr, 24);
        const g = gridOf(head, ndof);
        addon.applyAEbe3D(f64(kePtr, 576), i32(edofsPtr, g.nel * 24), f64(evalsPtr, g.nel),
            i32(activePtr, Math.max(activeCount, 0)), activeCount, f64(pPtr, ndof), f64(apPtr, ndof), ndof);
    }

    function ebePCG(densPtr, kePtr, edofsPtr, fPtr, uPtr, freedofsPtr, nel, edofSize, ndof, nfree,
        Emin, E0, penal, maxIter, tolerance, workPtr) {
        return addon.ebePCG(f64(densPtr, nel), f64(kePtr, edofSize * edofSize), i32(edofsPtr, nel * edofSize),
            f64(fPtr, ndof), f64(uPtr, ndof), i32(freedofsPtr, nfree), nel, edofSize, ndof, nfree,
            Emin, E0, penal, maxIter, tolerance);
    }

    // level descriptor: 20 x i32 at stride 80 (assembly/index.ts:743-757)
    function ebeMGPCG(levelsPtr, numLevels, kePtr, fPtr, uPtr, fixedMask0, ndof0, maxIter, tolerance, cgWorkPtr) {
        const levels = [];
        for (let li = 0; li < numLevels; li++) {
            const d = i32(levelsPtr + 80 * li, 20);
            const u32 = new Uint32Array(memory.buffer, levelsPtr + 80 * li, 20);
            const lv = {
                nelx: d[0], nely: d[1], nelz: d[2], ndof: d[3], nel: d[4], nfree: d[5],
                activeCount: d[6], hasDenseK: d[7],
            };
            lv.edofs = i32(u32[8], lv.nel * 24);
            lv.evals = f64(u32[9], lv.nel);
            lv.active = i32(u32[10], Math.max(lv.activeCount, 0));
            lv.fixedMask = u8(u32[11], lv.ndof);
            lv.freeDofs = i32(u32[12], lv.nfree);
            lv.invDiag = f64(u32[13], lv.ndof);
            lv.denseK = lv.hasDenseK ? f64(u32[14], lv.ndof * lv.ndof) : null;
            levels.push(lv);
        }
        return addon.ebeMGPCG(levels, f64(kePtr, 576), f64(fPtr, ndof0), f64(uPtr, ndof0), u8(fixedMask0, ndof0),
            ndof0, maxIter, tolerance);
    }

    // subdomain descriptor: BDDC_SUB_STRIDE = 32 bytes (assembly/index.ts:1204-1222)
    function ebeKSP_BDDC(subsPtr, numSubs, coarsePtr, numCoarse, kePtr, edofsPtr, evalsPtr, activePtr, activeCount,
        fixedMask, invDiagPtr, fPtr, uPtr, ndof, maxIter, tolerance, smootherIters, workPtr) {
        const g = gridOf(i32(edofsPtr, 24), ndof);
        const subs = [];
        for (let si = 0; si < numSubs; si++) {
            const d = i32(subsPtr + 32 * si, 8);
            const u32 = new Uint32Array(memory.buffer, subsPtr + 32 * si, 8);
            subs.push({
                numElements: d[0], numDofs: d[1], elements: i32(u32[2], Math.max(d[0], 0)),
                dofs: i32(u32[3], Math.max(d[1], 0)), localInvDiag: f64(u32[4], ndof),
            });
        }
        return addon.ebeKSP_BDDC(subs, i32(coarsePtr, Math.max(numCoarse, 0)), f64(kePtr, 576),
            i32(edofsPtr, g.nel * 24), f64(evalsPtr, g.nel), i32(activePtr, Math.max(activeCount, 0)),
            u8(fixedMask, ndof), f64(invDiagPtr, ndof), f64(fPtr, ndof), f64(uPtr, ndof), ndof, maxIter, tolerance,
            smootherIters);
    }

    return { memory, applyAEbe3D, ebePCG, ebeMGPCG, ebeKSP_BDDC };
}

module.exports = { createWasmFacade, gridOf };
