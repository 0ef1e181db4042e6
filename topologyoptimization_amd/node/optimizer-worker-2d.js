'use strict';
/**
 * 2-D worker of TopologySolver: the message protocol of the reference's
 * js/optimizer-worker.js (:2150-2173 in; 'progress' :724-743, 'paused'
 * :598, 'cancelled' :580, 'complete' :775-800 and the FEA-only 'complete'
 * :456-484 out) over the device-resident 2-D solver of libtopo_hip. The SIMP
 * loop (:583-700) runs on the GPU one iteration per solver2dStep(); between
 * iterations the worker yields to its event loop so pause / resume /
 * updateConfig / cancel are honoured at iteration boundaries. The AMR
 * grouping (useAMR), void filling (preventVoids), manufacturing constraints
 * and the genetic optimizer of the 2-D worker are not part of this path and
 * answer with an 'error' message naming the option.
 */
const path = require('path');
const { parentPort } = require('worker_threads');
const addon = require(path.join(__dirname, 'topo_addon.node'));
const { buildProblem2d, toTopoConfig2d } = require('./problem.js');

const FIELD_STRESS = 5;
let state = { cancelled: false, paused: false, resume: null, pending: null };

function post(msg, transfer) { parentPort.postMessage(msg, transfer); }

const yieldToLoop = () => new Promise((r) => setImmediate(r));

/** densities in the importer's row-major layout, repeated over nz layers
 * (:448-456 and :750-760) */
function densities3D(x, nelx, nely, nelz) {
    const d = new Float32Array(nelx * nely * nelz);
    for (let z = 0; z < nelz; z++)
        for (let y = 0; y < nely; y++)
            for (let xpos = 0; xpos < nelx; xpos++) d[xpos + y * nelx + z * nelx * nely] = x[y + xpos * nely];
    return d;
}

async function optimize(model, config) {
    if (config.optimizer === 'genetic') throw new Error("TopologySolver: the 'genetic' optimizer is not part of the MI355X SIMP path");
    for (const k of ['useAMR', 'preventVoids', 'manufacturingConstraint']) {
        if (config[k]) throw new Error(`TopologySolver: 2-D option '${k}' is not part of the MI355X SIMP path`);
    }
    const nelx = model.nx, nely = model.ny, nelz = model.nz;
    const nel = nelx * nely;
    const problem = buildProblem2d(addon, model, config);
    const cfg = toTopoConfig2d(addon, model, config);
    const h = addon.createSolver2d(cfg, problem.fixedMask, problem.F, problem.preserved, problem.voidMask);

    if (config.feaOnly) {
        // FEA-only (:384-485): model densities (solid > 0.5, else Emin), tolerance 1e-8
        const x = new Float32Array(nel);
        if (model.elements) {
            for (let ex = 0; ex < nelx; ex++)
                for (let ey = 0; ey < nely; ey++) x[ey + ex * nely] = model.elements[ex + ey * nelx] > 0.5 ? 1.0 : cfg.Emin;
        } else {
            x.fill(1.0);
        }
        const penal = config.penaltyFactor || 3;
        const t0 = Date.now();
        const info = addon.solver2dFea(h, x, penal, 1e-8);
        const elementStress = new Float32Array(nel);
        addon.solver2dGetField(h, FIELD_STRESS, elementStress);
        const totalTime = Date.now() - t0;
        post({ type: 'complete', result: {
            densities: densities3D(x, nelx, nely, nelz), finalCompliance: info.compliance, iterations: 0,
            volumeFraction: 1.0, nx: nelx, ny: nely, nz: nelz, meshData: null, maxStress: info.maxStress,
            elementStress, feaOnly: true,
            timing: { totalTime, avgIterationTime: totalTime, iterationTimes: [totalTime], usingWasm: false, usingGPU: true },
        } });
        return;
    }

    if (problem.x0) addon.solver2dSetDesign(h, problem.x0);
    let volfrac = cfg.volfrac;
    let maxIterations = cfg.max_iterations;
    let loop = 0;
    let change = 1;
    let c = 0;
    let maxStress = 0;
    const iterationTimes = [];
    const startTime = Date.now();
    while (change > 0.01 && loop < maxIterations) {
        await yieldToLoop();
        if (state.cancelled) {
            post({ type: 'cancelled', iteration: loop });
            return;
        }
        if (state.pending) {
            const p = state.pending;
            state.pending = null;
            if (p.volumeFraction !== undefined) volfrac = p.volumeFraction;
            if (p.maxIterations !== undefined) maxIterations = p.maxIterations;
            addon.solver2dUpdate(h, p.penaltyFactor !== undefined ? p.penaltyFactor : -1,
                p.filterRadius !== undefined ? p.filterRadius : -1,
                p.volumeFraction !== undefined ? p.volumeFraction : -1,
                p.maxIterations !== undefined ? p.maxIterations : -1);
        }
        if (state.paused) {
            post({ type: 'paused', iteration: loop });
            await new Promise((r) => { state.resume = r; });
        }
        const info = addon.solver2dStep(h);
        loop = info.iteration;
        change = info.change;
        c = info.compliance;
        maxStress = info.maxStress;
        iterationTimes.push(info.totalMs);
        const avg = iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length;
        post({
            type: 'progress', iteration: loop, compliance: c, meshData: null, maxStress, penal: info.penal,
            beta: info.beta,
            timing: { iterationTime: info.totalMs, avgIterationTime: avg, elapsedTime: Date.now() - startTime,
                usingWasm: false, usingGPU: true },
            solverStats: { cgIterations: info.cgIterations, tolerance: info.tolerance, backend: 'hip',
                change: info.change, volume: info.volume },
        });
    }
    if (state.cancelled) {
        post({ type: 'cancelled', iteration: loop });
        return;
    }
    const x = new Float32Array(nel);
    addon.solver2dGetDesign(h, x);
    const totalTime = Date.now() - startTime;
    const avg = iterationTimes.length ? iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length : 0;
    const densities = densities3D(x, nelx, nely, nelz);
    post({ type: 'complete', result: {
        densities, finalCompliance: c, iterations: loop, volumeFraction: volfrac, nx: nelx, ny: nely, nz: nelz,
        meshData: null, maxStress, amrStats: null, designVariables: x,
        timing: { totalTime, avgIterationTime: avg, iterationTimes, usingWasm: false, usingGPU: true },
    } }, [densities.buffer, x.buffer]);
}

parentPort.on('message', (data) => {
    switch (data.type) {
    case 'start':
        state = { cancelled: false, paused: false, resume: null, pending: null };
        optimize(data.model, data.config).catch((err) => post({ type: 'error', message: String(err && err.message || err) }));
        break;
    case 'cancel':
        state.cancelled = true;
        if (state.resume) { const r = state.resume; state.resume = null; r(); }
        break;
    case 'pause':
        state.paused = true;
        break;
    case 'resume':
        state.paused = false;
        if (state.resume) { const r = state.resume; state.resume = null; r(); }
        break;
    case 'updateConfig':
        // the 2-D worker keeps the latest patch only (:2169-2171)
        state.pending = data.config || null;
        break;
    default:
        break;
    }
});
