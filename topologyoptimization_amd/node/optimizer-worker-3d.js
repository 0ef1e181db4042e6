'use strict';
/**
 * Worker of TopologySolver: the message protocol of the reference's
 * js/optimizer-worker-3d.js (:4805-4826 in, :2079/:2118/:2283/:2323/:1599
 * out) over the device-resident solver of libtopo_hip. The SIMP loop
 * (:2074-2252) runs on the GPU one iteration per solverStep(); between
 * iterations the worker yields to its event loop so pause / resume /
 * updateConfig / cancel messages are honoured at iteration boundaries, as in
 * the reference.
 */
const path = require('path');
const { parentPort } = require('worker_threads');
const addon = require(path.join(__dirname, 'topo_addon.node'));
const { buildProblem, toTopoConfig } = require('./problem.js');


let state = { cancelled: false, paused: false, resume: null, pending: null };

function post(msg, transfer) { parentPort.postMessage(msg, transfer); }

const FIELD_STRESS = 5;

/* Volumetric snapshots (_buildVolumetricSnapshot :1579-1592 and the
 * volumetricOutputMode rules :1573-1577, :2116-2120, :2280-2289, :2319-2321):
 * the element stress field E_e w_e of the last iteration, read from the
 * device when a snapshot is due. */
const vol = { mode: 'on-stop', h: null, nel: 0, dims: null, iteration: 0, maxStress: 0 };

function normalizeVolumetricMode(mode) {
    if (mode === 'every-iteration') return 'every-iteration';
    if (mode === 'off') return 'off';
    return 'on-stop';
}

function volumetricSnapshot() {
    if (!vol.h || !vol.iteration) return null;
    const stress = new Float32Array(vol.nel);
    addon.solverGetField(vol.h, FIELD_STRESS, stress);
    return {
        payload: { nx: vol.dims.nx, ny: vol.dims.ny, nz: vol.dims.nz, iteration: vol.iteration,
            maxStress: vol.maxStress, stress },
        transfer: [stress.buffer],
    };
}

const yieldToLoop = () => new Promise((r) => setImmediate(r));

async function optimize(model, config) {
    const { nx, ny, nz } = model;
    const problem = buildProblem(addon, model, config);
    const cfg = toTopoConfig(addon, model, config);
    const nel = nx * ny * nz;
    const h = addon.createSolver(cfg, problem.fixedMask, problem.F, problem.preserved, problem.voidMask);
    if (problem.x0) addon.solverSetDesign(h, problem.x0);
    vol.mode = normalizeVolumetricMode(config.volumetricOutputMode);
    vol.h = h; vol.nel = nel; vol.dims = { nx, ny, nz }; vol.iteration = 0; vol.maxStress = 0;

    if (config.feaOnly) {
        // single FE analysis at the given densities (:1652-1883); tol -> 1e-8
        const x = new Float32Array(nel);
        for (let i = 0; i < nel; i++) {
            const v = model.elements ? model.elements[i] : 1;
            x[i] = v > 0 ? v : cfg.Emin;
        }
        const U = new Float64Array(problem.ndof);
        const t0 = Date.now();
        const info = addon.solverFe(h, x, cfg.penal, 1, 1, U);
        const w = new Float32Array(nel);
        addon.solverEnergy(h, U, w);
        const stress = new Float32Array(nel);
        let maxStress = 0;
        for (let i = 0; i < nel; i++) {
            const E = cfg.Emin + Math.pow(x[i], cfg.penal) * (cfg.E0 - cfg.Emin);
            stress[i] = E * w[i];
            if (stress[i] > maxStress) maxStress = stress[i];
        }
        post({ type: 'progress', iteration: 1, compliance: info.compliance, meshData: null,
            timing: { iterationTime: Date.now() - t0 }, maxStress, volumetricData: null });
        post({ type: 'complete', result: {
            densities: x, finalCompliance: info.compliance, iterations: 1, volumeFraction: cfg.volfrac,
            nx, ny, nz, meshData: null, maxStress, elementStress: stress, displacement: U, feaOnly: true,
            solverStats: { cgIterations: info.cgIterations, tolerance: info.tolerance },
            timing: { elapsedTime: Date.now() - t0, usingGPU: true, usingWasm: false },
        } });
        return;
    }

    let loop = 0;
    let change = 1;
    let c = 0;
    let maxIterations = cfg.max_iterations;
    let maxStress = 0;
    const iterationTimes = [];
    const start = Date.now();
    while (change > 0.01 && loop < maxIterations) {
        await yieldToLoop();
        if (state.cancelled) {
            post({ type: 'cancelled', iteration: loop });
            return;
        }
        if (state.pending) {
            const p = state.pending;
            state.pending = null;
            if (p.maxIterations !== undefined) maxIterations = p.maxIterations;
            if (p.volumetricOutputMode !== undefined) vol.mode = normalizeVolumetricMode(p.volumetricOutputMode);
            addon.solverUpdate(h, p.penaltyFactor !== undefined ? p.penaltyFactor : -1,
                p.filterRadius !== undefined ? p.filterRadius : -1,
                p.volumeFraction !== undefined ? p.volumeFraction : -1,
                p.maxIterations !== undefined ? p.maxIterations : -1);
        }
        if (state.paused) {
            const snap = vol.mode !== 'off' ? volumetricSnapshot() : null;
            if (snap) post({ type: 'paused', iteration: loop, volumetricData: snap.payload }, snap.transfer);
            else post({ type: 'paused', iteration: loop, volumetricData: null });
            await new Promise((r) => { state.resume = r; });
            if (state.cancelled) {
                post({ type: 'cancelled', iteration: loop });
                return;
            }
        }
        const info = addon.solverStep(h);
        loop = info.iteration;
        change = info.change;
        c = info.compliance;
        maxStress = info.maxStress;
        iterationTimes.push(info.totalMs);
        vol.iteration = loop;
        vol.maxStress = maxStress;
        const avg = iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length;
        const snap = vol.mode === 'every-iteration' ? volumetricSnapshot() : null;
        post({
            type: 'progress', iteration: loop, compliance: c, meshData: null, maxStress,
            volumetricData: snap ? snap.payload : null,
            timing: { iterationTime: info.totalMs, avgIterationTime: avg, elapsedTime: Date.now() - start,
                usingWasm: false, usingGPU: true },
            solverStats: { cgIterations: info.cgIterations, tolerance: info.tolerance, backend: 'hip',
                outerIterations: info.outerIterations, change: info.change, volume: info.volume },
        }, snap ? snap.transfer : undefined);
    }
    const densities = new Float32Array(nel);
    addon.solverGetDesign(h, densities);
    const avg = iterationTimes.length ? iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length : 0;
    const snap = vol.mode !== 'off' ? volumetricSnapshot() : null;
    post({ type: 'complete', result: {
        densities, finalCompliance: c, iterations: loop, volumeFraction: cfg.volfrac, nx, ny, nz,
        meshData: null, maxStress, volumetricData: snap ? snap.payload : null, amrStats: null,
        timing: { elapsedTime: Date.now() - start, avgIterationTime: avg, iterationTimes, usingWasm: false,
            usingGPU: true },
    } }, snap ? [densities.buffer].concat(snap.transfer) : [densities.buffer]);
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
        state.pending = Object.assign(state.pending || {}, data.config || {});
        break;
    case 'requestVolumetric': {
        // postLatestVolumetricSnapshot('on-demand') (:1595-1603)
        const snap = volumetricSnapshot();
        if (snap) post({ type: 'volumetric', reason: 'on-demand', volumetricData: snap.payload }, snap.transfer);
        else post({ type: 'volumetric', reason: 'on-demand', volumetricData: null });
        break;
    }
    default:
        break;
    }
});
