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

const yieldToLoop = () => new Promise((r) => setImmediate(r));

async function optimize(model, config) {
    const { nx, ny, nz } = model;
    const problem = buildProblem(addon, model, config);
    const cfg = toTopoConfig(addon, model, config);
    const nel = nx * ny * nz;
    const h = addon.createSolver(cfg, problem.fixedMask, problem.F, problem.preserved, problem.voidMask);
    if (problem.x0) addon.solverSetDesign(h, problem.x0);

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
            addon.solverUpdate(h, p.penaltyFactor !== undefined ? p.penaltyFactor : -1,
                p.filterRadius !== undefined ? p.filterRadius : -1,
                p.volumeFraction !== undefined ? p.volumeFraction : -1,
                p.maxIterations !== undefined ? p.maxIterations : -1);
        }
        if (state.paused) {
            post({ type: 'paused', iteration: loop, volumetricData: null });
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
        const avg = iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length;
        post({
            type: 'progress', iteration: loop, compliance: c, meshData: null, maxStress, volumetricData: null,
            timing: { iterationTime: info.totalMs, avgIterationTime: avg, elapsedTime: Date.now() - start,
                usingWasm: false, usingGPU: true },
            solverStats: { cgIterations: info.cgIterations, tolerance: info.tolerance, backend: 'hip',
                outerIterations: info.outerIterations, change: info.change, volume: info.volume },
        });
    }
    const densities = new Float32Array(nel);
    addon.solverGetDesign(h, densities);
    const avg = iterationTimes.length ? iterationTimes.reduce((a, b) => a + b, 0) / iterationTimes.length : 0;
    post({ type: 'complete', result: {
        densities, finalCompliance: c, iterations: loop, volumeFraction: cfg.volfrac, nx, ny, nz,
        meshData: null, maxStress, volumetricData: null, amrStats: null,
        timing: { elapsedTime: Date.now() - start, avgIterationTime: avg, iterationTimes, usingWasm: false,
            usingGPU: true },
    } }, [densities.buffer]);
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
    case 'requestVolumetric':
        post({ type: 'volumetric', reason: 'unavailable', volumetricData: null });
        break;
    default:
        break;
    }
});
