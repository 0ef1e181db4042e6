'use strict';
/**
 * TopologySolver over the MI355X path. Public surface of the reference's
 * library class (lib/topology-solver.js:132-518): optimize(model, config,
 * onProgress) -> Promise, analyzeFEA, pause, resume, updateConfig,
 * requestVolumetricSnapshot, cancel, the onPaused / onVolumetric hooks with
 * lastPausedData / lastVolumetricData, and TopologySolver.debug. CommonJS,
 * so it also loads on Node 12.
 *
 * Design. Each optimize() call is one Session: a worker_threads Worker
 * (optimizer-worker-3d.js or optimizer-worker-2d.js, which drive the
 * device-resident solver of libtopo_hip through the N-API addon) plus the
 * promise it settles. A session settles exactly once -- on the worker's
 * 'complete', 'cancelled' or 'error' message, on a worker crash, or when a
 * cancel is not acknowledged within CANCEL_GRACE_MS -- and settling always
 * stops the worker. Incoming messages are dispatched through a table of
 * handlers keyed by message type; outgoing commands (pause, resume,
 * updateConfig, requestVolumetric, cancel) go through Session.command, which
 * drops them once the session is over. The worker applies pause and config
 * patches at SIMP-iteration boundaries.
 *
 * Routing (config.solver, as the reference selects its workers): '2d', and
 * 'auto' / 'fea' on a model that is neither a cube nor more than one element
 * deep, use the 2-D worker; everything else the 3-D worker, with 'mgpcg',
 * 'cg' and 'petsc' choosing its linear solver and 'fea' a single analysis.
 * The genetic, nonlinear, fracture, dynamics and fatigue analyses are not on
 * this path: optimize() rejects them with an Error naming the solver.
 */
const path = require('path');

const RUNTIME = typeof process !== 'undefined' && process.versions && process.versions.node ? 'node' : 'unknown';
const OFF_PATH_SOLVERS = new Set(['genetic', 'nonlinear', 'fracture', 'dynamics', 'fatigue']);
const LINEAR_SOLVER_KEYS = new Set(['mgpcg', 'cg', 'petsc']);
const CANCELLED_MESSAGE = 'Optimization cancelled';
// a cancelled worker that has not answered by then is stopped by the host
const CANCEL_GRACE_MS = 300;
const WORKER_SCRIPT = { 3: 'optimizer-worker-3d.js', 2: 'optimizer-worker-2d.js' };

function offPathError(solver) {
    return new Error(`TopologySolver: solver '${solver}' is not part of the MI355X SIMP path`);
}

// Which worker a (model, config) pair runs on and the options it receives.
function planSession(model, config) {
    const solver = String(config.solver || 'auto').toLowerCase();
    if (OFF_PATH_SOLVERS.has(solver)) return { error: offPathError(solver) };
    const volumetric = model.type === 'cube' || (model.nz != null && model.nz > 1);
    let dims = 3;
    if (solver === '2d' || ((solver === 'auto' || solver === 'fea') && !volumetric)) dims = 2;
    const options = {};
    for (const key of Object.keys(config)) if (key !== 'solver') options[key] = config[key];
    if (LINEAR_SOLVER_KEYS.has(solver)) options.linearSolver = solver;
    if (solver === 'fea') options.feaOnly = true;
    return { script: WORKER_SCRIPT[dims], options };
}

// The 'start' command: the model fields the workers read, and the options.
function startCommand(model, options) {
    return {
        type: 'start',
        config: options,
        model: {
            nx: model.nx,
            ny: model.ny,
            nz: model.nz,
            type: model.type,
            meshType: model.meshType || 'box',
            elements: model.elements ? Array.from(model.elements) : undefined,
        },
    };
}

class Session {
    constructor(worker, settle, onProgress) {
        this.worker = worker;
        this.settleWith = settle;         // (ok, value) -> resolves or rejects the promise
        this.onProgress = onProgress;
        this.over = false;
        this.pauseAsked = false;
        this.cancelTimer = null;
    }

    command(type, fields) {
        if (this.over) return false;
        this.worker.postMessage(Object.assign({ type }, fields || {}));
        return true;
    }

    end(ok, value) {
        if (this.over) return false;
        this.over = true;
        if (this.cancelTimer !== null) clearTimeout(this.cancelTimer);
        this.cancelTimer = null;
        this.worker.terminate();
        this.settleWith(ok, value);
        return true;
    }
}

function callHook(fn, arg) {
    if (typeof fn === 'function') fn(arg);
}

// worker message type -> what the host does with it
const HANDLERS = {
    progress(host, session, m) {
        if (session.onProgress) {
            session.onProgress(m.iteration, m.compliance, m.meshData, m.timing, m.maxStress, m.volumetricData || null);
        }
    },
    paused(host, session, m) {
        host.lastPausedData = m;
        if (m.volumetricData) host.lastVolumetricData = m.volumetricData;
        callHook(host.onPaused, m);
    },
    volumetric(host, session, m) {
        host.lastVolumetricData = m.volumetricData || null;
        callHook(host.onVolumetric, m);
    },
    complete(host, session, m) {
        const result = m.result;
        host.lastVolumetricData = result && result.volumetricData ? result.volumetricData : null;
        host._finish(session, true, result);
    },
    cancelled(host, session) {
        host._finish(session, false, new Error(CANCELLED_MESSAGE));
    },
    error(host, session, m) {
        host._finish(session, false, new Error(m.message));
    },
};

class TopologySolver {
    constructor() {
        this._session = null;
        this.lastPausedData = null;
        this.lastVolumetricData = null;
        this.onPaused = null;
        this.onVolumetric = null;
        this.environment = { env: RUNTIME, gpu: 'unknown' };
    }

    /**
     * Run topology optimization in a background worker.
     * @param {object} model  { nx, ny, nz, elements?, type?, meshType? }
     * @param {object} config the reference's config keys plus this build's own
     *   (precision, mgLevels, mgOmega, coarseSolver, cgTol, filterType, ...)
     * @param {function} [progressCallback]
     *   (iteration, compliance, meshData, timing, maxStress, volumetricData)
     * @returns {Promise<object>} the result of the worker's 'complete' message
     */
    optimize(model, config, progressCallback) {
        const plan = planSession(model, config || {});
        if (plan.error) return Promise.reject(plan.error);
        return new Promise((resolve, reject) => {
            let worker;
            try {
                const { Worker } = require('worker_threads');
                worker = new Worker(path.join(__dirname, plan.script));
            } catch (err) {
                reject(new Error(`TopologySolver: cannot start the optimizer worker (${err.message})`));
                return;
            }
            const session = new Session(worker, (ok, value) => (ok ? resolve(value) : reject(value)),
                progressCallback || null);
            this._session = session;
            worker.on('message', (m) => {
                const handler = m && HANDLERS[m.type];
                if (handler && !session.over) handler(this, session, m);
            });
            worker.on('error', (err) => this._finish(session, false, new Error(err.message || String(err))));
            session.command('start', startCommand(model, plan.options));
        });
    }

    /** one FE analysis at the model's densities (solver 'fea') */
    analyzeFEA(model, config, progressCallback) {
        return this.optimize(model, Object.assign({}, config || {}, { solver: 'fea' }), progressCallback);
    }

    analyzeNonlinear() { return Promise.reject(offPathError('nonlinear')); }

    analyzeFracture() { return Promise.reject(offPathError('fracture')); }

    analyzeDynamics() { return Promise.reject(offPathError('dynamics')); }

    /** pause at the next iteration boundary; the worker answers 'paused' */
    pause() {
        const s = this._session;
        if (s && !s.pauseAsked && s.command('pause')) s.pauseAsked = true;
    }

    resume() {
        const s = this._session;
        if (s && s.pauseAsked && s.command('resume')) s.pauseAsked = false;
    }

    /** patch penaltyFactor, filterRadius, volumeFraction, maxIterations from the next iteration on */
    updateConfig(patch) {
        if (this._session) this._session.command('updateConfig', { config: patch });
    }

    /** the worker sends a 'volumetric' message; with no run going, the last snapshot is handed to onVolumetric */
    requestVolumetricSnapshot() {
        if (this._session && this._session.command('requestVolumetric')) return;
        if (this.lastVolumetricData) {
            callHook(this.onVolumetric, { type: 'volumetric', reason: 'cached', volumetricData: this.lastVolumetricData });
        }
    }

    /** cancel; the optimize() promise rejects with 'Optimization cancelled' */
    cancel() {
        const s = this._session;
        if (!s || !s.command('cancel') || s.cancelTimer !== null) return;
        s.cancelTimer = setTimeout(() => {
            s.cancelTimer = null;
            this._finish(s, false, new Error(CANCELLED_MESSAGE));
        }, CANCEL_GRACE_MS);
    }

    _finish(session, ok, value) {
        session.end(ok, value);
        if (this._session === session) this._session = null;
    }

    static get debug() {
        return {
            probeGPU() {
                const addon = require(path.join(__dirname, 'topo_addon.node'));
                return Promise.resolve(addon.deviceInfo().count > 0 ? 'hip' : 'none');
            },
            detectEnvironment() {
                return { env: RUNTIME, workerType: 'worker_threads' };
            },
            workerUrl(solverType) {
                const dims = String(solverType || '3d').toLowerCase() === '2d' ? 2 : 3;
                return 'file://' + path.join(__dirname, WORKER_SCRIPT[dims]);
            },
        };
    }
}

module.exports = { TopologySolver, default: TopologySolver };
