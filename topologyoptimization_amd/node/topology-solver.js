'use strict';
/**
 * TopologySolver -- the reference's library class (lib/topology-solver.js:132-518)
 * over the MI355X path: same constructor, methods, promise/progress/pause/
 * resume/updateConfig/cancel semantics and worker message protocol. Each
 * optimize() call runs in a worker_threads Worker (optimizer-worker-3d.js)
 * that drives the device-resident solver of libtopo_hip through the N-API
 * addon. CommonJS, so it also runs on Node 12.
 *
 * config.solver ('auto' | '2d' | '3d' | 'mgpcg' | 'cg' | 'petsc' | 'fea')
 * selects the worker and the linear solver as in the reference (:213-253):
 * '2d', and 'auto' / 'fea' on a model with nz <= 1 that is not a cube, run
 * the 2-D worker (optimizer-worker-2d.js); the others the 3-D worker. The
 * genetic, nonlinear, fracture, dynamics and fatigue solvers of the
 * reference are not part of this path and reject with an Error naming the
 * solver.
 */
const path = require('path');

const IS_NODE = typeof process !== 'undefined' && !!(process.versions && process.versions.node);

const UNSUPPORTED = ['genetic', 'nonlinear', 'fracture', 'dynamics', 'fatigue'];

class TopologySolver {
    constructor() {
        this._worker = null;
        this._paused = false;
        this._cancelReject = null;
        this.lastPausedData = null;
        this.lastVolumetricData = null;
        this.onPaused = null;
        this.onVolumetric = null;
        this.environment = { env: IS_NODE ? 'node' : 'unknown', gpu: 'unknown' };
    }

    /**
     * Run topology optimization in a background worker.
     * @param {object} model  { nx, ny, nz, elements?, type? }
     * @param {object} config reference config keys plus the build's own
     *   (precision, mgLevels, mgOmega, coarseSolver, cgTol, filterType, ...)
     * @param {function} [progressCallback]
     *   (iteration, compliance, meshData, timing, maxStress, volumetricData)
     * @returns {Promise<object>} result of the 'complete' message
     */
    optimize(model, config, progressCallback) {
        config = config || {};
        return new Promise((resolve, reject) => {
            this._cancelReject = reject;
            this._paused = false;
            const solverKey = String(config.solver || 'auto').toLowerCase();
            if (UNSUPPORTED.indexOf(solverKey) >= 0) {
                reject(new Error(`TopologySolver: solver '${solverKey}' is not part of the MI355X SIMP path`));
                return;
            }
            let linearSolver;
            let feaOnly = false;
            if (solverKey === 'mgpcg') linearSolver = 'mgpcg';
            else if (solverKey === 'cg') linearSolver = 'cg';
            else if (solverKey === 'petsc') linearSolver = 'petsc';
            else if (solverKey === 'fea') feaOnly = true;
            const byGeometry = model.type === 'cube' || (model.nz != null && model.nz > 1);
            const use3D = solverKey === '2d' ? false : (solverKey === 'auto' || solverKey === 'fea') ? byGeometry : true;
            let worker;
            try {
                const { Worker } = require('worker_threads');
                worker = new Worker(path.join(__dirname, use3D ? 'optimizer-worker-3d.js' : 'optimizer-worker-2d.js'));
            } catch (err) {
                reject(new Error(`TopologySolver: failed to create worker — ${err.message}`));
                return;
            }
            this._worker = worker;
            worker.on('message', (msg) => {
                switch (msg.type) {
                case 'progress':
                    if (progressCallback) {
                        progressCallback(msg.iteration, msg.compliance, msg.meshData, msg.timing, msg.maxStress,
                            msg.volumetricData || null);
                    }
                    break;
                case 'paused':
                    this.lastPausedData = msg;
                    this.lastVolumetricData = msg.volumetricData || this.lastVolumetricData;
                    if (typeof this.onPaused === 'function') this.onPaused(msg);
                    break;
                case 'volumetric':
                    this.lastVolumetricData = msg.volumetricData || null;
                    if (typeof this.onVolumetric === 'function') this.onVolumetric(msg);
                    break;
                case 'complete':
                    this.lastVolumetricData = (msg.result && msg.result.volumetricData) || null;
                    worker.terminate();
                    this._worker = null;
                    this._cancelReject = null;
                    resolve(msg.result);
                    break;
                case 'cancelled':
                    worker.terminate();
                    this._worker = null;
                    this._cancelReject = null;
                    reject(new Error('Optimization cancelled'));
                    break;
                case 'error':
                    worker.terminate();
                    this._worker = null;
                    reject(new Error(msg.message));
                    break;
                default:
                    break;
                }
            });
            worker.on('error', (err) => {
                worker.terminate();
                this._worker = null;
                reject(new Error(err.message || String(err)));
            });
            const workerConfig = Object.assign({}, config);
            delete workerConfig.solver;
            if (linearSolver) workerConfig.linearSolver = linearSolver;
            if (feaOnly) workerConfig.feaOnly = true;
            worker.postMessage({
                type: 'start',
                model: {
                    nx: model.nx, ny: model.ny, nz: model.nz, type: model.type,
                    meshType: model.meshType || 'box',
                    elements: model.elements ? Array.from(model.elements) : undefined,
                },
                config: workerConfig,
            });
        });
    }

    /** single FE analysis at the model's densities (solver 'fea') */
    analyzeFEA(model, config, progressCallback) {
        return this.optimize(model, Object.assign({}, config || {}, { solver: 'fea' }), progressCallback);
    }

    analyzeNonlinear() { return Promise.reject(new Error("TopologySolver: solver 'nonlinear' is not part of the MI355X SIMP path")); }

    analyzeFracture() { return Promise.reject(new Error("TopologySolver: solver 'fracture' is not part of the MI355X SIMP path")); }

    analyzeDynamics() { return Promise.reject(new Error("TopologySolver: solver 'dynamics' is not part of the MI355X SIMP path")); }

    /** pause at the next iteration boundary (worker answers 'paused') */
    pause() {
        if (this._worker) {
            this._paused = true;
            this._worker.postMessage({ type: 'pause' });
        }
    }

    resume() {
        if (this._worker && this._paused) {
            this._paused = false;
            this._worker.postMessage({ type: 'resume' });
        }
    }

    /** patch penaltyFactor, filterRadius, volumeFraction, maxIterations at the next iteration */
    updateConfig(patch) {
        if (this._worker) this._worker.postMessage({ type: 'updateConfig', config: patch });
    }

    requestVolumetricSnapshot() {
        if (this._worker) {
            this._worker.postMessage({ type: 'requestVolumetric' });
        } else if (this.lastVolumetricData && typeof this.onVolumetric === 'function') {
            this.onVolumetric({ type: 'volumetric', reason: 'cached', volumetricData: this.lastVolumetricData });
        }
    }

    /** cancel; the optimize() promise rejects with 'Optimization cancelled' */
    cancel() {
        if (this._worker) {
            this._worker.postMessage({ type: 'cancel' });
            setTimeout(() => {
                if (this._worker) {
                    this._worker.terminate();
                    this._worker = null;
                    if (this._cancelReject) {
                        this._cancelReject(new Error('Optimization cancelled'));
                        this._cancelReject = null;
                    }
                }
            }, 300);
        }
    }

    static get debug() {
        return {
            probeGPU() {
                const addon = require(path.join(__dirname, 'topo_addon.node'));
                const d = addon.deviceInfo();
                return Promise.resolve(d.count > 0 ? 'hip' : 'none');
            },
            detectEnvironment() {
                return { env: IS_NODE ? 'node' : 'unknown', workerType: 'worker_threads' };
            },
            workerUrl() {
                return 'file://' + path.join(__dirname, 'optimizer-worker-3d.js');
            },
        };
    }
}

module.exports = { TopologySolver, default: TopologySolver };
