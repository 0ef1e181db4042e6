'use strict';
/**
 * Reference config -> libtopo_hip problem: the setup the reference's worker
 * does before its SIMP loop (js/optimizer-worker-3d.js:1620-1650 material
 * and loop parameters, :1944-2022 supports / loads / preserved and void
 * elements), with the numbering done by the library's host functions
 * (topo_fixed_dofs, topo_load_vector, topo_preserved_elements,
 * topo_fixed_paint, topo_load_paint).
 */

const FIX = { left: 0, right: 1, bottom: 2, top: 3, 'bottom-corners': 4 };
const FIX_NONE = 5;
const LOAD = { down: 0, up: 1, left: 2, right: 3, 'top-center': 4 };
const LOAD_NONE = 5;
const FILTER = { sensitivity: 0, density: 1, helmholtz: 2 };

/** constraintDOFs ('all' or a string naming x/y/z, :4429-4434) -> bit mask */
function dofMask(constraintDOFs) {
    const mode = constraintDOFs || 'all';
    if (mode === 'all') return 7;
    return (mode.includes('x') ? 1 : 0) | (mode.includes('y') ? 2 : 0) | (mode.includes('z') ? 4 : 0);
}

/** painted face keys "vx,vy,vz[,face]" -> Int32Array n x 4 (face -1 = all nodes) */
function parseFaces(keys) {
    const out = [];
    for (const key of keys || []) {
        const parts = String(key).split(',');
        if (parts.length < 3) continue;
        const vx = parseInt(parts[0], 10);
        const vy = parseInt(parts[1], 10);
        const vz = parseInt(parts[2], 10);
        if (isNaN(vx) || isNaN(vy) || isNaN(vz)) continue;
        let fi = parts.length >= 4 ? parseInt(parts[3], 10) : -1;
        if (isNaN(fi) || fi < 0 || fi >= 6) fi = -1;
        out.push(vx, vy, vz, fi);
    }
    return Int32Array.from(out);
}

/** unit force direction of getLoadVectorFromPaint (:4673-4700) */
function paintDirection(direction, forceVector) {
    if (Array.isArray(forceVector) && forceVector.length >= 3) {
        const len = Math.sqrt(forceVector[0] ** 2 + forceVector[1] ** 2 + forceVector[2] ** 2);
        return len > 0 ? Float64Array.of(forceVector[0] / len, forceVector[1] / len, forceVector[2] / len)
            : Float64Array.of(0, 0, 0);
    }
    switch (direction) {
    case 'up': return Float64Array.of(0, 1, 0);
    case 'left': return Float64Array.of(-1, 0, 0);
    case 'right': return Float64Array.of(1, 0, 0);
    case 'top-center': return Float64Array.of(0, 0, 1);
    default: return Float64Array.of(0, -1, 0);
    }
}

function buildProblem(addon, model, config) {
    const nx = model.nx, ny = model.ny, nz = model.nz;
    const nel = nx * ny * nz;
    const ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1);
    const mask = dofMask(config.constraintDOFs);
    const pos = Object.prototype.hasOwnProperty.call(FIX, config.constraintPosition) ? FIX[config.constraintPosition] : FIX_NONE;
    const dir = Object.prototype.hasOwnProperty.call(LOAD, config.forceDirection) ? LOAD[config.forceDirection] : LOAD_NONE;
    const magnitude = config.forceMagnitude;
    const paintedC = config.paintedConstraints && config.paintedConstraints.length > 0;
    const paintedF = config.paintedForces && config.paintedForces.length > 0;

    const fixedMask = new Uint8Array(ndof);
    if (paintedC) addon.fixedPaint(nx, ny, nz, parseFaces(config.paintedConstraints), mask, fixedMask);
    else addon.fixedDofs(nx, ny, nz, pos, mask, fixedMask);

    const F = new Float32Array(ndof);
    const fv = Array.isArray(config.forceVector) && config.forceVector.length >= 3 ? Float64Array.from(config.forceVector.slice(0, 3)) : null;
    if (paintedF) addon.loadPaint(nx, ny, nz, parseFaces(config.paintedForces), paintDirection(config.forceDirection, config.forceVector), magnitude, F);
    else addon.loadVector(nx, ny, nz, dir, magnitude, fv, F);

    // preserved: every painted voxel in range, plus preset support / load
    // elements where nothing was painted (:1957-1986)
    const preserved = new Uint8Array(nel);
    const allKeys = [].concat(config.paintedConstraints || [], config.paintedForces || [], config.paintedKeep || []);
    const faces = parseFaces(allKeys);
    for (let i = 0; i < faces.length; i += 4) {
        const vx = faces[i], vy = faces[i + 1], vz = faces[i + 2];
        if (vx >= 0 && vx < nx && vy >= 0 && vy < ny && vz >= 0 && vz < nz) preserved[vx + vy * nx + vz * nx * ny] = 1;
    }
    addon.preservedElements(nx, ny, nz, paintedC ? FIX_NONE : pos, paintedF ? LOAD_NONE : dir, preserved);

    // void: outside the initial solid when constrainToSolid (:1988-1997)
    let voidMask = null;
    if (config.constrainToSolid && model.elements) {
        voidMask = new Uint8Array(nel);
        for (let i = 0; i < nel; i++) if (model.elements[i] < 0.5) voidMask[i] = 1;
    }

    // blended-curvature meshes start from the model's densities (:1932-1940)
    let x0 = null;
    if (model.meshType === 'blended-curvature' && model.elements) {
        const volfrac = config.volumeFraction;
        x0 = new Float32Array(nel);
        for (let i = 0; i < nel; i++) x0[i] = model.elements[i] > 0 ? Math.max(volfrac, model.elements[i]) : 0;
        for (let i = 0; i < nel; i++) if (preserved[i]) x0[i] = 1;
        if (voidMask) for (let i = 0; i < nel; i++) if (voidMask[i]) x0[i] = 0;
    }
    return { ndof, nel, fixedMask, F, preserved, voidMask, x0 };
}

/**
 * topo_config from the reference's config keys (volumeFraction,
 * penaltyFactor, filterRadius, maxIterations, youngsModulus, poissonsRatio,
 * penalStart, continuationIters, useProjection, betaMax, betaInterval,
 * linearSolver, petscPC) plus this library's own (precision
 * 'fp64'|'mixed'|'mixed-pc', coarseFp32, rhsTile,
 * filterType, mgLevels, mgOmega, mgOmegaCap, mgPreSmooth, mgPostSmooth,
 * coarseSolver 'sweeps'|'cholesky', cgTol, cgMaxIter, skipFactor,
 * exactApply, useGraphs, move).
 */
function toTopoConfig(addon, model, config) {
    const c = addon.configDefault(model.nx, model.ny, model.nz);
    const num = (v) => (typeof v === 'number' && isFinite(v));
    if (num(config.volumeFraction)) c.volfrac = config.volumeFraction;
    if (num(config.penaltyFactor)) c.penal = config.penaltyFactor;
    if (num(config.filterRadius)) c.rmin = config.filterRadius;
    if (num(config.maxIterations)) c.max_iterations = config.maxIterations;
    if (config.youngsModulus) c.E0 = config.youngsModulus;
    if (config.poissonsRatio !== undefined) c.nu = config.poissonsRatio;
    if (num(config.minimumDensity)) c.Emin = config.minimumDensity;
    if (config.penalStart != null) c.penal_start = config.penalStart;
    else c.penal_start = Math.min(c.penal, 1.5);
    if (config.continuationIters != null) c.continuation_iters = config.continuationIters;
    else c.continuation_iters = Math.max(20, Math.floor(c.max_iterations / 3));
    c.use_projection = config.useProjection !== false ? 1 : 0;
    if (config.betaMax != null) c.beta_max = config.betaMax;
    if (config.betaInterval != null) c.beta_interval = config.betaInterval;
    if (config.linearSolver === 'petsc') {
        // 'petsc' (FE() :4176-4181): the WASM KSP/BDDC solve for petscPC
        // 'bddc' (the default); the other preconditioners run the reference's
        // JS KSP loop, here the MG-PCG ('mg') or Jacobi-PCG ('jacobi') solves
        const pc = config.petscPC || 'bddc';
        c.linear_solver = pc === 'bddc' ? 2 : pc === 'mg' ? 1 : 0;
        if (pc === 'bddc') c.precision = 0;
    } else if (config.linearSolver) {
        c.linear_solver = config.linearSolver === 'mgpcg' ? 1 : 0;
    }
    if (config.precision && c.linear_solver !== 2)
        c.precision = config.precision === 'mixed' ? 1 : config.precision === 'mixed-pc' ? 2 : 0;
    if (config.filterType && Object.prototype.hasOwnProperty.call(FILTER, config.filterType)) c.filter_type = FILTER[config.filterType];
    if (num(config.mgLevels)) c.mg_levels = config.mgLevels;
    if (num(config.mgOmega)) c.mg_omega = config.mgOmega;
    if (num(config.mgOmegaCap)) c.mg_omega_cap = config.mgOmegaCap;
    if (num(config.mgPreSmooth)) c.mg_nu1 = config.mgPreSmooth;
    if (num(config.mgPostSmooth)) c.mg_nu2 = config.mgPostSmooth;
    if (config.coarseSolver) c.coarse_solver = config.coarseSolver === 'cholesky' ? 1 : 0;
    if (num(config.cgTol)) c.cg_tol = config.cgTol;
    if (num(config.cgMaxIter)) c.cg_max_iter = config.cgMaxIter;
    if (num(config.skipFactor)) c.skip_factor = config.skipFactor;
    if (config.exactApply != null) c.exact_apply = config.exactApply ? 1 : 0;
    if (config.useGraphs != null) c.use_graphs = config.useGraphs ? 1 : 0;
    if (num(config.move)) c.move = config.move;
    if (num(config.coarseFp32)) c.coarse_fp32 = config.coarseFp32;
    if (num(config.rhsTile)) c.rhs_tile = config.rhsTile;
    if (config.energyClamp != null) c.energy_clamp = config.energyClamp ? 1 : 0;
    if (config.mfmaApply != null) c.mfma_apply = config.mfmaApply ? 1 : 0;
    return c;
}

// ---- 2-D problems (js/optimizer-worker.js) ------------------------------------
const FIX2 = { left: 0, right: 1, bottom: 2, top: 3 };
const LOAD2 = { down: 0, up: 1, left: 2, right: 3 };

/** painted keys "vx,vy[,...]" -> Int32Array of (vx, vy) pairs (:2016-2041) */
function parseCells(keys) {
    const out = [];
    for (const key of keys || []) {
        const parts = String(key).split(',');
        if (parts.length < 2) continue;
        const vx = parseInt(parts[0], 10);
        const vy = parseInt(parts[1], 10);
        if (isNaN(vx) || isNaN(vy)) continue;
        out.push(vx, vy);
    }
    return Int32Array.from(out);
}

/** unit force direction of the 2-D getLoadVectorFromPaint (:2047-2063) */
function paintDirection2d(direction, forceVector) {
    if (Array.isArray(forceVector) && forceVector.length >= 2) {
        const len = Math.sqrt(forceVector[0] ** 2 + forceVector[1] ** 2);
        return len > 0 ? Float64Array.of(forceVector[0] / len, forceVector[1] / len) : Float64Array.of(0, 0);
    }
    switch (direction) {
    case 'up': return Float64Array.of(0, 1);
    case 'left': return Float64Array.of(-1, 0);
    case 'right': return Float64Array.of(1, 0);
    default: return Float64Array.of(0, -1);
    }
}

/**
 * The setup of the 2-D worker's optimize() before its loop (:517-553):
 * supports and loads from the presets or the painted cells, preserved
 * elements (painted cells, plus the preset support / load elements where
 * nothing was painted), void elements (constrainToSolid), and the initial
 * design. Element order e = ey + ex*nely; the importer's model.elements is
 * row-major (ex + ey*nelx).
 */
function buildProblem2d(addon, model, config) {
    const nx = model.nx, ny = model.ny;
    const nel = nx * ny;
    const ndof = 2 * (nx + 1) * (ny + 1);
    const mode = config.constraintDOFs || 'all';
    const mask = mode === 'all' ? 3 : (mode.includes('x') ? 1 : 0) | (mode.includes('y') ? 2 : 0);
    const pos = Object.prototype.hasOwnProperty.call(FIX2, config.constraintPosition) ? FIX2[config.constraintPosition] : FIX_NONE;
    const dir = Object.prototype.hasOwnProperty.call(LOAD2, config.forceDirection) ? LOAD2[config.forceDirection] : LOAD_NONE;
    const paintedC = config.paintedConstraints && config.paintedConstraints.length > 0;
    const paintedF = config.paintedForces && config.paintedForces.length > 0;

    const fixedMask = new Uint8Array(ndof);
    if (paintedC) addon.fixedPaint2d(nx, ny, parseCells(config.paintedConstraints), mask, fixedMask);
    else addon.fixedDofs2d(nx, ny, pos, mask, fixedMask);

    const F = new Float32Array(ndof);
    const fv = Array.isArray(config.forceVector) && config.forceVector.length >= 2 ? Float64Array.from(config.forceVector.slice(0, 2)) : null;
    if (paintedF) addon.loadPaint2d(nx, ny, parseCells(config.paintedForces), paintDirection2d(config.forceDirection, config.forceVector), config.forceMagnitude, F);
    else addon.loadVector2d(nx, ny, dir, config.forceMagnitude, fv, F);

    const preserved = new Uint8Array(nel);
    const cells = parseCells([].concat(config.paintedConstraints || [], config.paintedForces || [], config.paintedKeep || []));
    for (let i = 0; i < cells.length; i += 2) {
        const vx = cells[i], vy = cells[i + 1];
        if (vx >= 0 && vx < nx && vy >= 0 && vy < ny) preserved[vy + vx * ny] = 1;
    }
    addon.preservedElements2d(nx, ny, paintedC ? FIX_NONE : pos, paintedF ? LOAD_NONE : dir, preserved);

    let voidMask = null;
    if (config.constrainToSolid && model.elements) {
        voidMask = new Uint8Array(nel);
        for (let ex = 0; ex < nx; ex++)
            for (let ey = 0; ey < ny; ey++)
                if (model.elements[ex + ey * nx] < 0.5) voidMask[ey + ex * ny] = 1;
    }

    // blended-curvature meshes start from the model's densities (:521-526,
    // index for index as the reference reads them)
    let x0 = null;
    if (model.meshType === 'blended-curvature' && model.elements) {
        const volfrac = config.volumeFraction;
        x0 = new Float32Array(nel).fill(volfrac);
        for (let i = 0; i < nel; i++) x0[i] = model.elements[i] > 0 ? Math.max(volfrac, model.elements[i]) : 0;
        for (let i = 0; i < nel; i++) if (preserved[i]) x0[i] = 1;
        if (voidMask) for (let i = 0; i < nel; i++) if (voidMask[i]) x0[i] = 0;
    }
    return { ndof, nel, fixedMask, F, preserved, voidMask, x0 };
}

/** topo_config of the 2-D loop from the reference's config keys (:361-378,
 * :570-578) plus cgTol, cgMaxIter, skipFactor, useGraphs, move */
function toTopoConfig2d(addon, model, config) {
    const c = addon.configDefault2d(model.nx, model.ny);
    const num = (v) => (typeof v === 'number' && isFinite(v));
    if (num(config.volumeFraction)) c.volfrac = config.volumeFraction;
    if (num(config.penaltyFactor)) c.penal = config.penaltyFactor;
    if (num(config.filterRadius)) c.rmin = config.filterRadius;
    if (num(config.maxIterations)) c.max_iterations = config.maxIterations;
    if (config.youngsModulus) c.E0 = config.youngsModulus;
    if (config.poissonsRatio !== undefined) c.nu = config.poissonsRatio;
    if (config.penalStart != null) c.penal_start = config.penalStart;
    else c.penal_start = Math.min(c.penal, 1.5);
    if (config.continuationIters != null) c.continuation_iters = config.continuationIters;
    else c.continuation_iters = Math.max(20, Math.floor(c.max_iterations / 3));
    c.use_projection = config.useProjection !== false ? 1 : 0;
    if (config.betaMax != null) c.beta_max = config.betaMax;
    if (config.betaInterval != null) c.beta_interval = config.betaInterval;
    if (num(config.cgTol)) c.cg_tol = config.cgTol;
    if (num(config.cgMaxIter)) c.cg_max_iter = config.cgMaxIter;
    if (num(config.skipFactor)) c.skip_factor = config.skipFactor;
    if (config.useGraphs != null) c.use_graphs = config.useGraphs ? 1 : 0;
    if (num(config.move)) c.move = config.move;
    return c;
}

module.exports = { buildProblem, toTopoConfig, parseFaces, dofMask, paintDirection, buildProblem2d, toTopoConfig2d,
    parseCells, paintDirection2d };
