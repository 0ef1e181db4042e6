"""Problem generators: the reference's presets and the benchmark workloads.

All index arithmetic uses the reference numbering (node z fastest, element x
fastest). The cantilever, bridge and inverter workloads have no preset in the
reference; their boundary conditions are defined here and documented with
the seed of the random initial densities.
"""
import numpy as np

from . import (DOF_ALL, FIX_BOTTOM_CORNERS, FIX_LEFT, LOAD_TOP_CENTER, PREC_FP64, PREC_MIXED, PREC_MIXED_PC, SOLVER_MGPCG,
               SOLVER_PCG, COARSE_CHOLESKY, FILTER_HELMHOLTZ, FILTER_SENSITIVITY, OBJ_MECHANISM, default_config, dims,
               fixed_dofs, load_vector, preserved_elements)

SEED = 20261203


def splitmix64_uniform(n, seed=SEED, lo=0.05, hi=1.0):
    """x_i = lo + (hi-lo) * ((r_i >> 11) * 2^-53), r_i the splitmix64 stream
    of `seed`, rounded to float32, in element order e = ex + ey*nelx + ez*nelx*nely."""
    out = np.empty(n, np.float64)
    mask = (1 << 64) - 1
    s = seed & mask
    # vectorised splitmix64: state_i = seed + (i+1)*golden
    gamma = np.uint64(0x9E3779B97F4A7C15)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(s) + idx * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return (lo + (hi - lo) * out).astype(np.float32)


def node_index(nely, nelz, ix, iy, iz):
    return (ix * (nely + 1) + iy) * (nelz + 1) + iz


def cube_reference(n=5, **kw):
    """Plain reference method on an n^3 cube (the reference's own preset):
    4 bottom corners fixed in xyz, unit -z point force at the top-centre node
    ('top-center' with forceMagnitude -1), preserved constraint/force
    elements, volfrac 0.3, p = 3 from the first iteration, sensitivity
    filter rmin 1.5, OC move 0.2, projection off, Jacobi-PCG FP64 tol 1e-8."""
    cfg = default_config(n, n, n, volfrac=0.3, penal=3.0, penal_start=3.0, rmin=1.5, use_projection=0,
                         linear_solver=SOLVER_PCG, precision=PREC_FP64, cg_tol=1e-8, max_iterations=30,
                         filter_type=FILTER_SENSITIVITY)
    for k, v in kw.items():
        setattr(cfg, k, v)
    fixed = fixed_dofs(n, n, n, FIX_BOTTOM_CORNERS, DOF_ALL)
    F = load_vector(n, n, n, LOAD_TOP_CENTER, -1.0)
    pres = preserved_elements(n, n, n, FIX_BOTTOM_CORNERS, LOAD_TOP_CENTER)
    nel, _ = dims(n, n, n)
    x0 = np.full(nel, np.float32(cfg.volfrac), np.float32)
    x0[pres > 0] = 1.0
    return cfg, fixed, F, pres, None, x0


def cantilever(nelx, nely, nelz, **kw):
    """Cantilever: x = 0 face clamped (all DOFs); -z load of total 1.0 spread
    uniformly (float32 -1/(nely+1) per node) over the nodes of the bottom edge
    of the x = max face (ix = nelx, iz = 0, iy = 0..nely). volfrac 0.12,
    p = 3, sensitivity filter rmin 1.5, projection off, no preserved
    elements. Every element stays in the operator (skip_factor 0): with
    voided elements removed, disconnected parts make K singular and the CG
    iterates diverge along their rigid modes. MG smoothing weights are capped
    at 1.6 / lambda_max(D^-1 A) (damped Jacobi at omega 0.6 is unstable for
    these density contrasts). Element energies are clamped at 0
    (energy_clamp 1): from the U[0.05, 1] start the move limit voids large
    regions within three iterations, elements there move almost rigidly, and
    at 192x96x96 some energies u_e^T K0 u_e round to about -4e-13 in the
    fourth iteration; the reference's OC then takes the square root of a
    negative number and the design gets a NaN entry."""
    cfg = default_config(nelx, nely, nelz, volfrac=0.12, penal=3.0, penal_start=3.0, rmin=1.5, use_projection=0,
                         max_iterations=20, cg_tol=1e-6, skip_factor=0.0, mg_omega_cap=1.6, energy_clamp=1)
    for k, v in kw.items():
        setattr(cfg, k, v)
    nel, ndof = dims(nelx, nely, nelz)
    fixed = fixed_dofs(nelx, nely, nelz, FIX_LEFT, DOF_ALL)
    F = np.zeros(ndof, np.float32)
    per = np.float32(-1.0 / (nely + 1))
    for iy in range(nely + 1):
        F[3 * node_index(nely, nelz, nelx, iy, 0) + 2] = per
    x0 = np.full(nel, np.float32(cfg.volfrac), np.float32)
    return cfg, fixed, F, None, None, x0


def cantilever_c2(precision=PREC_FP64, **kw):
    """96x48x48 cantilever, Jacobi-PCG, tol 1e-6, 20 SIMP iterations."""
    return cantilever(96, 48, 48, linear_solver=SOLVER_PCG, precision=precision, **kw)


def cantilever_c3(precision=PREC_FP64, nlevels=4, **kw):
    """192x96x96 cantilever (5,447,811 DOF), 4-level MG-PCG, V(2,2) damped
    Jacobi omega 0.6 capped at 1.8 / lambda_max(D^-1 A), FP64 direct solve of
    the coarsest level, tol 1e-6, initial densities U[0.05, 1] from
    splitmix64(SEED)."""
    args = dict(linear_solver=SOLVER_MGPCG, precision=precision, mg_levels=nlevels, mg_omega=0.6,
                coarse_solver=COARSE_CHOLESKY, galerkin_max_ndof=0, mg_omega_cap=1.8)
    args.update(kw)
    cfg, fixed, F, pres, void, _ = cantilever(192, 96, 96, **args)
    nel, _ = dims(192, 96, 96)
    return cfg, fixed, F, pres, void, splitmix64_uniform(nel)


def bridge(nelx, nely, nelz, **kw):
    """Multi-load bridge: the two bottom edges along y at x = 0 and x = nelx
    (z = 0) pinned in xyz; 4 load cases, each a -z load of total 1.0 on an
    8x8-node patch (float32 -1/64 per node) of the top face (z = nelz) centred
    in y at x = 0.2, 0.4, 0.6, 0.8 of the span. volfrac 0.12, p = 3,
    sensitivity filter rmin 1.5, MG-PCG, tol 1e-6, every element kept,
    energies clamped at 0 (as the cantilever)."""
    cfg = default_config(nelx, nely, nelz, nload=4, volfrac=0.12, penal=3.0, penal_start=3.0, rmin=1.5,
                         use_projection=0, max_iterations=10, cg_tol=1e-6, linear_solver=SOLVER_MGPCG,
                         skip_factor=0.0, mg_omega_cap=1.6, energy_clamp=1)
    for k, v in kw.items():
        setattr(cfg, k, v)
    nel, ndof = dims(nelx, nely, nelz)
    fixed = np.zeros(ndof, np.uint8)
    for ix in (0, nelx):
        for iy in range(nely + 1):
            n = node_index(nely, nelz, ix, iy, 0)
            fixed[3 * n:3 * n + 3] = 1
    F = np.zeros((4, ndof), np.float32)
    for k, frac in enumerate((0.2, 0.4, 0.6, 0.8)):
        cx = int(round(frac * nelx))
        cy = nely // 2
        per = np.float32(-1.0 / 64)
        for ix in range(cx - 4, cx + 4):
            for iy in range(cy - 4, cy + 4):
                F[k, 3 * node_index(nely, nelz, ix, iy, nelz) + 2] = per
    x0 = np.full(nel, np.float32(cfg.volfrac), np.float32)
    return cfg, fixed, F.ravel(), None, None, x0


def bridge_c4(**kw):
    """The 128x64x64 bridge (1,635,075 DOF, 4 load cases), batched MG-PCG
    (4 levels, FP64 Cholesky coarsest, omega 0.6 capped), FP32 V-cycle inside
    FP64 CG reading an FP16 row-scaled copy of the coarse factor inverse,
    autotuned batch and element-tile shapes."""
    args = dict(mg_levels=4, mg_omega=0.6, coarse_solver=COARSE_CHOLESKY, galerkin_max_ndof=0,
                precision=PREC_MIXED_PC, rhs_tile=0, coarse_fp32=2)
    args.update(kw)
    return bridge(128, 64, 64, **args)


def inverter(nelx=160, nely=80, nelz=80, symmetric=True, **kw):
    """Force-inverter compliant mechanism. Input: +x force 1.0 at the centre
    of the x = 0 face, input spring k_in = 1 on that DOF; output: the x DOF
    at the centre of the x = max face with spring k_out = 0.001; the
    objective is that output displacement (minimised: the output moves
    against the input). Supports and symmetry (not fixed by the workload
    statement; chosen here): symmetric=True (the default) is the standard
    inverter set-up as a quarter model -- the grid is one quarter of a
    nelx x 2 nely x 2 nelz domain cut by the symmetry planes y = 0 and z = 0
    (u_y = 0 and u_z = 0 on them), so the face centres are the nodes
    (0, 0, 0) and (nelx, 0, 0) on their common edge, and every DOF of the
    x = 0 face nodes in the corner square of side nely/8 x nelz/8 farthest
    from both planes is fixed; OC move limit 0.1. symmetric=False models the
    full grid with the input and output at its face centres and every DOF of
    the x = 0 face fixed in its four corner squares. From the uniform start
    the output first follows the input (u_out > 0); the inverting lever
    forms after about 10-15 iterations (measured at 80 x 40 x 40: u_out
    -0.8 at iteration 15, -1.9 at 40; the full model needs about 20
    iterations and several thousand CG iterations per solve on the way).
    Helmholtz filter rmin 2 (PDE length 2 / (2 sqrt 3)), Heaviside beta 8,
    eta 0.5, volfrac 0.3, p = 3, FP32
    MG-PCG inside FP64 refinement (4 levels, Cholesky coarsest), tol 1e-6.
    The V-cycle reads the FP32 copy of the coarse factor inverse: with only
    the corner supports the coarse operator is ill-conditioned enough that
    the FP16 copy changes the solution (output displacement 12 % off the
    FP64 coarse solve's at the same CG tolerance) and doubles the CG
    iterations."""
    cfg = default_config(nelx, nely, nelz, volfrac=0.3, penal=3.0, penal_start=3.0, rmin=2.0,
                         filter_type=FILTER_HELMHOLTZ, use_projection=1, beta_fixed=8, eta=0.5, max_iterations=20,
                         cg_tol=1e-6, skip_factor=0.0, mg_omega_cap=1.6, objective=OBJ_MECHANISM,
                         linear_solver=SOLVER_MGPCG, precision=PREC_MIXED, mg_levels=4, mg_omega=0.6,
                         coarse_solver=COARSE_CHOLESKY, galerkin_max_ndof=0, mech_k_in=1.0, mech_k_out=0.001,
                         coarse_fp32=1, move=0.1 if symmetric else 0.2)
    for k, v in kw.items():
        setattr(cfg, k, v)
    nel, ndof = dims(nelx, nely, nelz)
    fixed = np.zeros(ndof, np.uint8)
    sy, sz = max(1, nely // 8), max(1, nelz // 8)
    if symmetric:
        # quarter model: y = 0 and z = 0 are symmetry planes (u_y = 0, u_z = 0
        # there); input and output on their common edge; supports: the
        # corner square of the x = 0 face farthest from both planes
        for ix in range(nelx + 1):
            for iy in range(nely + 1):
                fixed[3 * node_index(nely, nelz, ix, iy, 0) + 2] = 1
            for iz in range(nelz + 1):
                fixed[3 * node_index(nely, nelz, ix, 0, iz) + 1] = 1
        for iy in range(nely - sy, nely + 1):
            for iz in range(nelz - sz, nelz + 1):
                n = node_index(nely, nelz, 0, iy, iz)
                fixed[3 * n:3 * n + 3] = 1
        n_in = node_index(nely, nelz, 0, 0, 0)
        n_out = node_index(nely, nelz, nelx, 0, 0)
    else:
        for iy in range(nely + 1):
            for iz in range(nelz + 1):
                if min(iy, nely - iy) <= sy and min(iz, nelz - iz) <= sz:
                    n = node_index(nely, nelz, 0, iy, iz)
                    fixed[3 * n:3 * n + 3] = 1
        n_in = node_index(nely, nelz, 0, nely // 2, nelz // 2)
        n_out = node_index(nely, nelz, nelx, nely // 2, nelz // 2)
    assert not fixed[3 * n_in], "grid too small: the input node falls in a support square"
    cfg.mech_in_dof = 3 * n_in
    cfg.mech_out_dof = 3 * n_out
    F = np.zeros(ndof, np.float32)
    F[3 * n_in] = 1.0
    x0 = np.full(nel, np.float32(cfg.volfrac), np.float32)
    return cfg, fixed, F, None, None, x0
