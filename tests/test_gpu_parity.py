"""GPU parity: the HIP path through the C ABI against the oracle on the same
seeded inputs. Integer/index work and the FP64 kernels that keep the
reference's summation order are compared bit for bit; solver results within
the tolerances written at each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRIDS = [(1, 1, 1), (3, 3, 3), (5, 5, 5), (6, 4, 3), (4, 7, 5), (9, 2, 3)]


def _solver(T, cfg, fixed=None, F=None):
    nel, ndof = T.dims(cfg.nelx, cfg.nely, cfg.nelz)
    fixed = np.zeros(ndof, np.uint8) if fixed is None else fixed
    F = np.zeros(ndof * cfg.nload, np.float32) if F is None else F
    return T.Solver(cfg, fixed, F)


@pytest.mark.parametrize("dims", GRIDS)
def test_applyA_fp64_bitexact(gpu, oracle, dims):
    """reference-order kernel: FP64 identical to applyAEbe3D's sums"""
    T = gpu
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    nel, ndof = T.dims(*dims)
    # MG-PCG config: E via _powDensity (x*x*x), the reference's MG path
    cfg = T.default_config(nx, ny, nz, linear_solver=T.SOLVER_MGPCG, mg_levels=1, precision=T.PREC_MIXED,
                           exact_apply=1)
    S = _solver(T, cfg)
    x = rng.uniform(0.0, 1.0, nel).astype(np.float32)
    x[rng.uniform(size=nel) < 0.2] = 0.0           # some elements below the skip threshold
    S.set_density(x, 3.0)
    p = rng.normal(size=ndof)
    Ap = S.applyA(p)
    xd = x.astype(np.float64)
    E = 1e-9 + (xd * xd * xd) * (1 - 1e-9)   # _powDensity: x*x*x
    act = np.flatnonzero(E > 1e-6).astype(np.int32)
    ref = oracle.applyA(nx, ny, nz, oracle.k0(0.3), E, act, p)
    assert np.array_equal(Ap, ref), np.abs(Ap - ref).max()
    # FP32 operator
    Ap32 = S.applyA(p, prec=T.PREC_MIXED)
    assert np.abs(Ap32 - ref).max() <= 2e-6 * np.abs(ref).max()
    S.close()


FAST_GRIDS = [(1, 1, 1), (3, 3, 3), (6, 4, 3), (4, 7, 5), (9, 2, 3), (20, 13, 70), (7, 40, 11), (33, 9, 130)]


@pytest.mark.parametrize("dims", FAST_GRIDS)
def test_applyA_fast_kernel(gpu, oracle, dims):
    """Hadamard x-march kernel (all modes) against the oracle: FP64 within
    1e-14 and FP32 within 2e-6 of max |Ap|; grids cover several workgroups
    along (y,z) and x, odd sizes and nelz+1 > 64 (lane-0 wave crossings)."""
    T = gpu
    rng = np.random.default_rng(21)
    nx, ny, nz = dims
    nel, ndof = T.dims(*dims)
    cfg = T.default_config(nx, ny, nz, linear_solver=T.SOLVER_PCG, precision=T.PREC_MIXED)
    fixed = (rng.uniform(size=ndof) < 0.1).astype(np.uint8)
    S = T.Solver(cfg, fixed, np.zeros(ndof, np.float32))
    x = rng.uniform(0.0, 1.0, nel).astype(np.float32)
    x[rng.uniform(size=nel) < 0.2] = 0.0
    S.set_density(x, 3.0)
    p = rng.normal(size=ndof)
    xd = x.astype(np.float64)
    E = 1e-9 + (xd * xd * xd) * (1 - 1e-9)   # _powDensity: x*x*x
    act = np.flatnonzero(E > 1e-6).astype(np.int32)
    ref = oracle.applyA(nx, ny, nz, oracle.k0(0.3), E, act, p)
    scale = np.abs(ref).max()
    Ap = S.applyA(p)
    assert np.abs(Ap - ref).max() <= 1e-14 * scale, np.abs(Ap - ref).max() / scale
    Ap32 = S.applyA(p, prec=T.PREC_MIXED)
    assert np.abs(Ap32 - ref).max() <= 2e-6 * scale
    S.close()


def test_twin_applyAEbe3D_bitexact(gpu, oracle):
    T = gpu
    rng = np.random.default_rng(12)
    nx, ny, nz = 6, 5, 4
    nel, ndof = T.dims(nx, ny, nz)
    ke = oracle.k0(0.3)
    ed = oracle.edofs(nx, ny, nz)
    evals = rng.uniform(0.01, 2.0, nel)
    act = np.sort(rng.choice(nel, nel // 2, replace=False)).astype(np.int32)
    p = rng.normal(size=ndof)
    Ap = T.applyAEbe3D(ke, ed, evals, act, p)
    assert np.array_equal(Ap, oracle.applyA(nx, ny, nz, ke, evals, act, p))
    # the synthetic tridiagonal KE of tests/test-wasm-applyA.js:77-88 works too
    K2 = np.zeros((24, 24))
    for i in range(24):
        K2[i, i] = 4.0
        if i > 0:
            K2[i, i - 1] = K2[i - 1, i] = -1.0
    assert np.array_equal(T.applyAEbe3D(K2, ed, evals, act, p), oracle.applyA(nx, ny, nz, K2, evals, act, p))
    # zero input / no active element -> zero (test-wasm-applyA.js:270-311)
    assert not T.applyAEbe3D(ke, ed, evals, act, np.zeros(ndof)).any()
    assert not T.applyAEbe3D(ke, ed, evals, np.zeros(0, np.int32), p).any()
    bad = ed.copy()
    bad[3, 5] += 3
    with pytest.raises(T.TopoError):
        T.applyAEbe3D(ke, bad, evals, act, p)


def test_twin_ebePCG(gpu, oracle):
    T = gpu
    import topologyoptimization_amd.problems as P
    n = 6
    cfg, fixed, F, pres, void, x0 = P.cube_reference(n)
    nel, ndof = T.dims(n, n, n)
    rng = np.random.default_rng(13)
    dens = rng.uniform(0.2, 1.0, nel)
    free = np.flatnonzero(fixed == 0).astype(np.int32)
    ke = oracle.k0(0.3)
    U0 = rng.normal(size=ndof) * 0.01
    Ug = U0.copy()
    Uo = U0.copy()
    ig = T.ebePCG(dens, ke, oracle.edofs(n, n, n), F.astype(np.float64), Ug, free, 1e-9, 1.0, 3.0, 400, 1e-8)
    io = oracle.ebePCG(dens, ke, n, n, n, F.astype(np.float64), Uo, free, 1e-9, 1.0, 3.0, 400, 1e-8)
    assert abs(ig - io) <= 1
    assert np.abs(Ug - Uo).max() <= 1e-6 * np.abs(Uo).max()
    assert not Ug[fixed > 0].any()


def test_twin_ebeMGPCG(gpu, oracle):
    T = gpu
    n = 8
    nel, ndof = T.dims(n, n, n)
    fixed = oracle.fixed_dofs(n, n, n, 4)
    F = oracle.load_vector(n, n, n, 4, -1.0).astype(np.float64)
    rng = np.random.default_rng(14)
    x = rng.uniform(0.2, 1.0, nel).astype(np.float32)
    mg = oracle.MG(n, n, n, fixed)
    mg.update(x, 3.0)
    levels = []
    fm = fixed
    for li in range(mg.nlevels):
        d = mg.level_data(li)
        d["edofs"] = oracle.edofs(d["nelx"], d["nely"], d["nelz"])
        if li > 0:
            # injection of the fine mask (the reference's _downsampleFixedMaskBy2)
            f = levels[-1]
            nxF, nyF, nzF = f["nelx"] + 1, f["nely"] + 1, f["nelz"] + 1
            nxC, nyC, nzC = d["nelx"] + 1, d["nely"] + 1, d["nelz"] + 1
            fc = np.zeros(3 * nxC * nyC * nzC, np.uint8)
            for cx in range(nxC):
                for cy in range(nyC):
                    for cz in range(nzC):
                        nf = min(2 * cx, nxF - 1) * nyF * nzF + min(2 * cy, nyF - 1) * nzF + min(2 * cz, nzF - 1)
                        nc = cx * nyC * nzC + cy * nzC + cz
                        fc[3 * nc:3 * nc + 3] = fm[3 * nf:3 * nf + 3]
            fm = fc
        d["fixedMask"] = fm
        levels.append(d)
    Ug = np.zeros(ndof)
    it_g = T.ebeMGPCG(levels, oracle.k0(0.3), F, Ug, fixed, 200, 1e-8)
    Uo, it_o = mg.solve(F, maxIter=200, tol=1e-8)
    assert abs(it_g - it_o) <= 1
    assert np.abs(Ug - Uo).max() <= 1e-7 * np.abs(Uo).max()


def test_fe_pcg_and_mgpcg_vs_oracle(gpu, oracle):
    T = gpu
    import topologyoptimization_amd.problems as P
    for solver in (T.SOLVER_PCG, T.SOLVER_MGPCG):
        cfg, fixed, F, pres, void, x0 = P.cantilever(12, 6, 6, linear_solver=solver)
        nel, ndof = T.dims(12, 6, 6)
        rng = np.random.default_rng(15)
        x = rng.uniform(0.1, 1.0, nel).astype(np.float32)
        with T.Solver(cfg, fixed, F) as S:
            U, c, info = S.fe(x, 3.0, iteration=1, max_iterations=1)
        if solver == T.SOLVER_PCG:
            Uo = np.zeros(ndof)
            free = np.flatnonzero(fixed == 0).astype(np.int32)
            it = oracle.ebePCG(x.astype(np.float64), oracle.k0(0.3), 12, 6, 6, F.astype(np.float64), Uo, free, 1e-9,
                               1.0, 3.0, 400, 1e-6)
        else:
            mg = oracle.MG(12, 6, 6, fixed)
            mg.update(x, 3.0)
            Uo, it = mg.solve(F.astype(np.float64), maxIter=400, tol=1e-6)
        co = float(F.astype(np.float64) @ Uo)
        assert abs(info.cg_iters - it) <= 1, (solver, info.cg_iters, it)
        assert abs(c - co) <= 1e-9 * abs(co)
        # both stop at relative residual 1e-6; the operators differ by rounding
        assert np.abs(U - Uo).max() <= 1e-6 * np.abs(Uo).max()


def test_energy_filters_oc_bitexact(gpu, oracle):
    T = gpu
    rng = np.random.default_rng(16)
    nx, ny, nz = 7, 5, 4
    nel, ndof = T.dims(nx, ny, nz)
    cfg = T.default_config(nx, ny, nz, linear_solver=T.SOLVER_PCG)
    fixed = np.zeros(ndof, np.uint8)
    with T.Solver(cfg, fixed, np.zeros(ndof, np.float32)) as S:
        U = rng.normal(size=ndof)
        assert np.array_equal(S.energy(U), oracle.energy(nx, ny, nz, oracle.k0(0.3), U))
        x = rng.uniform(0, 1, nel).astype(np.float32)
        dc = -rng.uniform(0, 3, nel).astype(np.float32)
        assert np.array_equal(S.filter(T.FILTER_SENSITIVITY, x, dc), oracle.sens_filter(nx, ny, nz, 1.5, x, dc))
        xn, passes = S.oc(x, dc, 0.3)
        xo, po = oracle.oc(x, 0.3, dc)
        assert passes == po and np.array_equal(xn, xo)
    cfg = T.default_config(nx, ny, nz, filter_type=T.FILTER_DENSITY)
    with T.Solver(cfg, fixed, np.zeros(ndof, np.float32)) as S:
        v = rng.uniform(0, 1, nel).astype(np.float32)
        assert np.array_equal(S.filter(T.FILTER_DENSITY, None, v), oracle.density_filter(nx, ny, nz, 1.5, v))
        assert np.array_equal(S.filter_adjoint(T.FILTER_DENSITY, v),
                              oracle.density_filter(nx, ny, nz, 1.5, v, transpose=True))


@pytest.mark.parametrize("rmin", [3.0, 4.5])
def test_wide_filters_bitexact(gpu, oracle, rmin):
    """Sensitivity and density filters with radii up to the reference UI's
    5 (prepareFilter takes any rmin, js/optimizer-worker-3d.js:3006-3050):
    window +-floor(rmin) (sensitivity) and ceil(rmin) (density), bit-exact
    with the oracle, and three SIMP steps of a 10x6x6 cantilever with the
    wide sensitivity filter against the oracle loop."""
    T = gpu
    import topologyoptimization_amd.problems as P
    rng = np.random.default_rng(int(rmin * 10))
    nx, ny, nz = 12, 9, 8
    nel, ndof = T.dims(nx, ny, nz)
    fixed = np.zeros(ndof, np.uint8)
    cfg = T.default_config(nx, ny, nz, linear_solver=T.SOLVER_PCG, rmin=rmin)
    with T.Solver(cfg, fixed, np.zeros(ndof, np.float32)) as S:
        x = rng.uniform(0, 1, nel).astype(np.float32)
        dc = -rng.uniform(0, 3, nel).astype(np.float32)
        assert np.array_equal(S.filter(T.FILTER_SENSITIVITY, x, dc), oracle.sens_filter(nx, ny, nz, rmin, x, dc))
    cfg = T.default_config(nx, ny, nz, filter_type=T.FILTER_DENSITY, rmin=rmin)
    with T.Solver(cfg, fixed, np.zeros(ndof, np.float32)) as S:
        v = rng.uniform(0, 1, nel).astype(np.float32)
        assert np.array_equal(S.filter(T.FILTER_DENSITY, None, v), oracle.density_filter(nx, ny, nz, rmin, v))
        assert np.array_equal(S.filter_adjoint(T.FILTER_DENSITY, v),
                              oracle.density_filter(nx, ny, nz, rmin, v, transpose=True))
    cfg, fixed, F, pres, void, x0 = P.cantilever(10, 6, 6, linear_solver=T.SOLVER_PCG, rmin=rmin, cg_tol=1e-8,
                                                 cg_max_iter=4000, volfrac=0.3)
    ref = oracle.simp_run(cfg, fixed, F, None, None, 3, x0)
    with T.Solver(cfg, fixed, F) as S:
        comp = np.array([S.step().compliance for _ in range(3)])
        xg = S.design()
    assert np.abs(comp - ref["compliance"]).max() <= 1e-6 * np.abs(ref["compliance"]).max()
    assert np.abs(xg - ref["x"]).max() <= 1e-3


def test_filter_radius_beyond_window_rejected(gpu):
    T = gpu
    cfg = T.default_config(8, 4, 4, rmin=6.0)
    _, ndof = T.dims(8, 4, 4)
    with pytest.raises(T.TopoError) as e:
        T.Solver(cfg, np.zeros(ndof, np.uint8), np.zeros(ndof, np.float32))
    assert e.value.code == -8


def test_simp_c1_trajectory(gpu, oracle):
    """5x5x5 reference cube, 30 SIMP iterations, Jacobi-PCG FP64 tol 1e-8:
    compliance within 1e-6 relative at every iteration, final densities
    within 1e-3."""
    T = gpu
    import topologyoptimization_amd.problems as P
    cfg, fixed, F, pres, void, x0 = P.cube_reference(5)
    ref = oracle.simp_run(cfg, fixed, F, pres, void, 30, x0)
    comp = []
    with T.Solver(cfg, fixed, F, pres, void) as S:
        for k in range(ref["iterations"]):
            info = S.step()
            comp.append(info.compliance)
        x = S.design()
    comp = np.array(comp)
    rel = np.abs(comp - ref["compliance"]) / np.abs(ref["compliance"])
    assert rel.max() <= 1e-6, rel.max()
    assert np.abs(x - ref["x"]).max() <= 1e-3


@pytest.mark.parametrize("solver", [0, 1])
def test_mixed_matches_fp64(gpu, solver):
    """Mixed precision (FP32 inner solve, FP64 refinement) against the FP64
    path at matched tolerance: compliance within 1e-4 relative."""
    T = gpu
    import topologyoptimization_amd.problems as P
    out = {}
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        cfg, fixed, F, pres, void, x0 = P.cantilever(24, 12, 12, linear_solver=solver, precision=prec, cg_tol=1e-6)
        with T.Solver(cfg, fixed, F) as S:
            cs = [S.step().compliance for _ in range(3)]
        out[prec] = np.array(cs)
    rel = np.abs(out[1] - out[0]) / out[0]
    assert rel.max() <= 1e-4, rel


def test_mg_cholesky_capped_vs_oracle(gpu, oracle):
    """4-level MG-PCG with the FP64 direct coarse solve and the omega
    safeguard against the oracle on a random-density cantilever."""
    T = gpu
    import topologyoptimization_amd.problems as P
    nx, ny, nz = 32, 16, 16
    cfg, fixed, F, pres, void, x0 = P.cantilever(nx, ny, nz, linear_solver=T.SOLVER_MGPCG, mg_levels=4, mg_omega=0.6,
                                                 coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0,
                                                 mg_omega_cap=1.6, cg_tol=1e-6)
    assert cfg.skip_factor == 0.0
    nel, ndof = T.dims(nx, ny, nz)
    x = P.splitmix64_uniform(nel)
    with T.Solver(cfg, fixed, F) as S:
        U, c, info = S.fe(x, 3.0)
    mg = oracle.MG(nx, ny, nz, fixed, nlevels=4, omega=0.6, coarse=1, galerkin_max=0, omega_cap=1.6, skip=-1.0)
    mg.update(x, 3.0)
    Uo, it = mg.solve(F.astype(np.float64), maxIter=400, tol=1e-6)
    co = float(F.astype(np.float64) @ Uo)
    assert it < 400 and info.cg_iters < 400
    assert abs(info.cg_iters - it) <= 2, (info.cg_iters, it)
    assert abs(c - co) <= 1e-6 * abs(co)


def test_c3_style_simp_steps_vs_oracle(gpu, oracle):
    """Three SIMP steps of the C3 recipe (random densities, 4-level MG-PCG,
    omega 0.6 capped, FP64 direct coarse solve, every element kept) on a
    32x16x16 grid: FP64 and mixed compliance against the FP64 oracle."""
    T = gpu
    import topologyoptimization_amd.problems as P
    nx, ny, nz = 32, 16, 16
    nel, ndof = T.dims(nx, ny, nz)
    x0 = P.splitmix64_uniform(nel)
    mk = lambda prec: P.cantilever(nx, ny, nz, linear_solver=T.SOLVER_MGPCG, mg_levels=4, mg_omega=0.6,
                                   coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0, precision=prec)
    cfg, fixed, F, pres, void, _ = mk(T.PREC_FP64)
    ref = oracle.simp_run(cfg, fixed, F, None, None, 3, x0)
    for prec, tol in ((T.PREC_FP64, 1e-6), (T.PREC_MIXED, 1e-4)):
        cfg, fixed, F, pres, void, _ = mk(prec)
        with T.Solver(cfg, fixed, F) as S:
            S.set_design(x0)
            comp = np.array([S.step().compliance for _ in range(3)])
            x = S.design()
        rel = np.abs(comp - ref["compliance"]) / np.abs(ref["compliance"])
        assert rel.max() <= tol, (prec, rel)
        assert np.abs(x - ref["x"]).max() <= 1e-3
