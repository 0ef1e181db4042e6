"""The matrix-core FP32 operator (mfma_apply = 1: the element product as a
dense 32x32 contraction on v_mfma_f32_16x16x32_bf16, operands split into two
bfloat16 parts, FP32 accumulation) against the oracle, against the VALU
kernel inside the V-cycle and the solves, and in batched solves.

Its stated accuracy: each operand carries about 16 significant bits
(x = hi + lo), so Ap is within 2e-5 of max |Ap| of the FP64 oracle (the
VALU kernel: 2e-6); as a smoother inside the FP64 CG it leaves the
converged compliance unchanged to 1e-7 and the CG counts within 2. As the
operator of the FP32 inner CG (PREC_MIXED) it converges to the same
compliance but needs up to about twice the inner iterations (measured 84
against 46 at 48x24x24): the FP32 Krylov recurrences are sensitive to the
operator's rounding, the smoother is not."""
import numpy as np
import pytest

import topologyoptimization_amd.problems as P

pytestmark = pytest.mark.gpu

# nelz + 1 <= 127 (the matrix-core kernel runs with 512-thread workgroups);
# several workgroups along (y,z) and x, odd sizes, one-element grids
GRIDS = [(1, 1, 1), (3, 3, 3), (6, 4, 3), (4, 7, 5), (9, 2, 3), (20, 13, 70), (7, 40, 11), (40, 20, 20)]


@pytest.mark.parametrize("dims", GRIDS)
def test_mfma_applyA_vs_oracle(gpu, oracle, dims):
    T = gpu
    rng = np.random.default_rng(21)
    nx, ny, nz = dims
    nel, ndof = T.dims(*dims)
    fixed = (rng.uniform(size=ndof) < 0.1).astype(np.uint8)
    x = rng.uniform(0.0, 1.0, nel).astype(np.float32)
    x[rng.uniform(size=nel) < 0.2] = 0.0
    p = rng.normal(size=ndof)
    xd = x.astype(np.float64)
    E = 1e-9 + (xd * xd * xd) * (1 - 1e-9)   # _powDensity: x*x*x
    act = np.flatnonzero(E > 1e-6).astype(np.int32)
    ref = oracle.applyA(nx, ny, nz, oracle.k0(0.3), E, act, p)
    scale = np.abs(ref).max()
    out = {}
    for mf in (0, 1):
        cfg = T.default_config(nx, ny, nz, linear_solver=T.SOLVER_PCG, precision=T.PREC_MIXED, mfma_apply=mf)
        with T.Solver(cfg, fixed, np.zeros(ndof, np.float32)) as S:
            S.set_density(x, 3.0)
            out[mf] = S.applyA(p, prec=T.PREC_MIXED)
            if mf:
                # the timing op runs the same kernel whatever the solver's setting
                ms, by = S.time_op(T.TIME_APPLYA_MFMA, T.PREC_MIXED, 0, 2)
                assert ms > 0 and by > 0
    err = np.abs(out[1] - ref).max() / scale
    print("%s: matrix-core %.2e, VALU %.2e of max|Ap|" % (dims, err, np.abs(out[0] - ref).max() / scale))
    assert err <= 2e-5
    # a different kernel really ran (the split operands round differently)
    assert not np.array_equal(out[0], out[1])


@pytest.mark.parametrize("prec", [1, 2])
def test_mfma_smoother_in_solves(gpu, prec):
    """48x24x24 cantilever, 4-level MG-PCG with the FP64 direct coarsest
    solve, three SIMP steps: FP32 inner CG (1, the masked operator with its
    p.Ap partials; tolerance 1e-8) and FP64 CG with the FP32 V-cycle (2, the
    sweep variants; tolerance 1e-6, the benchmark's): same compliance as the
    VALU kernel to 1e-7, the solves converged, CG counts within 2 (2) or at
    most twice the VALU ones (1), and the V-cycle output within 1e-4 of the
    VALU one."""
    T = gpu
    dims = (48, 24, 24)
    nel, ndof = T.dims(*dims)
    x0 = P.splitmix64_uniform(nel)
    r = np.random.default_rng(5).normal(size=ndof)
    tol = 1e-8 if prec == T.PREC_MIXED else 1e-6
    res = {}
    for mf in (0, 1):
        cfg, fixed, F, _, _, _ = P.cantilever(*dims, precision=prec, mg_levels=4, mg_omega=0.6,
                                              coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0,
                                              cg_tol=tol, cg_max_iter=1000, mfma_apply=mf)
        with T.Solver(cfg, fixed, F) as S:
            S.set_design(x0)
            infos = [S.step() for _ in range(3)]
            z = S.vcycle(r, prec=T.PREC_MIXED)
        res[mf] = (infos, z)
    for a, b in zip(res[1][0], res[0][0]):
        print("prec %d: matrix-core c %.12e cg %d res %.2e | VALU c %.12e cg %d res %.2e" %
              (prec, a.compliance, a.cg_iters, a.residual, b.compliance, b.cg_iters, b.residual))
    for a, b in zip(res[1][0], res[0][0]):
        assert abs(a.compliance - b.compliance) <= 1e-7 * abs(b.compliance), (a.compliance, b.compliance)
        assert a.residual <= tol and a.cg_iters < 1000
        if prec == T.PREC_MIXED:
            assert a.cg_iters <= 2 * b.cg_iters, (a.cg_iters, b.cg_iters)
        else:
            assert abs(a.cg_iters - b.cg_iters) <= 2, (a.cg_iters, b.cg_iters)
    z1, z0 = res[1][1], res[0][1]
    assert np.abs(z1 - z0).max() <= 1e-4 * np.abs(z0).max()
    assert not np.array_equal(z1, z0)


def test_mfma_batched_equals_sequential(gpu):
    """Four load cases solved as one batched MG-PCG (one launch carries all
    right-hand sides) and one at a time: bit-identical with the matrix-core
    smoother too."""
    T = gpu
    dims = (32, 16, 16)
    nel, ndof = T.dims(*dims)
    x = P.splitmix64_uniform(nel, lo=0.2)
    out = {}
    for tile in (1, 4):
        cfg, fixed, F, _, _, _ = P.bridge(*dims, precision=2, rhs_tile=tile, mg_levels=3, mg_omega=0.6,
                                          coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0, cg_tol=1e-8,
                                          cg_max_iter=1000, mfma_apply=1)
        with T.Solver(cfg, fixed, F) as S:
            U, c, info = S.fe(x, 3.0)
            out[tile] = (U, c, info.cg_iters)
    assert np.array_equal(out[4][0], out[1][0])
    assert out[4][1] == out[1][1] and out[4][2] == out[1][2]
