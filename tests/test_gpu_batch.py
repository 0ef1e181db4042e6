"""Batched multi-load MG-PCG (several load cases per launch) and its
autotuner, on a reduced bridge (the 4-load workload at 32x16x16): batched
and one-at-a-time solves give bit-identical results; the summed compliance
matches the oracle's independent solves; the solve is linear in the loads."""
import numpy as np
import pytest

import topologyoptimization_amd.problems as P

pytestmark = pytest.mark.gpu
DIMS = (32, 16, 16)


def _bridge(T, **kw):
    args = dict(mg_levels=3, mg_omega=0.6, coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0, cg_tol=1e-8,
                cg_max_iter=1000)
    args.update(kw)
    return P.bridge(*DIMS, **args)


@pytest.mark.parametrize("prec", [0, 2])
def test_batched_equals_sequential(gpu, prec):
    T = gpu
    nel, ndof = T.dims(*DIMS)
    x = P.splitmix64_uniform(nel, lo=0.2)
    out = {}
    for tile in (1, 2, 4):
        cfg, fixed, F, _, _, _ = _bridge(T, precision=prec, rhs_tile=tile)
        with T.Solver(cfg, fixed, F) as S:
            U, c, info = S.fe(x, 3.0)
            out[tile] = (U, c, info.cg_iters)
            assert S.tuning()[0] == tile
    for tile in (2, 4):
        assert np.array_equal(out[tile][0], out[1][0]), tile
        assert out[tile][1] == out[1][1] and out[tile][2] == out[1][2]


@pytest.mark.parametrize("prec,cf", [(0, 0), (2, 2)])
def test_batched_equals_sequential_nested_dissection(gpu, monkeypatch, prec, cf):
    """The same with the coarsest level in nested-dissection order (forced
    at this size; the FP16 copy's row-envelope products with 1, 2 and 4
    right-hand sides): bit-identical to one load case at a time."""
    T = gpu
    monkeypatch.setenv("TOPO_ND_MIN", "0")
    monkeypatch.setenv("TOPO_ND_LEAF", "100")
    nel, ndof = T.dims(*DIMS)
    x = P.splitmix64_uniform(nel, lo=0.2)
    out = {}
    for tile in (1, 2, 4):
        cfg, fixed, F, _, _, _ = _bridge(T, precision=prec, rhs_tile=tile, coarse_fp32=cf)
        with T.Solver(cfg, fixed, F) as S:
            U, c, info = S.fe(x, 3.0)
            out[tile] = (U, c, info.cg_iters)
    assert out[1][2] > 0
    for tile in (2, 4):
        assert np.array_equal(out[tile][0], out[1][0]), tile
        assert out[tile][1] == out[1][1] and out[tile][2] == out[1][2]


def test_batched_vs_oracle_and_linearity(gpu, oracle):
    T = gpu
    nel, ndof = T.dims(*DIMS)
    cfg, fixed, F, _, _, _ = _bridge(T, precision=0, rhs_tile=4, cg_tol=1e-10, cg_max_iter=2000)
    x = P.splitmix64_uniform(nel, lo=0.2)
    with T.Solver(cfg, fixed, F) as S:
        U, c, info = S.fe(x, 3.0)
    Fk = F.reshape(4, ndof).astype(np.float64)
    Uk = U.reshape(4, ndof)
    assert abs(c - sum(Fk[k] @ Uk[k] for k in range(4))) <= 1e-12 * abs(c)
    mg = oracle.MG(*DIMS, fixed, nlevels=3, omega=0.6, coarse=1, galerkin_max=0, omega_cap=1.6, skip=-1.0)
    mg.update(x, 3.0)
    for k in range(4):
        Uo, _ = mg.solve(Fk[k], maxIter=2000, tol=1e-10)
        assert np.abs(Uk[k] - Uo).max() <= 1e-7 * np.abs(Uo).max(), k
    # linearity: one case with the sum of the loads = the sum of the solutions
    cfg1, _, _, _, _, _ = _bridge(T, precision=0, cg_tol=1e-10, cg_max_iter=2000)
    cfg1.nload = 1
    with T.Solver(cfg1, fixed, Fk.sum(0).astype(np.float32)) as S1:
        Us, cs, _ = S1.fe(x, 3.0)
    assert np.abs(Us - Uk.sum(0)).max() <= 1e-7 * np.abs(Us).max()


def test_autotuner_and_simp_steps(gpu):
    T = gpu
    cfg, fixed, F, _, _, x0 = _bridge(T, precision=2, rhs_tile=0, cg_tol=1e-6, cg_max_iter=400)
    with T.Solver(cfg, fixed, F) as S:
        infos = [S.step() for _ in range(3)]
        R, nt, cx, log = S.tuning()
    tiles, batches = log[log[:, 0] >= 0], log[log[:, 0] < 0]
    assert len(batches) == 3 and np.all(batches[:, 3] > 0)
    assert R == -int(batches[np.argmin(batches[:, 3]), 0])
    fine = tiles[tiles[:, 0] == 0]
    assert len(fine) == 14 and np.all(fine[:, 3] > 0)
    assert np.all((fine[:, 2] >= 0) & (fine[:, 2] <= 32))   # x-chunks (0 = the size rule)
    best = fine[np.argmin(fine[:, 3])]
    assert (nt, cx) == (int(best[1]), int(best[2]))
    # the mixed solver's FP64 CG operator is tuned on its own (same candidates,
    # the size rule logged as the x-chunk it gives)
    cg64 = tiles[tiles[:, 0] == 1000]
    assert len(cg64) == 14 and np.all(cg64[:, 3] > 0)
    assert np.all((cg64[:, 2] >= 1) & (cg64[:, 2] <= 32))
    print("autotuned: R=%d NT=%d CX=%d, fine operator %.4f ms" % (R, nt, cx, best[3]))
    cfg1, fixed, F, _, _, _ = _bridge(T, precision=2, rhs_tile=1, cg_tol=1e-6, cg_max_iter=400)
    with T.Solver(cfg1, fixed, F) as S:
        ref = [S.step() for _ in range(3)]
    for a, b in zip(infos, ref):
        assert abs(a.compliance - b.compliance) <= 1e-6 * abs(b.compliance)


@pytest.mark.parametrize("prec", [0, 2])
def test_workgroup_size_does_not_change_results(gpu, monkeypatch, prec):
    """256-, 512- and 1024-thread x-march workgroups (every level, fixed
    x-chunk) give the same bits for the operator on every level and for the
    V-cycle, FP64 and FP32: each node's sums are formed in the same order
    whatever the workgroup. (The CG's partial sums are per yz tile, so the
    fine level keeps one workgroup size; the autotuner varies it on the
    coarser levels only.)"""
    T = gpu
    nel, ndof = T.dims(40, 20, 20)
    x = P.splitmix64_uniform(nel, lo=0.2)
    r = np.random.default_rng(7).standard_normal(ndof)
    out = []
    for nt in (256, 512, 1024):
        monkeypatch.setenv("TOPO_XM_NT", str(nt))
        monkeypatch.setenv("TOPO_XM_CX", "4")
        cfg, fixed, F, _, _, _ = P.cantilever(40, 20, 20, linear_solver=T.SOLVER_MGPCG, precision=prec,
                                              mg_levels=3, cg_tol=1e-8, galerkin_max_ndof=0)
        with T.Solver(cfg, fixed, F) as S:
            S.fe(x, 3.0)   # builds the hierarchy for the design
            res = [S.vcycle(r)]
            precs = [T.PREC_FP64] + ([T.PREC_MIXED] if prec else [])
            for pr in precs:
                if pr != T.PREC_FP64:
                    res.append(S.vcycle(r, prec=pr))
                for lev, dims in enumerate(((40, 20, 20), (20, 10, 10), (10, 5, 5))):
                    p = np.random.default_rng(lev).standard_normal(T.dims(*dims)[1])
                    res.append(S.applyA(p, level=lev, prec=pr))
            out.append(res)
    for res in out[1:]:
        assert len(res) == len(out[0])
        for a, b in zip(out[0], res):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [0, 1])
def test_tile_shape_does_not_change_results(gpu, prec):
    """The autotuned element tiles give the same bits as the default ones:
    every node's sums are formed in the same order whatever the tiling.
    (E-average coarse levels: the dense Galerkin assembly of the reference
    mode adds with atomics and is reproducible to rounding only.)"""
    T = gpu
    nel, ndof = T.dims(40, 20, 20)
    out = []
    for tile in (0, 1):
        cfg, fixed, F, _, _, x0 = P.cantilever(40, 20, 20, linear_solver=T.SOLVER_MGPCG, precision=prec,
                                               rhs_tile=tile, mg_levels=3, cg_tol=1e-8, galerkin_max_ndof=0)
        with T.Solver(cfg, fixed, F) as S:
            U, c, _ = S.fe(P.splitmix64_uniform(nel, lo=0.2), 3.0)
            out.append((U, c, S.tuning()))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    assert len(out[0][2][3]) > 0 and len(out[1][2][3]) == 0
