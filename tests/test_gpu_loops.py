"""GPU SIMP loops against the oracle's trajectories (tests/golden/
trajectories.npz, written by tests/golden/make_trajectories.py from the FP64
CPU restatement of the reference loop; the cases are defined in
tests/trajectory_cases.py):

* the 192x96x96 benchmark recipe at 32x16x16, mixed precision (FP64 MG-PCG
  with the FP32 V-cycle reading the FP16 coarse copy), 10 iterations;
* the full 4-level recipe at 48x24x24, mixed precision, 10 iterations;
* the same recipe at 48x24x24 with 3 levels and the coarsest block factored
  in nested-dissection order (forced by TOPO_ND_MIN / TOPO_ND_LEAF), with
  the FP16 and the FP32 copies of its inverse factor, 6 iterations;
* the reference's default 3-D loop (projection with beta doubling, penalty
  ramp from 1.5, CG tolerance schedule, skip threshold 1000 Emin, omega 0.5,
  Galerkin coarse levels, 30 coarsest sweeps, cold-started MG-PCG), 15
  iterations, in FP64;
* the 96x48x48 Jacobi-PCG recipe at 24x12x12 in mixed precision (FP64
  refinement around FP32 Jacobi-PCG), 6 iterations.

Tolerances: compliance 1e-6 relative for FP64 runs and 1e-4 for mixed runs
at every iteration; final design max |dx| <= 1e-3; CG counts within the
stated slack."""
import os

import numpy as np
import pytest

import trajectory_cases as TC

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trajectories.npz"))


def _run(T, name, **cfg_over):
    cfg, fixed, F, pres, void, x0, niter = TC.build(name, "gpu")
    for k, v in cfg_over.items():
        setattr(cfg, k, v)
    infos = []
    with T.Solver(cfg, fixed, F, pres, void) as S:
        S.set_design(x0)
        for _ in range(niter):
            info = S.step()
            infos.append(info)
            print("%s: it %d c %.10e change %.4g vol %.6f cg %d ocp %d" % (
                name, info.iteration, info.compliance, info.change, info.volume, info.cg_iters, info.oc_passes),
                flush=True)
        x = S.design()
    return infos, x


def _check(name, infos, x, tol, cg_slack=None):
    ref_c = GOLD[name + "/compliance"]
    assert len(infos) == len(ref_c)
    c = np.array([i.compliance for i in infos])
    rel = np.abs(c - ref_c) / np.abs(ref_c)
    print(name, "compliance rel. diff per iteration:", np.array2string(rel, precision=2))
    assert rel.max() <= tol, (name, rel)
    assert np.abs(x - GOLD[name + "/x"]).max() <= 1e-3, np.abs(x - GOLD[name + "/x"]).max()
    ch = np.array([i.change for i in infos])
    assert np.allclose(ch, GOLD[name + "/change"], rtol=0, atol=1e-3)
    if cg_slack is not None:
        cg = np.array([i.cg_iters for i in infos])
        assert np.abs(cg - GOLD[name + "/cg"]).max() <= cg_slack, (cg, GOLD[name + "/cg"])


def test_c3_recipe_mixed_vs_oracle(gpu):
    """Benchmark recipe, mixed precision, 10 iterations: volume held at
    0.12 once reached, bisection of normal length, no NaN."""
    infos, x = _run(gpu, "c3_recipe_32")
    _check("c3_recipe_32", infos, x, 1e-4, cg_slack=2)
    assert np.isfinite(x).all()
    for i in infos[3:]:
        assert abs(i.volume - 0.12) <= 1e-3 and i.oc_passes < 100, (i.volume, i.oc_passes)


def test_c3_recipe_48_mixed_vs_oracle(gpu):
    """The full recipe (4 levels, FP64 Cholesky of the 6x3x3 coarsest
    block, FP16 copy in the FP32 V-cycle) at 48x24x24, 10 iterations."""
    infos, x = _run(gpu, "c3_recipe_48")
    _check("c3_recipe_48", infos, x, 1e-4, cg_slack=3)
    assert np.isfinite(x).all()
    for i in infos[3:]:
        assert abs(i.volume - 0.12) <= 1e-3 and i.oc_passes < 100, (i.volume, i.oc_passes)


@pytest.mark.parametrize("coarse_fp32", [2, 1])
def test_mixed_pc_nested_dissection_coarse_vs_oracle(gpu, monkeypatch, capfd, coarse_fp32):
    """The benchmark's mixed mode with the coarsest level in nested-dissection
    order and the FP16 (2) or FP32 (1) copy of Z = L^-1 in the FP32 V-cycle,
    6 iterations against the FP64 oracle (dense Cholesky coarsest)."""
    monkeypatch.setenv("TOPO_VERBOSE", "1")
    monkeypatch.setenv("TOPO_ND_MIN", "0")
    monkeypatch.setenv("TOPO_ND_LEAF", "200")
    capfd.readouterr()
    infos, x = _run(gpu, "c3_recipe_48_3lev", coarse_fp32=coarse_fp32)
    err = capfd.readouterr().err
    assert "nested-dissection Cholesky" in err and "Gauss-Jordan" not in err, err
    _check("c3_recipe_48_3lev", infos, x, 1e-4, cg_slack=3)


def test_reference_default_loop_vs_oracle(gpu):
    """The reference's default 3-D loop in FP64, 15 iterations."""
    T = gpu
    cfg, *_ = TC.build("reference_default_16", "gpu")
    assert cfg.use_projection == 1 and cfg.beta_fixed == 0 and cfg.penal_start == 1.5 and cfg.cg_tol == 0.0
    assert cfg.skip_factor == 1000.0 and cfg.mg_omega == 0.5 and cfg.galerkin_max_ndof == 3000
    assert cfg.coarse_solver == T.COARSE_SWEEPS and cfg.coarse_sweeps == 30 and cfg.warm_start == -1
    infos, x = _run(T, "reference_default_16")
    assert [i.beta for i in infos] == [2.0 ** (k // 5) for k in range(15)]
    assert abs(infos[-1].penal - (1.5 + 1.5 * 14 / 20)) < 1e-12
    _check("reference_default_16", infos, x, 1e-6, cg_slack=1)


def test_c2_recipe_mixed_jacobi_vs_oracle(gpu):
    """The C2 recipe's mixed leg (FP64 refinement around FP32 Jacobi-PCG)
    against the FP64 oracle, 6 iterations."""
    infos, x = _run(gpu, "c2_recipe_24")
    _check("c2_recipe_24", infos, x, 1e-4)
