"""CPU checks of the C-ABI library: it loads without a GPU, exports every
symbol include/topo.h declares, and its host-side setup (K0, numbering,
boundary-condition / load presets) is bit-exact with the oracle's
restatement of the reference."""
import os
import re

import numpy as np
import pytest

import topologyoptimization_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "topo.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(topo_[a-zA-Z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = T.lib()
    decl = _declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    assert set(T.SYMBOLS) == set(decl)
    assert b"gfx950" in L.topo_version()


def test_config_defaults_are_reference_defaults():
    c = T.default_config(10, 10, 10)
    assert (c.E0, c.Emin, c.nu, c.penal, c.rmin, c.move) == (1.0, 1e-9, 0.3, 3.0, 1.5, 0.2)
    assert c.linear_solver == T.SOLVER_MGPCG and c.cg_max_iter == 400 and c.mg_omega == 0.5
    assert c.mg_nu1 == 2 and c.mg_nu2 == 2 and c.coarse_sweeps == 30 and c.galerkin_max_ndof == 3000
    assert c.use_projection == 1 and c.beta_max == 64 and c.beta_interval == 5


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_k0_bitexact_with_oracle(oracle, nu):
    assert np.array_equal(T.k0(nu), oracle.k0(nu))


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 4), (5, 5, 5), (7, 3, 2)])
def test_numbering_and_presets_bitexact(oracle, dims):
    nx, ny, nz = dims
    assert np.array_equal(T.edofs(nx, ny, nz), oracle.edofs(nx, ny, nz))
    for pos in range(5):
        for mask in (T.DOF_ALL, T.DOF_X, T.DOF_Y | T.DOF_Z):
            assert np.array_equal(T.fixed_dofs(nx, ny, nz, pos, mask), oracle.fixed_dofs(nx, ny, nz, pos, mask))
        for d in range(5):
            assert np.array_equal(T.preserved_elements(nx, ny, nz, pos, d), oracle.preserved(nx, ny, nz, pos, d))
    for d in range(5):
        for mag in (1.0, -2.5):
            assert np.array_equal(T.load_vector(nx, ny, nz, d, mag), oracle.load_vector(nx, ny, nz, d, mag))
    fv = [1.0, -2.0, 0.5]
    assert np.array_equal(T.load_vector(nx, ny, nz, 0, 3.0, fv), oracle.load_vector(nx, ny, nz, 0, 3.0, fv))


def test_paint_load_counts_unique_nodes():
    # two faces sharing an edge: 4 + 4 - 2 unique nodes
    F = T.load_paint(4, 4, 4, [[3, 1, 1, 1], [3, 2, 1, 1]], [0, 0, -1], 6.0)
    nz = np.flatnonzero(F)
    assert len(nz) == 6 and np.allclose(F[nz], -1.0)
    assert np.all(nz % 3 == 2)


def test_problem_generators():
    import topologyoptimization_amd.problems as P
    cfg, fixed, F, pres, void, x0 = P.cantilever(96, 48, 48)
    nel, ndof = T.dims(96, 48, 48)
    assert ndof == 698691 and fixed.sum() == 3 * 49 * 49 == 7203
    assert abs(F.astype(np.float64).sum() + 1.0) < 1e-6 and np.count_nonzero(F) == 49
    x = P.splitmix64_uniform(1000)
    assert x.min() >= 0.05 and x.max() <= 1.0 and abs(x.mean() - 0.525) < 0.03
    assert np.array_equal(x, P.splitmix64_uniform(1000))
    _, _, F4, _, _, _ = P.bridge_c4()
    assert F4.size == 4 * 1635075
