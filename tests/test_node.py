"""The Node.js host (topologyoptimization_amd/node): the N-API addon, the
reference-config -> problem translation, TopologySolver's worker protocol
and the WASM-shaped facade. Each case runs a small script under tests/js in
a node subprocess and checks its JSON output against the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import topologyoptimization_amd as T
import topologyoptimization_amd.problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "tests", "js")
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "topologyoptimization_amd", "node", "topo_addon.node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_js(script, arg, timeout=120):
    p = subprocess.run([NODE, os.path.join(JS, script), json.dumps(arg)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def _mask(idx, n):
    m = np.zeros(n, np.uint8)
    m[np.asarray(idx, np.int64)] = 1
    return m


def _dense_F(Fd, n):
    F = np.zeros(n, np.float32)
    for k, v in Fd.items():
        F[int(k)] = v
    return F


def test_addon_exports():
    assert os.path.exists(ADDON), "addon not built"
    src = ("const a=require(%r);process.stdout.write(JSON.stringify({k:Object.keys(a).sort(),v:a.version(),"
           "k0:Array.from(a.k0(0.3,new Float64Array(576)))}))" % ADDON)
    p = subprocess.run([NODE, "-e", src], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    out = json.loads(p.stdout)
    for f in ("applyAEbe3D", "ebePCG", "ebeMGPCG", "createSolver", "solverStep", "solverFe", "solverEnergy",
              "fixedDofs", "loadVector", "preservedElements", "fixedPaint", "loadPaint", "configDefault"):
        assert f in out["k"], f
    assert "gfx950" in out["v"]
    assert np.array_equal(np.array(out["k0"]), T.k0(0.3).ravel())


PRESETS = [
    ("bottom-corners", "all", "top-center", -1.0),
    ("left", "all", "down", 2.0),
    ("right", "xz", "left", 1.0),
    ("bottom", "y", "up", 0.5),
    ("top", "all", "right", 3.0),
]


@pytest.mark.parametrize("pos,dofs,direction,mag", PRESETS)
def test_preset_problem_matches_oracle(oracle, pos, dofs, direction, mag):
    nx, ny, nz = 6, 4, 5
    nel, ndof = T.dims(nx, ny, nz)
    out = run_js("problem_dump.js", {"model": {"nx": nx, "ny": ny, "nz": nz},
                                     "config": {"constraintPosition": pos, "constraintDOFs": dofs,
                                                "forceDirection": direction, "forceMagnitude": mag,
                                                "volumeFraction": 0.3, "penaltyFactor": 3, "filterRadius": 1.5,
                                                "maxIterations": 60}})
    P_ = {"left": 0, "right": 1, "bottom": 2, "top": 3, "bottom-corners": 4}[pos]
    D_ = {"down": 0, "up": 1, "left": 2, "right": 3, "top-center": 4}[direction]
    mask = 7 if dofs == "all" else sum(b for c, b in (("x", 1), ("y", 2), ("z", 4)) if c in dofs)
    assert np.array_equal(_mask(out["fixed"], ndof), oracle.fixed_dofs(nx, ny, nz, P_, mask))
    assert np.array_equal(_dense_F(out["F"], ndof), oracle.load_vector(nx, ny, nz, D_, mag))
    assert np.array_equal(_mask(out["preserved"], nel), oracle.preserved(nx, ny, nz, P_, D_))
    c = out["config"]
    assert (c["volfrac"], c["penal"], c["rmin"], c["max_iterations"]) == (0.3, 3, 1.5, 60)
    # penalStart / continuationIters defaults of js/optimizer-worker-3d.js:2057-2059
    assert c["penal_start"] == 1.5 and c["continuation_iters"] == 20
    assert c["use_projection"] == 1


def test_painted_problem_matches_oracle(oracle):
    nx, ny, nz = 5, 4, 3
    nel, ndof = T.dims(nx, ny, nz)
    pc = ["0,0,0,0", "0,1,0,0", "0,1,1", "0,3,2,9", "bad", "1,2"]
    pf = ["4,3,2,1", "4,2,2,1", "4,2,2,3"]
    keep = ["2,2,2", "9,9,9"]
    for fv in (None, [1.0, -2.0, 0.5]):
        cfg = {"constraintPosition": "left", "constraintDOFs": "xy", "forceDirection": "right", "forceMagnitude": 2.5,
               "paintedConstraints": pc, "paintedForces": pf, "paintedKeep": keep, "volumeFraction": 0.4}
        if fv:
            cfg["forceVector"] = fv
        out = run_js("problem_dump.js", {"model": {"nx": nx, "ny": ny, "nz": nz}, "config": cfg})
        assert np.array_equal(_mask(out["fixed"], ndof), oracle.fixed_paint(nx, ny, nz, pc, "xy"))
        assert np.array_equal(_dense_F(out["F"], ndof),
                              oracle.load_paint(nx, ny, nz, pf, "right", 2.5, fv))
        pres = np.zeros(nel, np.uint8)
        for key in pc + pf + keep:
            parts = key.split(",")
            if len(parts) < 3:
                continue
            vx, vy, vz = (int(v) for v in parts[:3])
            if 0 <= vx < nx and 0 <= vy < ny and 0 <= vz < nz:
                pres[vx + vy * nx + vz * nx * ny] = 1
        assert np.array_equal(_mask(out["preserved"], nel), pres)


def test_oracle_paint_known_answer(oracle):
    # two +X faces sharing an edge -> 6 unique nodes, 1/6 of the load each;
    # all-corner key and an out-of-grid corner (dropped by the node bound)
    F = oracle.load_paint(4, 4, 4, ["3,1,1,1", "3,2,1,1"], "top-center", 6.0)
    nzi = np.flatnonzero(F)
    assert len(nzi) == 6 and np.all(F[nzi] == 1.0) and np.all(nzi % 3 == 2)
    m = oracle.fixed_paint(2, 2, 2, ["1,1,1"], "z")
    assert m.sum() == 8 and np.all(np.flatnonzero(m) % 3 == 2)
    assert oracle.fixed_paint(2, 2, 2, ["2,2,2"], "all").sum() == 3  # only node (2,2,2) is in range


def test_constrain_to_solid_void(oracle):
    nx, ny, nz = 4, 3, 2
    nel, _ = T.dims(nx, ny, nz)
    el = [1 if (i % 3) else 0 for i in range(nel)]
    out = run_js("problem_dump.js", {"model": {"nx": nx, "ny": ny, "nz": nz, "elements": el,
                                               "meshType": "blended-curvature"},
                                     "config": {"constraintPosition": "left", "forceDirection": "down",
                                                "forceMagnitude": 1, "constrainToSolid": True,
                                                "volumeFraction": 0.3}})
    assert out["void"] == [i for i in range(nel) if el[i] == 0]
    x0 = np.array(out["x0"], np.float32)
    pres = _mask(out["preserved"], nel)
    exp = np.where(np.array(el) > 0, 1.0, 0.0).astype(np.float32)
    exp[pres > 0] = 1.0
    exp[np.array(el) == 0] = 0.0
    assert np.array_equal(x0, exp)


# ---- GPU: the TopologySolver host end to end --------------------------------
def _c1_js_config(n):
    return {"model": {"nx": n, "ny": n, "nz": n, "type": "cube"},
            "config": {"solver": "cg", "volumeFraction": 0.3, "penaltyFactor": 3, "penalStart": 3,
                       "filterRadius": 1.5, "useProjection": False, "maxIterations": 30,
                       "constraintPosition": "bottom-corners", "constraintDOFs": "all",
                       "forceDirection": "top-center", "forceMagnitude": -1.0, "cgTol": 1e-8,
                       "precision": "fp64", "filterType": "sensitivity"}}


@pytest.mark.gpu
def test_topology_solver_c1_matches_oracle(gpu, oracle):
    """TopologySolver.optimize on the reference's 5^3 cube preset, through
    the worker and the N-API addon, against the oracle's SIMP loop."""
    n = 5
    out = run_js("run_optimize.js", _c1_js_config(n))
    assert out["ok"], out
    cfg, fixed, F, pres, void, x0 = P.cube_reference(n)
    ref = oracle.simp_run(cfg, fixed, F, pres, void, 30, x0)
    assert out["iterations"] == ref["iterations"]
    comp = np.array(out["compliance"])
    rel = np.abs(comp - ref["compliance"]) / np.abs(ref["compliance"])
    assert rel.max() <= 1e-6, rel.max()
    assert np.abs(np.array(out["densities"], np.float32) - ref["x"]).max() <= 1e-3


@pytest.mark.gpu
def test_topology_solver_controls(gpu):
    """pause -> 'paused' + resume, updateConfig(maxIterations), then cancel
    rejects with 'Optimization cancelled' (lib/topology-solver.js:456-511).
    Messages reach the worker between iterations, so a request made in the
    progress callback of iteration k takes effect after iteration k or k+1."""
    arg = _c1_js_config(6)
    arg["config"]["maxIterations"] = 200
    arg["actions"] = {"2": "pause", "3": {"maxIterations": 6}}
    out = run_js("run_optimize.js", arg)
    assert out["ok"], out
    paused = [e[1] for e in out["events"] if e[0] == "paused"]
    assert len(paused) == 1 and paused[0] in (2, 3), out["events"]
    assert ["resume", paused[0]] in out["events"]
    assert out["iterations"] <= 7 and len(out["compliance"]) == out["iterations"]
    arg["actions"] = {"2": "cancel"}
    arg["config"]["maxIterations"] = 200
    out = run_js("run_optimize.js", arg)
    assert not out["ok"] and out["error"] == "Optimization cancelled"
    assert 2 <= len(out["compliance"]) <= 4


@pytest.mark.gpu
def test_topology_solver_fea_only(gpu, oracle):
    n = 4
    nel, ndof = T.dims(n, n, n)
    arg = _c1_js_config(n)
    arg["config"]["solver"] = "fea"
    rng = np.random.default_rng(3)
    el = rng.uniform(0.2, 1.0, nel).round(3)
    arg["model"]["elements"] = el.tolist()
    out = run_js("run_optimize.js", arg)
    assert out["ok"] and out["feaOnly"], out
    cfg, fixed, F, pres, void, x0 = P.cube_reference(n)
    dens = el.astype(np.float32).astype(np.float64)
    free = np.flatnonzero(fixed == 0).astype(np.int32)
    U = np.zeros(ndof)
    oracle.ebePCG(dens, oracle.k0(0.3), n, n, n, F.astype(np.float64), U, free, 1e-9, 1.0, 3.0, 400, 1e-8)
    c_ref = float(F.astype(np.float64) @ U)
    assert abs(out["finalCompliance"] - c_ref) <= 1e-6 * abs(c_ref)


@pytest.mark.gpu
def test_wasm_facade_matches_oracle(gpu, oracle):
    """applyAEbe3D / ebePCG called by byte offset on the facade's
    WebAssembly.Memory: applyA bit-exact, PCG within solver tolerance."""
    n = 6
    nel, ndof = T.dims(n, n, n)
    rng = np.random.default_rng(21)
    evals = rng.uniform(0.1, 1.0, nel)
    evals[rng.random(nel) < 0.2] = 0.0
    p = rng.normal(size=ndof)
    fixed = oracle.fixed_dofs(n, n, n, 4)
    F = oracle.load_vector(n, n, n, 4, -1.0).astype(np.float64)
    dens = rng.uniform(0.2, 1.0, nel)
    free = np.flatnonzero(fixed == 0).astype(np.int32)
    out = run_js("facade_apply.js", {"n": n, "evals": evals.tolist(), "p": p.tolist(), "dens": dens.tolist(),
                                     "F": F.tolist(), "free": free.tolist()})
    active = np.flatnonzero(evals > 0).astype(np.int32)
    ref = oracle.applyA(n, n, n, oracle.k0(0.3), evals, active, p)
    assert np.array_equal(np.array(out["Ap"]), ref)
    U = np.zeros(ndof)
    it = oracle.ebePCG(dens, oracle.k0(0.3), n, n, n, F, U, free, 1e-9, 1.0, 3.0, 400, 1e-8)
    assert abs(out["iters"] - it) <= 1
    assert np.abs(np.array(out["U"]) - U).max() <= 1e-6 * np.abs(U).max()


@pytest.mark.gpu
def test_wasm_facade_ksp_bddc(gpu, oracle):
    """ebeKSP_BDDC called by byte offset on the facade's memory with the
    32-byte subdomain descriptors, against the oracle (30 iterations, random
    densities: the range over which the iterates are comparable)."""
    n = 6
    nel, ndof = T.dims(n, n, n)
    rng = np.random.default_rng(22)
    x = rng.uniform(0.0, 1.0, nel)
    K0 = oracle.k0(0.3)
    E = 1e-9 + x ** 3 * (1 - 1e-9)
    act = np.flatnonzero(E > 1e-6).astype(np.int32)
    fixed = oracle.fixed_dofs(n, n, n, 0)
    F = oracle.load_vector(n, n, n, 0, -1.0).astype(np.float64)
    ed = oracle.edofs(n, n, n)
    diag = np.zeros(ndof)
    for e in act:
        np.add.at(diag, ed[e], E[e] * np.diag(K0))
    invD = np.where((fixed == 0) & (diag > 1e-30), 1.0 / np.where(diag > 0, diag, 1.0), 0.0)
    b = oracle.Bddc(n, n, n)
    b.update(K0, E, fixed)
    subs = [tuple(a.tolist() for a in b.sub(si)) for si in range(b.nsubs)]
    evals = np.ones(nel)
    out = run_js("facade_apply.js", {"n": n, "evals": evals.tolist(), "p": [0.0] * ndof, "dens": [1.0] * nel,
                                     "F": F.tolist(), "free": np.flatnonzero(fixed == 0).tolist(),
                                     "bddc": {"subs": subs, "coarse": b.coarse().tolist(), "evals": E.tolist(),
                                              "active": act.tolist(), "fixed": fixed.tolist(), "invD": invD.tolist(),
                                              "F": F.tolist(), "maxIter": 30, "tol": 1e-8}})
    Uo, ito = b.solve(act, invD, F, 30, 1e-8, 3)
    assert out["bddc"]["iters"] == ito
    assert np.abs(np.array(out["bddc"]["U"]) - Uo).max() <= 1e-10 * np.abs(Uo).max()


@pytest.mark.gpu
def test_topology_solver_petsc_matches_oracle(gpu, oracle):
    """solver 'petsc' (petscPC 'bddc'): the KSP/BDDC solve behind
    TopologySolver on the 5^3 cube, 30 CG iterations per solve, against the
    oracle's SIMP loop for the first two iterations (with unconverged solves
    later iterations are sensitive to rounding)."""
    n = 5
    arg = _c1_js_config(n)
    arg["config"].update({"solver": "petsc", "maxIterations": 2, "cgMaxIter": 30})
    out = run_js("run_optimize.js", arg)
    assert out["ok"], out
    cfg, fixed, F, pres, void, x0 = P.cube_reference(n, linear_solver=T.SOLVER_KSP_BDDC, cg_max_iter=30,
                                                     max_iterations=2)
    ref = oracle.simp_run(cfg, fixed, F, pres, void, 2, x0)
    comp = np.array(out["compliance"])
    assert len(comp) == 2
    rel = np.abs(comp - ref["compliance"]) / np.abs(ref["compliance"])
    assert rel.max() <= 1e-8, rel


@pytest.mark.gpu
def test_topology_solver_volumetric_snapshots(gpu):
    """volumetricOutputMode (js/optimizer-worker-3d.js:1573-1603, :2280-2321):
    'every-iteration' sends the stress field E_e w_e with every progress
    message and the final result; requestVolumetricSnapshot() answers with a
    'volumetric' message; 'off' sends none."""
    n = 5
    arg = _c1_js_config(n)
    arg["config"].update({"maxIterations": 3, "volumetricOutputMode": "every-iteration"})
    arg["actions"] = {"1": "volumetric"}
    out = run_js("run_optimize.js", arg)
    assert out["ok"], out
    vol = out["volumetric"]
    prog = [v for v in vol if v[0] == "progress"]
    assert [v[1] for v in prog] == [1, 2, 3]
    for v in prog:
        assert v[4] == n ** 3 and v[2] > 0 and abs(v[3] - v[2]) <= 1e-6 * v[2]
    assert any(v[0] == "volumetric" and v[1] in (1, 2) for v in vol), vol
    comp = [v for v in vol if v[0] == "complete"]
    assert len(comp) == 1 and comp[0][1] == 3
    arg["config"]["volumetricOutputMode"] = "off"
    arg["actions"] = {}
    out = run_js("run_optimize.js", arg)
    assert out["ok"] and out["volumetric"] == [["complete", None]]


# ---- 2-D worker (js/optimizer-worker.js) -------------------------------------
@pytest.mark.parametrize("pos,dofs,direction,mag", [("left", "all", "down", 1.0), ("right", "x", "up", 2.0),
                                                     ("bottom", "y", "left", -1.5), ("top", "all", "right", 0.5),
                                                     ("bottom-corners", "all", "top-center", 1.0)])
def test_2d_preset_problem_matches_oracle(oracle, pos, dofs, direction, mag):
    nx, ny = 7, 4
    nel, ndof = T.dims2(nx, ny)
    out = run_js("problem_dump.js", {"twod": True, "model": {"nx": nx, "ny": ny, "nz": 1},
                                     "config": {"constraintPosition": pos, "constraintDOFs": dofs,
                                                "forceDirection": direction, "forceMagnitude": mag,
                                                "volumeFraction": 0.3, "penaltyFactor": 3, "filterRadius": 1.5,
                                                "maxIterations": 90}})
    assert np.array_equal(_mask(out["fixed"], ndof), oracle.fixed_dofs2(nx, ny, pos, dofs))
    assert np.array_equal(_dense_F(out["F"], ndof), oracle.load_vector2(nx, ny, direction, mag))
    assert np.array_equal(_mask(out["preserved"], nel), oracle.preserved2(nx, ny, pos, direction))
    c = out["config"]
    assert (c["nelx"], c["nely"], c["volfrac"], c["penal"], c["rmin"], c["max_iterations"]) == (nx, ny, 0.3, 3, 1.5, 90)
    # penalStart / continuationIters defaults of js/optimizer-worker.js:570-572
    assert c["penal_start"] == 1.5 and c["continuation_iters"] == 30
    assert c["cg_max_iter"] == 2000 and c["cg_tol"] == 0


def test_2d_painted_problem_matches_oracle(oracle):
    nx, ny = 6, 5
    nel, ndof = T.dims2(nx, ny)
    pc = ["0,0,0,4", "0,1", "0,4,2", "bad", "9,9"]
    pf = ["5,4,0,1", "5,3"]
    keep = ["2,2", "-1,2"]
    for fv in (None, [1.0, -2.0]):
        cfg = {"constraintPosition": "left", "constraintDOFs": "x", "forceDirection": "right", "forceMagnitude": 2.5,
               "paintedConstraints": pc, "paintedForces": pf, "paintedKeep": keep, "volumeFraction": 0.4}
        if fv:
            cfg["forceVector"] = fv
        out = run_js("problem_dump.js", {"twod": True, "model": {"nx": nx, "ny": ny, "nz": 1}, "config": cfg})
        cells = lambda keys: [tuple(int(v) for v in k.split(",")[:2]) for k in keys
                              if len(k.split(",")) >= 2 and all(p.lstrip("-").isdigit() for p in k.split(",")[:2])]
        assert np.array_equal(_mask(out["fixed"], ndof), oracle.fixed_paint2(nx, ny, cells(pc), "x"))
        assert np.array_equal(_dense_F(out["F"], ndof), oracle.load_paint2(nx, ny, cells(pf), "right", 2.5, fv))
        pres = np.zeros(nel, np.uint8)
        for vx, vy in cells(pc + pf + keep):
            if 0 <= vx < nx and 0 <= vy < ny:
                pres[vy + vx * ny] = 1
        assert np.array_equal(_mask(out["preserved"], nel), pres)


def _2d_js_config(nx, ny, it):
    return {"model": {"nx": nx, "ny": ny, "nz": 1},
            "config": {"solver": "auto", "volumeFraction": 0.4, "penaltyFactor": 3, "filterRadius": 1.5,
                       "maxIterations": it, "constraintPosition": "left", "forceDirection": "down",
                       "forceMagnitude": 1.0}}


@pytest.mark.gpu
def test_topology_solver_2d_matches_oracle(gpu, oracle):
    """TopologySolver.optimize on a 2-D model ('auto' with nz = 1 runs the
    2-D worker) against the oracle's 2-D loop: per-iteration compliance and
    the final densities (row-major, as the reference returns them)."""
    nx, ny, it = 30, 10, 8
    out = run_js("run_optimize.js", _2d_js_config(nx, ny, it))
    assert out["ok"], out
    cfg = T.default_config_2d(nx, ny, volfrac=0.4, max_iterations=it, penal_start=1.5, continuation_iters=20)
    fixed = oracle.fixed_dofs2(nx, ny, "left")
    F = oracle.load_vector2(nx, ny, "down", 1.0)
    pres = oracle.preserved2(nx, ny, "left", "down")
    x0 = np.full(nx * ny, 0.4, np.float32)
    x0[pres > 0] = 1
    ref = oracle.simp_run2(cfg, fixed, F, pres, None, it, x0)
    assert out["iterations"] == ref["iterations"]
    comp = np.array(out["compliance"])
    assert np.abs(comp - ref["compliance"]).max() <= 1e-5 * np.abs(ref["compliance"]).max()
    d = np.array(out["densities"], np.float32)
    xr = ref["x"].reshape(nx, ny).T.ravel()      # e = ey + ex*nely -> row-major ex + ey*nx
    assert d.size == nx * ny and np.abs(d - xr).max() <= 1e-3


@pytest.mark.gpu
def test_topology_solver_2d_fea_only(gpu, oracle):
    nx, ny = 20, 8
    arg = _2d_js_config(nx, ny, 1)
    arg["config"]["solver"] = "fea"
    el = (np.random.default_rng(2).random(nx * ny) > 0.3).astype(float)
    arg["model"]["elements"] = el.tolist()
    out = run_js("run_optimize.js", arg)
    assert out["ok"] and out["feaOnly"], out
    x = np.zeros(nx * ny, np.float32)
    for ex in range(nx):
        for ey in range(ny):
            x[ey + ex * ny] = 1.0 if el[ex + ey * nx] > 0.5 else 1e-9
    U, c, _ = oracle.fe2(nx, ny, oracle.k0_2d(0.3), x, 3.0, oracle.fixed_dofs2(nx, ny, "left"),
                         oracle.load_vector2(nx, ny, "down", 1.0), tol=1e-8)
    assert abs(out["finalCompliance"] - c) <= 1e-6 * abs(c)
    assert len(out["elementStress"]) == nx * ny and abs(out["maxStress"] - max(out["elementStress"])) == 0


@pytest.mark.gpu
def test_wasm_facade_ebePCG_2d(gpu, oracle):
    """ebePCG with edofSize 8 by byte offset on the facade's memory (the
    2-D use of tests/test-wasm-ebePCG.js, graded densities of its Test 4)."""
    nelx, nely = 10, 10
    nel, ndof = T.dims2(nelx, nely)
    fixed = oracle.fixed_dofs2(nelx, nely, "left")
    free = np.flatnonzero(fixed == 0).astype(np.int32)
    F = np.zeros(ndof)
    F[2 * ((nely + 1) * nelx + nely) + 1] = -1.0
    dens = 0.2 + 0.8 * (np.arange(nel) / nel)
    n = 2
    n3, d3 = T.dims(n, n, n)
    out = run_js("facade_apply.js", {"n": n, "evals": [1.0] * n3, "p": [0.0] * d3, "dens": [1.0] * n3,
                                     "F": [0.0] * d3, "free": list(range(d3)),
                                     "twod": {"nelx": nelx, "nely": nely, "dens": dens.tolist(), "F": F.tolist(),
                                              "free": free.tolist()}})
    U = np.zeros(ndof)
    it = oracle.ebePCG_ed(dens, oracle.k0_2d(0.3), oracle.edofs2(nelx, nely), F, U, free, 1e-9, 1.0, 3.0, 2000, 1e-8)
    assert out["twod"]["iters"] == it
    assert np.abs(np.array(out["twod"]["U"]) - U).max() <= 1e-8
