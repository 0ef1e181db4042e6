"""topologyoptimization_amd -- MI355X (gfx950) SIMP inner loop behind a C ABI.

Python view of include/topo.h (ctypes). The compute path is the HIP library
libtopo_hip.so built in-tree by `python -m topologyoptimization_amd.build`;
importing this package with the library missing raises, and every compute
call goes to the GPU (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtopo_hip.so")

# enums of include/topo.h
FIX_LEFT, FIX_RIGHT, FIX_BOTTOM, FIX_TOP, FIX_BOTTOM_CORNERS, FIX_NONE = range(6)
DOF_X, DOF_Y, DOF_Z, DOF_ALL = 1, 2, 4, 7
LOAD_DOWN, LOAD_UP, LOAD_LEFT, LOAD_RIGHT, LOAD_TOP_CENTER, LOAD_NONE = range(6)
SOLVER_PCG, SOLVER_MGPCG, SOLVER_KSP_BDDC = 0, 1, 2
PREC_FP64, PREC_MIXED, PREC_MIXED_PC = 0, 1, 2
COARSE_SWEEPS, COARSE_CHOLESKY = 0, 1
FILTER_SENSITIVITY, FILTER_DENSITY, FILTER_HELMHOLTZ = 0, 1, 2
OBJ_COMPLIANCE, OBJ_MECHANISM = 0, 1
FIELD_U, FIELD_XPHYS, FIELD_ENERGY, FIELD_DC, FIELD_DCN, FIELD_STRESS, FIELD_ADJOINT = range(7)
TIME_APPLYA, TIME_VCYCLE, TIME_ENERGY, TIME_FILTER, TIME_OC, TIME_COPY, TIME_APPLYA_MASKED, TIME_APPLYA_MFMA = range(8)

ERRORS = {0: "OK", -1: "E_ARG", -2: "E_HIP", -3: "E_OOM", -4: "E_BREAKDOWN", -5: "E_UNSTRUCTURED",
          -6: "E_NODEVICE", -7: "E_STATE", -8: "E_UNSUPPORTED", -9: "E_CANCELLED"}


class TopoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nelx", C.c_int32), ("nely", C.c_int32), ("nelz", C.c_int32), ("nload", C.c_int32),
        ("E0", C.c_double), ("Emin", C.c_double), ("nu", C.c_double),
        ("volfrac", C.c_double), ("penal", C.c_double), ("penal_start", C.c_double),
        ("continuation_iters", C.c_int32), ("max_iterations", C.c_int32),
        ("rmin", C.c_double), ("filter_type", C.c_int32), ("use_projection", C.c_int32),
        ("beta_max", C.c_double), ("beta_interval", C.c_int32), ("beta_fixed", C.c_int32),
        ("eta", C.c_double), ("linear_solver", C.c_int32), ("precision", C.c_int32),
        ("mg_levels", C.c_int32), ("mg_nu1", C.c_int32), ("mg_nu2", C.c_int32),
        ("mg_omega", C.c_double), ("coarse_solver", C.c_int32), ("coarse_sweeps", C.c_int32),
        ("galerkin_max_ndof", C.c_int32), ("cg_max_iter", C.c_int32), ("cg_tol", C.c_double),
        ("warm_start", C.c_int32), ("move", C.c_double), ("objective", C.c_int32),
        ("mech_in_dof", C.c_int32), ("mech_out_dof", C.c_int32),
        ("mech_k_in", C.c_double), ("mech_k_out", C.c_double),
        ("rhs_tile", C.c_int32), ("use_graphs", C.c_int32), ("skip_factor", C.c_double), ("exact_apply", C.c_int32),
        ("mg_omega_cap", C.c_double),
        ("coarse_fp32", C.c_int32),
        ("energy_clamp", C.c_int32),
        ("mfma_apply", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]

    def copy(self):
        c = Config()
        C.pointer(c)[0] = self
        return c


class StepInfo(C.Structure):
    _fields_ = [
        ("iteration", C.c_int32), ("cg_iters", C.c_int32), ("outer_iters", C.c_int32),
        ("applyA_fine", C.c_int32), ("vcycles", C.c_int32), ("oc_passes", C.c_int32),
        ("compliance", C.c_double), ("change", C.c_double), ("penal", C.c_double), ("beta", C.c_double),
        ("tolerance", C.c_double), ("volume", C.c_double), ("max_stress", C.c_double),
        ("residual", C.c_double), ("t_solve_ms", C.c_double), ("t_total_ms", C.c_double),
        ("t_setup_ms", C.c_double), ("t_post_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MGLevelDesc(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nelx", "nely", "nelz", "ndof", "nel", "nfree", "activeCount", "hasDenseK")] + [
        (n, C.c_void_p) for n in ("edofs", "evals", "active", "fixedMask", "freeDofs", "invDiag", "denseK", "Au",
                                  "res", "b", "x", "scratch")]


class BddcSubDesc(C.Structure):
    _fields_ = [("numElements", C.c_int32), ("numDofs", C.c_int32), ("elements", C.c_void_p), ("dofs", C.c_void_p),
                ("localInvDiag", C.c_void_p)]


_lib = None

# exported C symbols declared in include/topo.h
SYMBOLS = [
    "topo_config_default", "topo_version", "topo_last_error", "topo_device_info", "topo_k0", "topo_edofs",
    "topo_fixed_dofs", "topo_load_vector", "topo_preserved_elements", "topo_load_paint", "topo_fixed_paint",
    "topo_applyAEbe3D", "topo_ebePCG", "topo_ebeMGPCG", "topo_ebeKSP_BDDC", "topo_solver_create", "topo_solver_destroy",
    "topo_solver_set_design", "topo_solver_get_design", "topo_solver_set_iteration", "topo_solver_update",
    "topo_solver_step", "topo_solver_get_field", "topo_solver_fe", "topo_solver_set_density", "topo_op_applyA",
    "topo_op_vcycle", "topo_op_energy", "topo_op_filter", "topo_op_filter_adjoint", "topo_op_oc", "topo_time_op",
    "topo_solver_cancel", "topo_solver_forward", "topo_solver_backward", "topo_op_helmholtz", "topo_solver_tuning",
    "topo_k0_2d", "topo_edofs_2d", "topo_fixed_dofs_2d", "topo_load_vector_2d", "topo_preserved_elements_2d",
    "topo_fixed_paint_2d", "topo_load_paint_2d", "topo_config_default_2d", "topo_solver2d_create",
    "topo_solver2d_destroy", "topo_solver2d_set_design", "topo_solver2d_get_design", "topo_solver2d_update",
    "topo_solver2d_step", "topo_solver2d_fea", "topo_solver2d_get_field", "topo_solver2d_cancel",
    "topo_ebePCG_noalloc",
]


def lib():
    """Load libtopo_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libtopo_hip.so not built: run `python -m topologyoptimization_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    L.topo_version.restype = C.c_char_p
    L.topo_last_error.restype = C.c_char_p
    L.topo_config_default.argtypes = [vp, i32, i32, i32]
    L.topo_device_info.argtypes = [vp, C.c_char_p, i32]
    L.topo_k0.argtypes = [dbl, vp]
    L.topo_edofs.argtypes = [i32, i32, i32, vp]
    L.topo_fixed_dofs.argtypes = [i32, i32, i32, i32, i32, vp]
    L.topo_load_vector.argtypes = [i32, i32, i32, i32, dbl, vp, vp]
    L.topo_preserved_elements.argtypes = [i32, i32, i32, i32, i32, vp]
    L.topo_load_paint.argtypes = [i32, i32, i32, vp, i32, vp, dbl, vp]
    L.topo_fixed_paint.argtypes = [i32, i32, i32, vp, i32, i32, vp]
    L.topo_applyAEbe3D.argtypes = [vp, vp, vp, vp, i32, vp, vp, i32, vp]
    L.topo_ebePCG.argtypes = [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, dbl, dbl, dbl, i32, dbl, vp]
    L.topo_ebePCG_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, dbl, dbl, dbl, i32, dbl]
    L.topo_ebeMGPCG.argtypes = [vp, i32, vp, vp, vp, vp, i32, i32, dbl, vp]
    L.topo_ebeKSP_BDDC.argtypes = [vp, i32, vp, i32, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, i32, dbl, i32, vp]
    L.topo_solver_create.argtypes = [vp, vp, vp, vp, vp, vp]
    L.topo_solver_destroy.argtypes = [vp]
    L.topo_solver_set_design.argtypes = [vp, vp]
    L.topo_solver_get_design.argtypes = [vp, vp]
    L.topo_solver_set_iteration.argtypes = [vp, i32]
    L.topo_solver_update.argtypes = [vp, dbl, dbl, dbl, i32]
    L.topo_solver_step.argtypes = [vp, vp]
    L.topo_solver_get_field.argtypes = [vp, i32, vp]
    L.topo_solver_fe.argtypes = [vp, vp, dbl, i32, i32, vp, vp, vp]
    L.topo_solver_set_density.argtypes = [vp, vp, dbl]
    L.topo_op_applyA.argtypes = [vp, i32, i32, vp, vp]
    L.topo_op_vcycle.argtypes = [vp, i32, vp, vp]
    L.topo_op_energy.argtypes = [vp, vp, vp]
    L.topo_op_filter.argtypes = [vp, i32, vp, vp, vp]
    L.topo_op_filter_adjoint.argtypes = [vp, i32, vp, vp]
    L.topo_op_oc.argtypes = [vp, vp, vp, vp, dbl]
    L.topo_time_op.argtypes = [vp, i32, i32, i32, i32, vp, vp]
    L.topo_solver_cancel.argtypes = [vp]
    L.topo_solver_forward.argtypes = [vp, vp, vp, vp]
    L.topo_solver_backward.argtypes = [vp, vp, vp]
    L.topo_op_helmholtz.argtypes = [vp, vp, vp, vp]
    L.topo_solver_tuning.argtypes = [vp, vp, vp, vp, vp, i32]
    L.topo_k0_2d.argtypes = [dbl, vp]
    L.topo_edofs_2d.argtypes = [i32, i32, vp]
    L.topo_fixed_dofs_2d.argtypes = [i32, i32, i32, i32, vp]
    L.topo_load_vector_2d.argtypes = [i32, i32, i32, dbl, vp, vp]
    L.topo_preserved_elements_2d.argtypes = [i32, i32, i32, i32, vp]
    L.topo_fixed_paint_2d.argtypes = [i32, i32, vp, i32, i32, vp]
    L.topo_load_paint_2d.argtypes = [i32, i32, vp, i32, vp, dbl, vp]
    L.topo_config_default_2d.argtypes = [vp, i32, i32]
    L.topo_solver2d_create.argtypes = [vp, vp, vp, vp, vp, vp]
    L.topo_solver2d_destroy.argtypes = [vp]
    L.topo_solver2d_set_design.argtypes = [vp, vp]
    L.topo_solver2d_get_design.argtypes = [vp, vp]
    L.topo_solver2d_update.argtypes = [vp, dbl, dbl, dbl, i32]
    L.topo_solver2d_step.argtypes = [vp, vp]
    L.topo_solver2d_fea.argtypes = [vp, vp, dbl, dbl, vp]
    L.topo_solver2d_get_field.argtypes = [vp, i32, vp]
    L.topo_solver2d_cancel.argtypes = [vp]
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise TopoError(rc, lib().topo_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dims(nelx, nely, nelz):
    return nelx * nely * nelz, 3 * (nelx + 1) * (nely + 1) * (nelz + 1)


def default_config(nelx, nely, nelz, **kw):
    c = Config()
    _check(lib().topo_config_default(C.byref(c), nelx, nely, nelz))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def device_info():
    n = C.c_int32(0)
    buf = C.create_string_buffer(256)
    lib().topo_device_info(C.byref(n), buf, 256)
    return n.value, buf.value.decode()


def k0(nu=0.3):
    ke = np.zeros(576)
    _check(lib().topo_k0(nu, _p(ke)))
    return ke.reshape(24, 24)


def edofs(nelx, nely, nelz):
    nel, _ = dims(nelx, nely, nelz)
    e = np.zeros(nel * 24, np.int32)
    _check(lib().topo_edofs(nelx, nely, nelz, _p(e)))
    return e.reshape(nel, 24)


def fixed_dofs(nelx, nely, nelz, position, dof_mask=DOF_ALL, mask=None):
    _, ndof = dims(nelx, nely, nelz)
    m = np.zeros(ndof, np.uint8) if mask is None else mask
    _check(lib().topo_fixed_dofs(nelx, nely, nelz, position, dof_mask, _p(m)))
    return m


def load_vector(nelx, nely, nelz, direction, magnitude=1.0, force_vector=None):
    _, ndof = dims(nelx, nely, nelz)
    F = np.zeros(ndof, np.float32)
    fv = None if force_vector is None else np.asarray(force_vector, np.float64)
    _check(lib().topo_load_vector(nelx, nely, nelz, direction, magnitude, _p(fv), _p(F)))
    return F


def preserved_elements(nelx, nely, nelz, position, direction):
    nel, _ = dims(nelx, nely, nelz)
    m = np.zeros(nel, np.uint8)
    _check(lib().topo_preserved_elements(nelx, nely, nelz, position, direction, _p(m)))
    return m


def load_paint(nelx, nely, nelz, faces, direction3, magnitude, F=None):
    _, ndof = dims(nelx, nely, nelz)
    F = np.zeros(ndof, np.float32) if F is None else F
    fa = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 4))
    d = np.asarray(direction3, np.float64)
    _check(lib().topo_load_paint(nelx, nely, nelz, _p(fa), fa.shape[0], _p(d), magnitude, _p(F)))
    return F


# ---- WASM-op twins ---------------------------------------------------------
def applyAEbe3D(ke, edofs_arr, evals, active, p):
    ke = np.ascontiguousarray(ke, np.float64).ravel()
    ed = np.ascontiguousarray(edofs_arr, np.int32).ravel()
    evals = np.ascontiguousarray(evals, np.float64)
    active = np.ascontiguousarray(active, np.int32)
    p = np.ascontiguousarray(p, np.float64)
    Ap = np.zeros_like(p)
    _check(lib().topo_applyAEbe3D(_p(ke), _p(ed), _p(evals), _p(active), len(active), _p(p), _p(Ap), p.size, None))
    return Ap


def ebePCG(dens, ke, edofs_arr, F, U, freedofs, Emin, E0, penal, maxIter, tol):
    """ebePCG twin; the element DOF count (24 hex8, 8 quad4) is taken from
    ke (24x24 or 8x8)."""
    dens = np.ascontiguousarray(dens, np.float64)
    ke = np.ascontiguousarray(ke, np.float64).ravel()
    esz = int(round(np.sqrt(ke.size)))
    ed = np.ascontiguousarray(edofs_arr, np.int32).ravel()
    F = np.ascontiguousarray(F, np.float64)
    freedofs = np.ascontiguousarray(freedofs, np.int32)
    assert U.dtype == np.float64 and U.flags.c_contiguous
    return _check(lib().topo_ebePCG(_p(dens), _p(ke), _p(ed), _p(F), _p(U), _p(freedofs), dens.size, esz, F.size,
                                    freedofs.size, Emin, E0, penal, maxIter, tol, None))


def ebePCG_noalloc(dens, ke, edofs_arr, F, U, freedofs, Emin, E0, penal, maxIter, tol):
    """topo_ebePCG_noalloc (3-D hex8): ebePCG with no device allocation; U
    is the warm start and is overwritten."""
    dens = np.ascontiguousarray(dens, np.float64)
    ke = np.ascontiguousarray(ke, np.float64).ravel()
    ed = np.ascontiguousarray(edofs_arr, np.int32).ravel()
    F = np.ascontiguousarray(F, np.float64)
    freedofs = np.ascontiguousarray(freedofs, np.int32)
    assert U.dtype == np.float64 and U.flags.c_contiguous
    return _check(lib().topo_ebePCG_noalloc(_p(dens), _p(ke), _p(ed), _p(F), _p(U), _p(freedofs), dens.size, 24,
                                            F.size, freedofs.size, Emin, E0, penal, maxIter, tol))


def ebeMGPCG(levels, ke, F, U, fixed0, maxIter, tol):
    """levels: list of dicts with nelx,nely,nelz,edofs,evals,active,fixedMask,freeDofs,invDiag,denseK."""
    keep = []
    arr = (MGLevelDesc * len(levels))()
    for i, L in enumerate(levels):
        d = arr[i]
        nel, ndof = dims(L["nelx"], L["nely"], L["nelz"])
        d.nelx, d.nely, d.nelz, d.ndof, d.nel = L["nelx"], L["nely"], L["nelz"], ndof, nel
        for key, dt in (("edofs", np.int32), ("evals", np.float64), ("active", np.int32), ("fixedMask", np.uint8),
                        ("freeDofs", np.int32), ("invDiag", np.float64), ("denseK", np.float64)):
            v = L.get(key)
            if v is not None:
                v = np.ascontiguousarray(v, dt).ravel()
                keep.append(v)
                setattr(d, key, v.ctypes.data)
        d.activeCount = len(L["active"]) if L.get("active") is not None else 0
        d.nfree = len(L["freeDofs"]) if L.get("freeDofs") is not None else 0
        d.hasDenseK = int(L.get("denseK") is not None)
    ke = np.ascontiguousarray(ke, np.float64).ravel()
    F = np.ascontiguousarray(F, np.float64)
    fixed0 = np.ascontiguousarray(fixed0, np.uint8)
    return _check(lib().topo_ebeMGPCG(arr, len(levels), _p(ke), _p(F), _p(U), _p(fixed0), F.size, maxIter, tol, None))


def ebeKSP_BDDC(subs, coarse, ke, edofs_arr, evals, active, fixed, invDiag, F, maxIter, tol, smoother_iters=3):
    """subs: list of (elements, dofs, localInvDiag[ndof]) per subdomain;
    returns (U, CG iterations)."""
    keep = []
    arr = (BddcSubDesc * len(subs))()
    for i, (el, ds, li) in enumerate(subs):
        el = np.ascontiguousarray(el, np.int32)
        ds = np.ascontiguousarray(ds, np.int32)
        li = np.ascontiguousarray(li, np.float64)
        keep += [el, ds, li]
        arr[i].numElements, arr[i].numDofs = el.size, ds.size
        arr[i].elements, arr[i].dofs, arr[i].localInvDiag = el.ctypes.data, ds.ctypes.data, li.ctypes.data
    coarse = np.ascontiguousarray(coarse, np.int32)
    ke = np.ascontiguousarray(ke, np.float64).ravel()
    ed = np.ascontiguousarray(edofs_arr, np.int32).ravel()
    evals = np.ascontiguousarray(evals, np.float64)
    active = np.ascontiguousarray(active, np.int32)
    fixed = np.ascontiguousarray(fixed, np.uint8)
    invDiag = np.ascontiguousarray(invDiag, np.float64)
    F = np.ascontiguousarray(F, np.float64)
    U = np.zeros(F.size)
    it = _check(lib().topo_ebeKSP_BDDC(arr, len(subs), _p(coarse), coarse.size, _p(ke), _p(ed), _p(evals), _p(active),
                                       active.size, _p(fixed), _p(invDiag), _p(F), _p(U), F.size, maxIter, tol,
                                       smoother_iters, None))
    return U, it


# ---- device-resident solver ------------------------------------------------
class Solver:
    """Handle of topo_solver (device-resident SIMP state)."""

    def __init__(self, cfg, fixed, F, preserved=None, void=None):
        self.cfg = cfg.copy()
        self.nel, self.ndof = dims(cfg.nelx, cfg.nely, cfg.nelz)
        fixed = np.ascontiguousarray(fixed, np.uint8)
        F = np.ascontiguousarray(F, np.float32)
        assert fixed.size == self.ndof and F.size == self.ndof * cfg.nload
        pres = None if preserved is None else np.ascontiguousarray(preserved, np.uint8)
        vd = None if void is None else np.ascontiguousarray(void, np.uint8)
        h = C.c_void_p()
        _check(lib().topo_solver_create(C.byref(self.cfg), _p(fixed), _p(F), _p(pres), _p(vd), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().topo_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_design(self, x):
        x = np.ascontiguousarray(x, np.float32)
        _check(lib().topo_solver_set_design(self.h, _p(x)))

    def design(self):
        x = np.zeros(self.nel, np.float32)
        _check(lib().topo_solver_get_design(self.h, _p(x)))
        return x

    def step(self):
        info = StepInfo()
        _check(lib().topo_solver_step(self.h, C.byref(info)))
        return info

    def set_iteration(self, loop):
        """iterations already done (the next step() is iteration loop + 1)"""
        _check(lib().topo_solver_set_iteration(self.h, loop))

    def field(self, which):
        if which == FIELD_U:
            out = np.zeros(self.ndof * self.cfg.nload)
        elif which == FIELD_ADJOINT:
            out = np.zeros(self.ndof)
        else:
            out = np.zeros(self.nel, np.float32)
        _check(lib().topo_solver_get_field(self.h, which, _p(out)))
        return out

    def set_density(self, xphys, penal):
        xphys = np.ascontiguousarray(xphys, np.float32)
        _check(lib().topo_solver_set_density(self.h, _p(xphys), penal))

    def fe(self, xphys, penal, iteration=1, max_iterations=1, U=None):
        xphys = np.ascontiguousarray(xphys, np.float32)
        U = np.zeros(self.ndof * self.cfg.nload) if U is None else np.ascontiguousarray(U, np.float64).copy()
        c = C.c_double()
        info = StepInfo()
        _check(lib().topo_solver_fe(self.h, _p(xphys), penal, iteration, max_iterations, _p(U), C.byref(c),
                                    C.byref(info)))
        return U, c.value, info

    def applyA(self, p, level=0, prec=PREC_FP64):
        p = np.ascontiguousarray(p, np.float64)
        Ap = np.zeros_like(p)
        _check(lib().topo_op_applyA(self.h, level, prec, _p(p), _p(Ap)))
        return Ap

    def vcycle(self, r, prec=PREC_FP64):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        _check(lib().topo_op_vcycle(self.h, prec, _p(r), _p(z)))
        return z

    def energy(self, U):
        U = np.ascontiguousarray(U, np.float64)
        w = np.zeros(self.nel, np.float32)
        _check(lib().topo_op_energy(self.h, _p(U), _p(w)))
        return w

    def filter(self, ftype, x, v):
        x = None if x is None else np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        out = np.zeros_like(v)
        _check(lib().topo_op_filter(self.h, ftype, _p(x), _p(v), _p(out)))
        return out

    def filter_adjoint(self, ftype, v):
        v = np.ascontiguousarray(v, np.float32)
        out = np.zeros_like(v)
        _check(lib().topo_op_filter_adjoint(self.h, ftype, _p(v), _p(out)))
        return out

    def oc(self, x, dcn, volfrac):
        x = np.ascontiguousarray(x, np.float32)
        dcn = np.ascontiguousarray(dcn, np.float32)
        xnew = np.zeros_like(x)
        n = _check(lib().topo_op_oc(self.h, _p(x), _p(dcn), _p(xnew), volfrac))
        return xnew, n

    def time_op(self, op, prec=PREC_FP64, level=0, reps=20):
        ms = C.c_double()
        by = C.c_double()
        _check(lib().topo_time_op(self.h, op, prec, level, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def cancel(self):
        lib().topo_solver_cancel(self.h)

    def forward(self):
        """objective, mean(xPhys), StepInfo of the FP64 forward pass at the current design"""
        J = C.c_double()
        vol = C.c_double()
        info = StepInfo()
        _check(lib().topo_solver_forward(self.h, C.byref(J), C.byref(vol), C.byref(info)))
        return J.value, vol.value, info

    def backward(self):
        """dJ/dx and d mean(xPhys)/dx (float64, reference element order)"""
        dJ = np.zeros(self.nel)
        dV = np.zeros(self.nel)
        _check(lib().topo_solver_backward(self.h, _p(dJ), _p(dV)))
        return dJ, dV

    def tuning(self):
        """(rhs_tile, xm_nt, xm_cx, log[n x (R, NT, CX, ms per RHS)]) of the batched-solve autotuner"""
        r, nt, cx = C.c_int32(), C.c_int32(), C.c_int32()
        log = np.zeros((256, 4))
        n = _check(lib().topo_solver_tuning(self.h, C.byref(r), C.byref(nt), C.byref(cx), _p(log), 256))
        return r.value, nt.value, cx.value, log[:n]

    def helmholtz(self, v):
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros_like(v)
        it = C.c_int32()
        _check(lib().topo_op_helmholtz(self.h, _p(v), _p(out), C.byref(it)))
        return out, it.value


# ---- 2-D path (js/optimizer-worker.js) ---------------------------------------
def dims2(nelx, nely):
    return nelx * nely, 2 * (nelx + 1) * (nely + 1)


def default_config_2d(nelx, nely, **kw):
    c = Config()
    _check(lib().topo_config_default_2d(C.byref(c), nelx, nely))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def k0_2d(nu=0.3):
    ke = np.zeros(64)
    _check(lib().topo_k0_2d(nu, _p(ke)))
    return ke.reshape(8, 8)


def edofs_2d(nelx, nely):
    ed = np.zeros(nelx * nely * 8, np.int32)
    _check(lib().topo_edofs_2d(nelx, nely, _p(ed)))
    return ed.reshape(-1, 8)


def fixed_dofs_2d(nelx, nely, position, dof_mask=DOF_ALL, mask=None):
    _, ndof = dims2(nelx, nely)
    mask = np.zeros(ndof, np.uint8) if mask is None else mask
    _check(lib().topo_fixed_dofs_2d(nelx, nely, position, dof_mask, _p(mask)))
    return mask


def load_vector_2d(nelx, nely, direction, magnitude=1.0, force_vector=None):
    _, ndof = dims2(nelx, nely)
    F = np.zeros(ndof, np.float32)
    fv = None if force_vector is None else np.ascontiguousarray(force_vector, np.float64)
    _check(lib().topo_load_vector_2d(nelx, nely, direction, magnitude, _p(fv), _p(F)))
    return F


def preserved_elements_2d(nelx, nely, position, direction):
    pres = np.zeros(nelx * nely, np.uint8)
    _check(lib().topo_preserved_elements_2d(nelx, nely, position, direction, _p(pres)))
    return pres


def fixed_paint_2d(nelx, nely, cells, dof_mask=DOF_ALL, mask=None):
    _, ndof = dims2(nelx, nely)
    mask = np.zeros(ndof, np.uint8) if mask is None else mask
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 2)
    _check(lib().topo_fixed_paint_2d(nelx, nely, _p(cells), cells.shape[0], dof_mask, _p(mask)))
    return mask


def load_paint_2d(nelx, nely, cells, direction2, magnitude, F=None):
    _, ndof = dims2(nelx, nely)
    F = np.zeros(ndof, np.float32) if F is None else F
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 2)
    d = np.ascontiguousarray(direction2, np.float64)
    _check(lib().topo_load_paint_2d(nelx, nely, _p(cells), cells.shape[0], _p(d), magnitude, _p(F)))
    return F


class Solver2D:
    """Handle of topo_solver2d (device-resident 2-D SIMP state)."""

    def __init__(self, cfg, fixed, F, preserved=None, void=None):
        self.cfg = cfg.copy()
        self.nel, self.ndof = dims2(cfg.nelx, cfg.nely)
        fixed = np.ascontiguousarray(fixed, np.uint8)
        F = np.ascontiguousarray(F, np.float32)
        assert fixed.size == self.ndof and F.size == self.ndof
        pres = None if preserved is None else np.ascontiguousarray(preserved, np.uint8)
        vd = None if void is None else np.ascontiguousarray(void, np.uint8)
        h = C.c_void_p()
        _check(lib().topo_solver2d_create(C.byref(self.cfg), _p(fixed), _p(F), _p(pres), _p(vd), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().topo_solver2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_design(self, x):
        x = np.ascontiguousarray(x, np.float32)
        _check(lib().topo_solver2d_set_design(self.h, _p(x)))

    def design(self):
        x = np.zeros(self.nel, np.float32)
        _check(lib().topo_solver2d_get_design(self.h, _p(x)))
        return x

    def update(self, penal=0.0, rmin=0.0, volfrac=0.0, max_iterations=0):
        _check(lib().topo_solver2d_update(self.h, penal, rmin, volfrac, max_iterations))

    def step(self):
        info = StepInfo()
        _check(lib().topo_solver2d_step(self.h, C.byref(info)))
        return info

    def fea(self, x, penal=3.0, tol=1e-8):
        x = np.ascontiguousarray(x, np.float32)
        info = StepInfo()
        _check(lib().topo_solver2d_fea(self.h, _p(x), penal, tol, C.byref(info)))
        return info

    def field(self, which):
        out = np.zeros(self.ndof) if which == FIELD_U else np.zeros(self.nel, np.float32)
        _check(lib().topo_solver2d_get_field(self.h, which, _p(out)))
        return out

    def cancel(self):
        lib().topo_solver2d_cancel(self.h)
