// topo_addon.cpp -- Node-API binding of libtopo_hip (include/topo.h).
//
// Context-aware (NAPI_MODULE_INIT), so it loads in every worker_threads
// Worker, which is where the reference runs its solver
// (lib/topology-solver.js:96-106). Typed arrays are passed to the C ABI as
// plain pointers, without copies; errors of the library become JS
// exceptions carrying topo_last_error().
#define NAPI_VERSION 6
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "topo.h"

namespace {

#define NAPI_CALL(env, call)                                          \
    do {                                                              \
        napi_status _s = (call);                                      \
        if (_s != napi_ok) {                                          \
            napi_throw_error((env), nullptr, "N-API call failed: " #call); \
            return nullptr;                                           \
        }                                                             \
    } while (0)

napi_value throw_topo(napi_env env, int rc) {
    std::string msg = std::string("libtopo_hip error ") + std::to_string(rc) + ": " + topo_last_error();
    napi_throw_error(env, std::to_string(rc).c_str(), msg.c_str());
    return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
    if (argc < want) {
        napi_throw_type_error(env, nullptr, "too few arguments");
        return false;
    }
    return true;
}

int32_t to_i32(napi_env env, napi_value v) {
    int32_t r = 0;
    napi_get_value_int32(env, v, &r);
    return r;
}
double to_f64(napi_env env, napi_value v) {
    double r = 0;
    napi_get_value_double(env, v, &r);
    return r;
}
bool is_nullish(napi_env env, napi_value v) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    return t == napi_undefined || t == napi_null;
}
// data pointer and element count of a typed array (nullptr for null/undefined)
void* ta_ptr(napi_env env, napi_value v, size_t* len = nullptr) {
    if (is_nullish(env, v)) return nullptr;
    bool ok = false;
    napi_is_typedarray(env, v, &ok);
    if (!ok) {
        napi_throw_type_error(env, nullptr, "expected a TypedArray");
        return nullptr;
    }
    napi_typedarray_type type;
    size_t length = 0;
    void* data = nullptr;
    napi_value ab;
    size_t off = 0;
    napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &off);
    if (len) *len = length;
    return data;
}

napi_value num(napi_env env, double v) {
    napi_value r;
    napi_create_double(env, v, &r);
    return r;
}
napi_value i32v(napi_env env, int32_t v) {
    napi_value r;
    napi_create_int32(env, v, &r);
    return r;
}
void set(napi_env env, napi_value obj, const char* k, napi_value v) { napi_set_named_property(env, obj, k, v); }

napi_value get(napi_env env, napi_value obj, const char* k) {
    napi_value v;
    if (napi_get_named_property(env, obj, k, &v) != napi_ok) return nullptr;
    return v;
}

// ---------------------------------------------------------------------------
napi_value Version(napi_env env, napi_callback_info) {
    napi_value r;
    napi_create_string_utf8(env, topo_version(), NAPI_AUTO_LENGTH, &r);
    return r;
}

napi_value DeviceInfo(napi_env env, napi_callback_info) {
    int32_t n = 0;
    char name[256];
    topo_device_info(&n, name, sizeof name);
    napi_value o, s;
    napi_create_object(env, &o);
    set(env, o, "count", i32v(env, n));
    napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s);
    set(env, o, "name", s);
    return o;
}

// configDefault(nelx, nely, nelz) -> plain object with the topo_config fields
#define CFG_FIELDS(X)                                                                              \
    X(i, nelx) X(i, nely) X(i, nelz) X(i, nload) X(d, E0) X(d, Emin) X(d, nu) X(d, volfrac)           \
    X(d, penal) X(d, penal_start) X(i, continuation_iters) X(i, max_iterations) X(d, rmin)            \
    X(i, filter_type) X(i, use_projection) X(d, beta_max) X(i, beta_interval) X(i, beta_fixed)        \
    X(d, eta) X(i, linear_solver) X(i, precision) X(i, mg_levels) X(i, mg_nu1) X(i, mg_nu2)           \
    X(d, mg_omega) X(i, coarse_solver) X(i, coarse_sweeps) X(i, galerkin_max_ndof) X(i, cg_max_iter)  \
    X(d, cg_tol) X(i, warm_start) X(d, move) X(i, objective) X(i, mech_in_dof) X(i, mech_out_dof)     \
    X(d, mech_k_in) X(d, mech_k_out) X(i, rhs_tile) X(i, use_graphs) X(d, skip_factor)                \
    X(i, exact_apply) X(d, mg_omega_cap) X(i, coarse_fp32) X(i, energy_clamp) X(i, mfma_apply)

napi_value cfg_to_js(napi_env env, const topo_config& c) {
    napi_value o;
    napi_create_object(env, &o);
#define X(t, f) set(env, o, #f, num(env, (double)c.f));
    CFG_FIELDS(X)
#undef X
    return o;
}

void cfg_from_js(napi_env env, napi_value o, topo_config* c) {
#define X(t, f)                                                  \
    {                                                            \
        bool has = false;                                        \
        napi_has_named_property(env, o, #f, &has);               \
        if (has) {                                               \
            napi_value v = get(env, o, #f);                      \
            if (!is_nullish(env, v)) c->f = (decltype(c->f))to_f64(env, v); \
        }                                                        \
    }
    CFG_FIELDS(X)
#undef X
}

napi_value ConfigDefault(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return nullptr;
    topo_config c;
    int rc = topo_config_default(&c, to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]));
    if (rc < 0) return throw_topo(env, rc);
    return cfg_to_js(env, c);
}

// k0(nu, Float64Array(576))
napi_value K0(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_k0(to_f64(env, a[0]), (double*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

// edofs(nelx, nely, nelz, Int32Array)
napi_value Edofs(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!get_args(env, info, 4, a)) return nullptr;
    int rc = topo_edofs(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), (int32_t*)ta_ptr(env, a[3]));
    if (rc < 0) return throw_topo(env, rc);
    return a[3];
}

// fixedDofs(nelx, nely, nelz, position, dofMask, Uint8Array) -> count
napi_value FixedDofs(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    int rc = topo_fixed_dofs(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_i32(env, a[3]),
                             to_i32(env, a[4]), (uint8_t*)ta_ptr(env, a[5]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// loadVector(nelx, nely, nelz, direction, magnitude, forceVector|null, Float32Array)
napi_value LoadVector(napi_env env, napi_callback_info info) {
    napi_value a[7];
    if (!get_args(env, info, 7, a)) return nullptr;
    int rc = topo_load_vector(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_i32(env, a[3]),
                              to_f64(env, a[4]), (const double*)ta_ptr(env, a[5]), (float*)ta_ptr(env, a[6]));
    if (rc < 0) return throw_topo(env, rc);
    return a[6];
}

// preservedElements(nelx, nely, nelz, position, direction, Uint8Array)
napi_value Preserved(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    int rc = topo_preserved_elements(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_i32(env, a[3]),
                                     to_i32(env, a[4]), (uint8_t*)ta_ptr(env, a[5]));
    if (rc < 0) return throw_topo(env, rc);
    return a[5];
}

// loadPaint(nelx, nely, nelz, Int32Array faces(n*4), Float64Array dir3, magnitude, Float32Array F) -> nodes
napi_value LoadPaint(napi_env env, napi_callback_info info) {
    napi_value a[7];
    if (!get_args(env, info, 7, a)) return nullptr;
    size_t n = 0;
    const int32_t* faces = (const int32_t*)ta_ptr(env, a[3], &n);
    int rc = topo_load_paint(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), faces, (int32_t)(n / 4),
                             (const double*)ta_ptr(env, a[4]), to_f64(env, a[5]), (float*)ta_ptr(env, a[6]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// fixedPaint(nelx, nely, nelz, Int32Array faces, dofMask, Uint8Array) -> count
napi_value FixedPaint(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    size_t n = 0;
    const int32_t* faces = (const int32_t*)ta_ptr(env, a[3], &n);
    int rc = topo_fixed_paint(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), faces, (int32_t)(n / 4),
                              to_i32(env, a[4]), (uint8_t*)ta_ptr(env, a[5]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// ---- WASM-op twins ----------------------------------------------------------
// applyAEbe3D(ke, edofs, evals, active, activeCount, p, Ap, ndof)
napi_value ApplyAEbe3D(napi_env env, napi_callback_info info) {
    napi_value a[8];
    if (!get_args(env, info, 8, a)) return nullptr;
    int rc = topo_applyAEbe3D((const double*)ta_ptr(env, a[0]), (const int32_t*)ta_ptr(env, a[1]),
                              (const double*)ta_ptr(env, a[2]), (const int32_t*)ta_ptr(env, a[3]), to_i32(env, a[4]),
                              (const double*)ta_ptr(env, a[5]), (double*)ta_ptr(env, a[6]), to_i32(env, a[7]), nullptr);
    if (rc < 0) return throw_topo(env, rc);
    return a[6];
}

// ebePCG(dens, ke, edofs, F, U, freedofs, nel, edofSize, ndof, nfree, Emin, E0, penal, maxIter, tol) -> iters
napi_value EbePCG(napi_env env, napi_callback_info info) {
    napi_value a[15];
    if (!get_args(env, info, 15, a)) return nullptr;
    int rc = topo_ebePCG((const double*)ta_ptr(env, a[0]), (const double*)ta_ptr(env, a[1]),
                         (const int32_t*)ta_ptr(env, a[2]), (const double*)ta_ptr(env, a[3]), (double*)ta_ptr(env, a[4]),
                         (const int32_t*)ta_ptr(env, a[5]), to_i32(env, a[6]), to_i32(env, a[7]), to_i32(env, a[8]),
                         to_i32(env, a[9]), to_f64(env, a[10]), to_f64(env, a[11]), to_f64(env, a[12]),
                         to_i32(env, a[13]), to_f64(env, a[14]), nullptr);
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// ebeMGPCG(levels[], ke, F, U, fixedMask0, ndof0, maxIter, tol) -> iters; each
// level: {nelx,nely,nelz,ndof,nel,nfree,activeCount,hasDenseK,edofs,evals,
// active,fixedMask,freeDofs,invDiag,denseK}
napi_value EbeMGPCG(napi_env env, napi_callback_info info) {
    napi_value a[8];
    if (!get_args(env, info, 8, a)) return nullptr;
    uint32_t nl = 0;
    napi_get_array_length(env, a[0], &nl);
    std::vector<topo_mg_level> L(nl);
    for (uint32_t i = 0; i < nl; i++) {
        napi_value o;
        napi_get_element(env, a[0], i, &o);
        topo_mg_level& d = L[i];
        std::memset(&d, 0, sizeof d);
        d.nelx = to_i32(env, get(env, o, "nelx"));
        d.nely = to_i32(env, get(env, o, "nely"));
        d.nelz = to_i32(env, get(env, o, "nelz"));
        d.ndof = to_i32(env, get(env, o, "ndof"));
        d.nel = to_i32(env, get(env, o, "nel"));
        d.nfree = to_i32(env, get(env, o, "nfree"));
        d.activeCount = to_i32(env, get(env, o, "activeCount"));
        d.hasDenseK = to_i32(env, get(env, o, "hasDenseK"));
        d.edofs = (const int32_t*)ta_ptr(env, get(env, o, "edofs"));
        d.evals = (const double*)ta_ptr(env, get(env, o, "evals"));
        d.active = (const int32_t*)ta_ptr(env, get(env, o, "active"));
        d.fixedMask = (const uint8_t*)ta_ptr(env, get(env, o, "fixedMask"));
        d.freeDofs = (const int32_t*)ta_ptr(env, get(env, o, "freeDofs"));
        d.invDiag = (const double*)ta_ptr(env, get(env, o, "invDiag"));
        d.denseK = d.hasDenseK ? (const double*)ta_ptr(env, get(env, o, "denseK")) : nullptr;
    }
    int rc = topo_ebeMGPCG(L.data(), (int32_t)nl, (const double*)ta_ptr(env, a[1]), (const double*)ta_ptr(env, a[2]),
                           (double*)ta_ptr(env, a[3]), (const uint8_t*)ta_ptr(env, a[4]), to_i32(env, a[5]),
                           to_i32(env, a[6]), to_f64(env, a[7]), nullptr);
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// ebeKSP_BDDC(subs[{numElements, numDofs, elements, dofs, localInvDiag}],
//             coarse, ke, edofs, evals, active, fixedMask, invDiag, F, U,
//             ndof, maxIter, tolerance, smootherIters) -> iterations
napi_value EbeKSPBDDC(napi_env env, napi_callback_info info) {
    napi_value a[14];
    if (!get_args(env, info, 14, a)) return nullptr;
    uint32_t ns = 0;
    napi_get_array_length(env, a[0], &ns);
    std::vector<topo_bddc_sub> S(ns);
    for (uint32_t i = 0; i < ns; i++) {
        napi_value o;
        napi_get_element(env, a[0], i, &o);
        topo_bddc_sub& d = S[i];
        std::memset(&d, 0, sizeof d);
        d.numElements = to_i32(env, get(env, o, "numElements"));
        d.numDofs = to_i32(env, get(env, o, "numDofs"));
        d.elements = (const int32_t*)ta_ptr(env, get(env, o, "elements"));
        d.dofs = (const int32_t*)ta_ptr(env, get(env, o, "dofs"));
        d.localInvDiag = (const double*)ta_ptr(env, get(env, o, "localInvDiag"));
    }
    size_t ncoarse = 0, nactive = 0;
    const int32_t* coarse = (const int32_t*)ta_ptr(env, a[1], &ncoarse);
    const int32_t* active = (const int32_t*)ta_ptr(env, a[5], &nactive);
    int rc = topo_ebeKSP_BDDC(S.data(), (int32_t)ns, coarse, (int32_t)ncoarse, (const double*)ta_ptr(env, a[2]),
                              (const int32_t*)ta_ptr(env, a[3]), (const double*)ta_ptr(env, a[4]), active,
                              (int32_t)nactive, (const uint8_t*)ta_ptr(env, a[6]), (const double*)ta_ptr(env, a[7]),
                              (const double*)ta_ptr(env, a[8]), (double*)ta_ptr(env, a[9]), to_i32(env, a[10]),
                              to_i32(env, a[11]), to_f64(env, a[12]), to_i32(env, a[13]), nullptr);
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// ---- device-resident solver --------------------------------------------------
void solver_finalize(napi_env, void* data, void*) { topo_solver_destroy((topo_solver*)data); }

topo_solver* solver_of(napi_env env, napi_value v) {
    void* p = nullptr;
    napi_get_value_external(env, v, &p);
    if (!p) napi_throw_error(env, nullptr, "solver handle already destroyed");
    return (topo_solver*)p;
}

// createSolver(config, fixedMask, F, preserved|null, void|null) -> handle
napi_value CreateSolver(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    topo_config c;
    const int32_t nx = to_i32(env, get(env, a[0], "nelx")), ny = to_i32(env, get(env, a[0], "nely")),
                  nz = to_i32(env, get(env, a[0], "nelz"));
    int rc = topo_config_default(&c, nx, ny, nz);
    if (rc < 0) return throw_topo(env, rc);
    cfg_from_js(env, a[0], &c);
    topo_solver* s = nullptr;
    rc = topo_solver_create(&c, (const uint8_t*)ta_ptr(env, a[1]), (const float*)ta_ptr(env, a[2]),
                            (const uint8_t*)ta_ptr(env, a[3]), (const uint8_t*)ta_ptr(env, a[4]), &s);
    if (rc < 0) return throw_topo(env, rc);
    napi_value h;
    NAPI_CALL(env, napi_create_external(env, s, solver_finalize, nullptr, &h));
    return h;
}

napi_value info_to_js(napi_env env, const topo_step_info& i) {
    napi_value o;
    napi_create_object(env, &o);
    set(env, o, "iteration", i32v(env, i.iteration));
    set(env, o, "cgIterations", i32v(env, i.cg_iters));
    set(env, o, "outerIterations", i32v(env, i.outer_iters));
    set(env, o, "applyA", i32v(env, i.applyA_fine));
    set(env, o, "vcycles", i32v(env, i.vcycles));
    set(env, o, "ocPasses", i32v(env, i.oc_passes));
    set(env, o, "compliance", num(env, i.compliance));
    set(env, o, "change", num(env, i.change));
    set(env, o, "penal", num(env, i.penal));
    set(env, o, "beta", num(env, i.beta));
    set(env, o, "tolerance", num(env, i.tolerance));
    set(env, o, "volume", num(env, i.volume));
    set(env, o, "maxStress", num(env, i.max_stress));
    set(env, o, "solveMs", num(env, i.t_solve_ms));
    set(env, o, "totalMs", num(env, i.t_total_ms));
    set(env, o, "setupMs", num(env, i.t_setup_ms));
    set(env, o, "postMs", num(env, i.t_post_ms));
    return o;
}

napi_value SolverStep(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return nullptr;
    topo_step_info si;
    std::memset(&si, 0, sizeof si);
    int rc = topo_solver_step(solver_of(env, a[0]), &si);
    if (rc < 0) return throw_topo(env, rc);
    return info_to_js(env, si);
}

napi_value SolverSetDesign(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_solver_set_design(solver_of(env, a[0]), (const float*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

napi_value SolverGetDesign(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_solver_get_design(solver_of(env, a[0]), (float*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

// getField(h, field, TypedArray)
napi_value SolverGetField(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return nullptr;
    int rc = topo_solver_get_field(solver_of(env, a[0]), to_i32(env, a[1]), ta_ptr(env, a[2]));
    if (rc < 0) return throw_topo(env, rc);
    return a[2];
}

// update(h, penal, rmin, volfrac, maxIterations) (<= 0 keeps a value)
napi_value SolverUpdate(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    int rc = topo_solver_update(solver_of(env, a[0]), to_f64(env, a[1]), to_f64(env, a[2]), to_f64(env, a[3]),
                                to_i32(env, a[4]));
    if (rc < 0) return throw_topo(env, rc);
    return nullptr;
}

// fe(h, xPhys, penal, iteration, maxIterations, U) -> {compliance, ...}
napi_value SolverFe(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    double c = 0;
    topo_step_info si;
    std::memset(&si, 0, sizeof si);
    int rc = topo_solver_fe(solver_of(env, a[0]), (const float*)ta_ptr(env, a[1]), to_f64(env, a[2]), to_i32(env, a[3]),
                            to_i32(env, a[4]), (double*)ta_ptr(env, a[5]), &c, &si);
    if (rc < 0) return throw_topo(env, rc);
    return info_to_js(env, si);
}

// energy(h, U Float64Array, w Float32Array): element energies u_e^T k0 u_e
napi_value SolverEnergy(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return nullptr;
    int rc = topo_op_energy(solver_of(env, a[0]), (const double*)ta_ptr(env, a[1]), (float*)ta_ptr(env, a[2]));
    if (rc < 0) return throw_topo(env, rc);
    return a[2];
}

napi_value SolverSetIteration(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_solver_set_iteration(solver_of(env, a[0]), to_i32(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return nullptr;
}

napi_value SolverCancel(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return nullptr;
    topo_solver_cancel(solver_of(env, a[0]));
    return nullptr;
}

// ---- 2-D path -----------------------------------------------------------------
napi_value ConfigDefault2d(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    topo_config c;
    int rc = topo_config_default_2d(&c, to_i32(env, a[0]), to_i32(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return cfg_to_js(env, c);
}

// k0_2d(nu, Float64Array(64))
napi_value K0_2d(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_k0_2d(to_f64(env, a[0]), (double*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

// edofs2d(nelx, nely, Int32Array)
napi_value Edofs2d(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return nullptr;
    int rc = topo_edofs_2d(to_i32(env, a[0]), to_i32(env, a[1]), (int32_t*)ta_ptr(env, a[2]));
    if (rc < 0) return throw_topo(env, rc);
    return a[2];
}

// fixedDofs2d(nelx, nely, position, dofMask, Uint8Array) -> count
napi_value FixedDofs2d(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    int rc = topo_fixed_dofs_2d(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_i32(env, a[3]),
                                (uint8_t*)ta_ptr(env, a[4]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// loadVector2d(nelx, nely, direction, magnitude, forceVector|null, Float32Array)
napi_value LoadVector2d(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    int rc = topo_load_vector_2d(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_f64(env, a[3]),
                                 (const double*)ta_ptr(env, a[4]), (float*)ta_ptr(env, a[5]));
    if (rc < 0) return throw_topo(env, rc);
    return a[5];
}

// preservedElements2d(nelx, nely, position, direction, Uint8Array)
napi_value Preserved2d(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    int rc = topo_preserved_elements_2d(to_i32(env, a[0]), to_i32(env, a[1]), to_i32(env, a[2]), to_i32(env, a[3]),
                                        (uint8_t*)ta_ptr(env, a[4]));
    if (rc < 0) return throw_topo(env, rc);
    return a[4];
}

// fixedPaint2d(nelx, nely, Int32Array cells (vx,vy pairs), dofMask, Uint8Array) -> count
napi_value FixedPaint2d(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    size_t n = 0;
    const int32_t* cells = (const int32_t*)ta_ptr(env, a[2], &n);
    int rc = topo_fixed_paint_2d(to_i32(env, a[0]), to_i32(env, a[1]), cells, (int32_t)(n / 2), to_i32(env, a[3]),
                                 (uint8_t*)ta_ptr(env, a[4]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

// loadPaint2d(nelx, nely, Int32Array cells, Float64Array dir2, magnitude, Float32Array F) -> count
napi_value LoadPaint2d(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!get_args(env, info, 6, a)) return nullptr;
    size_t n = 0;
    const int32_t* cells = (const int32_t*)ta_ptr(env, a[2], &n);
    int rc = topo_load_paint_2d(to_i32(env, a[0]), to_i32(env, a[1]), cells, (int32_t)(n / 2),
                                (const double*)ta_ptr(env, a[3]), to_f64(env, a[4]), (float*)ta_ptr(env, a[5]));
    if (rc < 0) return throw_topo(env, rc);
    return i32v(env, rc);
}

void solver2d_finalize(napi_env, void* data, void*) { topo_solver2d_destroy((topo_solver2d*)data); }

topo_solver2d* solver2d_of(napi_env env, napi_value v) {
    void* p = nullptr;
    napi_get_value_external(env, v, &p);
    if (!p) napi_throw_error(env, nullptr, "solver handle already destroyed");
    return (topo_solver2d*)p;
}

// createSolver2d(config, fixedMask, F, preserved|null, void|null) -> handle
napi_value CreateSolver2d(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    topo_config c;
    int rc = topo_config_default_2d(&c, to_i32(env, get(env, a[0], "nelx")), to_i32(env, get(env, a[0], "nely")));
    if (rc < 0) return throw_topo(env, rc);
    cfg_from_js(env, a[0], &c);
    topo_solver2d* s = nullptr;
    rc = topo_solver2d_create(&c, (const uint8_t*)ta_ptr(env, a[1]), (const float*)ta_ptr(env, a[2]),
                              (const uint8_t*)ta_ptr(env, a[3]), (const uint8_t*)ta_ptr(env, a[4]), &s);
    if (rc < 0) return throw_topo(env, rc);
    napi_value h;
    NAPI_CALL(env, napi_create_external(env, s, solver2d_finalize, nullptr, &h));
    return h;
}

napi_value Solver2dStep(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return nullptr;
    topo_step_info si;
    std::memset(&si, 0, sizeof si);
    int rc = topo_solver2d_step(solver2d_of(env, a[0]), &si);
    if (rc < 0) return throw_topo(env, rc);
    return info_to_js(env, si);
}

// fea(h, x Float32Array, penal, tol) -> info
napi_value Solver2dFea(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!get_args(env, info, 4, a)) return nullptr;
    topo_step_info si;
    std::memset(&si, 0, sizeof si);
    int rc = topo_solver2d_fea(solver2d_of(env, a[0]), (const float*)ta_ptr(env, a[1]), to_f64(env, a[2]),
                               to_f64(env, a[3]), &si);
    if (rc < 0) return throw_topo(env, rc);
    return info_to_js(env, si);
}

napi_value Solver2dSetDesign(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_solver2d_set_design(solver2d_of(env, a[0]), (const float*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

napi_value Solver2dGetDesign(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return nullptr;
    int rc = topo_solver2d_get_design(solver2d_of(env, a[0]), (float*)ta_ptr(env, a[1]));
    if (rc < 0) return throw_topo(env, rc);
    return a[1];
}

napi_value Solver2dGetField(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return nullptr;
    int rc = topo_solver2d_get_field(solver2d_of(env, a[0]), to_i32(env, a[1]), ta_ptr(env, a[2]));
    if (rc < 0) return throw_topo(env, rc);
    return a[2];
}

napi_value Solver2dUpdate(napi_env env, napi_callback_info info) {
    napi_value a[5];
    if (!get_args(env, info, 5, a)) return nullptr;
    int rc = topo_solver2d_update(solver2d_of(env, a[0]), to_f64(env, a[1]), to_f64(env, a[2]), to_f64(env, a[3]),
                                  to_i32(env, a[4]));
    if (rc < 0) return throw_topo(env, rc);
    return nullptr;
}

napi_value Solver2dCancel(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return nullptr;
    topo_solver2d_cancel(solver2d_of(env, a[0]));
    return nullptr;
}

}  // namespace

NAPI_MODULE_INIT() {
    struct Fn {
        const char* name;
        napi_callback cb;
    };
    static const Fn fns[] = {
        {"version", Version}, {"deviceInfo", DeviceInfo}, {"configDefault", ConfigDefault}, {"k0", K0},
        {"edofs", Edofs}, {"fixedDofs", FixedDofs}, {"loadVector", LoadVector}, {"preservedElements", Preserved},
        {"loadPaint", LoadPaint}, {"fixedPaint", FixedPaint}, {"applyAEbe3D", ApplyAEbe3D}, {"ebePCG", EbePCG},
        {"ebeMGPCG", EbeMGPCG}, {"ebeKSP_BDDC", EbeKSPBDDC}, {"createSolver", CreateSolver}, {"solverStep", SolverStep},
        {"solverSetDesign", SolverSetDesign}, {"solverGetDesign", SolverGetDesign},
        {"solverGetField", SolverGetField}, {"solverUpdate", SolverUpdate}, {"solverFe", SolverFe},
        {"solverSetIteration", SolverSetIteration}, {"solverCancel", SolverCancel}, {"solverEnergy", SolverEnergy},
        {"configDefault2d", ConfigDefault2d}, {"k0_2d", K0_2d}, {"edofs2d", Edofs2d}, {"fixedDofs2d", FixedDofs2d},
        {"loadVector2d", LoadVector2d}, {"preservedElements2d", Preserved2d}, {"fixedPaint2d", FixedPaint2d},
        {"loadPaint2d", LoadPaint2d}, {"createSolver2d", CreateSolver2d}, {"solver2dStep", Solver2dStep},
        {"solver2dFea", Solver2dFea}, {"solver2dSetDesign", Solver2dSetDesign},
        {"solver2dGetDesign", Solver2dGetDesign}, {"solver2dGetField", Solver2dGetField},
        {"solver2dUpdate", Solver2dUpdate}, {"solver2dCancel", Solver2dCancel},
    };
    for (const Fn& f : fns) {
        napi_value v;
        napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &v);
        napi_set_named_property(env, exports, f.name, v);
    }
    return exports;
}
