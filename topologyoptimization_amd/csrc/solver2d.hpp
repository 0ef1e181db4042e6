// solver2d.hpp -- device-resident 2-D SIMP solver and the quad4 use of the
// ebePCG twin (js/optimizer-worker.js; assembly/index.ts:504-733 with
// edofSize 8). Included at the end of solver.hip: it drives the shared CG,
// filter and OC kernels with the quad4 operator of kernels_2d.hpp.
#pragma once
#include "kernels_2d.hpp"

namespace topo {

// the <= 2-row apply of the quad4 operator on stream s (modes of k_q4_apply);
// with pap (mode 1) the launch leaves its p.Ap partials there and returns
// their count (else 0)
static int q4_apply(hipStream_t s, int mode, const double* p, double* out, const double* E, const Grid2& g,
                    const uint8_t* fixed, const double* b, const int* gate, const double* ke, double* pap = nullptr) {
    const dim3 grid(std::min(nblk(g.nnodes()), kCgStride)), blk(256);
    if (mode == 0)
        hipLaunchKernelGGL(k_q4_apply<0>, grid, blk, 0, s, p, out, E, g, fixed, b, gate, ke, (double*)nullptr);
    else if (mode == 1)
        hipLaunchKernelGGL(k_q4_apply<1>, grid, blk, 0, s, p, out, E, g, fixed, b, gate, ke, pap);
    else
        hipLaunchKernelGGL(k_q4_apply<2>, grid, blk, 0, s, p, out, E, g, fixed, b, gate, ke, (double*)nullptr);
    return (mode == 1 && pap) ? (int)grid.x : 0;
}

// Recover (nelx, nely) from a 2-D edofs array and check every entry against
// _precomputeEdofs2D.
static bool edofs2d_structured(const int32_t* ed, int64_t nel, int64_t ndof, Grid2* out) {
    if (nel < 1 || ed[0] != 0 || ed[1] != 1 || ed[2] < 4 || (ed[2] & 1)) return false;
    const int64_t nely = ed[2] / 2 - 1;
    if (nely < 1 || nel % nely) return false;
    const int64_t nelx = nel / nely;
    if (2 * (nelx + 1) * (nely + 1) != ndof || nelx > INT32_MAX || nely > INT32_MAX) return false;
    std::vector<int32_t> ref(nel * 8);
    if (topo_edofs_2d((int32_t)nelx, (int32_t)nely, ref.data()) != TOPO_OK) return false;
    if (std::memcmp(ref.data(), ed, ref.size() * sizeof(int32_t)) != 0) return false;
    out->nelx = (int)nelx;
    out->nely = (int)nely;
    return true;
}

struct Solver2D {
    std::unique_ptr<Ctx> c;
    topo_config cfg;
    Grid2 g;
    int64_t nel = 0, ndof = 0, nfree = 0;
    double* ke = nullptr;          // 8x8 on the device
    uint8_t* fixed = nullptr;
    float* F32 = nullptr;
    double* F = nullptr;
    uint8_t* pres = nullptr;
    uint8_t* voidm = nullptr;
    float *x = nullptr, *xold = nullptr, *xnew = nullptr, *xphys = nullptr;
    float *dc = nullptr, *dcn = nullptr, *w = nullptr, *stress = nullptr, *Hs = nullptr;
    double *E = nullptr, *invD = nullptr, *U = nullptr;
    CGVecs v;
    FilterStencil* stencil = nullptr;
    int loop = 0;
    double change = 1.0;
    double penal_mesh = 3.0;       // this.penal of the worker (stress metric)
    volatile int cancel_flag = 0;
    int jcg_check = 32;            // Jacobi-PCG iterations per captured graph between host polls

    void init(const topo_config* cf, const uint8_t* fixedMask, const float* Fh, const uint8_t* preserved,
              const uint8_t* vm) {
        TOPO_CHECK(cf && fixedMask && Fh, TOPO_E_ARG, "topo_solver2d_create: null argument");
        cfg = *cf;
        TOPO_CHECK(cfg.nelx >= 1 && cfg.nely >= 1, TOPO_E_ARG, "2-D grid %d x %d", cfg.nelx, cfg.nely);
        TOPO_CHECK(cfg.nload == 1, TOPO_E_UNSUPPORTED, "the 2-D path solves one load case");
        TOPO_CHECK(cfg.filter_type == TOPO_FILTER_SENSITIVITY, TOPO_E_UNSUPPORTED,
                   "the 2-D path uses the sensitivity filter of the 2-D worker");
        TOPO_CHECK(cfg.objective == TOPO_OBJ_COMPLIANCE, TOPO_E_UNSUPPORTED, "the 2-D path minimises compliance");
        TOPO_CHECK(cfg.continuation_iters > 0 && cfg.beta_interval > 0, TOPO_E_ARG,
                   "continuation_iters and beta_interval must be positive");
        g.nelx = cfg.nelx;
        g.nely = cfg.nely;
        nel = g.nel();
        ndof = g.ndof();
        TOPO_CHECK(ndof < INT32_MAX, TOPO_E_ARG, "2-D grid too large (%lld DOFs)", (long long)ndof);
        c.reset(new Ctx());
        Ctx& C = *c;
        C.use_graphs = cfg.use_graphs != 0 && !graphs_disabled_by_env();
        if (const char* e = std::getenv("TOPO_JCG_CHECK")) jcg_check = std::max(1, std::min(64, std::atoi(e)));
        double keh[64];
        TOPO_CHECK(topo_k0_2d(cfg.nu, keh) == TOPO_OK, TOPO_E_ARG, "Poisson ratio %g", cfg.nu);
        ke = C.alloc<double>(64);
        TOPO_HIP(hipMemcpyAsync(ke, keh, sizeof keh, hipMemcpyHostToDevice, C.s));
        fixed = C.alloc<uint8_t>(ndof);
        F32 = C.alloc<float>(ndof);
        F = C.alloc<double>(ndof);
        TOPO_HIP(hipMemcpyAsync(fixed, fixedMask, ndof, hipMemcpyHostToDevice, C.s));
        TOPO_HIP(hipMemcpyAsync(F32, Fh, ndof * sizeof(float), hipMemcpyHostToDevice, C.s));
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(ndof)), dim3(256), 0, C.s, F, (const float*)F32, ndof);
        nfree = 0;
        for (int64_t i = 0; i < ndof; i++) nfree += fixedMask[i] ? 0 : 1;
        if (preserved) {
            pres = C.alloc<uint8_t>(nel);
            TOPO_HIP(hipMemcpyAsync(pres, preserved, nel, hipMemcpyHostToDevice, C.s));
        }
        if (vm) {
            voidm = C.alloc<uint8_t>(nel);
            TOPO_HIP(hipMemcpyAsync(voidm, vm, nel, hipMemcpyHostToDevice, C.s));
        }
        for (float** f : {&x, &xold, &xnew, &xphys, &dc, &dcn, &w, &stress, &Hs}) *f = C.alloc<float>(nel);
        E = C.alloc<double>(nel);
        invD = C.alloc<double>(ndof);
        U = C.alloc<double>(ndof);
        v.r = C.alloc<double>(ndof);
        v.z = C.alloc<double>(ndof);
        v.p = C.alloc<double>(ndof);
        v.Ap = C.alloc<double>(ndof);
        hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, U, 0.0, ndof, (const int*)nullptr);
        // x = volfrac, preserved -> 1, void -> 0 (:517-553)
        std::vector<float> x0(nel, (float)cfg.volfrac);
        for (int64_t e = 0; e < nel; e++) {
            if (preserved && preserved[e]) x0[e] = 1.0f;
            if (vm && vm[e]) x0[e] = 0.0f;
        }
        TOPO_HIP(hipMemcpyAsync(x, x0.data(), nel * sizeof(float), hipMemcpyHostToDevice, C.s));
        // prepareFilter :1391-1429: window +-floor(rmin), dist <= rmin, x-major
        FilterStencil st;
        std::memset(&st, 0, sizeof st);
        const int r = (int)std::floor(cfg.rmin);
        TOPO_CHECK(cfg.rmin > 0 && r <= kFilterMaxR, TOPO_E_UNSUPPORTED, "2-D filter radius %g (0 < rmin < %d)",
                   cfg.rmin, kFilterMaxR + 1);
        st.r = r;
        for (int kx = -r; kx <= r; kx++)
            for (int ky = -r; ky <= r; ky++) {
                const double dist = std::sqrt((double)(kx * kx + ky * ky));
                if (dist <= cfg.rmin) {
                    st.off[st.count][0] = kx; st.off[st.count][1] = ky; st.off[st.count][2] = 0;
                    st.w[st.count] = std::max(0.0, cfg.rmin - dist);
                    st.count++;
                }
            }
        stencil = C.alloc<FilterStencil>(1);
        TOPO_HIP(hipMemcpyAsync(stencil, &st, sizeof st, hipMemcpyHostToDevice, C.s));
        hipLaunchKernelGGL(k_filter_hs, dim3(nblk(nel)), dim3(256), 0, C.s, Hs, grid3(), (const FilterStencil*)stencil);
        C.check_launch();
        C.sync();
        penal_mesh = cfg.penal;
    }

    // the element grid as a one-element-thick 3-D grid: the zyx index of
    // (ex, ey, 0) is ex nely + ey, the 2-D worker's element index
    Grid grid3() const {
        Grid g3;
        g3.nelx = g.nelx;
        g3.nely = g.nely;
        g3.nelz = 1;
        return g3;
    }

    void set_design(const float* xh) {
        TOPO_HIP(hipMemcpyAsync(x, xh, nel * sizeof(float), hipMemcpyHostToDevice, c->s));
        c->sync();
    }

    // FE :1780-1797 with solveCG :1799-1873 at the densities xs: E, Jacobi
    // diagonal, PCG from U = 0 to ||r|| < tol, compliance F^T U
    CGResult fe(const float* xs, double penal, double tol, double* comp) {
        Ctx& C = *c;
        const double thr = cfg.skip_factor > 0 ? cfg.Emin * cfg.skip_factor : -1.0;
        hipLaunchKernelGGL(k_q4_moduli<float>, dim3(nblk(nel)), dim3(256), 0, C.s, xs, penal, cfg.E0, cfg.Emin, thr, E,
                           nel);
        hipLaunchKernelGGL(k_q4_invdiag, dim3(nblk(g.nnodes())), dim3(256), 0, C.s, (const double*)E,
                           (const uint8_t*)fixed, invD, g, (const double*)ke);
        C.check_launch();
        auto op = [&](const double* p, double* out, int mode, const double* b, const int* gate, double* pap) {
            return q4_apply(C.s, mode, p, out, E, g, fixed, b, gate, ke, pap);
        };
        const int maxit = (int)std::min<int64_t>(nfree, cfg.cg_max_iter);
        CGResult r = pcg_jacobi_op<double, double>(C, op, fixed, (const double*)invD, ndof, (const double*)F, U, false,
                                                   v, tol, maxit, &cancel_flag, {this, E, invD}, jcg_check, 1);
        TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
        hipLaunchKernelGGL((k_dot<float, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, (const float*)F32,
                           (const double*)U, ndof, C.dsc, C.partials, C.ticket);
        C.check_launch();
        *comp = C.read_d(C.dsc);
        return r;
    }

    // OC :1445-1492 with move 0.2 (the shared bisection kernels)
    int oc(double volfrac, double* vol) {
        return oc_bisect(*c, (const float*)x, (const float*)dcn, (const uint8_t*)pres, (const uint8_t*)voidm, cfg.move,
                         nel, volfrac, 0, xnew, vol);
    }

    // max over elements of a nonnegative field (k_change against no array)
    double field_max(const float* f) {
        Ctx& C = *c;
        hipLaunchKernelGGL(k_change, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, f, (const float*)nullptr, nel,
                           C.dsc + 4, C.partials, C.ticket, 1);
        C.check_launch();
        return C.read_d(C.dsc + 4);
    }

    // one iteration of the loop :583-700
    void step(topo_step_info* info) {
        Ctx& C = *c;
        auto t0 = std::chrono::steady_clock::now();
        cancel_flag = 0;
        loop++;
        hipLaunchKernelGGL((k_convert<float, float>), dim3(nblk(nel)), dim3(256), 0, C.s, xold, (const float*)x, nel);
        const double tol = cfg.cg_tol > 0 ? cfg.cg_tol : std::max(1e-8, change * 1e-2);
        const double pt = cfg.penal, ps = cfg.penal_start;
        const double penal = ps + (pt - ps) * std::min(1.0, (double)(loop - 1) / (double)cfg.continuation_iters);
        const int proj = cfg.use_projection != 0;
        const double beta =
            proj ? std::min(cfg.beta_max, std::pow(2.0, std::floor((double)(loop - 1) / (double)cfg.beta_interval)))
                 : 1.0;
        if (proj)
            hipLaunchKernelGGL(k_heaviside, dim3(nblk(nel)), dim3(256), 0, C.s, (const float*)x, xphys, nel, beta, 0.5);
        else
            hipLaunchKernelGGL((k_convert<float, float>), dim3(nblk(nel)), dim3(256), 0, C.s, xphys, (const float*)x,
                               nel);
        double comp = 0;
        auto ts = std::chrono::steady_clock::now();
        CGResult r = fe(xphys, penal, tol, &comp);
        auto te = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(k_q4_energy_sens, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)U, (const float*)xphys,
                           (const float*)x, penal, cfg.E0, cfg.Emin, 1, proj, beta, g, w, dc, (float*)nullptr,
                           (const double*)ke);
        hipLaunchKernelGGL(k_sens_filter, dim3(nblk(nel)), dim3(256), 0, C.s, (const float*)x, (const float*)dc,
                           (const float*)Hs, dcn, grid3(), (const FilterStencil*)stencil);
        C.check_launch();
        double vol = 0;
        const int passes = oc(cfg.volfrac, &vol);
        hipLaunchKernelGGL(k_change, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, (const float*)xnew, (const float*)xold,
                           nel, C.dsc + 4, C.partials, C.ticket, 0);
        C.check_launch();
        TOPO_HIP(hipMemcpyAsync(C.h_sc, C.dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, C.s));
        C.sync();
        change = C.h_sc[4];
        const double volume = C.h_sc[5];
        std::swap(x, xnew);
        hipLaunchKernelGGL(k_q4_stress, dim3(nblk(nel)), dim3(256), 0, C.s, (const float*)x, (const float*)w,
                           penal_mesh, cfg.E0, cfg.Emin, stress, nel);
        const double smax = field_max(stress);
        auto t1 = std::chrono::steady_clock::now();
        if (info) {
            std::memset(info, 0, sizeof *info);
            info->iteration = loop;
            info->cg_iters = r.iters;
            info->applyA_fine = r.applyA;
            info->oc_passes = passes;
            info->compliance = comp;
            info->change = change;
            info->penal = penal;
            info->beta = beta;
            info->tolerance = tol;
            info->volume = volume;
            info->max_stress = smax;
            info->residual = std::sqrt(r.rn);
            info->t_solve_ms = std::chrono::duration<double, std::milli>(te - ts).count();
            info->t_total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            info->t_post_ms = std::chrono::duration<double, std::milli>(t1 - te).count();
        }
        (void)vol;
    }

    // feaOnly :384-485
    void fea(const float* xh, double penal, double tol, topo_step_info* info) {
        Ctx& C = *c;
        auto t0 = std::chrono::steady_clock::now();
        cancel_flag = 0;
        TOPO_HIP(hipMemcpyAsync(xphys, xh, nel * sizeof(float), hipMemcpyHostToDevice, C.s));
        double comp = 0;
        CGResult r = fe(xphys, penal, tol, &comp);
        auto te = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(k_q4_energy_sens, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)U, (const float*)xphys,
                           (const float*)xphys, penal, cfg.E0, cfg.Emin, 0, 0, 1.0, g, w, (float*)nullptr, stress,
                           (const double*)ke);
        C.check_launch();
        const double smax = field_max(stress);
        auto t1 = std::chrono::steady_clock::now();
        if (info) {
            std::memset(info, 0, sizeof *info);
            info->cg_iters = r.iters;
            info->applyA_fine = r.applyA;
            info->compliance = comp;
            info->penal = penal;
            info->tolerance = tol;
            info->max_stress = smax;
            info->residual = std::sqrt(r.rn);
            info->t_solve_ms = std::chrono::duration<double, std::milli>(te - t0).count();
            info->t_total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
    }

    void get_f(const float* d, float* out) {
        TOPO_HIP(hipMemcpyAsync(out, d, nel * sizeof(float), hipMemcpyDeviceToHost, c->s));
        c->sync();
    }
};

}  // namespace topo

struct topo_solver2d {
    topo::Solver2D s;
};

using namespace topo;

extern "C" {

int topo_solver2d_create(const topo_config* cfg, const uint8_t* fixedMask, const float* F, const uint8_t* preserved,
                         const uint8_t* voidm, topo_solver2d** out) {
    TOPO_API_BEGIN
    TOPO_CHECK(cfg && out, TOPO_E_ARG, "topo_solver2d_create: null argument");
    std::unique_ptr<topo_solver2d> h(new topo_solver2d());
    h->s.init(cfg, fixedMask, F, preserved, voidm);
    *out = h.release();
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_destroy(topo_solver2d* s) {
    delete s;
    return TOPO_OK;
}

int topo_solver2d_set_design(topo_solver2d* h, const float* x) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && x, TOPO_E_ARG, "null argument");
    h->s.set_design(x);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_get_design(topo_solver2d* h, float* x) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && x, TOPO_E_ARG, "null argument");
    h->s.get_f(h->s.x, x);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_update(topo_solver2d* h, double penal, double rmin, double volfrac, int32_t max_iterations) {
    TOPO_API_BEGIN
    TOPO_CHECK(h, TOPO_E_ARG, "null solver");
    Solver2D& S = h->s;
    if (penal > 0) S.penal_mesh = penal;
    if (rmin > 0) S.cfg.rmin = rmin;
    if (volfrac > 0) S.cfg.volfrac = volfrac;
    if (max_iterations > 0) S.cfg.max_iterations = max_iterations;
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_step(topo_solver2d* h, topo_step_info* info) {
    TOPO_API_BEGIN
    TOPO_CHECK(h, TOPO_E_ARG, "null solver");
    h->s.step(info);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_fea(topo_solver2d* h, const float* x, double penal, double tol, topo_step_info* info) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && x && tol > 0, TOPO_E_ARG, "bad argument");
    h->s.fea(x, penal, tol, info);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_get_field(topo_solver2d* h, int32_t field, void* out) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && out, TOPO_E_ARG, "null argument");
    Solver2D& S = h->s;
    switch (field) {
    case TOPO_FIELD_U:
        TOPO_HIP(hipMemcpyAsync(out, S.U, S.ndof * sizeof(double), hipMemcpyDeviceToHost, S.c->s));
        S.c->sync();
        break;
    case TOPO_FIELD_XPHYS: S.get_f(S.xphys, (float*)out); break;
    case TOPO_FIELD_ENERGY: S.get_f(S.w, (float*)out); break;
    case TOPO_FIELD_DC: S.get_f(S.dc, (float*)out); break;
    case TOPO_FIELD_DCN: S.get_f(S.dcn, (float*)out); break;
    case TOPO_FIELD_STRESS: S.get_f(S.stress, (float*)out); break;
    default: TOPO_CHECK(false, TOPO_E_ARG, "field %d is not kept by the 2-D solver", field);
    }
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver2d_cancel(topo_solver2d* h) {
    if (!h) return TOPO_E_ARG;
    h->s.cancel_flag = 1;
    return TOPO_OK;
}

}  // extern "C"
