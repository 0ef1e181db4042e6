// solver.hip -- device-resident SIMP solver of libtopo_hip and the C ABI of
// include/topo.h (except the host-only setup functions in host_setup.cpp).
//
// Structure
//   Ctx        stream, reduction scratch, device CG/OC state, allocations
//   MGH        geometric multigrid hierarchy (level 0 = the fine grid);
//              used by Jacobi-PCG too (level 0's diagonal)
//   Solver     the topo_solver handle: design fields, U, CG vectors, SIMP step
// Precision: T = double (FP64 path) or float (inner solve of the mixed path);
// CG scalars are always FP64 and live on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <vector>
#include "kernels.hpp"
#include "kernels_xm.hpp"
#include "kernels_mf.hpp"
#include "kernels_adj.hpp"
#include "kernels_chol.hpp"
#include "kernels_bddc.hpp"
#include <rocblas/rocblas.h>

namespace topo {

static inline int nblk(int64_t n, int bs = 256) {
    int64_t b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > (1 << 20)) b = 1 << 20;
    return (int)b;
}
// nblk rounded up to a multiple of 8 blocks (the kernels using xcd_gtid)
static inline int nblk8(int64_t n, int bs = 256) { return (nblk(n, bs) + 7) & ~7; }


// ---------------------------------------------------------------------------
// dispatch of the batched kernels for nr = 1, 2 or 4 right-hand sides
#define TOPO_NR_SWITCH(nr, ...)                                                          \
    switch (nr) {                                                                        \
    case 1: { constexpr int NR = 1; __VA_ARGS__; } break;                                \
    case 2: { constexpr int NR = 2; __VA_ARGS__; } break;                                \
    case 4: { constexpr int NR = 4; __VA_ARGS__; } break;                                \
    default: TOPO_CHECK(false, TOPO_E_ARG, "%d right-hand sides per batch (1, 2 or 4)", nr); \
    }

constexpr int kMaxRHS = 4;   // right-hand sides one batched CG launch carries

// TOPO_USE_GRAPHS=0 turns graph capture off for every solver of the process
// (the CG iterations are then launched kernel by kernel; same results)
static bool graphs_disabled_by_env() {
    const char* e = std::getenv("TOPO_USE_GRAPHS");
    return e && std::atoi(e) == 0;
}

struct Ctx {
    hipStream_t s = nullptr;
    int device = 0;
    double* partials = nullptr;   // reduction partials (last-block and fence-free kernels)
    unsigned* ticket = nullptr;
    CGState* st = nullptr;        // device CG state
    CGState* st_alt = nullptr;    // second device CG state (MG-PCG books kept by the p update, by iteration parity)
    CGState* h_st = nullptr;      // pinned mirror
    OCState* oc = nullptr;
    OCState* h_oc = nullptr;
    double* dsc = nullptr;        // 16 device doubles of scratch
    double* h_sc = nullptr;       // pinned mirror
    std::vector<void*> allocs;
    bool use_graphs = false;
    struct GraphEntry {
        std::vector<const void*> key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    rocblas_handle blas = nullptr;
    rocblas_handle blas_handle() {
        if (!blas) {
            TOPO_CHECK(rocblas_create_handle(&blas) == rocblas_status_success, TOPO_E_HIP, "rocblas_create_handle failed");
            TOPO_CHECK(rocblas_set_stream(blas, s) == rocblas_status_success, TOPO_E_HIP, "rocblas_set_stream failed");
        }
        return blas;
    }
    // second stream (default priority) for work that overlaps a dependency
    // chain on s, with its own rocBLAS handle and a pool of ordering events;
    // a third one (s3, default priority) for the factorization's look-ahead
    hipStream_t s2 = nullptr;
    hipStream_t s3 = nullptr;
    rocblas_handle blas2 = nullptr;
    std::vector<hipEvent_t> evs;
    void side_ready(int nev) {
        if (!s2) TOPO_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        if (!s3) TOPO_HIP(hipStreamCreateWithFlags(&s3, hipStreamNonBlocking));
        if (!blas2) {
            TOPO_CHECK(rocblas_create_handle(&blas2) == rocblas_status_success, TOPO_E_HIP,
                       "rocblas_create_handle failed");
            TOPO_CHECK(rocblas_set_stream(blas2, s2) == rocblas_status_success, TOPO_E_HIP, "rocblas_set_stream failed");
        }
        while ((int)evs.size() < nev) {
            hipEvent_t e;
            TOPO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            evs.push_back(e);
        }
    }
    // `to` waits for the work issued so far on `from` (event k of the pool)
    void order(hipStream_t from, hipStream_t to, int k) {
        TOPO_HIP(hipEventRecord(evs[k], from));
        TOPO_HIP(hipStreamWaitEvent(to, evs[k], 0));
    }

    hipGraphExec_t graph_find(const std::vector<const void*>& key) {
        for (auto& g : graphs)
            if (g.key == key) return g.exec;
        return nullptr;
    }
    // capture `body` on the stream into a cached executable graph
    template <typename F>
    hipGraphExec_t graph_capture(const std::vector<const void*>& key, F&& body) {
        hipGraphExec_t ex = graph_find(key);
        if (ex) return ex;
        TOPO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        body();
        hipGraph_t gph = nullptr;
        TOPO_HIP(hipStreamEndCapture(s, &gph));
        TOPO_HIP(hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gph);
        graphs.push_back({key, ex});
        return ex;
    }

    void graph_clear() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    Ctx() {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            set_error("no HIP device available");
            throw TopoError{TOPO_E_NODEVICE};
        }
        TOPO_HIP(hipGetDevice(&device));
        TOPO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        partials = alloc<double>(std::max(32 * kRedBlocks, 3 * kCgRows * kCgStride));
        ticket = alloc<unsigned>(4);
        st = alloc<CGState>(kMaxRHS);
        st_alt = alloc<CGState>(kMaxRHS);
        oc = alloc<OCState>(1);
        dsc = alloc<double>(16);
        TOPO_HIP(hipHostMalloc((void**)&h_st, kMaxRHS * sizeof(CGState), hipHostMallocDefault));
        TOPO_HIP(hipHostMalloc((void**)&h_oc, sizeof(OCState), hipHostMallocDefault));
        TOPO_HIP(hipHostMalloc((void**)&h_sc, 16 * sizeof(double), hipHostMallocDefault));
        TOPO_HIP(hipMemsetAsync(ticket, 0, 4 * sizeof(unsigned), s));
        TOPO_HIP(hipMemsetAsync(st, 0, kMaxRHS * sizeof(CGState), s));
        TOPO_HIP(hipStreamSynchronize(s));
    }
    // A context over device and host memory the caller owns: no allocation,
    // the null stream, no graphs (the allocation-free entry point's).
    struct Borrowed {};
    bool borrowed = false;
    Ctx(Borrowed, double* partials_, unsigned* ticket_, CGState* st_, OCState* oc_, double* dsc_, CGState* h_st_,
        OCState* h_oc_, double* h_sc_)
        : partials(partials_), ticket(ticket_), st(st_), h_st(h_st_), oc(oc_), h_oc(h_oc_), dsc(dsc_), h_sc(h_sc_),
          borrowed(true) {
        TOPO_HIP(hipGetDevice(&device));
    }
    ~Ctx() {
        if (borrowed) {
            (void)hipStreamSynchronize(s);
            return;
        }
        if (s) (void)hipStreamSynchronize(s);
        if (s2) (void)hipStreamSynchronize(s2);
        if (s3) (void)hipStreamSynchronize(s3);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        if (blas) (void)rocblas_destroy_handle(blas);
        if (blas2) (void)rocblas_destroy_handle(blas2);
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
        if (s2) (void)hipStreamDestroy(s2);
        if (s3) (void)hipStreamDestroy(s3);
        for (void* p : allocs) (void)hipFree(p);
        if (h_st) (void)hipHostFree(h_st);
        if (h_oc) (void)hipHostFree(h_oc);
        if (h_sc) (void)hipHostFree(h_sc);
        if (s) (void)hipStreamDestroy(s);
    }
    template <typename T>
    T* alloc(int64_t n) {
        void* p = nullptr;
        if (n <= 0) n = 1;
        TOPO_HIP(hipMalloc(&p, (size_t)n * sizeof(T)));
        allocs.push_back(p);
        return (T*)p;
    }
    void free_one(void* p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) {
            (void)hipFree(p);
            allocs.erase(it);
        }
    }
    void sync() { TOPO_HIP(hipStreamSynchronize(s)); }
    void check_launch() { TOPO_HIP(hipGetLastError()); }
    // read the device CG state
    const CGState& read_st(int nr = 1) {
        TOPO_HIP(hipMemcpyAsync(h_st, st, nr * sizeof(CGState), hipMemcpyDeviceToHost, s));
        sync();
        return *h_st;
    }
    double read_d(const double* dptr) {
        TOPO_HIP(hipMemcpyAsync(h_sc, dptr, sizeof(double), hipMemcpyDeviceToHost, s));
        sync();
        return h_sc[0];
    }
};

// OC bisection (OC js/optimizer-worker-3d.js:3066-3101) on the device: each
// (k_oc_plan, k_oc_eval) pair takes up to kOcDepth bisection steps, and once
// the reference's stop test holds every later pair is a no-op, so a batch of
// pairs runs without host round trips; the host reads the state once per
// batch. A typical update takes 40-55 steps (one batch of 20 pairs); a NaN in
// the sensitivities (the volume test then never passes) walks lambda down to
// underflow in ~1100 steps, as the reference loop does. xnew is the last
// executed step's, and *vol the volume sum that step compared.
static int oc_bisect(Ctx& C, const float* x, const float* dcn, const uint8_t* pres, const uint8_t* voidm, double move,
                     int64_t nel, double volfrac, int mech, float* xnew, double* vol) {
    OCState h;
    std::memset(&h, 0, sizeof h);
    h.l1 = 0;
    h.l2 = 1e9;
    h.target = volfrac * (double)nel;
    *C.h_oc = h;
    TOPO_HIP(hipMemcpyAsync(C.oc, C.h_oc, sizeof(OCState), hipMemcpyHostToDevice, C.s));
    constexpr int kPairsPerBatch = 20;
    for (int batch = 0; batch < 60; batch++) {
        for (int k = 0; k < kPairsPerBatch; k++) {
            hipLaunchKernelGGL(k_oc_plan, dim3(1), dim3(64), 0, C.s, C.oc);
            hipLaunchKernelGGL(k_oc_eval, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, x, dcn, pres, voidm, move, nel,
                               C.oc, C.partials, C.ticket, mech);
        }
        C.check_launch();
        TOPO_HIP(hipMemcpyAsync(C.h_oc, C.oc, sizeof(OCState), hipMemcpyDeviceToHost, C.s));
        C.sync();
        if (C.h_oc->finished) break;
    }
    TOPO_CHECK(C.h_oc->finished, TOPO_E_BREAKDOWN, "OC bisection did not terminate");
    hipLaunchKernelGGL(k_oc_final, dim3(nblk(nel)), dim3(256), 0, C.s, x, dcn, pres, voidm, move, nel,
                       (const double*)&C.oc->lmid_last, xnew, mech);
    C.check_launch();
    if (vol) *vol = C.h_oc->vol_last;
    return C.h_oc->passes;
}

// Block coefficients of K0 in the Hadamard basis (see kernels_xm.hpp):
// B = H K0 H^T / 64 with modes q = (qx,qy,qz), index 3*(4qx + 2qy + qz) + d.
// Returns false when K0 does not have the cube-symmetric block structure
// (then the reference-order kernel is used).
static bool had_coefs(const double* K, HadCoef<double>* c) {
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    double Hm[24][24] = {{0}};
    for (int qx = 0; qx < 2; qx++)
        for (int qy = 0; qy < 2; qy++)
            for (int qz = 0; qz < 2; qz++) {
                const int qi = qx * 4 + qy * 2 + qz;
                for (int m = 0; m < 8; m++) {
                    const int sgn = ((qx * off[m][0] + qy * off[m][1] + qz * off[m][2]) & 1) ? -1 : 1;
                    for (int d = 0; d < 3; d++) Hm[3 * qi + d][3 * m + d] = sgn;
                }
            }
    double HK[24][24], B[24][24];
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0;
            for (int k = 0; k < 24; k++) s += Hm[i][k] * K[k * 24 + j];
            HK[i][j] = s;
        }
    double bmax = 0;
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0;
            for (int k = 0; k < 24; k++) s += HK[i][k] * Hm[j][k];
            B[i][j] = s / 64.0;
            bmax = std::max(bmax, std::fabs(B[i][j]));
        }
    auto I = [](int qx, int qy, int qz, int d) { return 3 * (qx * 4 + qy * 2 + qz) + d; };
    bool used[24][24] = {{false}};
    const double tol = 1e-12 * bmax;
    bool ok = true;
    auto tri = [&](const int m[3], double& am, double& b) {
        const double a = B[m[0]][m[0]];
        b = B[m[0]][m[1]];
        am = a - b;
        for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) {
                used[m[x]][m[y]] = true;
                if (std::fabs(B[m[x]][m[y]] - (x == y ? a : b)) > tol) ok = false;
            }
    };
    const int t0[3] = {I(1, 0, 0, 0), I(0, 1, 0, 1), I(0, 0, 1, 2)};
    const int t7[3] = {I(0, 1, 1, 0), I(1, 0, 1, 1), I(1, 1, 0, 2)};
    tri(t0, c->am0, c->b0);
    tri(t7, c->am7, c->b7);
    const int one[3][2] = {{I(1, 1, 0, 1), I(1, 0, 1, 2)}, {I(1, 1, 0, 0), I(0, 1, 1, 2)}, {I(1, 0, 1, 0), I(0, 1, 1, 1)}};
    const int two[3][2] = {{I(0, 0, 1, 1), I(0, 1, 0, 2)}, {I(0, 0, 1, 0), I(1, 0, 0, 2)}, {I(0, 1, 0, 0), I(1, 0, 0, 1)}};
    for (int k = 0; k < 3; k++) {
        const int a = one[k][0], b = one[k][1];
        c->oa[k] = B[a][a]; c->oc[k] = B[a][b]; c->ob[k] = B[b][b];
        used[a][a] = used[a][b] = used[b][a] = used[b][b] = true;
        if (std::fabs(B[a][b] - B[b][a]) > tol) ok = false;
        const int a2 = two[k][0], b2 = two[k][1];
        c->ta[k] = B[a2][a2]; c->tc[k] = B[a2][b2]; c->tb[k] = B[b2][b2];
        used[a2][a2] = used[a2][b2] = used[b2][a2] = used[b2][b2] = true;
        if (std::fabs(B[a2][b2] - B[b2][a2]) > tol) ok = false;
        const int f = I(1, 1, 1, k);
        c->tf[k] = B[f][f];
        used[f][f] = true;
    }
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++)
            if (!used[i][j] && std::fabs(B[i][j]) > tol) ok = false;
    return ok;
}

template <typename TD>
static HadCoef<TD> had_cast(const HadCoef<double>& c) {
    HadCoef<TD> r;
    r.am0 = (TD)c.am0; r.b0 = (TD)c.b0; r.am7 = (TD)c.am7; r.b7 = (TD)c.b7;
    for (int k = 0; k < 3; k++) {
        r.oa[k] = (TD)c.oa[k]; r.oc[k] = (TD)c.oc[k]; r.ob[k] = (TD)c.ob[k];
        r.ta[k] = (TD)c.ta[k]; r.tc[k] = (TD)c.tc[k]; r.tb[k] = (TD)c.tb[k];
        r.tf[k] = (TD)c.tf[k];
    }
    return r;
}

// bfloat16 bits of x, round to nearest even (finite x)
static uint16_t bf16_bits(double xd) {
    const float x = (float)xd;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static double bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// A-operand fragments of K0 for k_applyA_mf: the 24x24 K0 permuted to the
// 32x32 order 16 dx + 8 dy + 4 dz + c (local node m at offset off[m], the
// element numbering of the reference, `:1448-1494`; c = 3 rows and columns
// zero), split K = hi + lo in bfloat16; frag[(rb*2 + part)*64 + lane][i] =
// part of row 16 rb + lane % 16, column 8 (lane / 16) + i.
static void mf_k0_frags(const double* K, uint16_t (*frag)[8]) {
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    int ln[2][2][2];
    for (int m = 0; m < 8; m++) ln[off[m][0]][off[m][1]][off[m][2]] = m;
    auto kval = [&](int r, int k) {
        const int c = r & 3, c2 = k & 3;
        if (c == 3 || c2 == 3) return 0.0;
        const int m = ln[r >> 4][(r >> 3) & 1][(r >> 2) & 1];
        const int m2 = ln[k >> 4][(k >> 3) & 1][(k >> 2) & 1];
        return K[(3 * m + c) * 24 + 3 * m2 + c2];
    };
    for (int rb = 0; rb < 2; rb++)
        for (int lane = 0; lane < 64; lane++)
            for (int i = 0; i < 8; i++) {
                const double v = kval(16 * rb + lane % 16, 8 * (lane / 16) + i);
                const uint16_t hi = bf16_bits(v);
                frag[(rb * 2 + 0) * 64 + lane][i] = hi;
                frag[(rb * 2 + 1) * 64 + lane][i] = bf16_bits(v - bf16_value(hi));
            }
}

// Lets the k_applyA_mf variants use their dynamic LDS (up to ~82 KB with
// NT = 512, above the 64 KB a launch gets without asking); false when the
// device refuses, and the VALU kernel is used instead.
static bool mf_allow_lds() {
    const int bytes = (int)mf_lds_bytes(255, 512);
    const void* fns[6] = {(const void*)k_applyA_mf<512, 0>, (const void*)k_applyA_mf<512, 1>,
                          (const void*)k_applyA_mf<512, 2>, (const void*)k_applyA_mf<512, 3>,
                          (const void*)k_applyA_mf<512, 4>, (const void*)k_applyA_mf<512, 5>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    return true;
}

// The FP64 x-march kernels keep two node planes in LDS (up to ~74 KB with
// 512 threads, nnz <= 255; ~147 KB with 1024 threads, nnz <= 511), and the
// FP32 1024-thread ones need up to ~74 KB (nnz <= 511): grant it; a device
// that refuses cannot run them.
static void xm_allow_lds() {
    static bool done = false;
    if (done) return;
    const int b512 = (int)xm_lds_bytes(255, sizeof(double), 512);
    const int b1024 = (int)xm_lds_bytes(511, sizeof(double), 1024);
    const int s1024 = (int)xm_lds_bytes(511, sizeof(float), 1024);
    const void* g1024[6] = {(const void*)k_applyA_xm<float, float, 1024, 0>, (const void*)k_applyA_xm<float, float, 1024, 1>,
                            (const void*)k_applyA_xm<float, float, 1024, 2>, (const void*)k_applyA_xm<float, float, 1024, 3>,
                            (const void*)k_applyA_xm<float, float, 1024, 4>, (const void*)k_applyA_xm<float, float, 1024, 5>};
    for (const void* f : g1024) TOPO_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, s1024));
    const void* f512[6] = {(const void*)k_applyA_xm<double, double, 512, 0>, (const void*)k_applyA_xm<double, double, 512, 1>,
                           (const void*)k_applyA_xm<double, double, 512, 2>, (const void*)k_applyA_xm<double, double, 512, 3>,
                           (const void*)k_applyA_xm<double, double, 512, 4>, (const void*)k_applyA_xm<double, double, 512, 5>};
    const void* f1024[6] = {(const void*)k_applyA_xm<double, double, 1024, 0>, (const void*)k_applyA_xm<double, double, 1024, 1>,
                            (const void*)k_applyA_xm<double, double, 1024, 2>, (const void*)k_applyA_xm<double, double, 1024, 3>,
                            (const void*)k_applyA_xm<double, double, 1024, 4>, (const void*)k_applyA_xm<double, double, 1024, 5>};
    for (int i = 0; i < 6; i++) {
        TOPO_HIP(hipFuncSetAttribute(f512[i], hipFuncAttributeMaxDynamicSharedMemorySize, b512));
        TOPO_HIP(hipFuncSetAttribute(f1024[i], hipFuncAttributeMaxDynamicSharedMemorySize, b1024));
    }
    done = true;
}

// K0 in device memory (FP64 + FP32 copies), its Hadamard-block coefficients
// and its matrix-core fragments
struct DevKe {
    double* ke = nullptr;
    float* ke32 = nullptr;
    uint4* kf = nullptr;   // 256 x 16 B, see mf_k0_frags
    HadCoef<double> hc64;
    HadCoef<float> hc32;
    bool had_ok = false;
    void set(Ctx& c, const double* hke) {
        had_ok = had_coefs(hke, &hc64);
        hc32 = had_cast<float>(hc64);
        if (!ke) {
            ke = c.alloc<double>(576);
            ke32 = c.alloc<float>(576);
            kf = c.alloc<uint4>(256);
        }
        float k32[576];
        for (int i = 0; i < 576; i++) k32[i] = (float)hke[i];
        uint16_t frag[256][8];
        mf_k0_frags(hke, frag);
        TOPO_HIP(hipMemcpyAsync(ke, hke, 576 * sizeof(double), hipMemcpyHostToDevice, c.s));
        TOPO_HIP(hipMemcpyAsync(ke32, k32, 576 * sizeof(float), hipMemcpyHostToDevice, c.s));
        TOPO_HIP(hipMemcpyAsync(kf, frag, sizeof(frag), hipMemcpyHostToDevice, c.s));
        c.sync();
    }
};

// ---------------------------------------------------------------------------
struct MGLevel {
    Grid g;
    int64_t ndof = 0;
    double* E = nullptr;
    float* E32 = nullptr;
    bool ownE = true;
    uint8_t* fixed = nullptr;
    double* invD = nullptr;
    float* invD32 = nullptr;
    bool dense = false;
    double* K = nullptr;
    float* K32 = nullptr;
    double *b = nullptr, *x = nullptr, *x2 = nullptr, *res = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *x232 = nullptr, *res32 = nullptr;
    // direct coarse solve: explicit inverse of the free block
    bool chol = false;
    int64_t nf = 0, nfp = 0;       // free DOFs, padded to a multiple of 4
    int32_t* freeIdx = nullptr;
    int32_t* freeOf = nullptr;     // DOF -> free index or -1
    double* bf = nullptr;          // gathered right-hand side (padded)
    double* Kinv = nullptr;        // the free-DOF block (banded: its factor L in place), or the Gauss-Jordan inverse
    float* Kinv32 = nullptr;       // FP32 rounding of the Gauss-Jordan inverse (coarse_fp32)
    bool tri = false;              // coarse solve by x = Z^T (Z b), Z = L^-1 (else Kinv b)
    int bandw = 0;                 // max |i - j| of the free block's nonzeros
    double* Z = nullptr;           // L^-1 (lower triangular)
    float* Z32 = nullptr;
    _Float16* Z16 = nullptr;       // Z / s_i per row (coarse_fp32 = 2)
    double* zs2 = nullptr;         // s_i^2
    double* Linv = nullptr;        // inverses of the diagonal blocks of L
    double* ytmp = nullptr;        // Z b (nrhs x nfp)
    double* tpart = nullptr;       // chunk partials of the triangular products (nrhs x max(tnch, tnrb) x nfp)
    double* tpart1 = nullptr;      // row-chunk partials of Z b read by the column-tile pass (nrhs x tnch x nfp)
    int tcw = 0, tnch = 0;         // column chunk width (0: one wave per row quad) and count
    int trw = 8;                   // rows per wave of the chunked product (4 or 8)
    int tnrb = 0;                  // row blocks of the transposed FP16 product (column tiles of Z16)
    double* d0c = nullptr;         // original diagonal (pivot test)
    double* ztmp = nullptr;        // 64 x nfp scratch of the Z row blocks
    double *gj_d0 = nullptr, *gj_c = nullptr, *gj_r = nullptr;   // Gauss-Jordan fallback scratch (kept)
    double* work = nullptr;
    // nested-dissection order (nd): a separator tree in postorder over the
    // free DOFs, one dense front per node; Z = L^-1 is zero left of the
    // start of each row's subtree (rfirst)
    struct NDNode {
        int64_t o0 = 0, m = 0;    // own positions [o0, o0 + m), m a multiple of 64
        int64_t s0 = 0, f = 0;    // first position of the subtree; front size (own + ancestors)
        int parent = -1, kid[2] = {-1, -1}, height = 0;
        double* F = nullptr;      // front (f x f, row-major)
        double* Lb = nullptr;     // L of the ancestor rows ((f - m) x m)
        int32_t* g = nullptr;     // front index -> position
        double* Linv = nullptr;   // inverses of the own diagonal blocks
    };
    bool nd = false;
    std::vector<NDNode> ndn;
    std::vector<std::vector<int>> nd_h;   // nodes by height (leaves 0)
    std::vector<FrontDev*> nd_fd;         // their front descriptors (device)
    int32_t* rfirst = nullptr;            // per position: first nonzero column of its row of Z
    double* ndW = nullptr;                // scratch of the Z blocks (max (f - m) x m)
};

// row-major C (r x c) = alpha op(A) op(B) + beta C as the column-major
// C^T = alpha op(B)^T op(A)^T of rocBLAS
static void gemm_rm(rocblas_handle h, bool ta, bool tb, int64_t r, int64_t c, int64_t k, double alpha,
                    const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
    TOPO_CHECK(rocblas_dgemm(h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                             ta ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)c, (rocblas_int)r,
                             (rocblas_int)k, &alpha, B, (rocblas_int)ldb, A, (rocblas_int)lda, &beta, C,
                             (rocblas_int)ldc) == rocblas_status_success,
               TOPO_E_HIP, "rocblas_dgemm failed");
}

// Nested-dissection order of the free DOFs of a structured node grid: the
// node box is cut through its middle node plane across its longest axis,
// the plane's DOFs (the separator) are numbered after both halves, and the
// halves are cut the same way, to `depth` levels or until a part has at
// most `leafmax` free DOFs or fewer than 3 node planes on every axis. A part
// keeps the natural DOF order. Nodes come out in postorder (children before
// their parent), each own range padded to a multiple of 64 with positions
// that map to no DOF (-1). Returns the position -> DOF map.
static std::vector<int32_t> nd_order(const Grid& g, const std::vector<uint8_t>& fixed, int depth, int64_t leafmax,
                                     std::vector<MGLevel::NDNode>& nodes) {
    const int nnx = g.nnx(), nny = g.nny(), nnz = g.nnz();
    struct Box { int lo[3], hi[3]; };
    auto dofs = [&](const Box& b) {
        std::vector<int32_t> v;
        for (int ix = b.lo[0]; ix < b.hi[0]; ix++)
            for (int iy = b.lo[1]; iy < b.hi[1]; iy++)
                for (int iz = b.lo[2]; iz < b.hi[2]; iz++)
                    for (int d = 0; d < 3; d++) {
                        const int64_t dof = 3 * (((int64_t)ix * nny + iy) * nnz + iz) + d;
                        if (!fixed[dof]) v.push_back((int32_t)dof);
                    }
        return v;
    };
    std::vector<std::vector<int32_t>> own;
    nodes.clear();
    std::function<int(const Box&, int)> rec = [&](const Box& b, int d) -> int {
        int a = 0;
        for (int q = 1; q < 3; q++)
            if (b.hi[q] - b.lo[q] > b.hi[a] - b.lo[a]) a = q;
        std::vector<int32_t> all = dofs(b);
        if (d == 0 || (int64_t)all.size() <= leafmax || b.hi[a] - b.lo[a] < 3) {
            nodes.emplace_back();
            own.push_back(std::move(all));
            return (int)nodes.size() - 1;
        }
        const int mid = (b.lo[a] + b.hi[a]) / 2;
        Box b0 = b, b1 = b, bs = b;
        b0.hi[a] = mid;
        b1.lo[a] = mid + 1;
        bs.lo[a] = mid;
        bs.hi[a] = mid + 1;
        const int k0 = rec(b0, d - 1);
        const int k1 = rec(b1, d - 1);
        nodes.emplace_back();
        own.push_back(dofs(bs));
        const int me = (int)nodes.size() - 1;
        nodes[me].kid[0] = k0;
        nodes[me].kid[1] = k1;
        nodes[k0].parent = me;
        nodes[k1].parent = me;
        nodes[me].height = 1 + std::max(nodes[k0].height, nodes[k1].height);
        return me;
    };
    rec(Box{{0, 0, 0}, {nnx, nny, nnz}}, depth);
    std::vector<int32_t> fr;
    for (size_t i = 0; i < nodes.size(); i++) {
        MGLevel::NDNode& t = nodes[i];
        const int64_t m = (int64_t)own[i].size();
        t.o0 = (int64_t)fr.size();
        t.m = std::max<int64_t>(kCB, (m + kCB - 1) / kCB * kCB);
        fr.insert(fr.end(), own[i].begin(), own[i].end());
        fr.insert(fr.end(), (size_t)(t.m - m), -1);
        t.s0 = t.kid[0] >= 0 ? nodes[t.kid[0]].s0 : t.o0;
    }
    for (MGLevel::NDNode& t : nodes) {
        t.f = t.m;
        for (int p = t.parent; p >= 0; p = nodes[p].parent) t.f += nodes[p].m;
    }
    return fr;
}

// One triangular product of the coarse solve (k_tri_apply semantics): by
// column chunks (k_tri_part + k_tri_sum) when the level has a chunk width,
// else one wave per row quad.
// With sums = false and the chunked form, only the chunk partials are
// written (into l.tpart1) and true is returned: the caller's column-tile pass
// forms y from them (k_tri_colpart's ypart).
template <typename T, typename TK, typename TB, typename TY>
static bool tri_product(Ctx& c, const MGLevel& l, const TK* M, int upper, const TB* bf, TY* y, T* x, const int* gate,
                        const double* ys, int nr, bool sums = true) {
    if (l.tcw) {
        const int32_t* rf = upper ? (const int32_t*)nullptr : (const int32_t*)l.rfirst;
        double* part = sums ? l.tpart : l.tpart1;
        if (l.trw == 8) {
            const dim3 gp((unsigned)l.tnch, (unsigned)(l.nfp / 32));
            TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_part<NR, TK, TB, 8>), gp, dim3(256), 0, c.s, M, l.nfp, upper,
                                                  bf, part, l.tcw, l.tnch, gate, rf));
        } else {
            const dim3 gp((unsigned)l.tnch, (unsigned)(l.nfp / 16));
            TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_part<NR, TK, TB, 4>), gp, dim3(256), 0, c.s, M, l.nfp, upper,
                                                  bf, part, l.tcw, l.tnch, gate, rf));
        }
        if (!sums) return true;
        TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_sum<T, NR, TY>), dim3(nblk(4 * l.nfp)), dim3(256), 0, c.s,
                                              (const double*)l.tpart, l.tnch, l.nfp, y, x, (const int32_t*)l.freeIdx,
                                              l.nf, l.ndof, gate, ys));
    } else {
        const dim3 gt((unsigned)(l.nfp / 16));
        TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_apply<T, NR, TK, TB, TY>), gt, dim3(256), 0, c.s, M, l.nfp, upper,
                                              bf, y, x, (const int32_t*)l.freeIdx, l.nf, l.ndof, gate, ys));
    }
    return false;
}

// Column-tile product Z^T y of the coarse solve: nw = 8 waves per tile (32
// rows each) or 4 (64 rows each), the same for every batch width, so a
// batched solve gives the bits of one right-hand side at a time. With ypart
// (the ynch row-chunk partials of Z b), y is formed by the tiles themselves.
template <int NR, typename TB, typename TK>
static void tri_colpart(hipStream_t s, const TK* M, int64_t n, const TB* y, double* part, int nrb, const int* gate,
                        const int32_t* rfirst, int nw, const double* ypart = nullptr, int ynch = 0,
                        const double* ys = nullptr) {
    const dim3 gc((unsigned)((n + tc_cols<TK>() - 1) / tc_cols<TK>()), (unsigned)nrb);
    if (nw == 8)
        hipLaunchKernelGGL((k_tri_colpart<NR, TB, TK, 8>), gc, dim3(512), 0, s, M, n, y, part, nrb, gate, 0, rfirst,
                           ypart, ynch, ys);
    else
        hipLaunchKernelGGL((k_tri_colpart<NR, TB, TK, 4>), gc, dim3(256), 0, s, M, n, y, part, nrb, gate, 0, rfirst,
                           ypart, ynch, ys);
}

template <typename T> struct Sel;
template <> struct Sel<double> {
    static double* b(MGLevel& l) { return l.b; }
    static double* x(MGLevel& l) { return l.x; }
    static double* x2(MGLevel& l) { return l.x2; }
    static double* res(MGLevel& l) { return l.res; }
    static double* invD(MGLevel& l) { return l.invD; }
    static double* K(MGLevel& l) { return l.K; }
    static double* Kinv(MGLevel& l) { return l.Kinv; }
    static const double* E(MGLevel& l) { return l.E; }
    typedef double TE;
};
template <> struct Sel<float> {
    static float* b(MGLevel& l) { return l.b32; }
    static float* x(MGLevel& l) { return l.x32; }
    static float* x2(MGLevel& l) { return l.x232; }
    static float* res(MGLevel& l) { return l.res32; }
    static float* invD(MGLevel& l) { return l.invD32; }
    static float* K(MGLevel& l) { return l.K32; }

    static const float* E(MGLevel& l) { return l.E32; }
    typedef float TE;
};

// x = omega * invD * b (first sweep from x = 0), gated
template <typename T>
__global__ void k_jacobi_first(T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ invD,
                               const double* __restrict__ omp, int64_t n, const int* gate) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    x += batch_off(n);
    b += batch_off(n);
    const T omega = (T)*omp;
    for (int64_t i = gtid(); i < n; i += gstride()) x[i] = T(0) + omega * invD[i] * b[i];
}
// coarse direct solve with the explicit inverse of the free block (nf free
// DOFs, padded to ld = nfp, a multiple of 4, so rows are 16-byte aligned):
// bf[j] = b[free[j]] (zero for the padding), then x[free[i]] = sum_j
// Kinv[i][j] bf[j]; x = 0 on fixed DOFs.
// With xz (a buffer other than b), also xz = 0 over the level's ndof DOFs
// (the fixed DOFs of the solution keep that zero).
template <typename T, typename TB = double>
__global__ void k_gather_free(const T* __restrict__ b, const int32_t* __restrict__ freeIdx, int64_t nf, int64_t nfp,
                              TB* __restrict__ bf, const int* gate, int64_t ndof, T* __restrict__ xz = nullptr) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    b += batch_off(ndof);
    bf += batch_off(nfp);
    if (xz) xz += batch_off(ndof);
    const int64_t m = xz && ndof > nfp ? ndof : nfp;
    for (int64_t j = gtid(); j < m; j += gstride()) {
        if (j < nfp) bf[j] = (j < nf && freeIdx[j] >= 0) ? (TB)b[freeIdx[j]] : TB(0);
        if (xz && j < ndof) xz[j] = T(0);
    }
}
// one wave per 4 consecutive rows; each lane streams 16-byte vectors of the
// four rows against the same slices of the NR right-hand sides bf (read from
// L2), so the inverse is read from HBM once for all of them. The inverse and
// the accumulation are FP64 in every precision; x receives the result in the
// V-cycle's precision T. A converged right-hand side (gate) is not written.
template <typename T, int NR, typename TK>
__global__ void __launch_bounds__(256) k_inv_apply(const TK* __restrict__ Kinv, const int32_t* __restrict__ freeIdx,
                                                   int64_t nf, int64_t nfp, const double* __restrict__ bf,
                                                   T* __restrict__ x, const int* gate, int64_t ndof) {
    bool live[NR];
    bool any = false;
#pragma unroll
    for (int q = 0; q < NR; q++) {
        live[q] = !(gate && gate[(int64_t)q * kGateStride]);
        any = any || live[q];
    }
    if (!any) return;
    const int lane = threadIdx.x & 63;
    const int64_t row0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 4;
    if (row0 >= nf) return;
    double acc[4][NR];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < NR; q++) acc[r][q] = 0.0;
    const int nr = (int)std::min<int64_t>(4, nf - row0);
    if constexpr (std::is_same<TK, double>::value) {
        const int64_t nv = nfp / 2;
        const double2* bv = (const double2*)bf;
        const double2* r0 = (const double2*)(Kinv + row0 * nfp);
        for (int64_t j = lane; j < nv; j += 64) {
            double2 bj[NR];
#pragma unroll
            for (int q = 0; q < NR; q++) bj[q] = bv[q * nv + j];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (r < nr) {
                    const double2 a = r0[r * nv + j];
#pragma unroll
                    for (int q = 0; q < NR; q++) acc[r][q] += a.x * bj[q].x + a.y * bj[q].y;
                }
            }
        }
    } else {
        // FP32 copy of the FP64 inverse: 4 entries per 16-byte load, FP64 sums
        const int64_t nv = nfp / 4;
        const double2* bv = (const double2*)bf;
        const float4* r0 = (const float4*)(Kinv + row0 * nfp);
        for (int64_t j = lane; j < nv; j += 64) {
            double2 b0[NR], b1[NR];
#pragma unroll
            for (int q = 0; q < NR; q++) {
                b0[q] = bv[q * (nfp / 2) + 2 * j];
                b1[q] = bv[q * (nfp / 2) + 2 * j + 1];
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (r < nr) {
                    const float4 a = r0[r * nv + j];
#pragma unroll
                    for (int q = 0; q < NR; q++)
                        acc[r][q] += (double)a.x * b0[q].x + (double)a.y * b0[q].y + (double)a.z * b1[q].x +
                                     (double)a.w * b1[q].y;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < NR; q++)
            for (int off = 32; off > 0; off >>= 1) acc[r][q] += __shfl_down(acc[r][q], off, 64);
    if (lane == 0)
        for (int q = 0; q < NR; q++)
            if (live[q])
                for (int r = 0; r < nr; r++)
                    if (freeIdx[row0 + r] >= 0) x[q * ndof + freeIdx[row0 + r]] = (T)acc[r][q];
}

// Dense symmetric positive (semi-)definite inverse by Gauss-Jordan without
// pivoting, one launch per step k. A pivot below 1e-11 times the DOF's
// original diagonal marks a DOF that is (numerically) dependent on the ones
// eliminated before it -- a mechanism of the coarse operator after voided
// elements disconnect parts of the structure: its row and column are
// dropped (zero in the result), so the result is the inverse on the
// remaining subspace.
__global__ void k_gj_scale_row(double* A, int64_t n, int64_t k, const double* d0) {
    const double akk = A[k * n + k];
    const bool skip = !(akk > 1e-11 * d0[k]);
    const double inv = skip ? 0.0 : 1.0 / akk;
    for (int64_t j = gtid(); j < n; j += gstride()) {
        if (j == k) continue;
        A[k * n + j] *= inv;
    }
}
__global__ void k_gj_eliminate(double* A, int64_t n, int64_t k, const double* d0) {
    // rows i != k: A[i][j] -= A[i][k] * A[k][j] for j != k (row k already scaled)
    const double akk = A[k * n + k];
    if (!(akk > 1e-11 * d0[k])) return;
    for (int64_t w = gtid(); w < n * n; w += gstride()) {
        const int64_t i = w / n, j = w % n;
        if (i == k || j == k) continue;
        A[i * n + j] -= A[i * n + k] * A[k * n + j];
    }
}
__global__ void k_gj_fix_col(double* A, int64_t n, int64_t k, const double* d0) {
    const double akk = A[k * n + k];
    const bool skip = !(akk > 1e-11 * d0[k]);
    for (int64_t i = gtid(); i < n; i += gstride()) {
        if (i == k) continue;
        A[i * n + k] = skip ? 0.0 : -A[i * n + k] / akk;
    }
}
__global__ void k_gj_pivot(double* A, int64_t n, int64_t k, const double* d0) {
    if (gtid() == 0) {
        const double akk = A[k * n + k];
        A[k * n + k] = (akk > 1e-11 * d0[k]) ? 1.0 / akk : 0.0;
    }
}
__global__ void k_diag_copy(const double* A, int64_t n, double* d0) {
    for (int64_t i = gtid(); i < n; i += gstride()) d0[i] = A[i * n + i];
}
// Blocked form of the same Gauss-Jordan inverse (panels of 64 pivots):
//   D   = A_kk^-1            diagonal block, one workgroup, in LDS
//   A_kj <- D A_kj           row panel
//   A_ij <- A_ij - A_ik A_kj rank-64 update of everything outside the panel
//   A_ik <- -A_ik D          column panel
// with the same pivot-dropping rule inside D. 4 launches per panel.
constexpr int kGJB = 64;

__global__ void __launch_bounds__(256) k_bgj_diag(double* A, int64_t n, int64_t k0, int kb, const double* d0) {
    __shared__ double D[kGJB][kGJB + 1];
    __shared__ double rowk[kGJB], colk[kGJB], thr[kGJB];
    const int t = threadIdx.x;
    for (int e = t; e < kGJB * kGJB; e += 256) {
        const int i = e >> 6, j = e & 63;
        D[i][j] = (i < kb && j < kb) ? A[(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
    }
    if (t < kGJB) thr[t] = t < kb ? 1e-11 * d0[k0 + t] : 0.0;
    __syncthreads();
    // each thread owns 16 entries of the 64x64 block: rows r0..r0+3, cols c0..c0+3
    const int r0 = (t >> 4) * 4, c0 = (t & 15) * 4;
    for (int p = 0; p < kb; p++) {
        const double app = D[p][p];
        const bool keep = app > thr[p];
        if (t < kGJB) { rowk[t] = D[p][t]; colk[t] = D[t][p]; }
        __syncthreads();
        const double inv = keep ? 1.0 / app : 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = r0 + u, j = c0 + v;
                double val = D[i][j];
                if (!keep) {
                    if (i == p || j == p) val = 0.0;
                } else if (i == p && j == p) {
                    val = inv;
                } else if (i == p) {
                    val = rowk[j] * inv;
                } else if (j == p) {
                    val = -colk[i] * inv;
                } else {
                    val = val - colk[i] * rowk[j] * inv;
                }
                D[i][j] = val;
            }
        __syncthreads();
    }
    for (int e = t; e < kb * kb; e += 256) A[(k0 + e / kb) * n + k0 + e % kb] = D[e / kb][e % kb];
}

// row panel: A[k0+r][j] <- sum_q D[r][q] A[k0+q][j] for j outside the panel;
// one workgroup per 64 columns
__global__ void __launch_bounds__(256) k_bgj_row(double* A, int64_t n, int64_t k0, int kb) {
    __shared__ double D[kGJB][kGJB + 1];
    __shared__ double R[kGJB][kGJB + 1];
    const int t = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.x * kGJB;
    if (j0 == k0) return;
    const int jb = (int)min<int64_t>(kGJB, n - j0);
    for (int e = t; e < kb * kb; e += 256) D[e / kb][e % kb] = A[(k0 + e / kb) * n + k0 + e % kb];
    for (int e = t; e < kb * jb; e += 256) R[e / jb][e % jb] = A[(k0 + e / jb) * n + j0 + e % jb];
    __syncthreads();
    for (int e = t; e < kb * jb; e += 256) {
        const int r = e / jb, c = e % jb;
        double s = 0;
        for (int q = 0; q < kb; q++) s += D[r][q] * R[q][c];
        A[(k0 + r) * n + j0 + c] = s;
    }
}

// rank-kb update of the tiles outside the panel: C[i][j] -= Acol[i][q] Arow[q][j];
// 64x64 tile per workgroup, 4x4 outputs per thread
__global__ void __launch_bounds__(256) k_bgj_update(double* A, int64_t n, int64_t k0, int kb) {
    const int64_t i0 = (int64_t)blockIdx.y * kGJB, j0 = (int64_t)blockIdx.x * kGJB;
    if (i0 == k0 || j0 == k0) return;
    __shared__ double Cs[kGJB][kGJB + 1];   // column panel tile: rows i0.., cols k0..
    __shared__ double Rs[kGJB][kGJB + 1];   // row panel tile: rows k0.., cols j0..
    const int t = threadIdx.x;
    const int ib = (int)min<int64_t>(kGJB, n - i0), jb = (int)min<int64_t>(kGJB, n - j0);
    for (int e = t; e < kGJB * kGJB; e += 256) {
        const int r = e / kGJB, c = e % kGJB;
        Cs[r][c] = (r < ib && c < kb) ? A[(i0 + r) * n + k0 + c] : 0.0;
        Rs[r][c] = (r < kb && c < jb) ? A[(k0 + r) * n + j0 + c] : 0.0;
    }
    __syncthreads();
    const int tr = (t / 16) * 4, tc = (t % 16) * 4;
    double acc[4][4] = {{0}};
    for (int q = 0; q < kb; q++) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { a[u] = Cs[tr + u][q]; b[u] = Rs[q][tc + u]; }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            if (tr + u < ib && tc + v < jb) A[(i0 + tr + u) * n + j0 + tc + v] -= acc[u][v];
}

// operands of the rank-kb update as one plain GEMM A -= C R: C = A[:, panel]
// with the panel rows zeroed (n x kb), R = A[panel, :] with the panel
// columns zeroed (kb x n), so rows and columns of the panel are untouched
__global__ void k_bgj_panels(const double* __restrict__ A, int64_t n, int64_t k0, int kb, double* __restrict__ Cp,
                             double* __restrict__ Rp) {
    for (int64_t w = gtid(); w < n * kb; w += gstride()) {
        const int64_t i = w / kb, q = w % kb;
        Cp[w] = (i >= k0 && i < k0 + kb) ? 0.0 : A[i * n + k0 + q];
        const int64_t r = w / n, j = w % n;
        Rp[w] = (j >= k0 && j < k0 + kb) ? 0.0 : A[(k0 + r) * n + j];
    }
}

// column panel: A[i][k0+c] <- -sum_q A[i][k0+q] D[q][c] for i outside the panel
__global__ void __launch_bounds__(256) k_bgj_col(double* A, int64_t n, int64_t k0, int kb) {
    __shared__ double D[kGJB][kGJB + 1];
    __shared__ double Cq[kGJB][kGJB + 1];
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * kGJB;
    if (i0 == k0) return;
    const int ib = (int)min<int64_t>(kGJB, n - i0);
    for (int e = t; e < kb * kb; e += 256) D[e / kb][e % kb] = A[(k0 + e / kb) * n + k0 + e % kb];
    for (int e = t; e < ib * kb; e += 256) Cq[e / kb][e % kb] = A[(i0 + e / kb) * n + k0 + e % kb];
    __syncthreads();
    for (int e = t; e < ib * kb; e += 256) {
        const int r = e / kb, c = e % kb;
        double s = 0;
        for (int q = 0; q < kb; q++) s += Cq[r][q] * D[q][c];
        A[(i0 + r) * n + k0 + c] = -s;
    }
}

__global__ void k_extract_free(const double* __restrict__ K, int64_t n, const int32_t* __restrict__ freeIdx, int64_t nf,
                               int64_t ld, double* __restrict__ A) {
    for (int64_t w = gtid(); w < nf * nf; w += gstride()) {
        const int64_t i = w / nf, j = w % nf;
        A[i * ld + j] = (freeIdx[i] >= 0 && freeIdx[j] >= 0) ? K[(int64_t)freeIdx[i] * n + freeIdx[j]] : 0.0;
    }
}
// Free-DOF block of the assembled operator sum_e E_e P_e^T K0 P_e, written
// row by row without atomics: entry (a, b) gathers, in a fixed order, the
// <= 8 elements that hold both nodes (E in zyx element order, reference node
// numbering). freeOf maps a DOF to its free index or -1. A must be zero
// outside the 81 neighbour columns of each row.
__global__ void k_assemble_free(const double* __restrict__ E, double thr, Grid g, const int32_t* __restrict__ freeIdx,
                                const int32_t* __restrict__ freeOf, int64_t nf, int64_t ld,
                                const double* __restrict__ ke, double* __restrict__ A) {
    const int nny = g.nny(), nnz = g.nnz();
    for (int64_t w = gtid(); w < nf * 81; w += gstride()) {
        const int64_t fa = w / 81;
        const int q = (int)(w % 81), nb = q / 3, d2 = q % 3;
        const int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
        const int64_t a = freeIdx[fa];
        if (a < 0) continue;   // padding position of the nested-dissection order
        const int64_t na = a / 3;
        const int d1 = (int)(a % 3);
        const int iz = (int)(na % nnz);
        const int64_t t = na / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
        if (jx < 0 || jx > g.nelx || jy < 0 || jy > g.nely || jz < 0 || jz > g.nelz) continue;
        const int64_t bdof = 3 * (((int64_t)jx * nny + jy) * nnz + jz) + d2;
        const int32_t fb = freeOf[bdof];
        if (fb < 0) continue;
        double s = 0.0;
        for (int ex = max(max(ix, jx) - 1, 0); ex <= min(min(ix, jx), g.nelx - 1); ex++)
            for (int ey = max(max(iy, jy) - 1, 0); ey <= min(min(iy, jy), g.nely - 1); ey++)
                for (int ez = max(max(iz, jz) - 1, 0); ez <= min(min(iz, jz), g.nelz - 1); ez++) {
                    const double Ee = E[((int64_t)ex * g.nely + ey) * g.nelz + ez];
                    if (!(Ee > thr)) continue;
                    // local node = 4*c + {(0,0):0, (1,0):1, (1,1):2, (0,1):3}[(a,b)]
                    const int ai = ix - ex, bi = iy - ey, ci = iz - ez;
                    const int aj = jx - ex, bj = jy - ey, cj = jz - ez;
                    const int li = 4 * ci + (bi ? (ai ? 2 : 3) : ai);
                    const int lj = 4 * cj + (bj ? (aj ? 2 : 3) : aj);
                    s += Ee * ke[(3 * li + d1) * 24 + 3 * lj + d2];
                }
        A[fa * ld + fb] = s;
    }
}
// rows/cols of free DOFs without stiffness (and the padding rows) get a unit
// diagonal: one workgroup per row (grid nf)
__global__ void k_regularize(double* A, int64_t nf) {
    const int64_t i = blockIdx.x;
    if (i < nf && !(A[i * nf + i] > 1e-30)) {
        for (int64_t j = threadIdx.x; j < nf; j += blockDim.x) {
            A[i * nf + j] = 0;
            A[j * nf + i] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) A[i * nf + i] = 1.0;
    }
}

struct MGH {
    std::vector<MGLevel> L;
    double omega = 0.5;
    int nu1 = 2, nu2 = 2, coarse_sweeps = 30, coarse_solver = TOPO_COARSE_SWEEPS;
    int galerkin_max = 3000;
    double omega_cap = 0.0;
    std::vector<double> omegaL;
    double* d_omega = nullptr;    // per-level weights on the device (read by the kernels)
    double thr = 1e-6;   // skip threshold 1000 * Emin
    bool want32 = false;
    bool exact = false;   // reference-order kernel everywhere
    int xm_nt = 0;        // x-march workgroup size for every level (0 auto, 256, 512, 1024)
    int xm_cx = 0;        // x-march planes per workgroup for every level (0 auto)
    int xm_cx64 = 0;      // x-march planes per workgroup of every FP64 operator (0: autotuned / as FP32)
    int ncu = 256;        // compute units of the device (the x-chunk size rule)
    // the coarsest level's gathered right-hand side (bf) is zero at its
    // padding positions and its solution vectors at the fixed DOFs, so the
    // restriction may write bf directly (set by build once they are zeroed)
    bool bf_direct = false;
    // FP32 operator on the x-march levels with the element product on the
    // matrix cores (k_applyA_mf) instead of the VALU Hadamard kernel
    bool mfma = true;
    bool rz_fuse = true;  // r.z partials from the V-cycle's last sweep (else a separate kernel)
    std::vector<int> lv_nt, lv_cx;   // per-level choices of the autotuner (0 auto)
    // x-chunk of the FP64 operator where the autotuner timed it apart from
    // the level's FP32 one (the mixed solvers' CG operator; 0: lv_cx)
    std::vector<int> lv_cx64;
    DevKe* K0 = nullptr;
    Springs springs{};    // mechanism ports on the finest level
    int nrhs = 1;         // right-hand sides a batched V-cycle can carry (level vectors are nrhs x ndof)
    // FP32 V-cycles apply a rounded copy of the coarsest level's factor
    // inverse: 0 none (FP64), 1 FP32, 2 FP16 with one scale per row
    int coarse_fp32 = 0;
    // nested-dissection coarse factorization: tree depth (0: banded, natural
    // order), smallest block it is used for, largest part left uncut
    int nd_depth = 3;
    int64_t nd_min = 2048, nd_leaf = 768;
    int tc_nw = 8;   // waves per column tile of the coarse Z^T y product (4 or 8)
    // MG-PCG iterations per captured graph between host polls (converged
    // iterations are gated no-ops; 8 measured 0.5 % faster than 4 at C3)
    int cg_check = 8;
    // 'petsc' solver: the domain-decomposition preconditioner replaces the
    // V-cycle (one level; omega 1 for its global Jacobi sweeps)
    bool bddc = false;
    int bddc_sweeps = 3;                 // KSP_BDDC_SMOOTHER_ITERS (js :1031)
    double* bddc_val = nullptr;          // kBddcSlots x ndof local inverse diagonals
    uint8_t* bddc_cnt = nullptr;         // subdomains containing each DOF
    uint8_t* bddc_corner = nullptr;      // primal (corner) DOF flags
    int32_t *bddc_sx = nullptr, *bddc_sy = nullptr, *bddc_sz = nullptr;   // element -> subdomain, per axis
    int bddc_ns[3] = {0, 0, 0};
    // Jacobi-PCG iterations per captured graph between host polls (32
    // measured 5 % faster than 8 on the 2-D leg)
    int jcg_check = 32;
    template <typename T> const T* ke() const;
    template <typename T> HadCoef<T> had() const;

    // Every host-side choice a captured CG graph bakes in besides its buffer
    // pointers: the coarse-solve branch per level (triangular products or the
    // Gauss-Jordan inverse; FP16/FP32 copies present), the dense/E-average
    // operator kind, the autotuned element tiles, the kernel family and the
    // iterations per graph. Appended to the graph-cache key, so a design
    // update that changes one of them (a coarsest block that is no longer
    // positive definite falls back to Gauss-Jordan) captures a new graph
    // instead of replaying the old one.
    void graph_sig(std::vector<const void*>& key) const {
        auto put = [&](int64_t v) { key.push_back((const void*)(intptr_t)v); };
        put((int64_t)L.size());
        for (size_t li = 0; li < L.size(); li++) {
            const MGLevel& l = L[li];
            put((l.tri ? 1 : 0) | (l.dense ? 2 : 0) | (l.chol ? 4 : 0) | (l.nd ? 8 : 0) | (l.Z16 ? 16 : 0) |
                (l.Z32 ? 32 : 0) | (l.Kinv32 ? 64 : 0));
            put(li < lv_nt.size() ? lv_nt[li] : -1);
            put(li < lv_cx.size() ? lv_cx[li] : -1);
            put(li < lv_cx64.size() ? lv_cx64[li] : -1);
        }
        put((exact ? 1 : 0) | (rz_fuse ? 2 : 0) | (bddc ? 4 : 0) | (mfma ? 8 : 0));
        put(xm_nt);
        put(xm_cx);
        put(xm_cx64);
        put(cg_check);
        put(jcg_check);
        put(coarse_fp32);
    }

    // build levels: factor-2 coarsening of MGPrecond3D.ensure
    // (js/optimizer-worker-3d.js:408-442); nlevels <= 0 = reference rule
    void build(Ctx& c, Grid g0, const uint8_t* dfixed0, int nlevels, bool fp32) {
        want32 = fp32;
        {
            int dev = 0, n = 0;
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
                ncu = n;
        }
        xm_allow_lds();
        // the transfer kernels index nodes in 32 bits
        TOPO_CHECK(g0.nnodes() < ((int64_t)1 << 32), TOPO_E_ARG, "grid of %lld nodes too large",
                   (long long)g0.nnodes());
        L.clear();
        MGLevel l0;
        l0.g = g0;
        l0.ndof = g0.ndof();
        l0.fixed = const_cast<uint8_t*>(dfixed0);
        L.push_back(l0);
        const int maxL = nlevels > 0 ? nlevels : 6;
        int cx = g0.nelx, cy = g0.nely, cz = g0.nelz;
        for (int li = 1; li < maxL; li++) {
            const int nx = cx / 2, ny = cy / 2, nz = cz / 2;
            if (nx < 2 || ny < 2 || nz < 2) break;
            MGLevel l;
            l.g.nelx = nx; l.g.nely = ny; l.g.nelz = nz;
            l.ndof = l.g.ndof();
            L.push_back(l);
            cx = nx; cy = ny; cz = nz;
        }
        // fixed masks by injection (_downsampleFixedMaskBy2 :384-406), done on
        // the host from the fine mask
        std::vector<uint8_t> hf(L[0].ndof);
        TOPO_HIP(hipMemcpy(hf.data(), dfixed0, L[0].ndof, hipMemcpyDeviceToHost));
        for (size_t li = 1; li < L.size(); li++) {
            const Grid& gf = L[li - 1].g;
            const Grid& gc = L[li].g;
            std::vector<uint8_t> hc(L[li].ndof, 0);
            for (int ccz = 0; ccz < gc.nnz(); ccz++) {
                const int fz = std::min(ccz * 2, gf.nnz() - 1);
                for (int ccy = 0; ccy < gc.nny(); ccy++) {
                    const int fy = std::min(ccy * 2, gf.nny() - 1);
                    for (int ccx = 0; ccx < gc.nnx(); ccx++) {
                        const int fx = std::min(ccx * 2, gf.nnx() - 1);
                        const int64_t nF = ((int64_t)fx * gf.nny() + fy) * gf.nnz() + fz;
                        const int64_t nC = ((int64_t)ccx * gc.nny() + ccy) * gc.nnz() + ccz;
                        for (int d = 0; d < 3; d++) hc[3 * nC + d] = hf[3 * nF + d];
                    }
                }
            }
            L[li].fixed = c.alloc<uint8_t>(L[li].ndof);
            TOPO_HIP(hipMemcpy(L[li].fixed, hc.data(), L[li].ndof, hipMemcpyHostToDevice));
            hf.swap(hc);
        }
        for (size_t li = 0; li < L.size(); li++) {
            MGLevel& l = L[li];
            if (li > 0) {
                l.E = c.alloc<double>(l.g.nel());
                if (fp32) l.E32 = c.alloc<float>(l.g.nel());
            }
            l.invD = c.alloc<double>(l.ndof);
            const int64_t nv = l.ndof * nrhs;
            l.b = c.alloc<double>(nv);
            l.x = c.alloc<double>(nv);
            l.x2 = c.alloc<double>(nv);
            l.res = c.alloc<double>(nv);
            if (fp32) {
                l.invD32 = c.alloc<float>(l.ndof);
                l.b32 = c.alloc<float>(nv);
                l.x32 = c.alloc<float>(nv);
                l.x232 = c.alloc<float>(nv);
                l.res32 = c.alloc<float>(nv);
            }
        }
    }

    // per-density operator update (updateFromFine :558-582); level 0's E is
    // set by the caller.
    void update(Ctx& c) {
        MGLevel& l0 = L[0];
        hipLaunchKernelGGL(k_invdiag, dim3(nblk(l0.g.nnodes())), dim3(256), 0, c.s, l0.E, thr, l0.fixed, l0.invD,
                           l0.invD32, l0.g, (const double*)K0->ke);
        if (springs.n > 0)
            hipLaunchKernelGGL(k_spring_diag, dim3(1), dim3(64), 0, c.s, l0.invD, l0.invD32, springs);
        c.check_launch();
        for (size_t li = 1; li < L.size(); li++) {
            MGLevel& f = L[li - 1];
            MGLevel& l = L[li];
            if (l.ndof <= galerkin_max) {
                if (!l.K) {
                    l.K = c.alloc<double>(l.ndof * l.ndof);
                    if (want32) l.K32 = c.alloc<float>(l.ndof * l.ndof);
                }
                l.dense = true;
                hipLaunchKernelGGL(k_fill<double>, dim3(nblk(l.ndof * l.ndof)), dim3(256), 0, c.s, l.K, 0.0,
                                   l.ndof * l.ndof, (const int*)nullptr);
                if (!f.dense) {
                    hipLaunchKernelGGL(k_galerkin_elems, dim3(nblk(f.g.nel() * 576)), dim3(256), 0, c.s, f.E, thr, f.g,
                                       l.g, l.K, (const double*)K0->ke);
                } else {
                    double* T = c.alloc<double>(f.ndof * l.ndof);
                    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(f.ndof * l.ndof)), dim3(256), 0, c.s, T, 0.0,
                                       f.ndof * l.ndof, (const int*)nullptr);
                    hipLaunchKernelGGL(k_galerkin_KP, dim3(nblk(f.ndof * f.ndof)), dim3(256), 0, c.s, f.K, f.ndof, f.g,
                                       l.g, T);
                    hipLaunchKernelGGL(k_galerkin_PtT, dim3(nblk(f.ndof * l.ndof)), dim3(256), 0, c.s, T, f.ndof, f.g,
                                       l.g, l.K);
                    c.sync();
                    c.free_one(T);
                }
                hipLaunchKernelGGL(k_dense_invdiag, dim3(nblk(l.ndof)), dim3(256), 0, c.s, l.K, l.ndof, l.fixed, l.invD,
                                   l.invD32);
                if (want32)
                    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(l.ndof * l.ndof)), dim3(256), 0, c.s, l.K32,
                                       l.K, l.ndof * l.ndof);
            } else {
                l.dense = false;
                hipLaunchKernelGGL(k_coarsen_E, dim3(nblk(l.g.nel())), dim3(256), 0, c.s, f.E, l.E, l.E32, f.g, l.g);
                hipLaunchKernelGGL(k_invdiag, dim3(nblk(l.g.nnodes())), dim3(256), 0, c.s, l.E, thr, l.fixed, l.invD,
                                   l.invD32, l.g, (const double*)K0->ke);
            }
            c.check_launch();
        }
        omegaL.assign(L.size(), omega);
        const int nl = (int)std::min<size_t>(L.size(), 16);
        // the power iterations of every level run on the second stream,
        // beside the coarse factorization (which does not depend on the
        // smoother weights): the second stream waits for the work issued so
        // far, and their launches are issued once the factorization's are
        // (issued first, their 60 launches per level held the start of the
        // factorization's chain back by their issue time); their results
        // are read after it
        std::function<void()> power;
        if (omega_cap > 0) {
            if (!d_psc) d_psc = c.alloc<double>(4 * 16);
            c.side_ready(NB_EV_MIN);
            c.order(c.s, c.s2, 0);
            power = [this, &c, nl] {
                std::swap(c.s, c.s2);
                for (int li = 0; li < nl; li++) power_launch(c, li, d_psc + 4 * li);
                std::swap(c.s, c.s2);
                c.check_launch();
            };
        }
        if (coarse_solver == TOPO_COARSE_CHOLESKY)
            factor_coarsest(c, power ? &power : nullptr);
        else if (power)
            power();
        if (omega_cap > 0) {
            TOPO_HIP(hipStreamSynchronize(c.s2));
            double h[64];
            TOPO_HIP(hipMemcpy(h, d_psc, sizeof(double) * 4 * nl, hipMemcpyDeviceToHost));
            for (int li = 0; li < nl; li++) {
                // v is unit-norm after the first step, so rho = ||D^-1 A v||
                const double rho = std::sqrt(h[4 * li + 1]);
                if (rho > 0) omegaL[li] = std::min(omega, omega_cap / rho);
            }
        }
        upload_omega(c);
        // the coarsest direct solve's buffers: bf all zero (only free
        // positions are written later), x and x32 zero (the solve writes the
        // free DOFs; the power iterations used x as scratch)
        bf_direct = false;
        {
            MGLevel& lc = L.back();
            if (L.size() > 1 && lc.chol && lc.bf && lc.freeOf) {
                hipLaunchKernelGGL(k_fill<double>, dim3(nblk(lc.nfp * nrhs)), dim3(256), 0, c.s, lc.bf, 0.0,
                                   lc.nfp * nrhs, (const int*)nullptr);
                hipLaunchKernelGGL(k_fill<double>, dim3(nblk(lc.ndof * nrhs)), dim3(256), 0, c.s, lc.x, 0.0,
                                   lc.ndof * nrhs, (const int*)nullptr);
                if (lc.x32)
                    hipLaunchKernelGGL(k_fill<float>, dim3(nblk(lc.ndof * nrhs)), dim3(256), 0, c.s, lc.x32, 0.0f,
                                       lc.ndof * nrhs, (const int*)nullptr);
                c.check_launch();
                bf_direct = true;
            }
        }
        if (bddc) bddc_build(c);
    }

    double* d_psc = nullptr;   // power-iteration scratch, 4 doubles per level
    static constexpr int NB_EV_MIN = 2;
    void upload_omega(Ctx& c) {
        if (!d_omega) d_omega = c.alloc<double>(16);
        double h[16] = {0};
        for (size_t li = 0; li < L.size() && li < 16; li++) h[li] = omegaL.empty() ? omega : omegaL[li];
        TOPO_HIP(hipMemcpyAsync(d_omega, h, sizeof h, hipMemcpyHostToDevice, c.s));
        c.sync();
    }
    const double* omp(int li) const { return d_omega ? d_omega + li : nullptr; }

    // smoother safeguard: rho ~ lambda_max(D^-1 A) of level li by 20 power
    // iterations from the hashed vector v_i = frac(i*2654435761/2^32) - 1/2 on
    // free DOFs (the oracle's power_rho); FP64.
    // (launches only; sc[1] receives rho^2 for the level, sc 4 doubles)
    void power_launch(Ctx& c, int li, double* sc) {
        MGLevel& l = L[li];
        hipLaunchKernelGGL(k_power_init, dim3(nblk(l.ndof)), dim3(256), 0, c.s, l.x, (const uint8_t*)l.fixed, l.ndof,
                           sc);
        for (int it = 0; it < 20; it++) {
            apply<double>(c, li, l.x, l.x2, 0, nullptr, nullptr);
            hipLaunchKernelGGL(k_power_scale, dim3(kRedBlocks), dim3(kRedThreads), 0, c.s, l.x2,
                               (const double*)l.invD, l.ndof, sc, c.partials, c.ticket);
            hipLaunchKernelGGL(k_power_norm, dim3(nblk(l.ndof)), dim3(256), 0, c.s, l.x, (const double*)l.x2, l.ndof,
                               (const double*)sc);
        }
    }

    // FP64 direct solve of the coarsest level: the free-DOF block is
    // assembled (deterministically, no atomics), factored by the banded
    // blocked Cholesky and Z = L^-1 formed once per density update; the
    // V-cycle applies x = Z^T (Z b). A block that is not positive definite
    // (voided elements left out of the operator make mechanisms) falls back
    // to Gauss-Jordan with pivot dropping and one GEMV with the inverse.
    void coarse_block(Ctx& c, MGLevel& l) {
        const int64_t nf = l.nf, ld = l.nfp;
        hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ld * ld)), dim3(256), 0, c.s, l.Kinv, 0.0, ld * ld,
                           (const int*)nullptr);
        if (l.dense)
            hipLaunchKernelGGL(k_extract_free, dim3(nblk(nf * nf)), dim3(256), 0, c.s, (const double*)l.K, l.ndof,
                               l.freeIdx, nf, ld, l.Kinv);
        else
            hipLaunchKernelGGL(k_assemble_free, dim3(nblk(nf * 81)), dim3(256), 0, c.s, (const double*)l.E, thr, l.g,
                               l.freeIdx, l.freeOf, nf, ld, (const double*)K0->ke, l.Kinv);
        hipLaunchKernelGGL(k_regularize, dim3((unsigned)ld), dim3(256), 0, c.s, l.Kinv, ld);
    }
    // `overlap` (may be null) is called once, after the nested-dissection
    // factorization's launches are issued and before the host waits for
    // it; on the other paths (which use the side streams) before they start
    void factor_coarsest(Ctx& c, const std::function<void()>* overlap = nullptr) {
        MGLevel& l = L.back();
        bool ov_done = false;
        auto run_overlap = [&] {
            if (overlap && !ov_done) {
                ov_done = true;
                (*overlap)();
            }
        };
        if (!l.freeIdx) {
            std::vector<uint8_t> hf(l.ndof);
            TOPO_HIP(hipMemcpy(hf.data(), l.fixed, l.ndof, hipMemcpyDeviceToHost));
            std::vector<int32_t> fr, of(l.ndof, -1);
            for (int64_t i = 0; i < l.ndof; i++)
                if (!hf[i]) {
                    of[i] = (int32_t)fr.size();
                    fr.push_back((int32_t)i);
                }
            l.nf = (int64_t)fr.size();
            TOPO_CHECK(l.nf > 0 && l.nf <= 20000, TOPO_E_UNSUPPORTED,
                       "coarsest level has %lld free DOFs; the direct solve supports 1..20000", (long long)l.nf);
            // bandwidth of the free block: nonzeros couple the DOFs of nodes
            // at most one element apart
            const int nny = l.g.nny(), nnz = l.g.nnz();
            int w = 0;
            for (int64_t fa = 0; fa < l.nf; fa++) {
                const int64_t na = fr[fa] / 3;
                const int iz = (int)(na % nnz), iy = (int)((na / nnz) % nny), ix = (int)(na / ((int64_t)nnz * nny));
                for (int dx = -1; dx <= 1; dx++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dz = -1; dz <= 1; dz++) {
                            const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                            if (jx < 0 || jx > l.g.nelx || jy < 0 || jy > l.g.nely || jz < 0 || jz > l.g.nelz) continue;
                            for (int d = 0; d < 3; d++) {
                                const int32_t fb = of[3 * (((int64_t)jx * nny + jy) * nnz + jz) + d];
                                if (fb >= 0) w = std::max<int>(w, (int)std::abs(fb - fa));
                            }
                        }
            }
            l.bandw = w;
            // nested-dissection order for larger blocks (positions include
            // the padding of each tree node's own range)
            if (nd_depth > 0 && l.nf >= nd_min) {
                std::vector<int32_t> frn = nd_order(l.g, hf, nd_depth, nd_leaf, l.ndn);
                if (l.ndn.size() > 1) {
                    fr.swap(frn);
                    std::fill(of.begin(), of.end(), -1);
                    for (size_t p = 0; p < fr.size(); p++)
                        if (fr[p] >= 0) of[fr[p]] = (int32_t)p;
                    l.nf = (int64_t)fr.size();
                    l.nd = true;
                } else {
                    l.ndn.clear();
                }
            }
            l.nfp = (l.nf + kCB - 1) / kCB * kCB;
            l.freeIdx = c.alloc<int32_t>(l.nf);
            TOPO_HIP(hipMemcpy(l.freeIdx, fr.data(), l.nf * sizeof(int32_t), hipMemcpyHostToDevice));
            l.freeOf = c.alloc<int32_t>(l.ndof);
            TOPO_HIP(hipMemcpy(l.freeOf, of.data(), l.ndof * sizeof(int32_t), hipMemcpyHostToDevice));
            l.Kinv = c.alloc<double>(l.nfp * l.nfp);
            l.bf = c.alloc<double>(l.nfp * nrhs);
            l.ytmp = c.alloc<double>(l.nfp * nrhs);
            {
                const char* e = std::getenv("TOPO_TRI_CW");
                l.tcw = e ? std::atoi(e) : 2048;
                const char* er = std::getenv("TOPO_TRI_RW");
                l.trw = er && std::atoi(er) == 4 ? 4 : 8;
                if (l.tcw < 0 || l.tcw % 64) l.tcw = 2048;
                l.tnch = l.tcw ? (int)((l.nfp + l.tcw - 1) / l.tcw) : 0;
                l.tnrb = (int)((l.nfp + kTcRows - 1) / kTcRows);
                l.tpart = c.alloc<double>((int64_t)nrhs * std::max(l.tnch, l.tnrb) * l.nfp);
                if (l.tnch) l.tpart1 = c.alloc<double>((int64_t)nrhs * l.tnch * l.nfp);
            }
            l.d0c = c.alloc<double>(l.nfp);
            if (want32 && coarse_fp32) l.Kinv32 = c.alloc<float>(l.nfp * l.nfp);
            if (l.nd) nd_alloc(c, l);
        }
        const int64_t n = l.nfp;
        const int NB = (int)(n / kCB);
        const int wb = std::max(1, (l.bandw + kCB - 1) / kCB);
        coarse_block(c, l);
        // nested-dissection or banded Cholesky when the band is narrow
        // enough to pay off
        bool done = false;
        if (l.nd) {
            if (!l.Z) {
                l.Z = c.alloc<double>(n * n);
                // written only inside the factor's structure: the rest stays 0
                hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n * n)), dim3(256), 0, c.s, l.Z, 0.0, n * n,
                                   (const int*)nullptr);
                if (want32 && coarse_fp32 == 1) l.Z32 = c.alloc<float>(n * n);
                if (want32 && coarse_fp32 == 2) {
                    l.Z16 = c.alloc<_Float16>(n * n);
                    l.zs2 = c.alloc<double>(n);
                }
            }
            int* dflag = (int*)(c.dsc + 14);   // device scratch word
            hipLaunchKernelGGL(k_diag_copy, dim3(nblk(n)), dim3(256), 0, c.s, (const double*)l.Kinv, n, l.d0c);
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c.s, dflag, 0, (int64_t)1, (const int*)nullptr);
            nd_factor(c, l, dflag);
            c.check_launch();
            run_overlap();
            TOPO_HIP(hipMemcpyAsync(&c.h_sc[0], dflag, sizeof(int), hipMemcpyDeviceToHost, c.s));
            c.sync();
            int flag = 0;
            std::memcpy(&flag, &c.h_sc[0], sizeof(int));
            if (!flag) {
                z_copies(c, l);
                done = true;
                l.tri = true;
            }
        } else if (NB > 2 * wb) {
            run_overlap();
            if (!l.Z) {
                l.Z = c.alloc<double>(n * n);
                l.Linv = c.alloc<double>((int64_t)NB * kCB * kCB);
                if (want32 && coarse_fp32 == 1) l.Z32 = c.alloc<float>(n * n);
                if (want32 && coarse_fp32 == 2) {
                    l.Z16 = c.alloc<_Float16>(n * n);
                    l.zs2 = c.alloc<double>(n);
                }
            }
            int* dflag = (int*)(c.dsc + 14);   // device scratch word
            hipLaunchKernelGGL(k_diag_copy, dim3(nblk(n)), dim3(256), 0, c.s, (const double*)l.Kinv, n, l.d0c);
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c.s, dflag, 0, (int64_t)1, (const int*)nullptr);
            // Z = L^-1 by row blocks on the second stream: row block I of L
            // is final once the diagonal step I is done (the panels of the
            // earlier steps wrote its band), and the factorization goes on
            // below it, so the DGEMMs of row I overlap steps I+1, ...
            // events: [0, NB) diagonal steps -> Z rows, [NB, 2 NB) panels ->
            // the rest of their windows, [2 NB, 3 NB) those rests done,
            // 3 NB and 3 NB + 1 the joins
            c.side_ready(3 * NB + 2);
            if (!l.ztmp) l.ztmp = c.alloc<double>(kCB * n);
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n * n)), dim3(256), 0, c.s, l.Z, 0.0, n * n,
                               (const int*)nullptr);
            // look-ahead: the update of step K is split into the next block
            // column (on s: all that the diagonal and panel steps K+1 read)
            // and the rest of the window (on s3, beside those steps). The
            // rest of step K writes blocks of columns >= K+2, so s waits for
            // it before the column update of step K+1, the next write there.
            for (int K = 0; K < NB; K++) {
                hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(64), 0, c.s, l.Kinv, n, K, l.Linv,
                                   (const double*)l.d0c, 1e-11, dflag);
                c.order(c.s, c.s2, K);
                zrow(c, l, K, wb, c.s2, c.blas2);
                const int nwin = std::min(wb, NB - 1 - K);
                if (nwin > 0) {
                    hipLaunchKernelGGL(k_chol_panel, dim3(nwin), dim3(256), 0, c.s, l.Kinv, n, K, NB, wb,
                                       (const double*)l.Linv);
                    if (nwin > 1) {
                        c.order(c.s, c.s3, NB + K);
                        hipLaunchKernelGGL(k_chol_update, dim3(nwin * (nwin - 1) / 2), dim3(256), 0, c.s3, l.Kinv, n,
                                           K, NB, wb, 2);
                    }
                }
                TOPO_HIP(hipEventRecord(c.evs[2 * NB + K], c.s3));
                if (nwin > 0) {
                    if (K > 0) TOPO_HIP(hipStreamWaitEvent(c.s, c.evs[2 * NB + K - 1], 0));
                    hipLaunchKernelGGL(k_chol_update, dim3(nwin), dim3(256), 0, c.s, l.Kinv, n, K, NB, wb, 1);
                }
            }
            c.order(c.s3, c.s, 3 * NB);
            c.order(c.s2, c.s, 3 * NB + 1);
            c.check_launch();
            TOPO_HIP(hipMemcpyAsync(&c.h_sc[0], dflag, sizeof(int), hipMemcpyDeviceToHost, c.s));
            c.sync();
            int flag = 0;
            std::memcpy(&flag, &c.h_sc[0], sizeof(int));
            if (!flag) {
                z_copies(c, l);
                done = true;
                l.tri = true;
            } else {
                coarse_block(c, l);   // the factorization overwrote the block
            }
        }
        run_overlap();
        if (!done) {
            // not positive definite (or too small a block to band): dense
            // Gauss-Jordan inverse with pivot dropping
            gauss_jordan(c, l, l.Kinv, n);
            if (l.Kinv32)
                hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(n * n)), dim3(256), 0, c.s, l.Kinv32,
                                   (const double*)l.Kinv, n * n);
            l.tri = false;
        }
        c.check_launch();
        l.chol = true;
        if (std::getenv("TOPO_VERBOSE"))
            std::fprintf(stderr, "[topo] coarsest level: %lld positions, %s\n", (long long)l.nfp,
                         l.tri ? (l.nd ? "nested-dissection Cholesky" : "banded Cholesky")
                               : "Gauss-Jordan inverse (factorization met a non-positive pivot)");
    }
    // the rounded copies of Z the FP32 V-cycle reads (FP16 with its row
    // scales, or FP32); both triangular products read the same matrix, the
    // second by column tiles, so no transposed copy is kept
    void z_copies(Ctx& c, MGLevel& l) {
        const int64_t n = l.nfp;
        if (l.Z16)
            hipLaunchKernelGGL(k_rowscale_f16, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.s, (const double*)l.Z, n,
                               l.Z16, l.zs2, (const int32_t*)l.rfirst);
        if (l.Z32)
            hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(n * n)), dim3(256), 0, c.s, l.Z32,
                               (const double*)l.Z, n * n);
    }

    // fronts, gather maps, descriptors and row starts of the
    // nested-dissection factorization (once per hierarchy)
    void nd_alloc(Ctx& c, MGLevel& l) {
        const int64_t n = l.nfp;
        int H = 0;
        for (const MGLevel::NDNode& t : l.ndn) H = std::max(H, t.height + 1);
        l.nd_h.assign(H, {});
        int64_t wmax = 1;
        std::vector<int32_t> rf(n, 0);
        for (size_t i = 0; i < l.ndn.size(); i++) {
            MGLevel::NDNode& t = l.ndn[i];
            l.nd_h[t.height].push_back((int)i);
            t.F = c.alloc<double>(t.f * t.f);
            const int64_t u = t.f - t.m;
            if (u > 0) t.Lb = c.alloc<double>(u * t.m);
            wmax = std::max(wmax, u * t.m);
            std::vector<int32_t> gi;
            for (int64_t k = 0; k < t.m; k++) gi.push_back((int32_t)(t.o0 + k));
            for (int p = t.parent; p >= 0; p = l.ndn[p].parent)
                for (int64_t k = 0; k < l.ndn[p].m; k++) gi.push_back((int32_t)(l.ndn[p].o0 + k));
            TOPO_CHECK((int64_t)gi.size() == t.f, TOPO_E_ARG, "front map size");
            t.g = c.alloc<int32_t>(t.f);
            TOPO_HIP(hipMemcpy(t.g, gi.data(), t.f * sizeof(int32_t), hipMemcpyHostToDevice));
            t.Linv = c.alloc<double>(t.m * kCB);
            for (int64_t k = 0; k < t.m; k++) rf[t.o0 + k] = (int32_t)t.s0;
        }
        l.ndW = c.alloc<double>(wmax);
        l.rfirst = c.alloc<int32_t>(n);
        TOPO_HIP(hipMemcpy(l.rfirst, rf.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        l.nd_fd.assign(H, nullptr);
        for (int h = 0; h < H; h++) {
            std::vector<FrontDev> fd;
            for (int id : l.nd_h[h]) {
                const MGLevel::NDNode& t = l.ndn[id];
                fd.push_back(FrontDev{t.F, t.f, t.Linv, l.d0c + t.o0, (int)(t.m / kCB)});
            }
            l.nd_fd[h] = (FrontDev*)c.alloc<char>(fd.size() * sizeof(FrontDev));
            TOPO_HIP(hipMemcpy(l.nd_fd[h], fd.data(), fd.size() * sizeof(FrontDev), hipMemcpyHostToDevice));
        }
    }

    // Multifrontal Cholesky of the block A (l.Kinv, nested-dissection order)
    // and Z = L^-1. Bottom-up by height, independent nodes together: the
    // front F = [own, ancestors] is A's rows/columns of the own DOFs plus the
    // children's update blocks; its own block is factored by the chain
    // kernels (L_DD), Z_DD = L_DD^-1 (rocBLAS trtri, into Z), the ancestor
    // rows L_aD = F_aD Z_DD^T and the update block F_aa -= L_aD L_aD^T (for
    // the parent's front). Then top-down (ancestors first): Z's blocks below
    // the diagonal, Z_PD = -sum_{Q on the path from P down to the parent of
    // D} Z_PQ (L_QD Z_DD), the rows of L on D's ancestor path being L_aD.
    // Every other block of Z is zero (Z_ij != 0 only for i an ancestor of
    // j), so each row of Z starts at its node's subtree.
    void nd_factor(Ctx& c, MGLevel& l, int* dflag) {
        const int64_t n = l.nfp;
        rocblas_handle h = c.blas_handle();
        for (size_t ht = 0; ht < l.nd_h.size(); ht++) {
            int maxNB = 0;
            for (int id : l.nd_h[ht]) {
                MGLevel::NDNode& t = l.ndn[id];
                const double* Cp[2] = {nullptr, nullptr};
                int64_t Cl[2] = {0, 0};
                for (int k = 0; k < 2; k++)
                    if (t.kid[k] >= 0) {
                        const MGLevel::NDNode& ch = l.ndn[t.kid[k]];
                        Cp[k] = ch.F + ch.m * ch.f + ch.m;
                        Cl[k] = ch.f;
                    }
                hipLaunchKernelGGL(k_front_assemble, dim3(nblk(t.f * t.f)), dim3(256), 0, c.s, t.F, t.f, t.m,
                                   (const int32_t*)t.g, (const double*)l.Kinv, n, Cp[0], Cl[0], Cp[1], Cl[1]);
                maxNB = std::max(maxNB, (int)(t.m / kCB));
            }
            const FrontDev* fd = l.nd_fd[ht];
            const unsigned nfr = (unsigned)l.nd_h[ht].size();
            for (int K = 0; K < maxNB; K++) {
                hipLaunchKernelGGL(k_chol_diag, dim3(1, nfr), dim3(64), 0, c.s, (double*)nullptr, (int64_t)0, K,
                                   (double*)nullptr, (const double*)nullptr, 1e-11, dflag, fd);
                const int nwin = maxNB - 1 - K;
                if (nwin > 0) {
                    hipLaunchKernelGGL(k_chol_panel, dim3(nwin, nfr), dim3(256), 0, c.s, (double*)nullptr,
                                       (int64_t)0, K, 0, 0, (const double*)nullptr, fd);
                    hipLaunchKernelGGL(k_chol_update, dim3(nwin * (nwin + 1) / 2, nfr), dim3(256), 0, c.s,
                                       (double*)nullptr, (int64_t)0, K, 0, 0, 0, fd);
                }
            }
            for (int id : l.nd_h[ht]) {
                MGLevel::NDNode& t = l.ndn[id];
                double* Zd = l.Z + t.o0 * n + t.o0;
                // row-major lower L_DD = column-major upper
                TOPO_CHECK(rocblas_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, (rocblas_int)t.m, t.F,
                                          (rocblas_int)t.f, Zd, (rocblas_int)n) == rocblas_status_success,
                           TOPO_E_HIP, "rocblas_dtrtri failed");
                const int64_t u = t.f - t.m;
                // plain GEMMs: faster here than rocBLAS's triangular (trmm) and
                // symmetric (syrk) forms at half the flops
                if (u > 0) {
                    gemm_rm(h, false, true, u, t.m, t.m, 1.0, t.F + t.m * t.f, t.f, Zd, n, 0.0, t.Lb, t.m);
                    gemm_rm(h, false, true, u, u, t.m, -1.0, t.Lb, t.m, t.Lb, t.m, 1.0, t.F + t.m * t.f + t.m, t.f);
                }
            }
        }
        for (int id = (int)l.ndn.size() - 1; id >= 0; id--) {
            const MGLevel::NDNode& t = l.ndn[id];
            const int64_t u = t.f - t.m;
            if (u == 0) continue;
            const double* Zd = l.Z + t.o0 * n + t.o0;
            // W = L_aD Z_DD (rows in the order of the ancestor path)
            gemm_rm(h, false, false, u, t.m, t.m, 1.0, t.Lb, t.m, Zd, n, 0.0, l.ndW, t.m);
            std::vector<int> path;
            std::vector<int64_t> offs;
            int64_t off = 0;
            for (int p = t.parent; p >= 0; p = l.ndn[p].parent) {
                path.push_back(p);
                offs.push_back(off);
                off += l.ndn[p].m;
            }
            for (size_t i = 0; i < path.size(); i++) {
                const MGLevel::NDNode& P = l.ndn[path[i]];
                double* C = l.Z + P.o0 * n + t.o0;
                // diagonal term first (Z_PP lower triangular, zero above)
                gemm_rm(h, false, false, P.m, t.m, P.m, -1.0, l.Z + P.o0 * n + P.o0, n, l.ndW + offs[i] * t.m, t.m,
                        0.0, C, n);
                for (size_t j = 0; j < i; j++) {
                    const MGLevel::NDNode& Q = l.ndn[path[j]];
                    gemm_rm(h, false, false, P.m, t.m, Q.m, -1.0, l.Z + P.o0 * n + Q.o0, n, l.ndW + offs[j] * t.m,
                            t.m, 1.0, C, n);
                }
            }
        }
    }

    // Row block I of Z = L^-1: Z_II = L_II^-1 and, for the columns left of
    // the diagonal block, Z_I,0:I = -L_II^-1 (L_I,band Z_band,0:I), band = the
    // wb row blocks above I. Both products are plain rocBLAS DGEMMs on
    // row-major sub-matrices (a row-major C = A B is the column-major
    // C^T = B^T A^T). ztmp is scratch of 64 x n; the rows above I must be
    // done on stream `st` before.
    void zrow(Ctx& c, MGLevel& l, int I, int wb, hipStream_t st, rocblas_handle h) {
        const int64_t n = l.nfp;
        {
            // Z_II
            hipLaunchKernelGGL(k_copy_block, dim3(1), dim3(256), 0, st, (const double*)l.Linv + (int64_t)I * kCB * kCB,
                               (int64_t)kCB, l.Z + (int64_t)I * kCB * n + (int64_t)I * kCB, n);
            if (I == 0) return;
            const int K0 = std::max(0, I - wb);
            const int kk = (I - K0) * kCB;     // inner dimension
            const int nc = I * kCB;             // columns 0 .. I*64
            // T (64 x nc, ld n) = L[I, K0:I] (64 x kk, ld n) * Z[K0:I, 0:nc] (kk x nc, ld n)
            const double one = 1.0, zero = 0.0, mone = -1.0;
            const double* Lrow = l.Kinv + (int64_t)I * kCB * n + (int64_t)K0 * kCB;
            const double* Zb = l.Z + (int64_t)K0 * kCB * n;
            TOPO_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, kCB, kk, &one, Zb,
                                     (rocblas_int)n, Lrow, (rocblas_int)n, &zero, l.ztmp, (rocblas_int)n) ==
                           rocblas_status_success, TOPO_E_HIP, "rocblas_dgemm failed");
            // Z[I, 0:nc] = -Linv_I (64 x 64) * T (64 x nc)
            TOPO_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, kCB, kCB, &mone,
                                     l.ztmp, (rocblas_int)n, l.Linv + (int64_t)I * kCB * kCB, kCB, &zero,
                                     l.Z + (int64_t)I * kCB * n, (rocblas_int)n) == rocblas_status_success,
                       TOPO_E_HIP, "rocblas_dgemm failed");
        }
    }

    // in-place inverse of the n x n block A (row-major, ld n) by blocked
    // Gauss-Jordan with pivot dropping
    // the scratch (original diagonal, panel copies) is allocated on the
    // level's first fallback and kept for later ones
    void gauss_jordan(Ctx& c, MGLevel& l, double* A, int64_t nf) {
        if (!l.gj_d0) l.gj_d0 = c.alloc<double>(nf);
        double* d0 = l.gj_d0;
        hipLaunchKernelGGL(k_diag_copy, dim3(nblk(nf)), dim3(256), 0, c.s, (const double*)A, nf, d0);
        const int nt = (int)((nf + kGJB - 1) / kGJB);
        // large blocks: the rank-64 update as a plain rocBLAS DGEMM
        const bool use_blas = nf > 1024;
        if (use_blas && !l.gj_c) {
            l.gj_c = c.alloc<double>(nf * kGJB);
            l.gj_r = c.alloc<double>(nf * kGJB);
        }
        double* Cp = use_blas ? l.gj_c : nullptr;
        double* Rp = use_blas ? l.gj_r : nullptr;
        for (int64_t k0 = 0; k0 < nf; k0 += kGJB) {
            const int kb = (int)std::min<int64_t>(kGJB, nf - k0);
            hipLaunchKernelGGL(k_bgj_diag, dim3(1), dim3(256), 0, c.s, A, nf, k0, kb, (const double*)d0);
            hipLaunchKernelGGL(k_bgj_row, dim3(nt), dim3(256), 0, c.s, A, nf, k0, kb);
            if (use_blas) {
                hipLaunchKernelGGL(k_bgj_panels, dim3(nblk(nf * kb)), dim3(256), 0, c.s, (const double*)A, nf, k0,
                                   kb, Cp, Rp);
                // row-major A -= C R  <=>  column-major A^T -= R^T C^T
                const double alpha = -1.0, beta = 1.0;
                TOPO_CHECK(rocblas_dgemm(c.blas_handle(), rocblas_operation_none, rocblas_operation_none, (rocblas_int)nf,
                                         (rocblas_int)nf, kb, &alpha, Rp, (rocblas_int)nf, Cp, kb, &beta, A,
                                         (rocblas_int)nf) == rocblas_status_success,
                           TOPO_E_HIP, "rocblas_dgemm failed");
            } else {
                hipLaunchKernelGGL(k_bgj_update, dim3(nt, nt), dim3(256), 0, c.s, A, nf, k0, kb);
            }
            hipLaunchKernelGGL(k_bgj_col, dim3(nt), dim3(256), 0, c.s, A, nf, k0, kb);
        }
        c.sync();
    }

    // Box partition of the structured grid (KSPSolver._buildSubdomains
    // js/optimizer-worker-3d.js:1155-1225): subdomain counts per axis, the
    // element -> subdomain maps and the corner DOF flags (:1230-1254).
    void bddc_setup(Ctx& c) {
        MGLevel& l = L[0];
        const Grid& g = l.g;
        const double tot = (double)g.nelx * g.nely * g.nelz;
        const double target = std::max(8.0, std::min(64.0, std::ceil(std::cbrt(tot / 100))));
        auto count = [&](int n, int a, int b) {
            const double v = std::round(std::cbrt(target * ((double)n * n) / ((double)a * b)));
            return (int)std::max(2.0, std::min((double)n, v));
        };
        bddc_ns[0] = count(g.nelx, g.nely, g.nelz);
        bddc_ns[1] = count(g.nely, g.nelx, g.nelz);
        bddc_ns[2] = count(g.nelz, g.nelx, g.nely);
        const int ne[3] = {g.nelx, g.nely, g.nelz};
        int32_t** sub[3] = {&bddc_sx, &bddc_sy, &bddc_sz};
        uint8_t* bnd[3];
        for (int a = 0; a < 3; a++) {
            const int n = ne[a], ns = bddc_ns[a];
            std::vector<int32_t> hs(n);
            std::vector<uint8_t> hb(n + 1, 0);
            for (int k = 0; k < ns; k++) {
                // floor(k n / ns) .. floor((k+1) n / ns) (:1172-1181)
                const int a0 = (int)((int64_t)k * n / ns), a1 = (int)((int64_t)(k + 1) * n / ns);
                for (int e = a0; e < a1; e++) hs[e] = k;
                hb[a0] = 1;
                hb[a1] = 1;
            }
            *sub[a] = c.alloc<int32_t>(n);
            TOPO_HIP(hipMemcpy(*sub[a], hs.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
            bnd[a] = c.alloc<uint8_t>(n + 1);
            TOPO_HIP(hipMemcpy(bnd[a], hb.data(), n + 1, hipMemcpyHostToDevice));
        }
        bddc_val = c.alloc<double>((int64_t)kBddcSlots * l.ndof);
        bddc_cnt = c.alloc<uint8_t>(l.ndof);
        bddc_corner = c.alloc<uint8_t>(l.ndof);
        hipLaunchKernelGGL(k_bddc_corner_flags, dim3(nblk(g.nnodes())), dim3(256), 0, c.s, g, (const uint8_t*)bnd[0],
                           (const uint8_t*)bnd[1], (const uint8_t*)bnd[2], bddc_corner);
        c.sync();
        for (int a = 0; a < 3; a++) c.free_one(bnd[a]);
        bddc = true;
    }
    // local inverse diagonals of the current moduli (_updateBDDC :1231-1278)
    void bddc_build(Ctx& c) {
        MGLevel& l = L[0];
        hipLaunchKernelGGL(k_bddc_build, dim3(nblk(l.g.nnodes())), dim3(256), 0, c.s, (const double*)l.E,
                           (const uint8_t*)l.fixed, l.g, (const int32_t*)bddc_sx, (const int32_t*)bddc_sy,
                           (const int32_t*)bddc_sz, bddc_ns[0], bddc_ns[1], (const double*)K0->ke, bddc_val,
                           bddc_cnt);
        c.check_launch();
    }
    // z = M^-1 r (bddcApply assembly/index.ts:1272-1371), FP64 on level 0;
    // returns the buffer holding z (zero at the fixed DOFs)
    double* bddc_apply(Ctx& c, const double* r, const int* gate) {
        MGLevel& l = L[0];
        double* z = l.x;
        double* z2 = l.x2;
        hipLaunchKernelGGL(k_bddc_local, dim3(nblk(l.ndof)), dim3(256), 0, c.s, r, z, (const double*)bddc_val,
                           (const uint8_t*)bddc_cnt, l.ndof, gate);
        for (int s = 1; s < bddc_sweeps; s++) {
            apply<double>(c, 0, z, z2, 3, r, gate);
            std::swap(z, z2);
        }
        apply<double>(c, 0, z, l.res, 0, nullptr, gate);
        hipLaunchKernelGGL(k_bddc_corner, dim3(nblk(l.ndof)), dim3(256), 0, c.s, z, r, (const double*)l.res,
                           (const double*)l.invD, (const uint8_t*)bddc_corner, (const uint8_t*)l.fixed, l.ndof, gate);
        return z;
    }

    // x-march workgroup size of level li's operator in precision T
    template <typename T>
    int xm_pick_nt(int li) const {
        const int H = L[li].g.nnz() + 1;
        const int nt_sel = xm_nt ? xm_nt : (li < (int)lv_nt.size() ? lv_nt[li] : 0);
        if (nt_sel == 256 && 2 * H <= 256) return 256;
        return (nt_sel == 1024 || (nt_sel == 0 && 4 * H > 512)) ? 1024 : 512;
    }

    // the size rule for the x-planes per workgroup of level li's x-march
    // operator with NT-thread workgroups (before clamping to 1..32): as many
    // x-chunks per yz tile as fit the chip in one round of workgroups at
    // four waves per SIMD (ncu * 1024 / NT of them), the planes split evenly
    int xm_auto_cx(int li, int NT) const {
        const Grid& g = L[li].g;
        const int H = g.nnz() + 1;
        const int gx = (g.nny() * g.nnz() + (NT - H) - 1) / (NT - H);
        const int gy = std::max(1, (int)((int64_t)ncu * 1024 / NT / gx));
        const int cx = (g.nnx() + gy - 1) / gy;
        return std::max(1, std::min(32, cx));
    }

    // x-planes per workgroup of level li's x-march operator in precision T (0 auto)
    template <typename T>
    int xm_pick_cx(int li) const {
        if (sizeof(T) == 8 && xm_cx64) return xm_cx64;
        if (xm_cx) return xm_cx;
        if (sizeof(T) == 8 && li < (int)lv_cx64.size() && lv_cx64[li]) return lv_cx64[li];
        return li < (int)lv_cx.size() ? lv_cx[li] : 0;
    }

    // whether level li's operator in precision T is the x-march kernel (mode 4 needs it)
    template <typename T>
    bool xm_level(int li) const {
        const MGLevel& l = L[li];
        if (l.dense || exact || !K0->had_ok) return false;
        return 2 * (l.g.nnz() + 1) <= xm_pick_nt<T>(li);
    }

    // out = A in (mode 0), masked (1), b - A in (2), Jacobi sweep (3), the
    // first two sweeps from x = 0 (4: `in` unused; x-march levels only), a
    // Jacobi sweep that also leaves the r.z partials of its output with r =
    // rdot (5). With pap (mode 1 or 5, x-march level, no springs) the launch
    // leaves its partials there and returns their count (0 when it cannot:
    // then mode 5 runs as mode 3).
    template <typename T>
    int apply(Ctx& c, int li, const T* in, T* out, int mode, const T* b, const int* gate, int nr = 1,
              double* pap = nullptr, const double* rdot = nullptr) {
        int npart = 0;
        MGLevel& l = L[li];
        if (mode == 5 && !(rdot && pap && xm_level<T>(li) && !(li == 0 && springs.n > 0))) mode = 3;
        const double* om = mode >= 3 ? omp(li) : nullptr;
        TOPO_CHECK(mode != 4 || (xm_level<T>(li) && !(li == 0 && springs.n > 0)), TOPO_E_UNSUPPORTED,
                   "fused first sweep on a level without the x-march operator");
        if (l.dense) {
            hipLaunchKernelGGL(k_dense_apply<T>, dim3(nblk(l.ndof * 64), nr), dim3(256), 0, c.s, Sel<T>::K(l), in, out, l.ndof,
                               l.fixed, mode, b, Sel<T>::invD(l), om, gate);
        } else {
            typedef typename Sel<T>::TE TE;
            const int H = l.g.nnz() + 1;
            const int cx_sel = xm_pick_cx<T>(li);
            const int NT = xm_pick_nt<T>(li);
            if (!exact && K0->had_ok && 2 * H <= NT) {
                const int nyz = l.g.nny() * l.g.nnz();
                const int gx = (nyz + (NT - H) - 1) / (NT - H);
                int CX = cx_sel > 0 ? cx_sel : xm_auto_cx(li, NT);
                CX = std::max(1, std::min(32, CX));
                // CG partials: one per (yz tile, node plane), whatever CX
                const int64_t nplane = (int64_t)gx * l.g.nnx();
                double* part = ((mode == 1 || mode == 5) && pap && nplane <= kCgStride &&
                                !(li == 0 && springs.n > 0)) ? pap : nullptr;
                if (mode == 5 && !part) mode = 3;
                if (part) npart = (int)nplane;
                const int gy = (l.g.nnx() + CX - 1) / CX;
                const dim3 grid(8 * ((gx * gy + 7) / 8), nr);
                bool on_mfma = false;
                if constexpr (std::is_same<T, float>::value) {
                    if (mfma && NT == 512) {
                        on_mfma = true;
                        const size_t lds = mf_lds_bytes(l.g.nnz(), 512);
#define TOPO_MF_LAUNCH(MV)                                                                                        \
    hipLaunchKernelGGL((k_applyA_mf<512, MV>), grid, dim3(512), lds, c.s, in, out, (const float*)Sel<T>::E(l), thr, \
                       l.g, (const uint8_t*)l.fixed, b, (const float*)Sel<T>::invD(l), om, gate, (const uint4*)K0->kf, \
                       CX, gx, gy, part, rdot)
                        switch (mode) {
                        case 0: TOPO_MF_LAUNCH(0); break;
                        case 1: TOPO_MF_LAUNCH(1); break;
                        case 2: TOPO_MF_LAUNCH(2); break;
                        case 3: TOPO_MF_LAUNCH(3); break;
                        case 5: TOPO_MF_LAUNCH(5); break;
                        default: TOPO_MF_LAUNCH(4); break;
                        }
#undef TOPO_MF_LAUNCH
                    }
                }
                if (!on_mfma) {
                const size_t lds = xm_lds_bytes(l.g.nnz(), sizeof(T), NT);
#define TOPO_XM_LAUNCH(NTV, MV)                                                                                  \
    hipLaunchKernelGGL((k_applyA_xm<T, TE, NTV, MV>), grid, dim3(NTV), lds, c.s, in, out,                      \
                       (const TE*)Sel<T>::E(l), thr, l.g, (const uint8_t*)l.fixed, b, (const T*)Sel<T>::invD(l), \
                       om, gate, had<T>(), CX, gx, gy, part, rdot)
#define TOPO_XM_MODES(NTV)                                 \
    switch (mode) {                                        \
    case 0: TOPO_XM_LAUNCH(NTV, 0); break;                 \
    case 1: TOPO_XM_LAUNCH(NTV, 1); break;                 \
    case 2: TOPO_XM_LAUNCH(NTV, 2); break;                 \
    case 3: TOPO_XM_LAUNCH(NTV, 3); break;                 \
    case 5: TOPO_XM_LAUNCH(NTV, 5); break;                 \
    default: TOPO_XM_LAUNCH(NTV, 4); break;                \
    }
                if (NT == 512) {
                    TOPO_XM_MODES(512)
                } else if (NT == 256) {
                    TOPO_XM_MODES(256)
                } else {
                    TOPO_XM_MODES(1024)
                }
#undef TOPO_XM_MODES
#undef TOPO_XM_LAUNCH
                }
            } else {
                hipLaunchKernelGGL((k_applyA_exact<T, TE>), dim3(nblk(l.g.nnodes()), nr), dim3(256), 0, c.s, in, out,
                                   (const TE*)Sel<T>::E(l), thr, l.g, l.fixed, mode, b, Sel<T>::invD(l), om, gate,
                                   ke<T>());
            }
        }
        if (li == 0 && springs.n > 0)
            hipLaunchKernelGGL(k_spring<T>, dim3(1, nr), dim3(64), 0, c.s, in, out, springs, mode, (const uint8_t*)l.fixed,
                               (const T*)Sel<T>::invD(l), om, gate, l.ndof);
        return npart;
    }

    // one V-cycle (mgVcycle assembly/index.ts:837-1048) for A x = b from
    // x = 0 on level li; returns the buffer holding x. With rdot (level 0)
    // the last post-smoothing sweep also leaves the partials of r.z, r =
    // rdot and z the V-cycle's output, in rzpart and their count in *nrz (0
    // when it could not).
    template <typename T>
    T* vcycle(Ctx& c, int li, const T* b, const int* gate, int nr = 1, const double* rdot = nullptr,
              double* rzpart = nullptr, int* nrz = nullptr, bool b_in_bf = false) {
        MGLevel& l = L[li];
        T* x = Sel<T>::x(l);
        T* x2 = Sel<T>::x2(l);
        const bool coarsest = li == (int)L.size() - 1;
        if (coarsest && l.chol) {
            // the FP16 coarse copy works on FP32 vectors (float views of bf,
            // ytmp); the gather also zeroes x (b is another buffer)
            const bool f16 = std::is_same<T, float>::value && l.Z16 && l.tri;
            const int gn = nblk(std::max(l.nfp, l.ndof));
            // (b_in_bf: the restriction wrote bf itself, and x is zero at
            // the fixed DOFs from the build)
            if (b_in_bf) {
            } else if (f16)
                hipLaunchKernelGGL((k_gather_free<T, float>), dim3(gn, nr), dim3(256), 0, c.s, b, l.freeIdx, l.nf,
                                   l.nfp, (float*)l.bf, gate, l.ndof, x);
            else
                hipLaunchKernelGGL((k_gather_free<T, double>), dim3(gn, nr), dim3(256), 0, c.s, b, l.freeIdx, l.nf,
                                   l.nfp, l.bf, gate, l.ndof, x);
            const dim3 gi((unsigned)((l.nf + 15) / 16));
            const bool f32 = std::is_same<T, float>::value && l.Kinv32;
            if (l.tri) {
                // x = Z^T (Z b): lower-triangular Z into ytmp, then the upper
                // triangular Z^T scattered to the free DOFs; the FP32 V-cycle
                // reads the rounded copies (coarse_fp32: FP32, or FP16 with
                // the row scales applied to ytmp)
                if (f16) {
                    // S^2 Zq b by row chunks, then Zq^T y by column tiles of
                    // the same matrix (just read: the second pass finds much
                    // of it in the memory-side cache)
                    // (the column tiles sum the row chunks of S^2 Zq b themselves)
                    const bool yp = tri_product<T, _Float16, float, float>(c, l, l.Z16, 0, (const float*)l.bf,
                                                                           (float*)l.ytmp, (T*)nullptr, gate, l.zs2,
                                                                           nr, false);
                    TOPO_NR_SWITCH(nr, (tri_colpart<NR, float, _Float16>(c.s, (const _Float16*)l.Z16, l.nfp,
                                                                         (const float*)l.ytmp, l.tpart, l.tnrb, gate,
                                                                         (const int32_t*)l.rfirst, tc_nw,
                                                                         yp ? l.tpart1 : nullptr, l.tnch, l.zs2)));
                    TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_sum<T, NR, float>), dim3(nblk(4 * l.nfp)), dim3(256),
                                                          0, c.s, (const double*)l.tpart, l.tnrb, l.nfp,
                                                          (float*)nullptr, x, (const int32_t*)l.freeIdx, l.nf, l.ndof,
                                                          gate, (const double*)nullptr));
                } else if (f32 && l.Z32) {
                    // Z b by rows, Z^T y by column tiles of the same FP32 copy
                    const bool yp = tri_product<T, float, double, double>(c, l, l.Z32, 0, (const double*)l.bf,
                                                                          l.ytmp, (T*)nullptr, gate, nullptr, nr,
                                                                          false);
                    TOPO_NR_SWITCH(nr, (tri_colpart<NR, double, float>(c.s, (const float*)l.Z32, l.nfp,
                                                                        (const double*)l.ytmp, l.tpart, l.tnrb, gate,
                                                                        (const int32_t*)l.rfirst, tc_nw,
                                                                        yp ? l.tpart1 : nullptr, l.tnch, nullptr)));
                    TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_sum<T, NR, double>), dim3(nblk(4 * l.nfp)), dim3(256),
                                                          0, c.s, (const double*)l.tpart, l.tnrb, l.nfp,
                                                          (double*)nullptr, x, (const int32_t*)l.freeIdx, l.nf, l.ndof,
                                                          gate, (const double*)nullptr));
                } else {
                    // FP64: Z b by rows, Z^T y by column tiles of Z (FP64 sums)
                    const bool yp = tri_product<T, double, double, double>(c, l, l.Z, 0, (const double*)l.bf,
                                                                           l.ytmp, (T*)nullptr, gate, nullptr, nr,
                                                                           false);
                    TOPO_NR_SWITCH(nr, (tri_colpart<NR, double, double>(c.s, (const double*)l.Z, l.nfp,
                                                                         (const double*)l.ytmp, l.tpart, l.tnrb, gate,
                                                                         (const int32_t*)l.rfirst, tc_nw,
                                                                         yp ? l.tpart1 : nullptr, l.tnch, nullptr)));
                    TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_tri_sum<T, NR, double>), dim3(nblk(4 * l.nfp)), dim3(256),
                                                          0, c.s, (const double*)l.tpart, l.tnrb, l.nfp,
                                                          (double*)nullptr, x, (const int32_t*)l.freeIdx, l.nf, l.ndof,
                                                          gate, (const double*)nullptr));
                }
                return x;
            }
            // the FP32 V-cycle reads the FP32 rounding of the FP64 inverse
            // (half the bytes) when coarse_fp32 is set
            if (f32) {
                const float* Ki = l.Kinv32;
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_inv_apply<T, NR, float>), gi, dim3(256), 0, c.s, Ki, l.freeIdx,
                                                      l.nf, l.nfp, (const double*)l.bf, x, gate, l.ndof));
            } else {
                const double* Ki = l.Kinv;
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_inv_apply<T, NR, double>), gi, dim3(256), 0, c.s, Ki, l.freeIdx,
                                                      l.nf, l.nfp, (const double*)l.bf, x, gate, l.ndof));
            }
            return x;
        }
        const int nsweep_pre = coarsest ? nu1 + coarse_sweeps : nu1;
        int s0 = 1;
        if (nsweep_pre >= 2 && xm_level<T>(li) && !(li == 0 && springs.n > 0)) {
            // the first two sweeps in one pass (x0 = omega invD b formed on the fly)
            apply<T>(c, li, (const T*)b, x2, 4, b, gate, nr);
            std::swap(x, x2);
            s0 = 2;
        } else if (nsweep_pre > 0) {
            hipLaunchKernelGGL(k_jacobi_first<T>, dim3(nblk(l.ndof), nr), dim3(256), 0, c.s, x, b,
                               (const T*)Sel<T>::invD(l), omp(li), l.ndof, gate);
        } else {
            hipLaunchKernelGGL(k_fill<T>, dim3(nblk(l.ndof), nr), dim3(256), 0, c.s, x, T(0), l.ndof, gate);
        }
        for (int s = s0; s < nsweep_pre; s++) {
            apply<T>(c, li, x, x2, 3, b, gate, nr);
            std::swap(x, x2);
        }
        if (coarsest) return x;
        apply<T>(c, li, x, Sel<T>::res(l), 2, b, gate, nr);
        MGLevel& lc = L[li + 1];
        // into the coarsest direct solve's gathered right-hand side when the
        // next level is that solve (its gather kernel is then not needed)
        const bool to_bf = bf_direct && li + 1 == (int)L.size() - 1 && lc.chol && lc.bf && lc.freeOf;
        if (to_bf && std::is_same<T, float>::value && lc.Z16 && lc.tri)
            hipLaunchKernelGGL((k_restrict<T, float>), dim3(nblk8(lc.g.nnodes()), nr), dim3(256), 0, c.s,
                               (const T*)Sel<T>::res(l), Sel<T>::b(lc), l.g, lc.g, lc.fixed, gate, (float*)lc.bf,
                               (const int32_t*)lc.freeOf, lc.nfp);
        else if (to_bf)
            hipLaunchKernelGGL((k_restrict<T, double>), dim3(nblk8(lc.g.nnodes()), nr), dim3(256), 0, c.s,
                               (const T*)Sel<T>::res(l), Sel<T>::b(lc), l.g, lc.g, lc.fixed, gate, lc.bf,
                               (const int32_t*)lc.freeOf, lc.nfp);
        else
            hipLaunchKernelGGL(k_restrict<T>, dim3(nblk8(lc.g.nnodes()), nr), dim3(256), 0, c.s,
                               (const T*)Sel<T>::res(l), Sel<T>::b(lc), l.g, lc.g, lc.fixed, gate);
        T* xc = vcycle<T>(c, li + 1, Sel<T>::b(lc), gate, nr, nullptr, nullptr, nullptr, to_bf);
        hipLaunchKernelGGL(k_prolong_add<T>, dim3(nblk8(lc.g.nnodes()), nr), dim3(256), 0, c.s, x, (const T*)xc, l.g,
                           lc.g, l.fixed, gate);
        for (int s = 0; s < nu2; s++) {
            const bool last = rdot && li == 0 && s == nu2 - 1;
            const int np = apply<T>(c, li, x, x2, last ? 5 : 3, b, gate, nr, last ? rzpart : nullptr,
                                    last ? rdot : nullptr);
            if (last && nrz) *nrz = np;
            std::swap(x, x2);
        }
        return x;
    }
};

template <> const double* MGH::ke<double>() const { return K0->ke; }
template <> const float* MGH::ke<float>() const { return K0->ke32; }
template <> HadCoef<double> MGH::had<double>() const { return K0->hc64; }
template <> HadCoef<float> MGH::had<float>() const { return K0->hc32; }

// ---------------------------------------------------------------------------
// CG drivers. Host polls the device state every `check` iterations; update
// kernels are no-ops once converged, so the iterate equals the one a
// per-iteration test would stop at.
struct CGVecs {
    void *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr;
};

struct CGResult {
    int iters = 0;
    int nonfinite = 0;
    int vcycles = 0;
    int applyA = 0;
    double relres = 0;   // largest final relative residual of the solves summed here
    double rn = 0;       // final ||r||^2 (Jacobi-PCG)
};

static void cg_set(Ctx& c, double tol, int maxit, double guard = 1e-12, int nr = 1, int abs_tol = 0) {
    CGState h;
    std::memset(&h, 0, sizeof h);
    h.tol = tol;
    h.maxit = maxit;
    h.guard = guard;
    h.abs_tol = abs_tol;
    for (int q = 0; q < nr; q++) c.h_st[q] = h;
    TOPO_HIP(hipMemcpyAsync(c.st, c.h_st, nr * sizeof(CGState), hipMemcpyHostToDevice, c.s));
}



// Jacobi-PCG (ebePCG assembly/index.ts:504-733 in full space with fixed DOFs
// masked) over an operator `op(p, out, mode, b, gate)`: mode 1 out = A p with
// fixed DOFs zeroed, mode 2 out = b - A p at free DOFs (0 at fixed). U is the
// warm start (TU precision), rhs in T. abs_tol: stop at ||r|| < tol instead
// of ||r|| < tol ||r0||. key identifies the operator for the graph cache.
template <typename T, typename TU, typename Op>
static CGResult pcg_jacobi_op(Ctx& c, Op&& op, const uint8_t* fixed, const T* invD, int64_t n, const T* rhs, TU* U,
                              bool warm, CGVecs& v, double tol, int maxit, volatile int* cancel,
                              std::vector<const void*> key, int jcg_check, int abs_tol = 0) {
    T* r = (T*)v.r;
    T* z = (T*)v.z;
    T* p = (T*)v.p;
    T* Ap = (T*)v.Ap;
    CGResult res;
    if (warm) {
        // r = rhs - A U on free DOFs (U in T precision for the operator)
        if (std::is_same<T, TU>::value) {
            op((const T*)U, r, 2, rhs, (const int*)nullptr, (double*)nullptr);
        } else {
            hipLaunchKernelGGL((k_convert<T, TU>), dim3(nblk(n)), dim3(256), 0, c.s, z, (const TU*)U, n);
            op((const T*)z, r, 2, rhs, (const int*)nullptr, (double*)nullptr);
        }
        res.applyA++;
    } else {
        hipLaunchKernelGGL(k_fill<TU>, dim3(nblk(n)), dim3(256), 0, c.s, U, TU(0), n, (const int*)nullptr);
        hipLaunchKernelGGL((k_convert<T, T>), dim3(nblk(n)), dim3(256), 0, c.s, r, rhs, n);
        hipLaunchKernelGGL(k_mask<T>, dim3(nblk(n)), dim3(256), 0, c.s, r, fixed, n);
    }
    hipLaunchKernelGGL(k_jacobi_apply<T>, dim3(nblk(n)), dim3(256), 0, c.s, z, (const T*)r, invD, n);
    hipLaunchKernelGGL((k_convert<T, T>), dim3(nblk(n)), dim3(256), 0, c.s, p, (const T*)z, n);
    cg_set(c, tol, maxit, 1e-12, 1, abs_tol);
    hipLaunchKernelGGL(k_cg_start<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, c.s, (const T*)r, (const T*)z, n, c.st,
                       c.partials, c.ticket);
    c.check_launch();
    const int* gate = &c.st->done;
    // iterations per captured graph between host polls (converged ones are
    // gated no-ops)
    const int check = std::max(1, std::min(64, jcg_check));
    // z = invD r is never stored: the residual update forms it for r.z,
    // the p update again (same operands, same bits); U += alpha p rides on
    // the p update, which reads p anyway; the operator may leave the p.Ap
    // partials itself
    auto body = [&]() {
        for (int k = 0; k < check; k++) {
            int nb0 = op((const T*)p, Ap, 1, (const T*)nullptr, gate, c.partials);
            if (nb0 == 0) {
                hipLaunchKernelGGL((k_cgp_pAp<T, 1>), dim3(kCgBlocks), dim3(kRedThreads), 0, c.s, (const T*)p,
                                   (const T*)Ap, n, c.st, c.partials);
                nb0 = kCgBlocks;
            }
            hipLaunchKernelGGL((k_cgp_update_rj<T>), dim3(kCgBlocks), dim3(kRedThreads), 0, c.s, r, (const T*)Ap, invD,
                               n, c.st, c.partials, nb0);
            hipLaunchKernelGGL((k_cgp_pupdate_rj<T, TU>), dim3(kCgBlocks), dim3(kRedThreads), 0, c.s, p, (const T*)r,
                               invD, U, n, c.st, (const double*)c.partials, nb0);
            hipLaunchKernelGGL(k_cgp_finish<1>, dim3(1), dim3(kRedThreads), 0, c.s, c.st, (const double*)c.partials,
                               nb0);
        }
    };
    hipGraphExec_t ex = nullptr;
    if (c.use_graphs)
    {
        key.insert(key.begin(), {(const void*)&pcg_jacobi_op<T, TU, Op>, U, r, z, p, Ap, rhs});
        ex = c.graph_capture(key, body);
    }
    for (;;) {
        const CGState& h = c.read_st();
        if (h.done) {
            res.iters = h.iters;
            res.relres = std::sqrt(h.rn / (h.r0 > 1e-300 ? h.r0 : 1e-300));
            res.rn = h.rn;
            res.nonfinite = h.nonfinite;
            break;
        }
        if (cancel && *cancel) throw TopoError{TOPO_E_CANCELLED};
        if (ex) TOPO_HIP(hipGraphLaunch(ex, c.s));
        else body();
        c.check_launch();
    }
    res.applyA += res.iters;
    return res;
}

// Jacobi-PCG on the fine level of a hierarchy (its level-0 operator)
template <typename T, typename TU>
static CGResult pcg_jacobi(Ctx& c, MGH& mg, const T* rhs, TU* U, bool warm, CGVecs& v, double tol, int maxit,
                           volatile int* cancel) {
    MGLevel& l = mg.L[0];
    // (the p.Ap partials of the 3-D operator are left to k_cgp_pAp: 0)
    auto op = [&](const T* p, T* out, int mode, const T* b, const int* gate, double*) {
        mg.apply<T>(c, 0, p, out, mode, b, gate);
        return 0;
    };
    std::vector<const void*> key{&mg};
    mg.graph_sig(key);
    return pcg_jacobi_op<T, TU>(c, op, l.fixed, (const T*)Sel<T>::invD(l), l.ndof, rhs, U, warm, v, tol, maxit, cancel,
                                key, mg.jcg_check);
}

// The V-cycle part of precondition() without the widening copy: returns the
// V-cycle's own output buffer (precision PT, fixed DOFs not yet zeroed). With
// r_lo_ready the FP32 residual has already been written (k_cg_update_ur).
template <typename T, typename PT>
static const PT* precondition_raw(Ctx& c, MGH& mg, const T* r, CGVecs& v32, const int* gate, int nr,
                                  bool r_lo_ready, double* rzpart = nullptr, int* nrz = nullptr) {
    const int64_t n = mg.L[0].ndof;
    // r.z partials from the V-cycle's last sweep need the FP64 residual
    const double* rdot = nullptr;
    if constexpr (std::is_same<T, double>::value) rdot = rzpart ? r : nullptr;
    if constexpr (std::is_same<T, PT>::value) {
        (void)n; (void)v32; (void)r_lo_ready;
        if constexpr (std::is_same<T, double>::value)
            if (mg.bddc) return mg.bddc_apply(c, r, gate);
        TOPO_CHECK(!mg.bddc, TOPO_E_UNSUPPORTED, "the domain-decomposition preconditioner is FP64 only");
        return mg.vcycle<T>(c, 0, r, gate, nr, rdot, rzpart, nrz);
    } else {
        PT* r32 = (PT*)v32.r;
        if (!r_lo_ready)
            hipLaunchKernelGGL((k_convert_gated<PT, T>), dim3(nblk(n), nr), dim3(256), 0, c.s, r32, r, n, gate);
        return mg.vcycle<PT>(c, 0, r32, gate, nr, rdot, rzpart, nrz);
    }
}

// z = M^-1 r by one V-cycle in precision PT (T when equal; an FP32 V-cycle
// inside an FP64 CG when PT = float: r is rounded to FP32, the V-cycle runs on
// the FP32 operators, z comes back in FP64). Returns the buffer holding z.
template <typename T, typename PT>
static T* precondition(Ctx& c, MGH& mg, const T* r, T* z, CGVecs& v32, const int* gate, int nr) {
    const int64_t n = mg.L[0].ndof;
    if constexpr (std::is_same<T, PT>::value) {
        const T* zz = precondition_raw<T, PT>(c, mg, r, v32, gate, nr, false);
        hipLaunchKernelGGL((k_convert<T, T>), dim3(nblk(n), nr), dim3(256), 0, c.s, z, zz, n);
        return z;
    } else {
        PT* r32 = (PT*)v32.r;
        hipLaunchKernelGGL((k_convert_gated<PT, T>), dim3(nblk(n), nr), dim3(256), 0, c.s, r32, r, n, gate);
        PT* zz = mg.vcycle<PT>(c, 0, r32, gate, nr);
        hipLaunchKernelGGL((k_convert_gated<T, PT>), dim3(nblk(n), nr), dim3(256), 0, c.s, z, (const PT*)zz, n, gate);
        return z;
    }
}

// MG-PCG on the fine level (ebeMGPCG assembly/index.ts:1076-1193) for nr
// right-hand sides at once (rhs, U and the CG vectors hold nr vectors of n
// entries; each has its own CG scalars and stops on its own); PT is the
// precision of the V-cycle preconditioner.
template <typename T, typename TU, typename PT = T>
static CGResult pcg_mg(Ctx& c, MGH& mg, const T* rhs, TU* U, bool warm, CGVecs& v, double tol, int maxit,
                       volatile int* cancel, CGVecs* v32 = nullptr, int nr = 1) {
    MGLevel& l = mg.L[0];
    const int64_t n = l.ndof;
    TOPO_CHECK(nr >= 1 && nr <= kMaxRHS && nr <= mg.nrhs, TOPO_E_ARG, "batch of %d right-hand sides too large", nr);
    T* r = (T*)v.r;
    T* z = (T*)v.z;
    T* p = (T*)v.p;
    T* Ap = (T*)v.Ap;
    CGVecs& w32 = *(v32 ? v32 : &v);
    const dim3 gv(nblk(n), nr);
    CGResult res;
    if (warm) {
        if (std::is_same<T, TU>::value) {
            mg.apply<T>(c, 0, (const T*)U, r, 2, rhs, nullptr, nr);
        } else {
            hipLaunchKernelGGL((k_convert<T, TU>), gv, dim3(256), 0, c.s, z, (const TU*)U, n);
            mg.apply<T>(c, 0, z, r, 2, rhs, nullptr, nr);
        }
        hipLaunchKernelGGL(k_mask<TU>, gv, dim3(256), 0, c.s, U, l.fixed, n);
        res.applyA += nr;
    } else {
        hipLaunchKernelGGL(k_fill<TU>, gv, dim3(256), 0, c.s, U, TU(0), n, (const int*)nullptr);
        hipLaunchKernelGGL((k_convert<T, T>), gv, dim3(256), 0, c.s, r, rhs, n);
        hipLaunchKernelGGL(k_mask<T>, gv, dim3(256), 0, c.s, r, l.fixed, n);
    }
    precondition<T, PT>(c, mg, r, z, w32, nullptr, nr);
    res.vcycles += nr;
    hipLaunchKernelGGL(k_mask<T>, gv, dim3(256), 0, c.s, z, l.fixed, n);
    hipLaunchKernelGGL((k_convert<T, T>), gv, dim3(256), 0, c.s, p, (const T*)z, n);
    cg_set(c, tol, maxit, 1e-12, nr);
    TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cg_start<T, NR>), dim3(kRedBlocks), dim3(kRedThreads), 0, c.s,
                                           (const T*)r, (const T*)z, n, c.st, c.partials, c.ticket));
    c.check_launch();
    // The p update keeps the books (k_cgp_pupdate_books) where the context
    // has the second state: iteration k reads state k & 1 and leaves the
    // next in the other, so a block of iterations is even (the state is
    // back in c.st for the host).
    const bool books = c.st_alt != nullptr && std::is_same<T, TU>::value;
    const int check = books ? (mg.cg_check + 1) & ~1 : mg.cg_check;
    auto body = [&]() {
        for (int k = 0; k < check; k++) {
            CGState* sk = (books && (k & 1)) ? c.st_alt : c.st;
            CGState* sn = (k & 1) ? c.st : c.st_alt;
            const int* gate = &sk->done;
            // fence-free iteration: each reduction's partials are summed by
            // the kernel that needs them, the p update (or k_cgp_finish)
            // keeps the books; the applyA leaves the p.Ap partials itself
            // when it can
            int nb0 = mg.apply<T>(c, 0, p, Ap, 1, nullptr, gate, nr, c.partials);
            if (nb0 == 0) {
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_pAp<T, NR>), dim3(kCgBlocks), dim3(kRedThreads), 0, c.s,
                                                       (const T*)p, (const T*)Ap, n, (const CGState*)sk, c.partials));
                nb0 = kCgBlocks;
            }
            // r_lo: the V-cycle's FP32 residual, written by the update itself
            PT* r_lo = std::is_same<T, PT>::value ? nullptr : (PT*)w32.r;
            // (U += alpha p is done by the p update below)
            TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_update_ur<T, TU, NR, PT>), dim3(kCgBlocks),
                                                   dim3(kRedThreads), 0, c.s, (TU*)nullptr, r, (const T*)p,
                                                   (const T*)Ap, n, (const CGState*)sk, c.partials, r_lo, nb0));
            // z is never formed: r.z and p = z + beta p read the V-cycle's
            // output, which is zero at the fixed DOFs (the mask of ebeMGPCG
            // :1170-1172 is a no-op on it)
            // the V-cycle's last sweep leaves the r.z partials when it can
            int nb2 = 0;
            const PT* zz = precondition_raw<T, PT>(c, mg, r, w32, gate, nr, true,
                                                   mg.rz_fuse ? c.partials + cg_row(2, 0) : nullptr, &nb2);
            if (nb2 == 0) {
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_rz<T, PT, NR, false>), dim3(kCgBlocks), dim3(kRedThreads),
                                                       0, c.s, (const T*)r, zz, (const uint8_t*)l.fixed, n,
                                                       (const CGState*)sk, c.partials));
                nb2 = kCgBlocks;
            }
            if (books) {
                if constexpr (std::is_same<T, TU>::value)
                    TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_pupdate_books<T, PT, NR, TU>), dim3(kCgBlocks),
                                                           dim3(kRedThreads), 0, c.s, p, zz, n, (const CGState*)sk, sn,
                                                           (const double*)c.partials, U, nb0, nb2));
            } else {
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_pupdate<T, PT, NR, false, TU>), dim3(kCgBlocks),
                                                       dim3(kRedThreads), 0, c.s, p, zz, (const uint8_t*)l.fixed, n,
                                                       c.st, (const double*)c.partials, U, nb0, nb2));
                TOPO_NR_SWITCH(nr, hipLaunchKernelGGL((k_cgp_finish<NR>), dim3(1), dim3(kRedThreads), 0, c.s, c.st,
                                                       (const double*)c.partials, nb0, nb2));
            }
        }
    };
    hipGraphExec_t ex = nullptr;
    if (c.use_graphs)
    {
        std::vector<const void*> key{(const void*)&pcg_mg<T, TU, PT>, U, r, z, p, Ap, rhs, &mg,
                                     (const void*)(intptr_t)nr, w32.r};
        mg.graph_sig(key);
        ex = c.graph_capture(key, body);
    }
    for (;;) {
        c.read_st(nr);
        bool done = true;
        for (int q = 0; q < nr; q++) done = done && c.h_st[q].done;
        if (done) {
            for (int q = 0; q < nr; q++) {
                const CGState& h = c.h_st[q];
                res.iters += h.iters;
                res.relres = std::max(res.relres, std::sqrt(h.rn / (h.r0 > 1e-300 ? h.r0 : 1e-300)));
                res.nonfinite |= h.nonfinite;
            }
            break;
        }
        if (cancel && *cancel) throw TopoError{TOPO_E_CANCELLED};
        if (ex) TOPO_HIP(hipGraphLaunch(ex, c.s));
        else body();
        c.check_launch();
    }
    res.applyA += res.iters;
    res.vcycles += res.iters;
    return res;
}

// ---------------------------------------------------------------------------
struct Solver {
    topo_config cfg;
    Grid g;
    int64_t nel = 0, ndof = 0;
    std::unique_ptr<Ctx> c;
    MGH mg;
    uint8_t* fixed = nullptr;
    double* F = nullptr;          // nload * ndof
    uint8_t* pres = nullptr;      // zyx, or null
    uint8_t* voidm = nullptr;
    float *x = nullptr, *xold = nullptr, *xnew = nullptr, *xphys = nullptr, *xf = nullptr;
    float *dc = nullptr, *dcn = nullptr, *w = nullptr, *stress = nullptr, *Hs = nullptr, *tmpf = nullptr;
    double* sw = nullptr;
    double* E = nullptr;          // fine moduli (level 0)
    float* E32 = nullptr;
    double* U = nullptr;          // nload * ndof
    double* adj = nullptr;
    bool have_U = false;
    // CG work
    CGVecs v64, v32;
    double* r64 = nullptr;
    float* d32 = nullptr;
    float* rhs32 = nullptr;
    int64_t nfree = 0;
    int loop = 0;
    double last_c = 0;
    volatile int cancel_flag = 0;
    bool stencil_ready = false;
    DevKe K0;
    FilterStencil* stencil = nullptr;
    // design map and adjoint in FP64 (Helmholtz filter, forward/backward)
    double *xd64 = nullptr, *xf64 = nullptr, *xp64 = nullptr, *g64 = nullptr, *gv64 = nullptr, *t64 = nullptr;
    double *hb = nullptr, *hy = nullptr, *hr = nullptr, *hz = nullptr, *hp = nullptr, *hAp = nullptr, *hinvD = nullptr;
    HelmKe helm{};
    double* Lvec = nullptr;       // mechanism: unit vector at the output DOF
    double fb_penal = 3, fb_beta = 1;
    bool fb_ready = false;        // a forward pass has run at the current design
    int helm_iters = 0;
    double mixed_eta = 0;         // inner tolerance of the mixed solve (0: clamp(sqrt(tol), 1e-4, 1e-1))
    int fallback_solves = 0;      // mixed solves finished in FP64 after the FP32 inner solve stalled
    int batch_cap = 1;            // right-hand sides the batched buffers hold
    int rhs_tile = 0;             // right-hand sides per batched solve (0: not tuned yet)
    bool tuned = false;

    void init(const topo_config* cfgin, const uint8_t* hfixed, const float* hF, const uint8_t* hpres,
              const uint8_t* hvoid) {
        cfg = *cfgin;
        TOPO_CHECK(cfg.nelx >= 1 && cfg.nely >= 1 && cfg.nelz >= 1, TOPO_E_ARG, "grid dims must be >= 1");
        TOPO_CHECK(cfg.nload >= 1 && cfg.nload <= 64, TOPO_E_ARG, "nload must be in 1..64");
        TOPO_CHECK(cfg.linear_solver == TOPO_SOLVER_PCG || cfg.linear_solver == TOPO_SOLVER_MGPCG ||
                       cfg.linear_solver == TOPO_SOLVER_KSP_BDDC, TOPO_E_ARG, "unknown linear_solver %d",
                   cfg.linear_solver);
        TOPO_CHECK(cfg.linear_solver != TOPO_SOLVER_KSP_BDDC || cfg.precision == TOPO_PREC_FP64, TOPO_E_ARG,
                   "the KSP/BDDC solver runs in FP64 (precision 0)");
        TOPO_CHECK(cfg.precision == TOPO_PREC_FP64 || cfg.precision == TOPO_PREC_MIXED ||
                       cfg.precision == TOPO_PREC_MIXED_PC, TOPO_E_ARG, "unknown precision %d", cfg.precision);
        TOPO_CHECK(hfixed && hF, TOPO_E_ARG, "fixedMask and F are required");
        TOPO_CHECK(cfg.objective == TOPO_OBJ_COMPLIANCE || cfg.objective == TOPO_OBJ_MECHANISM, TOPO_E_ARG,
                   "unknown objective %d", cfg.objective);
        TOPO_CHECK(cfg.filter_type >= TOPO_FILTER_SENSITIVITY && cfg.filter_type <= TOPO_FILTER_HELMHOLTZ, TOPO_E_ARG,
                   "unknown filter_type %d", cfg.filter_type);
        g.nelx = cfg.nelx; g.nely = cfg.nely; g.nelz = cfg.nelz;
        nel = g.nel();
        ndof = g.ndof();
        c.reset(new Ctx());
        Ctx& C = *c;
        C.use_graphs = cfg.use_graphs != 0 && !graphs_disabled_by_env();
        {
            double hke[576];
            k0_matrix(cfg.nu, hke);
            K0.set(C, hke);
        }
        mg.K0 = &K0;
        if (cfg.objective == TOPO_OBJ_MECHANISM) {
            TOPO_CHECK(cfg.nload == 1, TOPO_E_ARG, "the mechanism objective takes one load case (the input force)");
            const int64_t din = cfg.mech_in_dof, dout = cfg.mech_out_dof;
            TOPO_CHECK(din >= 0 && din < g.ndof() && dout >= 0 && dout < g.ndof() && din != dout, TOPO_E_ARG,
                       "mechanism DOFs out of range");
            TOPO_CHECK(!hfixed[din] && !hfixed[dout], TOPO_E_ARG, "mechanism DOFs must be free");
            Springs sp{};
            if (cfg.mech_k_in > 0) { sp.dof[sp.n] = din; sp.k[sp.n] = cfg.mech_k_in; sp.n++; }
            if (cfg.mech_k_out > 0) { sp.dof[sp.n] = dout; sp.k[sp.n] = cfg.mech_k_out; sp.n++; }
            mg.springs = sp;
        }
        stencil = C.alloc<FilterStencil>(1);
        fixed = C.alloc<uint8_t>(ndof);
        TOPO_HIP(hipMemcpy(fixed, hfixed, ndof, hipMemcpyHostToDevice));
        nfree = 0;
        for (int64_t i = 0; i < ndof; i++) nfree += hfixed[i] ? 0 : 1;
        F = C.alloc<double>(ndof * cfg.nload);
        {
            std::vector<double> tmp(ndof * cfg.nload);
            for (int64_t i = 0; i < ndof * cfg.nload; i++) tmp[i] = (double)hF[i];
            TOPO_HIP(hipMemcpy(F, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        uint8_t* tmpu = C.alloc<uint8_t>(nel);
        if (hpres) {
            pres = C.alloc<uint8_t>(nel);
            TOPO_HIP(hipMemcpy(tmpu, hpres, nel, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_u8_ref_to_zyx, dim3(nblk(nel)), dim3(256), 0, C.s, tmpu, pres, g);
            C.sync();
        }
        if (hvoid) {
            voidm = C.alloc<uint8_t>(nel);
            TOPO_HIP(hipMemcpy(tmpu, hvoid, nel, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_u8_ref_to_zyx, dim3(nblk(nel)), dim3(256), 0, C.s, tmpu, voidm, g);
            C.sync();
        }
        C.free_one(tmpu);
        x = C.alloc<float>(nel); xold = C.alloc<float>(nel); xnew = C.alloc<float>(nel);
        xphys = C.alloc<float>(nel); xf = C.alloc<float>(nel);
        dc = C.alloc<float>(nel); dcn = C.alloc<float>(nel); w = C.alloc<float>(nel);
        stress = C.alloc<float>(nel); Hs = C.alloc<float>(nel); tmpf = C.alloc<float>(nel);
        sw = C.alloc<double>(nel);
        E = C.alloc<double>(nel);
        const bool mixed = cfg.precision != TOPO_PREC_FP64;
        if (mixed) E32 = C.alloc<float>(nel);
        xd64 = C.alloc<double>(nel); xf64 = C.alloc<double>(nel); xp64 = C.alloc<double>(nel);
        g64 = C.alloc<double>(nel); gv64 = C.alloc<double>(nel); t64 = C.alloc<double>(nel);
        if (cfg.filter_type == TOPO_FILTER_HELMHOLTZ) {
            const int64_t nn = g.nnodes();
            hb = C.alloc<double>(nn); hy = C.alloc<double>(nn); hr = C.alloc<double>(nn); hz = C.alloc<double>(nn);
            hp = C.alloc<double>(nn); hAp = C.alloc<double>(nn); hinvD = C.alloc<double>(nn);
        }
        if (cfg.objective == TOPO_OBJ_MECHANISM) {
            adj = C.alloc<double>(ndof);
            Lvec = C.alloc<double>(ndof);
            std::vector<double> hl(ndof, 0.0);
            hl[cfg.mech_out_dof] = 1.0;
            TOPO_HIP(hipMemcpy(Lvec, hl.data(), ndof * sizeof(double), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, adj, 0.0, ndof, (const int*)nullptr);
        }
        U = C.alloc<double>(ndof * cfg.nload);
        hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ndof * cfg.nload)), dim3(256), 0, C.s, U, 0.0, ndof * cfg.nload,
                           (const int*)nullptr);
        // load cases solved together by one batched MG-PCG (FP64, or FP64 with
        // the FP32 V-cycle): at most kMaxRHS per batch
        batch_cap = 1;
        if (cfg.nload > 1 && cfg.linear_solver == TOPO_SOLVER_MGPCG && cfg.precision != TOPO_PREC_MIXED &&
            cfg.rhs_tile != 1) {
            TOPO_CHECK(cfg.rhs_tile == 0 || cfg.rhs_tile == 2 || cfg.rhs_tile == 4, TOPO_E_ARG,
                       "rhs_tile must be 0 (autotune), 1, 2 or 4");
            batch_cap = cfg.rhs_tile > 0 ? cfg.rhs_tile : kMaxRHS;
            while (batch_cap > cfg.nload) batch_cap /= 2;
        }
        const int64_t nvb = ndof * batch_cap;
        v64.r = C.alloc<double>(nvb); v64.z = C.alloc<double>(nvb);
        v64.p = C.alloc<double>(nvb); v64.Ap = C.alloc<double>(nvb);
        if (mixed) {
            v32.r = C.alloc<float>(nvb); v32.z = C.alloc<float>(ndof);
            v32.p = C.alloc<float>(ndof); v32.Ap = C.alloc<float>(ndof);
            r64 = C.alloc<double>(ndof);
            d32 = C.alloc<float>(ndof);
            rhs32 = C.alloc<float>(ndof);
        }
        // multigrid hierarchy (level 0 always exists: its diagonal is the
        // Jacobi preconditioner)
        mg.omega = cfg.mg_omega;
        mg.nu1 = cfg.mg_nu1;
        mg.nu2 = cfg.mg_nu2;
        mg.coarse_sweeps = cfg.coarse_sweeps;
        // Jacobi-PCG uses level 0 for its diagonal only: no coarse factor
        mg.coarse_solver = cfg.linear_solver == TOPO_SOLVER_MGPCG ? cfg.coarse_solver : TOPO_COARSE_SWEEPS;
        mg.galerkin_max = cfg.galerkin_max_ndof;
        // skip threshold (skipThreshold = Emin * 1000, js/optimizer-worker-3d.js:762);
        // -1 keeps every element (E >= Emin > 0)
        mg.thr = cfg.skip_factor > 0 ? cfg.Emin * cfg.skip_factor : -1.0;
        mg.omega_cap = cfg.mg_omega_cap;
        mg.exact = cfg.exact_apply != 0;
        mf_ok = mf_allow_lds();
        mg.mfma = cfg.mfma_apply != 0 && mf_ok;
        if (cfg.mfma_apply && !mg.mfma && std::getenv("TOPO_VERBOSE"))
            std::fprintf(stderr, "[topo] matrix-core apply unavailable (LDS limit): FP32 VALU kernel\n");
        if (const char* e = getenv("TOPO_XM_NT")) mg.xm_nt = atoi(e);
        if (const char* e = getenv("TOPO_XM_CX")) mg.xm_cx = atoi(e);
        if (const char* e = getenv("TOPO_XM_CX64")) mg.xm_cx64 = atoi(e);
        if (const char* e = getenv("TOPO_RZ_FUSE")) mg.rz_fuse = atoi(e) != 0;
        if (const char* e = getenv("TOPO_ND_DEPTH")) mg.nd_depth = atoi(e);
        if (const char* e = getenv("TOPO_ND_MIN")) mg.nd_min = atoll(e);
        if (const char* e = getenv("TOPO_ND_LEAF")) mg.nd_leaf = atoll(e);
        if (const char* e = getenv("TOPO_TC_NW")) mg.tc_nw = atoi(e) == 4 ? 4 : 8;
        if (const char* e = getenv("TOPO_CG_CHECK")) mg.cg_check = std::max(1, std::min(16, atoi(e)));
        if (const char* e = getenv("TOPO_JCG_CHECK")) mg.jcg_check = std::max(1, std::min(64, atoi(e)));

        if (const char* e = getenv("TOPO_MIXED_ETA")) mixed_eta = atof(e);
        const int nlev = cfg.linear_solver == TOPO_SOLVER_MGPCG ? cfg.mg_levels : 1;
        mg.nrhs = batch_cap;
        mg.coarse_fp32 = cfg.coarse_fp32;
        if (cfg.linear_solver == TOPO_SOLVER_KSP_BDDC) {
            // plain global Jacobi sweeps inside the preconditioner
            mg.omega = 1.0;
            mg.omega_cap = 0.0;
        }
        mg.build(C, g, fixed, cfg.linear_solver == TOPO_SOLVER_MGPCG ? nlev : 1, mixed);
        mg.L[0].E = E;
        mg.L[0].E32 = E32;
        if (cfg.linear_solver == TOPO_SOLVER_KSP_BDDC) mg.bddc_setup(C);
        // filter stencil + normalisers
        setup_filter(cfg.rmin);
        // initial design: volfrac, preserved -> 1, void -> 0 (:1928-2022)
        std::vector<float> hx(nel, (float)cfg.volfrac);
        if (hpres) for (int64_t i = 0; i < nel; i++) if (hpres[i]) hx[i] = 1.0f;
        if (hvoid) for (int64_t i = 0; i < nel; i++) if (hvoid[i]) hx[i] = 0.0f;
        set_design(hx.data());
        C.sync();
    }

    // Helmholtz filter element matrix r^2 K_L + M of the unit cube with the
    // PDE length r = rmin / (2 sqrt 3) (the equivalence of Lazarov & Sigmund
    // 2011 between the PDE filter and a linear filter of radius rmin)
    void setup_helmholtz(double rmin) {
        const double rp = rmin / (2.0 * std::sqrt(3.0));
        const double m1[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
        const double k1[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) {
                int ai, bi, ci, aj, bj, cj;
                local_offsets(i, ai, bi, ci);
                local_offsets(j, aj, bj, cj);
                const double mx = m1[ai][aj], my = m1[bi][bj], mz = m1[ci][cj];
                const double kl = k1[ai][aj] * my * mz + mx * k1[bi][bj] * mz + mx * my * k1[ci][cj];
                helm.a[i * 8 + j] = rp * rp * kl + mx * my * mz;
            }
        Ctx& C = *c;
        hipLaunchKernelGGL(k_helm_invdiag, dim3(nblk(g.nnodes())), dim3(256), 0, C.s, hinvD, g, helm.a[0]);
        C.sync();
        C.check_launch();
        stencil_ready = true;
    }

    void setup_filter(double rmin) {
        if (cfg.filter_type == TOPO_FILTER_HELMHOLTZ) {
            setup_helmholtz(rmin);
            return;
        }
        FilterStencil st;
        std::memset(&st, 0, sizeof st);
        if (cfg.filter_type == TOPO_FILTER_DENSITY) {
            // applyDensityFilter window ceil(r), d^2 <= r^2, order z, y, x
            const int r = (int)std::ceil(rmin);
            TOPO_CHECK(rmin > 0 && r <= kFilterMaxR, TOPO_E_UNSUPPORTED,
                       "density filter radius %g (0 < rmin <= %d)", rmin, kFilterMaxR);
            st.r = r;
            const double r2 = rmin * rmin;
            for (int kz = -r; kz <= r; kz++)
                for (int ky = -r; ky <= r; ky++)
                    for (int kx = -r; kx <= r; kx++) {
                        const double d2 = (double)(kx * kx + ky * ky + kz * kz);
                        if (d2 <= r2) {
                            st.off[st.count][0] = kx; st.off[st.count][1] = ky; st.off[st.count][2] = kz;
                            st.w[st.count] = rmin - std::sqrt(d2);
                            st.count++;
                        }
                    }
        } else {
            // prepareFilter window +-floor(rmin), dist <= rmin, order x, y, z
            const int r = (int)std::floor(rmin);
            TOPO_CHECK(rmin > 0 && r <= kFilterMaxR, TOPO_E_UNSUPPORTED,
                       "sensitivity filter radius %g (0 < rmin < %d)", rmin, kFilterMaxR + 1);
            st.r = r;
            for (int kx = -r; kx <= r; kx++)
                for (int ky = -r; ky <= r; ky++)
                    for (int kz = -r; kz <= r; kz++) {
                        const double dist = std::sqrt((double)(kx * kx + ky * ky + kz * kz));
                        if (dist <= rmin) {
                            st.off[st.count][0] = kx; st.off[st.count][1] = ky; st.off[st.count][2] = kz;
                            st.w[st.count] = std::max(0.0, rmin - dist);
                            st.count++;
                        }
                    }
        }
        Ctx& C = *c;
        TOPO_HIP(hipMemcpyAsync(stencil, &st, sizeof st, hipMemcpyHostToDevice, C.s));
        if (cfg.filter_type == TOPO_FILTER_DENSITY)
            hipLaunchKernelGGL(k_density_sw, dim3(nblk(nel)), dim3(256), 0, C.s, sw, g, (const FilterStencil*)stencil);
        else
            hipLaunchKernelGGL(k_filter_hs, dim3(nblk(nel)), dim3(256), 0, C.s, Hs, g, (const FilterStencil*)stencil);
        C.sync();
        C.check_launch();
        stencil_ready = true;
    }

    void set_design(const float* hx) {
        Ctx& C = *c;
        TOPO_HIP(hipMemcpyAsync(tmpf, hx, nel * sizeof(float), hipMemcpyHostToDevice, C.s));
        hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(nel)), dim3(256), 0, C.s, tmpf, x, g);
        C.sync();
        fb_ready = false;
    }
    void get_field_f(const float* dsrc, float* hout) {
        Ctx& C = *c;
        hipLaunchKernelGGL(k_zyx_to_ref, dim3(nblk(nel)), dim3(256), 0, C.s, dsrc, tmpf, g);
        TOPO_HIP(hipMemcpyAsync(hout, tmpf, nel * sizeof(float), hipMemcpyDeviceToHost, C.s));
        C.sync();
    }

    // E and the operators of all levels from the physical densities (zyx)
    void set_density_dev(const float* dxphys, double penal) {
        Ctx& C = *c;
        // ebePCG uses Math.pow (assembly/index.ts:544), so does the KSP/BDDC
        // wrapper (js/optimizer-worker-3d.js:3824-3830); the MG path _powDensity
        const int powmode = cfg.linear_solver == TOPO_SOLVER_MGPCG ? 0 : 1;
        hipLaunchKernelGGL(k_moduli<float>, dim3(nblk(nel)), dim3(256), 0, C.s, dxphys, penal, cfg.E0, cfg.Emin, powmode, E,
                           E32, nel);
        C.check_launch();
        mg.update(C);
    }

    double tolerance_for(int iteration, int maxIt) const {
        if (cfg.cg_tol > 0) return cfg.cg_tol;
        // CG tolerance schedule (js/optimizer-worker-3d.js:3597-3602)
        const double progress = maxIt > 1 ? (double)(iteration - 1) / (maxIt - 1) : 1.0;
        return std::exp(std::log(1e-3) + progress * (std::log(1e-8) - std::log(1e-3)));
    }

    // solve K U_k = F_k for every load case (FE() :4155-4371)
    void fe(int iteration, int maxIt, topo_step_info* info) {
        const double tol = tolerance_for(iteration, maxIt);
        const int64_t nfree_cap = cfg.cg_max_iter;
        const bool mgp = cfg.linear_solver == TOPO_SOLVER_MGPCG;
        // warm start: PCG warm (:3589-3595), WASM MGPCG and KSP/BDDC cold
        // (assembly/index.ts:1101, :1465)
        bool warm = iteration > 1 && have_U;
        if (cfg.warm_start == 0) warm = false;
        if (cfg.warm_start == -1 && (mgp || mg.bddc)) warm = false;
        // maxIter = min(nfree, MAX_CG_ITERATIONS) (:3609, :3779)
        const int maxit = (int)std::min<int64_t>(nfree_cap, nfree);
        CGResult tot;
        int outer = 0;
        if (!tuned) {
            if (cfg.rhs_tile == 0) autotune();
            else rhs_tile = std::min(std::max(cfg.rhs_tile, 1), batch_cap);
            tuned = true;
        }
        if (batch_cap > 1) {
            for (int k = 0; k < cfg.nload;) {
                int r = rhs_tile;
                while (r > cfg.nload - k) r /= 2;
                solve_batch(F + (int64_t)k * ndof, U + (int64_t)k * ndof, r, warm, tol, maxit, tot);
                k += r;
            }
        } else {
            for (int k = 0; k < cfg.nload; k++)
                solve_one(F + (int64_t)k * ndof, U + (int64_t)k * ndof, warm, tol, maxit, tot, outer);
        }
        // the mechanism's adjoint: K lam = L (L = unit vector at the output DOF)
        if (cfg.objective == TOPO_OBJ_MECHANISM) solve_one(Lvec, adj, warm, tol, maxit, tot, outer);
        have_U = true;
        if (info) {
            info->cg_iters = tot.iters;
            info->vcycles = tot.vcycles;
            info->applyA_fine = tot.applyA;
            info->outer_iters = outer;
            info->tolerance = tol;
            info->residual = tot.relres;
        }
    }

    // nr load cases at once: batched MG-PCG (FP64, or FP64 with the FP32
    // V-cycle preconditioner)
    void solve_batch(const double* rhs, double* sol, int nr, bool warm, double tol, int maxit, CGResult& tot) {
        Ctx& C = *c;
        CGResult r = cfg.precision == TOPO_PREC_MIXED_PC
                         ? pcg_mg<double, double, float>(C, mg, rhs, sol, warm, v64, tol, maxit, &cancel_flag, &v32, nr)
                         : pcg_mg<double, double>(C, mg, rhs, sol, warm, v64, tol, maxit, &cancel_flag, nullptr, nr);
        TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
        tot.iters += r.iters; tot.vcycles += r.vcycles; tot.applyA += r.applyA;
        tot.relres = std::max(tot.relres, r.relres);
    }

    // Autotuner (rhs_tile 0), run once on the operators of the first solve
    // with HIP events on the solver's stream:
    //  1. element-tile shape of the x-march operator on the two finest
    //     levels: x-planes per workgroup CX in {auto, 2, 3, 4, 5, 6, 8, 10,
    //     12, 16} (the workgroup size keeps its rule: another one is another
    //     compiled kernel whose contractions may round differently), timed
    //     on the operator in the V-cycle's precision; CX only moves work
    //     between workgroups, every node's sums keep their order, so the
    //     results are bit-identical whatever the choice;
    //  2. with several load cases, the batch size R in {1, 2, 4} up to
    //     batch_cap, timed on one fine operator plus one V-cycle of R
    //     right-hand sides, normalised per right-hand side.
    // The cached graphs are dropped afterwards.
    void autotune() {
        Ctx& C = *c;
        const bool pc = cfg.precision == TOPO_PREC_MIXED_PC;
        const bool f32 = cfg.precision != TOPO_PREC_FP64;
        for (int q = 0; q < batch_cap; q++)
            TOPO_HIP(hipMemcpyAsync((double*)v64.r + q * ndof, F + (int64_t)(q % cfg.nload) * ndof,
                                    ndof * sizeof(double), hipMemcpyDeviceToDevice, C.s));
        hipEvent_t e0, e1;
        TOPO_HIP(hipEventCreate(&e0));
        TOPO_HIP(hipEventCreate(&e1));
        auto time_it = [&](auto&& fn, int reps) {
            fn();   // warm-up (code object load, caches)
            TOPO_HIP(hipEventRecord(e0, C.s));
            for (int k = 0; k < reps; k++) fn();
            TOPO_HIP(hipEventRecord(e1, C.s));
            TOPO_HIP(hipEventSynchronize(e1));
            float ms = 0;
            TOPO_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        tune_log.clear();
        // candidate 0 is the size rule (one round of workgroups); around it
        // every chunk from 4 to 12 planes (a launch's time goes with its
        // largest chunk, and one plane more or less moves the round's fill),
        // and the large ones that fit the chip in one round at fewer
        // workgroups per CU
        constexpr int kNcx = 14;
        const int cx_c[kNcx] = {0, 4, 5, 6, 7, 8, 9, 10, 11, 12, 16, 20, 24, 32};
        // the coarser levels: also one to three planes per workgroup, for
        // latency-bound small grids
        const int cx_s[kNcx] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16, 20, 32};
        // per candidate the least of two timings of 5 launches, the two
        // rounds over all candidates one after the other (clock and cache
        // drift between candidates evens out)
        auto sweep_cx = [&](double (&ms)[kNcx], const int (&cand)[kNcx], auto&& launch) {
            for (int b = 0; b < kNcx; b++) ms[b] = 1e300;
            for (int round = 0; round < 2; round++)
                for (int b = 0; b < kNcx; b++)
                    ms[b] = std::min(ms[b], time_it([&]() { launch(cand[b]); }, 5));
        };
        const int nl = (int)mg.L.size();
        mg.lv_nt.assign(nl, 0);
        mg.lv_cx.assign(nl, 0);
        mg.lv_cx64.assign(nl, 0);
        if (!mg.xm_nt && !mg.xm_cx && !mg.exact) {
            for (int li = 0; li < std::min(nl, 3); li++) {
                MGLevel& l = mg.L[li];
                if (l.dense || l.chol) continue;
                const int (&cand)[kNcx] = li == 0 ? cx_c : cx_s;
                // operands: the level's own vectors, filled from the rhs copy
                hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, l.x,
                                   (const double*)v64.r, l.ndof);
                if (f32)
                    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, l.x32,
                                       (const double*)v64.r, l.ndof);
                const int H = l.g.nnz() + 1;
                const int nt_rule = 4 * H > 512 ? 1024 : 512;
                // the coarser levels also try 256-thread workgroups (more
                // workgroups on grids too small to fill the chip; the sums
                // are the same bits for every workgroup size)
                const int nnt = (li > 0 && 2 * H <= 256) ? 2 : 1;
                double best = 1e300;
                int best_nt = 0, best_cx = 0;
                for (int ti = 0; ti < nnt; ti++) {
                    const int ntv = ti == 0 ? nt_rule : 256;
                    mg.lv_nt[li] = ti == 0 ? 0 : 256;
                    double ms[kNcx];
                    sweep_cx(ms, cand, [&](int cx) {
                        mg.lv_cx[li] = cx;
                        if (f32) mg.apply<float>(C, li, l.x32, l.x232, 1, nullptr, nullptr);
                        else mg.apply<double>(C, li, l.x, l.x2, 1, nullptr, nullptr);
                    });
                    for (int b = 0; b < kNcx; b++) {
                        tune_log.push_back({li, ntv, cand[b], ms[b]});
                        if (ms[b] < best) {
                            best = ms[b];
                            best_nt = mg.lv_nt[li];
                            best_cx = cand[b];
                        }
                    }
                }
                mg.lv_nt[li] = best_nt;
                if (li < 2) {
                    tuned_nt[li] = best_nt ? best_nt : nt_rule;
                    tuned_cx[li] = best_cx;
                }
                mg.lv_cx[li] = best_cx;
            }
            // the mixed solvers' CG operator (FP64, masked, finest level):
            // its own x-chunk, as FP64 arithmetic and registers give it
            // another best tile than the FP32 smoothers (log entries 1000)
            MGLevel& l = mg.L[0];
            if (f32 && !mg.xm_cx64 && !l.dense && mg.xm_level<double>(0)) {
                const int nt = mg.xm_pick_nt<double>(0);
                double ms[kNcx];
                // candidate 0 is the size rule, made explicit (0 in lv_cx64
                // would mean "as the FP32 operator")
                const int cx_rule = mg.xm_auto_cx(0, nt);
                sweep_cx(ms, cx_c, [&](int cx) {
                    mg.lv_cx64[0] = cx ? cx : cx_rule;
                    mg.apply<double>(C, 0, l.x, l.x2, 1, nullptr, nullptr);
                });
                int bi = 0;
                for (int b = 0; b < kNcx; b++) {
                    tune_log.push_back({1000, nt, cx_c[b] ? cx_c[b] : cx_rule, ms[b]});
                    if (ms[b] < ms[bi]) bi = b;
                }
                tuned_cx64 = cx_c[bi] ? cx_c[bi] : cx_rule;
                mg.lv_cx64[0] = tuned_cx64;
            }
        }
        rhs_tile = 1;
        if (batch_cap > 1) {
            double best = 1e300;
            for (int R = 1; R <= batch_cap; R *= 2) {
                const double ms = time_it([&]() {
                    mg.apply<double>(C, 0, (const double*)v64.r, (double*)v64.Ap, 1, nullptr, nullptr, R);
                    if (pc)
                        precondition<double, float>(C, mg, (const double*)v64.r, (double*)v64.z, v32, nullptr, R);
                    else
                        precondition<double, double>(C, mg, (const double*)v64.r, (double*)v64.z, v64, nullptr, R);
                }, 3);
                tune_log.push_back({-R, 0, 0, ms / R});
                if (ms / R < best) { best = ms / R; rhs_tile = R; }
            }
        }
        TOPO_HIP(hipEventDestroy(e0));
        TOPO_HIP(hipEventDestroy(e1));
        C.graph_clear();
        tuned = true;
    }
    int tuned_nt[2] = {0, 0}, tuned_cx[2] = {0, 0}, tuned_cx64 = 0;
    bool mf_ok = false;   // the matrix-core applyA may run (its LDS granted)
    struct TuneEntry {
        int what;          // >= 0: tile timing of that MG level; < 0: -R, batch of R right-hand sides
        int nt, cx;
        double ms;         // per launch (tiles) or per right-hand side (batches)
    };
    std::vector<TuneEntry> tune_log;

    // K sol = rhs with the configured solver and precision
    void solve_one(const double* rhs, double* sol, bool warm, double tol, int maxit, CGResult& tot, int& outer) {
        Ctx& C = *c;
        const bool mgp = cfg.linear_solver == TOPO_SOLVER_MGPCG;
        if (cfg.precision == TOPO_PREC_MIXED_PC && mgp) {
            CGResult r = pcg_mg<double, double, float>(C, mg, rhs, sol, warm, v64, tol, maxit, &cancel_flag, &v32);
            TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
            tot.iters += r.iters; tot.vcycles += r.vcycles; tot.applyA += r.applyA;
            tot.relres = std::max(tot.relres, r.relres);
        } else if (cfg.precision == TOPO_PREC_FP64 || cfg.precision == TOPO_PREC_MIXED_PC) {
            // KSP/BDDC: the same PCG with the domain-decomposition preconditioner
            CGResult r = (mgp || mg.bddc) ? pcg_mg<double, double>(C, mg, rhs, sol, warm, v64, tol, maxit, &cancel_flag)
                             : pcg_jacobi<double, double>(C, mg, rhs, sol, warm, v64, tol, maxit, &cancel_flag);
            TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
            tot.iters += r.iters; tot.vcycles += r.vcycles; tot.applyA += r.applyA;
            tot.relres = std::max(tot.relres, r.relres);
        } else {
            CGResult r = solve_mixed(rhs, sol, warm, tol, maxit, mgp, &outer);
            tot.iters += r.iters; tot.vcycles += r.vcycles; tot.applyA += r.applyA;
            tot.relres = std::max(tot.relres, r.relres);
        }
    }

    // Helmholtz filter of an element field (FP64, zyx): b = T in, Jacobi-PCG
    // on (r^2 K_L + M) y = b to 1e-12 relative, out = T^T y
    void helm_solve(const double* in, double* out) {
        Ctx& C = *c;
        const int64_t nn = g.nnodes();
        hipLaunchKernelGGL(k_helm_rhs<double>, dim3(nblk(nn)), dim3(256), 0, C.s, in, hb, g);
        hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nn)), dim3(256), 0, C.s, hy, 0.0, nn, (const int*)nullptr);
        hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nn)), dim3(256), 0, C.s, hr, (const double*)hb, nn);
        hipLaunchKernelGGL(k_jacobi_apply<double>, dim3(nblk(nn)), dim3(256), 0, C.s, hz, (const double*)hr,
                           (const double*)hinvD, nn);
        hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nn)), dim3(256), 0, C.s, hp, (const double*)hz, nn);
        cg_set(C, 1e-13, 5000, 0.0);
        hipLaunchKernelGGL(k_cg_start<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, (const double*)hr,
                           (const double*)hz, nn, C.st, C.partials, C.ticket);
        const int* gate = &C.st->done;
        for (;;) {
            const CGState& h = C.read_st();
            if (h.done) {
                TOPO_CHECK(!h.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar in the Helmholtz filter");
                helm_iters = h.iters;
                break;
            }
            for (int k = 0; k < 8; k++) {
                hipLaunchKernelGGL(k_helm_apply, dim3(nblk(nn)), dim3(256), 0, C.s, (const double*)hp, hAp, g, helm, 0,
                                   (const double*)nullptr, gate);
                hipLaunchKernelGGL(k_cg_alpha<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, (const double*)hp,
                                   (const double*)hAp, nn, C.st, C.partials, C.ticket);
                hipLaunchKernelGGL((k_cg_update_jacobi<double, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s,
                                   hy, hr, hz, (const double*)hp, (const double*)hAp, (const double*)hinvD, nn, C.st,
                                   C.partials, C.ticket);
                hipLaunchKernelGGL(k_cg_pupdate<double>, dim3(nblk(nn)), dim3(256), 0, C.s, hp, (const double*)hz, nn,
                                   C.st);
            }
            C.check_launch();
        }
        hipLaunchKernelGGL(k_helm_out, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)hy, out, g);
        C.check_launch();
    }

    // penalty continuation (:2056-2064, :2130) and beta continuation (:2133)
    // at SIMP iteration `it`
    void schedule(int it, double* penal, double* beta) const {
        *penal = cfg.penal_start +
                 (cfg.penal - cfg.penal_start) * std::min(1.0, (double)(it - 1) / std::max(1, cfg.continuation_iters));
        *beta = 1.0;
        if (cfg.use_projection)
            *beta = cfg.beta_fixed > 0 ? (double)cfg.beta_fixed
                                       : std::min(cfg.beta_max, std::pow(2.0, std::floor((double)(it - 1) / cfg.beta_interval)));
    }

    // Forward pass in FP64 at the current design x and SIMP iteration `it`:
    // x -> filter (density or Helmholtz; the sensitivity filter acts on the
    // gradient instead) -> Heaviside (if on) -> E -> solve(s) -> objective
    // (compliance sum_k F_k^T U_k, or the mechanism's u_out = L^T U). Returns
    // the objective; *vol = mean(xPhys).
    double forward(int it, topo_step_info* info, double* vol) {
        Ctx& C = *c;
        schedule(it, &fb_penal, &fb_beta);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(nel)), dim3(256), 0, C.s, xd64, (const float*)x, nel);
        if (cfg.filter_type == TOPO_FILTER_DENSITY)
            hipLaunchKernelGGL(k_density_filter_d, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)xd64, xf64, sw, g, 0,
                               (const FilterStencil*)stencil);
        else if (cfg.filter_type == TOPO_FILTER_HELMHOLTZ)
            helm_solve(xd64, xf64);
        else
            hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nel)), dim3(256), 0, C.s, xf64, (const double*)xd64,
                               nel);
        if (cfg.use_projection)
            hipLaunchKernelGGL(k_heaviside_d, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)xf64, xp64, nel, fb_beta,
                               cfg.eta);
        else
            hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nel)), dim3(256), 0, C.s, xp64, (const double*)xf64,
                               nel);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(nel)), dim3(256), 0, C.s, xphys, (const double*)xp64, nel);
        hipLaunchKernelGGL(k_moduli<double>, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)xp64, fb_penal, cfg.E0,
                           cfg.Emin, 0, E, E32, nel);
        C.check_launch();
        mg.update(C);
        C.sync();
        fe(it, cfg.max_iterations, info);
        double J;
        if (cfg.objective == TOPO_OBJ_MECHANISM) {
            TOPO_HIP(hipMemcpyAsync(C.h_sc, U + cfg.mech_out_dof, sizeof(double), hipMemcpyDeviceToHost, C.s));
            C.sync();
            J = C.h_sc[0];
        } else {
            J = compliance();
        }
        hipLaunchKernelGGL((k_dot<double, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, (const double*)xp64,
                           (const double*)nullptr, nel, C.dsc + 3, C.partials, C.ticket);
        if (vol) *vol = C.read_d(C.dsc + 3) / (double)nel;
        fb_ready = true;
        return J;
    }

    // Backward pass (hand-written adjoint of forward()): g64 = dJ/dx and
    // gv64 = d mean(xPhys)/dx, zyx, FP64.
    void backward() {
        Ctx& C = *c;
        TOPO_CHECK(fb_ready, TOPO_E_ARG, "backward needs a forward pass at the current design");
        const double* lam = cfg.objective == TOPO_OBJ_MECHANISM ? adj : nullptr;
        for (int k = 0; k < cfg.nload; k++) {
            const double* Uk = U + (int64_t)k * ndof;
            hipLaunchKernelGGL(k_adj_sens, dim3(nblk(nel)), dim3(256), 0, C.s, Uk, lam ? lam : Uk, (const double*)xp64,
                               fb_penal, cfg.E0, cfg.Emin, g, g64, k > 0 ? 1 : 0, (const double*)K0.ke);
        }
        hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nel)), dim3(256), 0, C.s, gv64, 1.0 / (double)nel, nel,
                           (const int*)nullptr);
        if (cfg.use_projection) {
            hipLaunchKernelGGL(k_heaviside_grad, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)xf64, g64, nel, fb_beta,
                               cfg.eta);
            hipLaunchKernelGGL(k_heaviside_grad, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)xf64, gv64, nel,
                               fb_beta, cfg.eta);
        }
        for (double* v : {g64, gv64}) {
            if (cfg.filter_type == TOPO_FILTER_DENSITY) {
                hipLaunchKernelGGL(k_density_filter_d, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)v, t64, sw, g, 1,
                                   (const FilterStencil*)stencil);
                hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nel)), dim3(256), 0, C.s, v, (const double*)t64,
                                   nel);
            } else if (cfg.filter_type == TOPO_FILTER_HELMHOLTZ) {
                // the Helmholtz map is symmetric: its adjoint is itself
                helm_solve(v, t64);
                hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(nel)), dim3(256), 0, C.s, v, (const double*)t64,
                                   nel);
            }
        }
        C.check_launch();
    }

    // FP64 iterative refinement around an FP32 inner solve: the outer
    // residual r = F - K U, its norm and the stopping test
    // ||r||^2 < tol^2 ||r0||^2 are FP64; the inner MG-/Jacobi-PCG runs on FP32
    // vectors and an FP32 operator with FP64 CG scalars.
    CGResult solve_mixed(const double* Fk, double* Uk, bool warm, double tol, int maxit, bool mgp, int* outer_out) {
        Ctx& C = *c;
        CGResult tot;
        if (!warm)
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, Uk, 0.0, ndof, (const int*)nullptr);
        else
            hipLaunchKernelGGL(k_mask<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, Uk, fixed, ndof);
        double* r = r64;
        mg.apply<double>(C, 0, Uk, r, 2, Fk, nullptr);
        tot.applyA++;
        hipLaunchKernelGGL((k_dot<double, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, r, r, ndof, C.dsc,
                           C.partials, C.ticket);
        const double rn0 = C.read_d(C.dsc);
        const double tolSq = tol * tol * (rn0 > 1e-30 ? rn0 : 1e-30);
        double rn = rn0;
        double eta = std::min(1e-1, std::max(1e-4, std::sqrt(tol)));
        if (mixed_eta > 0) eta = mixed_eta;
        int outer = 0;
        bool stalled = false;
        while (rn >= tolSq && outer < 20 && tot.iters < maxit) {
            const double scale = std::sqrt(rn);
            C.h_sc[7] = scale;   // pinned staging word (every read_d synchronises before it is rewritten)
            TOPO_HIP(hipMemcpyAsync(C.dsc + 1, C.h_sc + 7, sizeof(double), hipMemcpyHostToDevice, C.s));
            hipLaunchKernelGGL(k_scale_to<float>, dim3(nblk(ndof)), dim3(256), 0, C.s, rhs32, r, C.dsc + 1, ndof);
            // inner solve A d = r/scale from d = 0
            CGResult ir = mgp ? pcg_mg<float, float>(C, mg, rhs32, d32, false, v32, eta, maxit - tot.iters, &cancel_flag)
                              : pcg_jacobi<float, float>(C, mg, rhs32, d32, false, v32, eta, maxit - tot.iters,
                                                         &cancel_flag);
            hipLaunchKernelGGL(k_axpy_refine<float>, dim3(nblk(ndof)), dim3(256), 0, C.s, Uk, (const float*)d32, scale,
                               ndof);
            tot.iters += ir.iters; tot.vcycles += ir.vcycles; tot.applyA += ir.applyA;
            mg.apply<double>(C, 0, Uk, r, 2, Fk, nullptr);
            tot.applyA++;
            hipLaunchKernelGGL((k_dot<double, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, r, r, ndof, C.dsc,
                               C.partials, C.ticket);
            const double rn_prev = rn;
            rn = C.read_d(C.dsc);
            outer++;
            TOPO_CHECK(std::isfinite(rn), TOPO_E_BREAKDOWN, "non-finite residual in mixed refinement");
            if (ir.iters == 0) break;
            // an FP32 inner solve that no longer cuts the FP64 residual norm by
            // 10x has hit its attainable accuracy (high-contrast, near-singular
            // operators): finish in FP64
            if (rn > 1e-2 * rn_prev) {
                stalled = true;
                break;
            }
        }
        if (stalled && rn >= tolSq && tot.iters < maxit) {
            // FP64 PCG warm-started from U; its tolerance is relative to the
            // current residual, scaled to reach tol relative to the first one
            const double tol64 = tol * std::sqrt(rn0 / rn);
            CGResult fr = mgp ? pcg_mg<double, double>(C, mg, Fk, Uk, true, v64, tol64, maxit - tot.iters, &cancel_flag)
                              : pcg_jacobi<double, double>(C, mg, Fk, Uk, true, v64, tol64, maxit - tot.iters,
                                                           &cancel_flag);
            TOPO_CHECK(!fr.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
            tot.iters += fr.iters; tot.vcycles += fr.vcycles; tot.applyA += fr.applyA;
            fallback_solves++;
            rn = fr.relres * fr.relres * rn;
        }
        *outer_out += outer;
        tot.relres = std::sqrt(rn / (rn0 > 1e-300 ? rn0 : 1e-300));
        return tot;
    }

    double compliance() {
        Ctx& C = *c;
        hipLaunchKernelGGL((k_dot<double, double>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, F, U, ndof * cfg.nload,
                           C.dsc + 2, C.partials, C.ticket);
        return C.read_d(C.dsc + 2);
    }

    int oc_run(const float* dx, const float* ddcn, float* dxnew, double volfrac, double* vol, int mech = 0) {
        return oc_bisect(*c, dx, ddcn, pres, voidm, cfg.move, nel, volfrac, mech, dxnew, vol);
    }

    void filter_fwd(int type, const float* dxd, const float* din, float* dout) {
        Ctx& C = *c;
        if (type == TOPO_FILTER_SENSITIVITY)
            hipLaunchKernelGGL(k_sens_filter, dim3(nblk(nel)), dim3(256), 0, C.s, dxd, din, Hs, dout, g,
                               (const FilterStencil*)stencil);
        else
            hipLaunchKernelGGL(k_density_filter, dim3(nblk(nel)), dim3(256), 0, C.s, din, dout, sw, g, 0,
                               (const FilterStencil*)stencil);
        C.check_launch();
    }

    // SIMP iteration through the FP64 forward/backward pair: used for the
    // mechanism objective and for the Helmholtz filter (no reference
    // counterpart). The gradient is cast to float for the OC update (the
    // mechanism's damped OC, or the reference OC for compliance); with the
    // sensitivity filter it is filtered as in the reference first.
    void step_fb(topo_step_info* info) {
        Ctx& C = *c;
        auto t0 = std::chrono::steady_clock::now();
        topo_step_info loc;
        std::memset(&loc, 0, sizeof loc);
        if (cancel_flag) throw TopoError{TOPO_E_CANCELLED};
        loop++;
        TOPO_HIP(hipMemcpyAsync(xold, x, nel * sizeof(float), hipMemcpyDeviceToDevice, C.s));
        double vphys = 0;
        auto ts = std::chrono::steady_clock::now();
        const double J = forward(loop, &loc, &vphys);
        auto te = std::chrono::steady_clock::now();
        backward();
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(nel)), dim3(256), 0, C.s, dc, (const double*)g64, nel);
        if (cfg.filter_type == TOPO_FILTER_SENSITIVITY)
            hipLaunchKernelGGL(k_sens_filter, dim3(nblk(nel)), dim3(256), 0, C.s, x, dc, Hs, dcn, g,
                               (const FilterStencil*)stencil);
        else
            TOPO_HIP(hipMemcpyAsync(dcn, dc, nel * sizeof(float), hipMemcpyDeviceToDevice, C.s));
        C.check_launch();
        double vol = 0;
        loc.oc_passes = oc_run(x, dcn, xnew, cfg.volfrac, &vol, cfg.objective == TOPO_OBJ_MECHANISM ? 1 : 0);
        hipLaunchKernelGGL(k_change, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, xnew, xold, nel, C.dsc + 4, C.partials,
                           C.ticket, 0);
        C.check_launch();
        TOPO_HIP(hipMemcpyAsync(C.h_sc, C.dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, C.s));
        C.sync();
        const double change = C.h_sc[4];
        const double volume = C.h_sc[5];
        std::swap(x, xnew);
        fb_ready = false;
        auto t1 = std::chrono::steady_clock::now();
        loc.iteration = loop;
        loc.compliance = J;
        loc.change = change;
        loc.penal = fb_penal;
        loc.beta = fb_beta;
        loc.volume = volume / (double)nel;
        loc.t_solve_ms = std::chrono::duration<double, std::milli>(te - ts).count();
        loc.t_total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        loc.t_post_ms = std::chrono::duration<double, std::milli>(t1 - te).count();
        if (info) *info = loc;
    }

    // one SIMP iteration (optimize() loop body js/optimizer-worker-3d.js:2124-2252)
    void step(topo_step_info* info) {
        if (cfg.objective == TOPO_OBJ_MECHANISM || cfg.filter_type == TOPO_FILTER_HELMHOLTZ) {
            step_fb(info);
            return;
        }
        Ctx& C = *c;
        auto t0 = std::chrono::steady_clock::now();
        topo_step_info loc;
        std::memset(&loc, 0, sizeof loc);
        if (cancel_flag) throw TopoError{TOPO_E_CANCELLED};
        loop++;
        TOPO_HIP(hipMemcpyAsync(xold, x, nel * sizeof(float), hipMemcpyDeviceToDevice, C.s));
        const double penal = cfg.penal_start + (cfg.penal - cfg.penal_start) *
                                                   std::min(1.0, (double)(loop - 1) / std::max(1, cfg.continuation_iters));
        const bool proj = cfg.use_projection != 0;
        double beta = 1.0;
        if (proj)
            beta = cfg.beta_fixed > 0 ? (double)cfg.beta_fixed
                                      : std::min(cfg.beta_max, std::pow(2.0, std::floor((double)(loop - 1) / cfg.beta_interval)));
        const float* xd = x;
        if (cfg.filter_type == TOPO_FILTER_DENSITY) {
            filter_fwd(TOPO_FILTER_DENSITY, nullptr, x, xf);
            xd = xf;
        }
        const float* xp = xd;
        if (proj) {
            hipLaunchKernelGGL(k_heaviside, dim3(nblk(nel)), dim3(256), 0, C.s, xd, xphys, nel, beta, cfg.eta);
            xp = xphys;
        } else if (xd != xphys) {
            TOPO_HIP(hipMemcpyAsync(xphys, xd, nel * sizeof(float), hipMemcpyDeviceToDevice, C.s));
        }
        auto tsu = std::chrono::steady_clock::now();
        set_density_dev(xp, penal);
        C.sync();
        auto ts = std::chrono::steady_clock::now();
        fe(loop, cfg.max_iterations, &loc);
        C.sync();
        auto te = std::chrono::steady_clock::now();
        last_c = compliance();
        // energies + sensitivities (first load case carries the energies for
        // nload=1; for several cases the sensitivities are summed)
        for (int k = 0; k < cfg.nload; k++) {
            hipLaunchKernelGGL(k_energy_sens, dim3(nblk(nel)), dim3(256), 0, C.s, U + (int64_t)k * ndof, xp, xd, penal,
                               cfg.E0, cfg.Emin, (int)proj, beta, cfg.eta, g, w, stress, k == 0 ? dc : tmpf,
                               (const double*)K0.ke, cfg.energy_clamp);
            if (k > 0) hipLaunchKernelGGL(k_add_f, dim3(nblk(nel)), dim3(256), 0, C.s, dc, (const float*)tmpf, nel);
        }
        C.check_launch();
        if (cfg.filter_type == TOPO_FILTER_DENSITY)
            hipLaunchKernelGGL(k_density_filter, dim3(nblk(nel)), dim3(256), 0, C.s, dc, dcn, sw, g, 1,
                               (const FilterStencil*)stencil);
        else
            hipLaunchKernelGGL(k_sens_filter, dim3(nblk(nel)), dim3(256), 0, C.s, x, dc, Hs, dcn, g,
                               (const FilterStencil*)stencil);
        C.check_launch();
        double vol = 0;
        loc.oc_passes = oc_run(x, dcn, xnew, cfg.volfrac, &vol);
        hipLaunchKernelGGL(k_change, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, xnew, xold, nel, C.dsc + 4, C.partials,
                           C.ticket, 0);
        C.check_launch();
        TOPO_HIP(hipMemcpyAsync(C.h_sc, C.dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, C.s));
        C.sync();
        const double change = C.h_sc[4];
        const double volume = C.h_sc[5];
        std::swap(x, xnew);
        // max stress
        hipLaunchKernelGGL(k_change, dim3(kRedBlocks), dim3(kRedThreads), 0, C.s, stress, (const float*)nullptr, nel,
                           C.dsc + 6, C.partials, C.ticket, 1);
        TOPO_HIP(hipMemcpyAsync(C.h_sc, C.dsc, 8 * sizeof(double), hipMemcpyDeviceToHost, C.s));
        C.sync();
        auto t1 = std::chrono::steady_clock::now();
        loc.iteration = loop;
        loc.compliance = last_c;
        loc.change = change;
        loc.penal = penal;
        loc.beta = beta;
        loc.volume = volume / (double)nel;
        loc.max_stress = C.h_sc[6];
        loc.t_solve_ms = std::chrono::duration<double, std::milli>(te - ts).count();
        loc.t_total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        loc.t_setup_ms = std::chrono::duration<double, std::milli>(ts - tsu).count();
        loc.t_post_ms = std::chrono::duration<double, std::milli>(t1 - te).count();
        if (info) *info = loc;
    }
};

}  // namespace topo

using namespace topo;

struct topo_solver {
    Solver s;
};

// ---------------------------------------------------------------------------
extern "C" {

int topo_solver_create(const topo_config* cfg, const uint8_t* fixedMask, const float* F, const uint8_t* preserved,
                       const uint8_t* voidm, topo_solver** out) {
    TOPO_API_BEGIN
    TOPO_CHECK(cfg && out, TOPO_E_ARG, "topo_solver_create: null argument");
    std::unique_ptr<topo_solver> h(new topo_solver());
    h->s.init(cfg, fixedMask, F, preserved, voidm);
    *out = h.release();
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_destroy(topo_solver* s) {
    delete s;
    return TOPO_OK;
}

int topo_solver_set_design(topo_solver* h, const float* x) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && x, TOPO_E_ARG, "null argument");
    h->s.set_design(x);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_get_design(topo_solver* h, float* x) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && x, TOPO_E_ARG, "null argument");
    h->s.get_field_f(h->s.x, x);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_set_iteration(topo_solver* h, int32_t loop) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && loop >= 0, TOPO_E_ARG, "bad argument");
    h->s.loop = loop;
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_update(topo_solver* h, double penal, double rmin, double volfrac, int32_t max_iterations) {
    TOPO_API_BEGIN
    TOPO_CHECK(h, TOPO_E_ARG, "null solver");
    Solver& S = h->s;
    if (penal > 0) S.cfg.penal = penal;
    if (volfrac > 0) S.cfg.volfrac = volfrac;
    if (max_iterations > 0) S.cfg.max_iterations = max_iterations;
    if (rmin > 0 && rmin != S.cfg.rmin) {
        S.cfg.rmin = rmin;
        S.setup_filter(rmin);
    }
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_step(topo_solver* h, topo_step_info* info) {
    TOPO_API_BEGIN
    TOPO_CHECK(h, TOPO_E_ARG, "null solver");
    h->s.step(info);
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_get_field(topo_solver* h, int32_t field, void* out) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && out, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    switch (field) {
    case TOPO_FIELD_U:
        TOPO_HIP(hipMemcpyAsync(out, S.U, S.ndof * S.cfg.nload * sizeof(double), hipMemcpyDeviceToHost, S.c->s));
        S.c->sync();
        break;
    case TOPO_FIELD_XPHYS: S.get_field_f(S.xphys, (float*)out); break;
    case TOPO_FIELD_ENERGY: S.get_field_f(S.w, (float*)out); break;
    case TOPO_FIELD_DC: S.get_field_f(S.dc, (float*)out); break;
    case TOPO_FIELD_DCN: S.get_field_f(S.dcn, (float*)out); break;
    case TOPO_FIELD_STRESS: S.get_field_f(S.stress, (float*)out); break;
    case TOPO_FIELD_ADJOINT:
        TOPO_CHECK(S.adj, TOPO_E_ARG, "no adjoint field (mechanism objective only)");
        TOPO_HIP(hipMemcpyAsync(out, S.adj, S.ndof * sizeof(double), hipMemcpyDeviceToHost, S.c->s));
        S.c->sync();
        break;
    default: TOPO_CHECK(false, TOPO_E_ARG, "unknown field %d", field);
    }
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_tuning(topo_solver* h, int32_t* rhs_tile, int32_t* xm_nt, int32_t* xm_cx, double* log,
                       int32_t max_entries) {
    TOPO_API_BEGIN
    TOPO_CHECK(h, TOPO_E_ARG, "null solver");
    Solver& S = h->s;
    if (rhs_tile) *rhs_tile = S.batch_cap > 1 ? S.rhs_tile : 1;
    if (xm_nt) *xm_nt = S.mg.xm_nt ? S.mg.xm_nt : S.tuned_nt[0];
    if (xm_cx) *xm_cx = S.mg.xm_cx ? S.mg.xm_cx : S.tuned_cx[0];
    const int n = (int)S.tune_log.size();
    if (log)
        for (int i = 0; i < n && i < max_entries; i++) {
            log[4 * i] = S.tune_log[i].what;
            log[4 * i + 1] = S.tune_log[i].nt;
            log[4 * i + 2] = S.tune_log[i].cx;
            log[4 * i + 3] = S.tune_log[i].ms;
        }
    return n;
    TOPO_API_END
}

int topo_solver_forward(topo_solver* h, double* objective, double* volume, topo_step_info* info) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && objective, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    topo_step_info loc;
    std::memset(&loc, 0, sizeof loc);
    double vol = 0;
    *objective = S.forward(std::max(1, S.loop), &loc, &vol);
    loc.compliance = *objective;
    loc.penal = S.fb_penal;
    loc.beta = S.fb_beta;
    loc.volume = vol;
    if (volume) *volume = vol;
    if (info) *info = loc;
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_backward(topo_solver* h, double* dJdx, double* dVdx) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && dJdx, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    Ctx& C = *S.c;
    S.backward();
    for (int q = 0; q < 2; q++) {
        double* dst = q == 0 ? dJdx : dVdx;
        if (!dst) continue;
        hipLaunchKernelGGL(k_zyx_to_ref_t<double>, dim3(nblk(S.nel)), dim3(256), 0, C.s,
                           (const double*)(q == 0 ? S.g64 : S.gv64), S.t64, S.g);
        TOPO_HIP(hipMemcpyAsync(dst, S.t64, S.nel * sizeof(double), hipMemcpyDeviceToHost, C.s));
        C.sync();
    }
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_helmholtz(topo_solver* h, const double* in, double* out, int32_t* iters) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && in && out, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_CHECK(S.cfg.filter_type == TOPO_FILTER_HELMHOLTZ, TOPO_E_ARG, "solver was not created with the Helmholtz filter");
    Ctx& C = *S.c;
    TOPO_HIP(hipMemcpyAsync(S.t64, in, S.nel * sizeof(double), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_ref_to_zyx_t<double>, dim3(nblk(S.nel)), dim3(256), 0, C.s, (const double*)S.t64, S.g64, S.g);
    S.helm_solve(S.g64, S.gv64);
    hipLaunchKernelGGL(k_zyx_to_ref_t<double>, dim3(nblk(S.nel)), dim3(256), 0, C.s, (const double*)S.gv64, S.t64, S.g);
    TOPO_HIP(hipMemcpyAsync(out, S.t64, S.nel * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    S.fb_ready = false;
    if (iters) *iters = S.helm_iters;
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_set_density(topo_solver* h, const float* xPhys, double penal) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && xPhys, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_HIP(hipMemcpyAsync(S.tmpf, xPhys, S.nel * sizeof(float), hipMemcpyHostToDevice, S.c->s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, S.c->s, S.tmpf, S.xphys, S.g);
    S.set_density_dev(S.xphys, penal);
    S.c->sync();
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_fe(topo_solver* h, const float* xPhys, double penal, int32_t iteration, int32_t max_iterations,
                   double* Uio, double* cout, topo_step_info* info) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && xPhys && Uio, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_HIP(hipMemcpyAsync(S.tmpf, xPhys, S.nel * sizeof(float), hipMemcpyHostToDevice, S.c->s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, S.c->s, S.tmpf, S.xphys, S.g);
    S.set_density_dev(S.xphys, penal);
    TOPO_HIP(hipMemcpyAsync(S.U, Uio, S.ndof * S.cfg.nload * sizeof(double), hipMemcpyHostToDevice, S.c->s));
    S.have_U = true;
    topo_step_info loc;
    std::memset(&loc, 0, sizeof loc);
    S.fe(iteration, max_iterations, &loc);
    TOPO_HIP(hipMemcpyAsync(Uio, S.U, S.ndof * S.cfg.nload * sizeof(double), hipMemcpyDeviceToHost, S.c->s));
    S.c->sync();
    const double cc = S.compliance();
    if (cout) *cout = cc;
    loc.compliance = cc;
    if (info) *info = loc;
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_applyA(topo_solver* h, int32_t level, int32_t prec, const double* p, double* Ap) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && p && Ap, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_CHECK(level >= 0 && level < (int)S.mg.L.size(), TOPO_E_ARG, "level %d out of range", level);
    TOPO_CHECK(prec == TOPO_PREC_FP64 || (prec == TOPO_PREC_MIXED && S.cfg.precision != TOPO_PREC_FP64), TOPO_E_ARG,
               "FP32 operator needs a mixed-precision solver");
    MGLevel& l = S.mg.L[level];
    Ctx& C = *S.c;
    if (prec == TOPO_PREC_FP64) {
        TOPO_HIP(hipMemcpyAsync(l.x, p, l.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
        S.mg.apply<double>(C, level, l.x, l.x2, 0, nullptr, nullptr);
        TOPO_HIP(hipMemcpyAsync(Ap, l.x2, l.ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    } else {
        TOPO_HIP(hipMemcpyAsync(l.x, p, l.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, l.x32, l.x, l.ndof);
        S.mg.apply<float>(C, level, l.x32, l.x232, 0, nullptr, nullptr);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, l.x2, l.x232, l.ndof);
        TOPO_HIP(hipMemcpyAsync(Ap, l.x2, l.ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    }
    C.sync();
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_vcycle(topo_solver* h, int32_t prec, const double* r, double* z) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && r && z, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    Ctx& C = *S.c;
    MGLevel& l = S.mg.L[0];
    double* rd = (double*)S.v64.r;
    TOPO_HIP(hipMemcpyAsync(rd, r, l.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    if (prec == TOPO_PREC_FP64) {
        double* zz = S.mg.vcycle<double>(C, 0, rd, nullptr);
        TOPO_HIP(hipMemcpyAsync(z, zz, l.ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    } else {
        TOPO_CHECK(S.cfg.precision != TOPO_PREC_FP64, TOPO_E_ARG, "FP32 V-cycle needs a mixed-precision solver");
        float* r32 = (float*)S.v32.r;
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, r32, rd, l.ndof);
        float* zz = S.mg.vcycle<float>(C, 0, r32, nullptr);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, (double*)S.v64.z, zz,
                           l.ndof);
        TOPO_HIP(hipMemcpyAsync(z, S.v64.z, l.ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    }
    C.sync();
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_energy(topo_solver* h, const double* Uh, float* wout) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && Uh && wout, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    Ctx& C = *S.c;
    double* Ud = (double*)S.v64.z;
    TOPO_HIP(hipMemcpyAsync(Ud, Uh, S.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_energy_sens, dim3(nblk(S.nel)), dim3(256), 0, C.s, Ud, S.xphys, S.xphys, 3.0, S.cfg.E0,
                       S.cfg.Emin, 0, 1.0, 0.5, S.g, S.w, (float*)nullptr, S.dc, (const double*)S.K0.ke,
                       S.cfg.energy_clamp);
    S.get_field_f(S.w, wout);
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_filter(topo_solver* h, int32_t type, const float* xh, const float* in, float* out) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && in && out, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_CHECK(S.cfg.filter_type != TOPO_FILTER_HELMHOLTZ, TOPO_E_ARG, "use topo_op_helmholtz for the Helmholtz filter");
    TOPO_CHECK(type == (S.cfg.filter_type == TOPO_FILTER_DENSITY ? TOPO_FILTER_DENSITY : TOPO_FILTER_SENSITIVITY),
               TOPO_E_ARG, "filter type %d does not match the solver's filter", type);
    TOPO_CHECK(type != TOPO_FILTER_SENSITIVITY || xh, TOPO_E_ARG, "sensitivity filter needs x");
    Ctx& C = *S.c;
    if (xh) {
        TOPO_HIP(hipMemcpyAsync(S.tmpf, xh, S.nel * sizeof(float), hipMemcpyHostToDevice, C.s));
        hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.tmpf, S.xnew, S.g);
    }
    TOPO_HIP(hipMemcpyAsync(S.tmpf, in, S.nel * sizeof(float), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.tmpf, S.dc, S.g);
    S.filter_fwd(type, S.xnew, S.dc, S.dcn);
    S.get_field_f(S.dcn, out);
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_filter_adjoint(topo_solver* h, int32_t type, const float* in, float* out) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && in && out, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    TOPO_CHECK(type == TOPO_FILTER_DENSITY && S.cfg.filter_type == TOPO_FILTER_DENSITY, TOPO_E_ARG,
               "adjoint defined for the density filter");
    Ctx& C = *S.c;
    TOPO_HIP(hipMemcpyAsync(S.tmpf, in, S.nel * sizeof(float), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.tmpf, S.dc, S.g);
    hipLaunchKernelGGL(k_density_filter, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.dc, S.dcn, S.sw, S.g, 1,
                       (const FilterStencil*)S.stencil);
    S.get_field_f(S.dcn, out);
    return TOPO_OK;
    TOPO_API_END
}

int topo_op_oc(topo_solver* h, const float* xh, const float* dcnh, float* xnewh, double volfrac) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && xh && dcnh && xnewh, TOPO_E_ARG, "null argument");
    Solver& S = h->s;
    Ctx& C = *S.c;
    TOPO_HIP(hipMemcpyAsync(S.tmpf, xh, S.nel * sizeof(float), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.tmpf, S.xold, S.g);
    TOPO_HIP(hipMemcpyAsync(S.tmpf, dcnh, S.nel * sizeof(float), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_ref_to_zyx, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.tmpf, S.dc, S.g);
    const int passes = S.oc_run(S.xold, S.dc, S.xnew, volfrac, nullptr);
    S.get_field_f(S.xnew, xnewh);
    return passes;
    TOPO_API_END
}

int topo_time_op(topo_solver* h, int32_t op, int32_t prec, int32_t level, int32_t reps, double* avg_ms, double* bytes) {
    TOPO_API_BEGIN
    TOPO_CHECK(h && avg_ms && reps > 0, TOPO_E_ARG, "bad argument");
    Solver& S = h->s;
    Ctx& C = *S.c;
    TOPO_CHECK(level >= 0 && level < (int)S.mg.L.size(), TOPO_E_ARG, "level out of range");
    MGLevel& l = S.mg.L[level];
    const bool f32 = prec == TOPO_PREC_MIXED;
    TOPO_CHECK(!f32 || S.cfg.precision != TOPO_PREC_FP64, TOPO_E_ARG, "FP32 timing needs a mixed solver");
    hipEvent_t e0, e1;
    TOPO_HIP(hipEventCreate(&e0));
    TOPO_HIP(hipEventCreate(&e1));
    double by = 0;
    auto run = [&]() {
        switch (op) {
        case TOPO_TIME_APPLYA:
            if (f32) S.mg.apply<float>(C, level, l.x32, l.x232, 0, nullptr, nullptr);
            else S.mg.apply<double>(C, level, l.x, l.x2, 0, nullptr, nullptr);
            break;
        case TOPO_TIME_APPLYA_MASKED:
            // the CG's operator: masked, with the p.Ap partials
            if (f32) S.mg.apply<float>(C, level, l.x32, l.x232, 1, nullptr, nullptr, 1, C.partials);
            else S.mg.apply<double>(C, level, l.x, l.x2, 1, nullptr, nullptr, 1, C.partials);
            break;
        case TOPO_TIME_APPLYA_MFMA: {
            TOPO_CHECK(f32 && S.mf_ok, TOPO_E_UNSUPPORTED, "matrix-core applyA needs a mixed solver and its LDS");
            const bool keep = S.mg.mfma;
            S.mg.mfma = true;
            S.mg.apply<float>(C, level, l.x32, l.x232, 0, nullptr, nullptr);
            S.mg.mfma = keep;
            break;
        }
        case TOPO_TIME_VCYCLE:
            if (f32) S.mg.vcycle<float>(C, 0, S.mg.L[0].b32, nullptr);
            else S.mg.vcycle<double>(C, 0, S.mg.L[0].b, nullptr);
            break;
        case TOPO_TIME_ENERGY:
            hipLaunchKernelGGL(k_energy_sens, dim3(nblk(S.nel)), dim3(256), 0, C.s, S.U, S.xphys, S.xphys, 3.0, S.cfg.E0,
                               S.cfg.Emin, 0, 1.0, 0.5, S.g, S.w, S.stress, S.dc, (const double*)S.K0.ke,
                               S.cfg.energy_clamp);
            break;
        case TOPO_TIME_FILTER:
            TOPO_CHECK(S.cfg.filter_type != TOPO_FILTER_HELMHOLTZ, TOPO_E_ARG, "no timing op for the Helmholtz filter");
            S.filter_fwd(S.cfg.filter_type == TOPO_FILTER_DENSITY ? TOPO_FILTER_DENSITY : TOPO_FILTER_SENSITIVITY, S.x,
                         S.dc, S.dcn);
            break;
        case TOPO_TIME_COPY:
            hipLaunchKernelGGL((k_convert<double, double>), dim3(nblk(l.ndof)), dim3(256), 0, C.s, l.x2, (const double*)l.x,
                               l.ndof);
            break;
        default: TOPO_CHECK(false, TOPO_E_ARG, "unknown op %d", op);
        }
    };
    // algorithmic bytes of one launch
    const double sv = f32 ? 4.0 : 8.0;
    if (op == TOPO_TIME_APPLYA || op == TOPO_TIME_APPLYA_MFMA) by = 2.0 * l.ndof * sv + (double)l.g.nel() * sv;
    else if (op == TOPO_TIME_APPLYA_MASKED) by = 2.0 * l.ndof * sv + (double)l.g.nel() * sv + (double)l.ndof;
    else if (op == TOPO_TIME_COPY) by = 2.0 * l.ndof * 8.0;
    else if (op == TOPO_TIME_ENERGY) by = (double)S.ndof * 8 + (double)S.nel * 4 * 4;
    else if (op == TOPO_TIME_FILTER) by = (double)S.nel * 4 * 3;
    run();
    C.sync();
    TOPO_HIP(hipEventRecord(e0, C.s));
    for (int i = 0; i < reps; i++) run();
    TOPO_HIP(hipEventRecord(e1, C.s));
    TOPO_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TOPO_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    if (bytes) *bytes = by;
    return TOPO_OK;
    TOPO_API_END
}

int topo_solver_cancel(topo_solver* h) {
    if (!h) return TOPO_E_ARG;
    h->s.cancel_flag = 1;
    return TOPO_OK;
}

}  // extern "C"

// the 2-D SIMP path (quad4 operator, 2-D solver handle)
#include "solver2d.hpp"

// ---------------------------------------------------------------------------
// WASM-op twins. One process-wide context, serialised by a mutex (the
// AssemblyScript exports are single-threaded and not re-entrant either).
#include <mutex>

namespace topo {

// ---------------------------------------------------------------------------
// Scratch of the allocation-free ebePCG (topo_ebePCG_noalloc): one
// file-scope device array, byte offsets below. Each GPU holds its own copy
// for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.
constexpr int64_t kNaMaxDof = 262144;   // DOFs (3 x nodes), e.g. a 40 x 40 x 40 grid has 206,763
constexpr int64_t kNaMaxNel = 131072;   // elements
constexpr size_t kNaAlign = 256;
constexpr size_t na_up(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaOffKe = 0;
constexpr size_t kNaOffDens = na_up(kNaOffKe + 576 * sizeof(double));
constexpr size_t kNaOffDensZ = na_up(kNaOffDens + kNaMaxNel * sizeof(double));
constexpr size_t kNaOffE = na_up(kNaOffDensZ + kNaMaxNel * sizeof(double));
constexpr size_t kNaOffVec = na_up(kNaOffE + kNaMaxNel * sizeof(double));       // F U invD r z p Ap
constexpr size_t kNaOffPart = na_up(kNaOffVec + 7 * kNaMaxDof * sizeof(double));
constexpr size_t kNaPartDoubles = 3 * (size_t)kCgRows * kCgStride > 32 * (size_t)kRedBlocks
                                      ? 3 * (size_t)kCgRows * kCgStride : 32 * (size_t)kRedBlocks;
constexpr size_t kNaOffDsc = na_up(kNaOffPart + kNaPartDoubles * sizeof(double));
constexpr size_t kNaOffSt = na_up(kNaOffDsc + 16 * sizeof(double));
constexpr size_t kNaOffOc = na_up(kNaOffSt + kMaxRHS * sizeof(CGState));
constexpr size_t kNaOffTicket = na_up(kNaOffOc + sizeof(OCState));
constexpr size_t kNaOffFixed = na_up(kNaOffTicket + 4 * sizeof(unsigned));
constexpr size_t kNaOffFree = na_up(kNaOffFixed + kNaMaxDof);
constexpr size_t kNaBytes = na_up(kNaOffFree + kNaMaxDof * sizeof(int32_t));

static __device__ __align__(256) unsigned char scratch[kNaBytes];

static std::mutex g_na_mu;
constexpr int kNaMaxDevices = 64;
static unsigned char* g_na_ptr[kNaMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once
// per device; the array's name is not a device address in host code)
static unsigned char* na_scratch_ptr() {
    int dev = 0;
    TOPO_HIP(hipGetDevice(&dev));
    TOPO_CHECK(dev >= 0 && dev < kNaMaxDevices, TOPO_E_ARG, "device %d out of range", dev);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        TOPO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TOPO_CHECK(p != nullptr, TOPO_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_na_ptr[dev] = (unsigned char*)p;
    }
    return g_na_ptr[dev];
}

// fixed[freedofs[i]] = 0 (entries outside [0, ndof) are skipped)
__global__ void k_mark_free(const int32_t* __restrict__ freedofs, int64_t nfree, int64_t ndof,
                            uint8_t* __restrict__ fixed) {
    for (int64_t i = gtid(); i < nfree; i += gstride()) {
        const int64_t d = freedofs[i];
        if (d >= 0 && d < ndof) fixed[d] = 0;
    }
}

struct TwinCtx {
    std::unique_ptr<Ctx> c;
    DevKe K0;
    double last_ke[576];
    bool have_ke = false;
    // structured-grid check cache: pointer, size and a sample of entries
    const int32_t* ed_ptr = nullptr;
    int64_t ed_ndof = 0;
    int32_t ed_sample[64];
    Grid ed_grid;

    void ensure() {
        if (!c) c.reset(new Ctx());
    }
    void set_ke(const double* ke) {
        if (!have_ke || std::memcmp(ke, last_ke, sizeof last_ke) != 0) {
            K0.set(*c, ke);
            std::memcpy(last_ke, ke, sizeof last_ke);
            have_ke = true;
        }
    }
    Grid grid_of(const int32_t* ed, int64_t ndof) {
        TOPO_CHECK(ed, TOPO_E_ARG, "edofs is null");
        int32_t smp[64];
        bool cached = ed == ed_ptr && ndof == ed_ndof;
        if (cached) {
            const int64_t nel = ed_grid.nel();
            for (int i = 0; i < 64; i++) {
                smp[i] = ed[(int64_t)(((double)i / 63.0) * (nel * 24 - 1))];
                if (smp[i] != ed_sample[i]) cached = false;
            }
        }
        if (cached) return ed_grid;
        Grid gg;
        TOPO_CHECK(edofs_structured(ed, ndof, &gg), TOPO_E_UNSTRUCTURED,
                   "edofs is not a structured hex8 grid in the reference numbering");
        ed_ptr = ed;
        ed_ndof = ndof;
        ed_grid = gg;
        const int64_t nel = gg.nel();
        for (int i = 0; i < 64; i++) ed_sample[i] = ed[(int64_t)(((double)i / 63.0) * (nel * 24 - 1))];
        return gg;
    }
};

static std::mutex g_twin_mu;
static TwinCtx g_twin;

// E in zyx order with inactive elements (not in `active`) set to 0
static std::vector<double> masked_E_zyx(const Grid& gg, const double* evals, const int32_t* active, int32_t cnt) {
    const int64_t nel = gg.nel();
    std::vector<double> Ez(nel, 0.0);
    for (int32_t a = 0; a < cnt; a++) {
        const int64_t e = active[a];
        TOPO_CHECK(e >= 0 && e < nel, TOPO_E_ARG, "active element %lld out of range", (long long)e);
        const int64_t ex = e % gg.nelx, t = e / gg.nelx, ey = t % gg.nely, ez = t / gg.nely;
        Ez[(ex * gg.nely + ey) * gg.nelz + ez] = evals[e];
    }
    return Ez;
}

// ebePCG with edofSize 8 (assembly/index.ts:504-733 over the quad4 grid of
// _precomputeEdofs2D): E by Math.pow, active above 1000 Emin, Jacobi
// diagonal, warm start from U at free DOFs, relative tolerance, cgMax =
// min(maxIter, nfree); U returned with fixed DOFs 0.
static int ebepcg_2d(const double* dens, const double* ke, const int32_t* edofs, const double* Fh, double* Uh,
                     const int32_t* freedofs, int32_t nel, int32_t ndof, int32_t nfree, double Emin, double E0,
                     double penal, int32_t maxIter, double tolerance) {
    TOPO_CHECK(edofs, TOPO_E_ARG, "edofs is null");
    Grid2 g;
    TOPO_CHECK(edofs2d_structured(edofs, nel, ndof, &g), TOPO_E_UNSTRUCTURED,
               "edofs is not a structured quad4 grid in the 2-D worker's numbering");
    std::lock_guard<std::mutex> lk(g_twin_mu);
    TwinCtx& T = g_twin;
    T.ensure();
    Ctx& C = *T.c;
    std::vector<uint8_t> fixedh(ndof, 1);
    for (int32_t i = 0; i < nfree; i++) {
        TOPO_CHECK(freedofs[i] >= 0 && freedofs[i] < ndof, TOPO_E_ARG, "freedofs entry out of range");
        fixedh[freedofs[i]] = 0;
    }
    std::vector<void*> tmp;
    auto A = [&](size_t bytes) {
        void* q = C.alloc<uint8_t>((int64_t)bytes);
        tmp.push_back(q);
        return q;
    };
    uint8_t* dfixed = (uint8_t*)A(ndof);
    double* dke = (double*)A(64 * sizeof(double));
    double* ddens = (double*)A(nel * sizeof(double));
    double* dE = (double*)A(nel * sizeof(double));
    double* dinvD = (double*)A(ndof * sizeof(double));
    double* dF = (double*)A(ndof * sizeof(double));
    double* dU = (double*)A(ndof * sizeof(double));
    CGVecs v;
    v.r = A(ndof * sizeof(double)); v.z = A(ndof * sizeof(double));
    v.p = A(ndof * sizeof(double)); v.Ap = A(ndof * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(dfixed, fixedh.data(), ndof, hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dke, ke, 64 * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(ddens, dens, nel * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dF, Fh, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dU, Uh, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL(k_q4_moduli<double>, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)ddens, penal, E0, Emin,
                       Emin * 1000.0, dE, (int64_t)nel);
    hipLaunchKernelGGL(k_q4_invdiag, dim3(nblk(g.nnodes())), dim3(256), 0, C.s, (const double*)dE,
                       (const uint8_t*)dfixed, dinvD, g, (const double*)dke);
    // the warm start reads only free DOFs (Uf = U[free], :583-587)
    hipLaunchKernelGGL(k_mask<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, dU, (const uint8_t*)dfixed, (int64_t)ndof);
    C.check_launch();
    auto op = [&](const double* p, double* out, int mode, const double* b, const int* gate, double* pap) {
        return q4_apply(C.s, mode, p, out, dE, g, dfixed, b, gate, dke, pap);
    };
    const int cgMax = maxIter < nfree ? maxIter : nfree;
    CGResult r = pcg_jacobi_op<double, double>(C, op, dfixed, (const double*)dinvD, ndof, (const double*)dF, dU, true,
                                               v, tolerance, cgMax, nullptr,
                                               {&T, dE, dinvD, dfixed, dke, (const void*)(intptr_t)g.nelx,
                                                (const void*)(intptr_t)g.nely}, 32);
    TOPO_HIP(hipMemcpyAsync(Uh, dU, ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    for (void* q : tmp) C.free_one(q);
    TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
    return r.iters;
}

}  // namespace topo

extern "C" {

int topo_applyAEbe3D(const double* ke, const int32_t* edofs, const double* evals, const int32_t* active,
                     int32_t activeCount, const double* p, double* Ap, int32_t ndof, double* scratch) {
    (void)scratch;
    TOPO_API_BEGIN
    TOPO_CHECK(ke && evals && p && Ap && (activeCount == 0 || active) && activeCount >= 0 && ndof > 0, TOPO_E_ARG,
               "topo_applyAEbe3D: bad argument");
    std::lock_guard<std::mutex> lk(g_twin_mu);
    TwinCtx& T = g_twin;
    T.ensure();
    Ctx& C = *T.c;
    const Grid gg = T.grid_of(edofs, ndof);
    T.set_ke(ke);
    std::vector<double> Ez = masked_E_zyx(gg, evals, active, activeCount);
    double* dE = C.alloc<double>(gg.nel());
    double* dp = C.alloc<double>(ndof);
    double* dAp = C.alloc<double>(ndof);
    TOPO_HIP(hipMemcpyAsync(dE, Ez.data(), Ez.size() * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dp, p, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    hipLaunchKernelGGL((k_applyA_exact<double, double>), dim3(nblk(gg.nnodes())), dim3(256), 0, C.s, (const double*)dp,
                       dAp, (const double*)dE, 0.0, gg, (const uint8_t*)nullptr, 0, (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, (const int*)nullptr, (const double*)T.K0.ke);
    C.check_launch();
    TOPO_HIP(hipMemcpyAsync(Ap, dAp, ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    C.free_one(dE);
    C.free_one(dp);
    C.free_one(dAp);
    return TOPO_OK;
    TOPO_API_END
}

int topo_ebePCG(const double* dens, const double* ke, const int32_t* edofs, const double* Fh, double* Uh,
                const int32_t* freedofs, int32_t nel, int32_t edofSize, int32_t ndof, int32_t nfree, double Emin,
                double E0, double penal, int32_t maxIter, double tolerance, double* work) {
    (void)work;
    TOPO_API_BEGIN
    TOPO_CHECK(dens && ke && Fh && Uh && (nfree == 0 || freedofs) && ndof > 0 && nel > 0 && nfree >= 0 && nfree <= ndof,
               TOPO_E_ARG, "topo_ebePCG: bad argument");
    TOPO_CHECK(edofSize == 24 || edofSize == 8, TOPO_E_UNSUPPORTED,
               "topo_ebePCG: edofSize %d (24: 3-D hex8, 8: 2-D quad4)", edofSize);
    if (edofSize == 8)
        return ebepcg_2d(dens, ke, edofs, Fh, Uh, freedofs, nel, ndof, nfree, Emin, E0, penal, maxIter, tolerance);
    std::lock_guard<std::mutex> lk(g_twin_mu);
    TwinCtx& T = g_twin;
    T.ensure();
    Ctx& C = *T.c;
    const Grid gg = T.grid_of(edofs, ndof);
    TOPO_CHECK(gg.nel() == nel, TOPO_E_ARG, "nel %d does not match edofs (%lld)", nel, (long long)gg.nel());
    T.set_ke(ke);
    std::vector<uint8_t> fixedh(ndof, 1);
    for (int32_t i = 0; i < nfree; i++) {
        TOPO_CHECK(freedofs[i] >= 0 && freedofs[i] < ndof, TOPO_E_ARG, "freedofs entry out of range");
        fixedh[freedofs[i]] = 0;
    }
    // densities to zyx order
    std::vector<double> dz(nel);
    for (int64_t e = 0; e < nel; e++) {
        const int64_t ex = e % gg.nelx, t = e / gg.nelx, ey = t % gg.nely, ez = t / gg.nely;
        dz[(ex * gg.nely + ey) * gg.nelz + ez] = dens[e];
    }
    std::vector<void*> tmp;
    auto A = [&](size_t bytes) {
        void* q = C.alloc<uint8_t>((int64_t)bytes);
        tmp.push_back(q);
        return q;
    };
    uint8_t* dfixed = (uint8_t*)A(ndof);
    double* ddens = (double*)A(nel * sizeof(double));
    double* dE = (double*)A(nel * sizeof(double));
    double* dF = (double*)A(ndof * sizeof(double));
    double* dU = (double*)A(ndof * sizeof(double));
    CGVecs v;
    v.r = A(ndof * sizeof(double)); v.z = A(ndof * sizeof(double));
    v.p = A(ndof * sizeof(double)); v.Ap = A(ndof * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(dfixed, fixedh.data(), ndof, hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(ddens, dz.data(), nel * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dF, Fh, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    TOPO_HIP(hipMemcpyAsync(dU, Uh, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    // E via Math.pow (assembly/index.ts:541-550)
    hipLaunchKernelGGL(k_moduli<double>, dim3(nblk(nel)), dim3(256), 0, C.s, (const double*)ddens, penal, E0, Emin, 1, dE,
                       (float*)nullptr, (int64_t)nel);
    // the warm start reads only free DOFs (Uf = U[free], :583-587)
    hipLaunchKernelGGL(k_mask<double>, dim3(nblk(ndof)), dim3(256), 0, C.s, dU, (const uint8_t*)dfixed, (int64_t)ndof);
    MGH mg;
    mg.K0 = &T.K0;
    mg.thr = Emin * 1000.0;
    mg.exact = true;   // reference summation order (drop-in)
    mg.build(C, gg, dfixed, 1, false);
    mg.L[0].E = dE;
    mg.update(C);
    const int cgMax = maxIter < nfree ? maxIter : nfree;
    CGResult r = pcg_jacobi<double, double>(C, mg, dF, dU, true, v, tolerance, cgMax, nullptr);
    TOPO_HIP(hipMemcpyAsync(Uh, dU, ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    for (MGLevel& l : mg.L) {
        if (l.fixed != dfixed) C.free_one(l.fixed);
        C.free_one(l.invD); C.free_one(l.b); C.free_one(l.x); C.free_one(l.x2); C.free_one(l.res);
    }
    for (void* q : tmp) C.free_one(q);
    TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
    return r.iters;
    TOPO_API_END
}

// ebePCG without allocation: every device buffer (inputs, CG vectors,
// reduction partials, CG state) lives in one file-scope device array; see
// include/topo.h for the contract and the size limits.
int topo_ebePCG_noalloc(const double* dens, const double* ke, const int32_t* edofs, const double* Fh, double* Uh,
                        const int32_t* freedofs, int32_t nel, int32_t edofSize, int32_t ndof, int32_t nfree,
                        double Emin, double E0, double penal, int32_t maxIter, double tolerance) {
    TOPO_API_BEGIN
    TOPO_CHECK(dens && ke && edofs && Fh && Uh && (nfree == 0 || freedofs) && ndof > 0 && nel > 0 && nfree >= 0 &&
                   nfree <= ndof, TOPO_E_ARG, "topo_ebePCG_noalloc: bad argument");
    TOPO_CHECK(edofSize == 24, TOPO_E_UNSUPPORTED, "topo_ebePCG_noalloc: edofSize %d (the 3-D hex8 path, 24)",
               edofSize);
    TOPO_CHECK(ndof <= kNaMaxDof && nel <= kNaMaxNel, TOPO_E_UNSUPPORTED,
               "topo_ebePCG_noalloc: %d DOFs / %d elements exceed the scratch limits (%lld DOFs, %lld elements)", ndof,
               nel, (long long)kNaMaxDof, (long long)kNaMaxNel);
    Grid gg;
    TOPO_CHECK(edofs_structured(edofs, ndof, &gg), TOPO_E_UNSTRUCTURED,
               "edofs is not a structured hex8 grid in the reference numbering");
    TOPO_CHECK(gg.nel() == nel, TOPO_E_ARG, "nel %d does not match edofs (%lld)", nel, (long long)gg.nel());
    for (int32_t i = 0; i < nfree; i++)
        TOPO_CHECK(freedofs[i] >= 0 && freedofs[i] < ndof, TOPO_E_ARG, "freedofs entry out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    unsigned char* base = na_scratch_ptr();
    double* d_ke = (double*)(base + kNaOffKe);
    double* d_dens = (double*)(base + kNaOffDens);
    double* d_densz = (double*)(base + kNaOffDensZ);
    double* d_E = (double*)(base + kNaOffE);
    double* d_F = (double*)(base + kNaOffVec);
    double* d_U = d_F + kNaMaxDof;
    double* d_invD = d_U + kNaMaxDof;
    CGVecs v;
    v.r = d_invD + kNaMaxDof;
    v.z = (double*)v.r + kNaMaxDof;
    v.p = (double*)v.z + kNaMaxDof;
    v.Ap = (double*)v.p + kNaMaxDof;
    double* d_part = (double*)(base + kNaOffPart);
    double* d_dsc = (double*)(base + kNaOffDsc);
    CGState* d_st = (CGState*)(base + kNaOffSt);
    OCState* d_oc = (OCState*)(base + kNaOffOc);
    unsigned* d_ticket = (unsigned*)(base + kNaOffTicket);
    uint8_t* d_fixed = (uint8_t*)(base + kNaOffFixed);
    int32_t* d_free = (int32_t*)(base + kNaOffFree);
    // inputs into the scratch (the reference's order), copies of the caller's arrays
    TOPO_HIP(hipMemcpy(d_ke, ke, 576 * sizeof(double), hipMemcpyHostToDevice));
    TOPO_HIP(hipMemcpy(d_dens, dens, (size_t)nel * sizeof(double), hipMemcpyHostToDevice));
    TOPO_HIP(hipMemcpy(d_F, Fh, (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    TOPO_HIP(hipMemcpy(d_U, Uh, (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    if (nfree > 0) TOPO_HIP(hipMemcpy(d_free, freedofs, (size_t)nfree * sizeof(int32_t), hipMemcpyHostToDevice));
    // what the previous call left: the reduction ticket must start at zero
    hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(64), 0, nullptr, d_ticket, 0u, (int64_t)4, (const int*)nullptr);
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3(nblk(ndof)), dim3(256), 0, nullptr, d_fixed, (uint8_t)1, (int64_t)ndof,
                       (const int*)nullptr);
    if (nfree > 0)
        hipLaunchKernelGGL(k_mark_free, dim3(nblk(nfree)), dim3(256), 0, nullptr, (const int32_t*)d_free,
                           (int64_t)nfree, (int64_t)ndof, d_fixed);
    hipLaunchKernelGGL(k_ref_to_zyx_t<double>, dim3(nblk(nel)), dim3(256), 0, nullptr, (const double*)d_dens, d_densz,
                       gg);
    // E via Math.pow (assembly/index.ts:541-550), the same kernels as topo_ebePCG
    hipLaunchKernelGGL(k_moduli<double>, dim3(nblk(nel)), dim3(256), 0, nullptr, (const double*)d_densz, penal, E0,
                       Emin, 1, d_E, (float*)nullptr, (int64_t)nel);
    hipLaunchKernelGGL(k_mask<double>, dim3(nblk(ndof)), dim3(256), 0, nullptr, d_U, (const uint8_t*)d_fixed,
                       (int64_t)ndof);
    const double thr = Emin * 1000.0;
    hipLaunchKernelGGL(k_invdiag, dim3(nblk(gg.nnodes())), dim3(256), 0, nullptr, (const double*)d_E, thr,
                       (const uint8_t*)d_fixed, d_invD, (float*)nullptr, gg, (const double*)d_ke);
    TOPO_HIP(hipGetLastError());
    static CGState h_st[kMaxRHS];
    static OCState h_oc;
    static double h_sc[16];
    Ctx C(Ctx::Borrowed{}, d_part, d_ticket, d_st, d_oc, d_dsc, h_st, &h_oc, h_sc);
    auto op = [&](const double* p, double* out, int mode, const double* b, const int* gate, double*) {
        hipLaunchKernelGGL((k_applyA_exact<double, double>), dim3(nblk(gg.nnodes()), 1), dim3(256), 0, nullptr, p, out,
                           (const double*)d_E, thr, gg, (const uint8_t*)d_fixed, mode, b, (const double*)d_invD,
                           (const double*)nullptr, gate, (const double*)d_ke);
        return 0;
    };
    const int cgMax = maxIter < nfree ? maxIter : nfree;
    CGResult r = pcg_jacobi_op<double, double>(C, op, d_fixed, (const double*)d_invD, ndof, (const double*)d_F, d_U,
                                               true, v, tolerance, cgMax, nullptr, {}, 32);
    TOPO_HIP(hipMemcpy(Uh, d_U, (size_t)ndof * sizeof(double), hipMemcpyDeviceToHost));
    TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
    return r.iters;
    TOPO_API_END
}

int topo_ebeMGPCG(const topo_mg_level* levels, int32_t numLevels, const double* ke, const double* Fh, double* Uh,
                  const uint8_t* fixedMask0, int32_t ndof0, int32_t maxIter, double tolerance, double* cgWork) {
    (void)cgWork;
    TOPO_API_BEGIN
    TOPO_CHECK(levels && numLevels >= 1 && ke && Fh && Uh && fixedMask0 && ndof0 > 0, TOPO_E_ARG,
               "topo_ebeMGPCG: bad argument");
    std::lock_guard<std::mutex> lk(g_twin_mu);
    TwinCtx& T = g_twin;
    T.ensure();
    Ctx& C = *T.c;
    T.set_ke(ke);
    std::vector<void*> tmp;
    auto A = [&](size_t bytes) {
        void* q = C.alloc<uint8_t>((int64_t)bytes);
        tmp.push_back(q);
        return q;
    };
    MGH mg;
    mg.K0 = &T.K0;
    mg.thr = 0.0;   // inactive elements carry E = 0
    mg.exact = true;
    mg.omega = 0.5; mg.nu1 = 2; mg.nu2 = 2; mg.coarse_sweeps = 30;
    for (int li = 0; li < numLevels; li++) {
        const topo_mg_level& d = levels[li];
        MGLevel l;
        l.g.nelx = d.nelx; l.g.nely = d.nely; l.g.nelz = d.nelz;
        l.ndof = l.g.ndof();
        TOPO_CHECK(d.ndof == l.ndof && d.nel == l.g.nel(), TOPO_E_ARG, "level %d: ndof/nel inconsistent with dims", li);
        TOPO_CHECK(li > 0 || d.ndof == ndof0, TOPO_E_ARG, "level 0 ndof != ndof0");
        TOPO_CHECK(d.fixedMask && d.invDiag, TOPO_E_ARG, "level %d: fixedMask/invDiag missing", li);
        l.fixed = (uint8_t*)A(l.ndof);
        TOPO_HIP(hipMemcpyAsync(l.fixed, li == 0 ? fixedMask0 : d.fixedMask, l.ndof, hipMemcpyHostToDevice, C.s));
        l.invD = (double*)A(l.ndof * sizeof(double));
        TOPO_HIP(hipMemcpyAsync(l.invD, d.invDiag, l.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
        if (d.hasDenseK) {
            TOPO_CHECK(d.denseK, TOPO_E_ARG, "level %d: denseK missing", li);
            l.dense = true;
            l.K = (double*)A(l.ndof * l.ndof * sizeof(double));
            TOPO_HIP(hipMemcpyAsync(l.K, d.denseK, l.ndof * l.ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
        } else {
            const Grid gg = T.grid_of(d.edofs, l.ndof);
            TOPO_CHECK(gg.nelx == d.nelx && gg.nely == d.nely && gg.nelz == d.nelz, TOPO_E_UNSTRUCTURED,
                       "level %d: edofs does not match its dims", li);
            TOPO_CHECK(d.evals && (d.activeCount == 0 || d.active), TOPO_E_ARG, "level %d: evals/active missing", li);
            std::vector<double> Ez = masked_E_zyx(gg, d.evals, d.active, d.activeCount);
            l.E = (double*)A(Ez.size() * sizeof(double));
            TOPO_HIP(hipMemcpyAsync(l.E, Ez.data(), Ez.size() * sizeof(double), hipMemcpyHostToDevice, C.s));
            C.sync();
        }
        l.b = (double*)A(l.ndof * sizeof(double));
        l.x = (double*)A(l.ndof * sizeof(double));
        l.x2 = (double*)A(l.ndof * sizeof(double));
        l.res = (double*)A(l.ndof * sizeof(double));
        mg.L.push_back(l);
    }
    mg.upload_omega(C);
    for (int li = 1; li < numLevels; li++)
        TOPO_CHECK(mg.L[li].g.nelx == mg.L[li - 1].g.nelx / 2 && mg.L[li].g.nely == mg.L[li - 1].g.nely / 2 &&
                       mg.L[li].g.nelz == mg.L[li - 1].g.nelz / 2,
                   TOPO_E_UNSUPPORTED, "level %d is not the factor-2 coarsening of level %d", li, li - 1);
    double* dF = (double*)A(ndof0 * sizeof(double));
    double* dU = (double*)A(ndof0 * sizeof(double));
    CGVecs v;
    v.r = A(ndof0 * sizeof(double)); v.z = A(ndof0 * sizeof(double));
    v.p = A(ndof0 * sizeof(double)); v.Ap = A(ndof0 * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(dF, Fh, ndof0 * sizeof(double), hipMemcpyHostToDevice, C.s));
    CGResult r = pcg_mg<double, double>(C, mg, dF, dU, false, v, tolerance, maxIter, nullptr);
    TOPO_HIP(hipMemcpyAsync(Uh, dU, ndof0 * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    for (void* q : tmp) C.free_one(q);
    TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
    return r.iters;
    TOPO_API_END
}

int topo_ebeKSP_BDDC(const topo_bddc_sub* subs, int32_t numSubs, const int32_t* coarse, int32_t numCoarse,
                     const double* ke, const int32_t* edofs, const double* evals, const int32_t* active,
                     int32_t activeCount, const uint8_t* fixedMask, const double* invDiag, const double* Fh,
                     double* Uh, int32_t ndof, int32_t maxIter, double tolerance, int32_t smootherIters,
                     double* work) {
    (void)work;
    TOPO_API_BEGIN
    TOPO_CHECK(subs && numSubs >= 1 && (numCoarse == 0 || coarse) && ke && edofs && evals && fixedMask && invDiag &&
                   Fh && Uh && ndof > 0 && smootherIters >= 1 && (activeCount == 0 || active),
               TOPO_E_ARG, "topo_ebeKSP_BDDC: bad argument");
    std::lock_guard<std::mutex> lk(g_twin_mu);
    TwinCtx& T = g_twin;
    T.ensure();
    Ctx& C = *T.c;
    T.set_ke(ke);
    const Grid gg = T.grid_of(edofs, ndof);
    // per-DOF slots of the subdomains' local inverse diagonals, in ascending
    // subdomain order (the accumulation order of bddcApply :1294-1310)
    std::vector<double> hval((size_t)kBddcSlots * ndof, 0.0);
    std::vector<uint8_t> hcnt(ndof, 0), hcor(ndof, 0);
    for (int si = 0; si < numSubs; si++) {
        const topo_bddc_sub& d = subs[si];
        TOPO_CHECK(d.numDofs >= 0 && (d.numDofs == 0 || (d.dofs && d.localInvDiag)), TOPO_E_ARG,
                   "subdomain %d: dofs/localInvDiag missing", si);
        for (int k = 0; k < d.numDofs; k++) {
            const int32_t g = d.dofs[k];
            TOPO_CHECK(g >= 0 && g < ndof, TOPO_E_ARG, "subdomain %d: DOF %d out of range", si, g);
            TOPO_CHECK(hcnt[g] < kBddcSlots, TOPO_E_UNSUPPORTED, "DOF %d belongs to more than %d subdomains", g,
                       kBddcSlots);
            hval[(size_t)hcnt[g] * ndof + g] = d.localInvDiag[g];
            hcnt[g]++;
        }
    }
    for (int k = 0; k < numCoarse; k++) {
        TOPO_CHECK(coarse[k] >= 0 && coarse[k] < ndof, TOPO_E_ARG, "coarse DOF %d out of range", coarse[k]);
        hcor[coarse[k]] = 1;
    }
    std::vector<void*> tmp;
    auto A = [&](size_t bytes) {
        void* q = C.alloc<uint8_t>((int64_t)bytes);
        tmp.push_back(q);
        return q;
    };
    MGH mg;
    mg.K0 = &T.K0;
    mg.thr = 0.0;   // inactive elements carry E = 0
    mg.exact = true;
    mg.omega = 1.0;
    MGLevel l;
    l.g = gg;
    l.ndof = ndof;
    l.fixed = (uint8_t*)A(ndof);
    TOPO_HIP(hipMemcpyAsync(l.fixed, fixedMask, ndof, hipMemcpyHostToDevice, C.s));
    l.invD = (double*)A(ndof * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(l.invD, invDiag, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    std::vector<double> Ez = masked_E_zyx(gg, evals, active, activeCount);
    l.E = (double*)A(Ez.size() * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(l.E, Ez.data(), Ez.size() * sizeof(double), hipMemcpyHostToDevice, C.s));
    l.b = (double*)A(ndof * sizeof(double));
    l.x = (double*)A(ndof * sizeof(double));
    l.x2 = (double*)A(ndof * sizeof(double));
    l.res = (double*)A(ndof * sizeof(double));
    mg.L.push_back(l);
    mg.bddc = true;
    mg.bddc_sweeps = smootherIters;
    mg.bddc_val = (double*)A(hval.size() * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(mg.bddc_val, hval.data(), hval.size() * sizeof(double), hipMemcpyHostToDevice, C.s));
    mg.bddc_cnt = (uint8_t*)A(ndof);
    TOPO_HIP(hipMemcpyAsync(mg.bddc_cnt, hcnt.data(), ndof, hipMemcpyHostToDevice, C.s));
    mg.bddc_corner = (uint8_t*)A(ndof);
    TOPO_HIP(hipMemcpyAsync(mg.bddc_corner, hcor.data(), ndof, hipMemcpyHostToDevice, C.s));
    C.sync();
    mg.upload_omega(C);
    double* dF = (double*)A(ndof * sizeof(double));
    double* dU = (double*)A(ndof * sizeof(double));
    CGVecs v;
    v.r = A(ndof * sizeof(double)); v.z = A(ndof * sizeof(double));
    v.p = A(ndof * sizeof(double)); v.Ap = A(ndof * sizeof(double));
    TOPO_HIP(hipMemcpyAsync(dF, Fh, ndof * sizeof(double), hipMemcpyHostToDevice, C.s));
    CGResult r = pcg_mg<double, double>(C, mg, dF, dU, false, v, tolerance, maxIter, nullptr);
    TOPO_HIP(hipMemcpyAsync(Uh, dU, ndof * sizeof(double), hipMemcpyDeviceToHost, C.s));
    C.sync();
    for (void* q : tmp) C.free_one(q);
    TOPO_CHECK(!r.nonfinite, TOPO_E_BREAKDOWN, "non-finite CG scalar");
    return r.iters;
    TOPO_API_END
}

}  // extern "C"
