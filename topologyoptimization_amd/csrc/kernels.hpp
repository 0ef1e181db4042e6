// kernels.hpp -- device kernels of libtopo_hip (gfx950). Included by exactly
// one translation unit (solver.hip).
//
// Layout in HBM
//   DOF vectors    reference order 3n+d, n = (ix*nny + iy)*nnz + iz
//   element arrays "zyx" order ei = (ex*nely + ey)*nelz + ez, so that
//                  consecutive lanes (consecutive iz) read consecutive
//                  elements; converted from/to the reference order
//                  (x fastest) at the host boundary.
#pragma once
#include "common.hpp"

namespace topo {

// Sum of a double over the 64 lanes of a wavefront, left in lane 63 (the
// other lanes hold partial sums): DPP row shifts by 1, 2, 4, 8 (inclusive
// sums within each row of 16 lanes), then the row broadcasts of lane 15 and
// lane 31 carry the row totals up. VALU only, no LDS round trips. The
// order of additions is fixed by the lane positions.
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_d(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, RMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, RMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_d<0x111, 0xf>(v);   // row_shr:1
    v += dpp_d<0x112, 0xf>(v);   // row_shr:2
    v += dpp_d<0x114, 0xf>(v);   // row_shr:4
    v += dpp_d<0x118, 0xf>(v);   // row_shr:8
    v += dpp_d<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_d<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

// K0 (24x24 row-major, FP64 and FP32 copies) and the filter stencil live in
// per-solver device buffers passed by pointer; every lane reads the same
// address, so the loads are scalar.

// ---------------------------------------------------------------------------
// small helpers
__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }
// gtid() in XCD-aware block order, for a grid of a multiple of 8 blocks (the
// identity otherwise): the hardware places block L on XCD L % 8, and block L
// takes block (L % 8) * (gridDim.x / 8) + L / 8 of the items, so each XCD
// works on one contiguous range and neighbouring items, which share operands
// (the transfers' stencils), meet in one L2. Same stride as gtid().
__device__ __forceinline__ int64_t xcd_gtid() {
    const unsigned G = gridDim.x, L = blockIdx.x;
    const unsigned vb = (G & 7u) ? L : (L & 7u) * (G >> 3) + (L >> 3);
    return (int64_t)vb * blockDim.x + threadIdx.x;
}

// local index of the element node with offsets (a,b,c) in {0,1}^3
// (order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1))
__host__ __device__ __forceinline__ int local_node(int a, int b, int c) {
    return c * 4 + (b ? (a ? 2 : 3) : (a ? 1 : 0));
}
__host__ __device__ __forceinline__ void local_offsets(int L, int& a, int& b, int& c) {
    c = L >> 2;
    const int q = L & 3;
    a = (q == 1 || q == 2);
    b = (q >= 2);
}


// JS Math.max / Math.min semantics (NaN propagates)
__device__ __forceinline__ double js_max(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b); }
__device__ __forceinline__ double js_min(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b); }

// CG scalars living on the device (no host round trip per iteration)
struct CGState {
    double rz, rz_new, pAp, alpha, beta;
    double rn;        // ||r||^2 at the last update
    double r0;        // ||r0||^2
    double tolSq;
    double tol;
    int done;         // converged: later update kernels are no-ops
    int iters;        // iterations completed
    int maxit;
    int nonfinite;
    double guard;     // denominator guard: 1e-12 in the reference's PCGs, 0 elsewhere
    int abs_tol;      // 1: stop at ||r|| < tol (the 2-D worker's solveCG), 0: ||r|| < tol ||r0||
};

// Batched launches (several right-hand sides in one launch): blockIdx.y
// selects the right-hand side q; its vectors start q * n entries after the
// first (n = the level's DOF count) and its gate is the `done` flag of the
// q-th CGState of an array. With gridDim.y = 1 nothing changes.
constexpr int kGateStride = (int)(sizeof(CGState) / sizeof(int));
__device__ __forceinline__ const int* batch_gate(const int* gate) {
    return gate ? gate + (int64_t)blockIdx.y * kGateStride : nullptr;
}
__device__ __forceinline__ int64_t batch_off(int64_t n) { return (int64_t)blockIdx.y * n; }

// ---------------------------------------------------------------------------
// element order conversions
template <typename T>
__global__ void k_ref_to_zyx_t(const T* __restrict__ in, T* __restrict__ out, Grid g) {
    for (int64_t i = gtid(); i < g.nel(); i += gstride()) {
        const int64_t ez = i % g.nelz, t = i / g.nelz, ey = t % g.nely, ex = t / g.nely;
        out[i] = in[ex + g.nelx * (ey + (int64_t)g.nely * ez)];
    }
}
template <typename T>
__global__ void k_zyx_to_ref_t(const T* __restrict__ in, T* __restrict__ out, Grid g) {
    for (int64_t i = gtid(); i < g.nel(); i += gstride()) {
        const int64_t ez = i % g.nelz, t = i / g.nelz, ey = t % g.nely, ex = t / g.nely;
        out[ex + g.nelx * (ey + (int64_t)g.nely * ez)] = in[i];
    }
}
__global__ void k_ref_to_zyx(const float* __restrict__ in, float* __restrict__ out, Grid g) {
    for (int64_t i = gtid(); i < g.nel(); i += gstride()) {
        const int64_t ez = i % g.nelz, t = i / g.nelz, ey = t % g.nely, ex = t / g.nely;
        out[i] = in[ex + g.nelx * (ey + (int64_t)g.nely * ez)];
    }
}
__global__ void k_zyx_to_ref(const float* __restrict__ in, float* __restrict__ out, Grid g) {
    for (int64_t i = gtid(); i < g.nel(); i += gstride()) {
        const int64_t ez = i % g.nelz, t = i / g.nelz, ey = t % g.nely, ex = t / g.nely;
        out[ex + g.nelx * (ey + (int64_t)g.nely * ez)] = in[i];
    }
}
__global__ void k_u8_ref_to_zyx(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, Grid g) {
    for (int64_t i = gtid(); i < g.nel(); i += gstride()) {
        const int64_t ez = i % g.nelz, t = i / g.nelz, ey = t % g.nely, ex = t / g.nely;
        out[i] = in[ex + g.nelx * (ey + (int64_t)g.nely * ez)];
    }
}

// ---------------------------------------------------------------------------
// Element moduli E_e = Emin + x^p (E0 - Emin).
// powmode 0: _powDensity fast path for integer p (js/optimizer-worker-3d.js:293-306)
// powmode 1: Math.pow as in ebePCG (assembly/index.ts:544)
__device__ __forceinline__ double pow_density(double x, double p, int powmode) {
#pragma clang fp contract(off)
    if (powmode == 0 && p == __builtin_floor(p) && p >= 0 && p <= 6) {
        switch ((int)p) {
        case 0: return 1.0;
        case 1: return x;
        case 2: return x * x;
        case 3: return x * x * x;
        case 4: { double x2 = x * x; return x2 * x2; }
        case 5: { double x2 = x * x; return x2 * x2 * x; }
        case 6: { double x2 = x * x; return x2 * x2 * x2; }
        }
    }
    return pow(x, p);
}

template <typename TX>
__global__ void k_moduli(const TX* __restrict__ xphys, double penal, double E0, double Emin, int powmode,
                         double* __restrict__ E, float* __restrict__ E32, int64_t nel) {
#pragma clang fp contract(off)
    const double dE = E0 - Emin;
    for (int64_t i = gtid(); i < nel; i += gstride()) {
        const double v = Emin + pow_density((double)xphys[i], penal, powmode) * dE;
        E[i] = v;
        if (E32) E32[i] = (float)v;
    }
}

// Jacobi diagonal: diag[3n+d] = sum over active elements around n of
// E_e K0_ii, added in ascending reference element order like
// _updateOperatorLevel (js/optimizer-worker-3d.js:773-784); invDiag = 1/d
// on free DOFs with d > 1e-30, else 0 (:787-795).
__global__ void k_invdiag(const double* __restrict__ E, double thr, const uint8_t* __restrict__ fixed,
                          double* __restrict__ invD, float* __restrict__ invD32, Grid g, const double* __restrict__ ke) {
#pragma clang fp contract(off)
    const int nny = g.nny(), nnz = g.nnz();
    for (int64_t n = gtid(); n < g.nnodes(); n += gstride()) {
        const int iz = (int)(n % nnz);
        const int64_t t = n / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        double acc[3] = {0.0, 0.0, 0.0};
        for (int c = 1; c >= 0; c--) {
            const int ez = iz - c;
            if (ez < 0 || ez >= g.nelz) continue;
            for (int b = 1; b >= 0; b--) {
                const int ey = iy - b;
                if (ey < 0 || ey >= g.nely) continue;
                for (int a = 1; a >= 0; a--) {
                    const int ex = ix - a;
                    if (ex < 0 || ex >= g.nelx) continue;
                    const double Ee = E[((int64_t)ex * g.nely + ey) * g.nelz + ez];
                    if (!(Ee > thr)) continue;
                    const int L = local_node(a, b, c);
                    for (int d = 0; d < 3; d++) acc[d] += Ee * ke[(3 * L + d) * 25];
                }
            }
        }
        for (int d = 0; d < 3; d++) {
            const int64_t i = 3 * n + d;
            const double v = (!fixed[i] && acc[d] > 1e-30) ? 1.0 / acc[d] : 0.0;
            invD[i] = v;
            if (invD32) invD32[i] = (float)v;
        }
    }
}

// ---------------------------------------------------------------------------
// applyA, reference-order gather ("exact" kernel).
// Ap[3n+d] = sum over active elements e around n, in ascending reference
// element order, of E_e * (sum_j K0[3L+d][j] p[edof_e[j]]) with j ascending:
// the addition sequence of applyAEbe3D (assembly/index.ts:403-439), without
// FMA contraction, so FP64 results are identical to the reference's.
// mode: 0 plain Ap; 1 masked (Ap = 0 on fixed DOFs); 2 residual
// out = b - Ap on free DOFs, 0 on fixed; 3 damped-Jacobi sweep
// out = x + omega*invD*(b - Ap) on free DOFs, x on fixed.
template <typename T, typename TE>
__global__ void __launch_bounds__(256) k_applyA_exact(const T* __restrict__ p, T* __restrict__ out,
                                                      const TE* __restrict__ E, double thr, Grid g,
                                                      const uint8_t* __restrict__ fixed, int mode,
                                                      const T* __restrict__ b, const T* __restrict__ invD,
                                                      const double* __restrict__ omp, const int* gate,
                                                      const T* __restrict__ ke) {
#pragma clang fp contract(off)
    gate = batch_gate(gate);
    if (gate && *gate) return;
    {
        const int64_t bo = batch_off(g.ndof());
        p += bo; out += bo;
        if (b) b += bo;
    }
    const T omega = omp ? (T)*omp : T(0);
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t nyz = (int64_t)nny * nnz;
    for (int64_t n = gtid(); n < g.nnodes(); n += gstride()) {
        const int iz = (int)(n % nnz);
        const int64_t t = n / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        T acc[3] = {T(0), T(0), T(0)};
        for (int c = 1; c >= 0; c--) {
            const int ez = iz - c;
            if (ez < 0 || ez >= g.nelz) continue;
            for (int bb = 1; bb >= 0; bb--) {
                const int ey = iy - bb;
                if (ey < 0 || ey >= g.nely) continue;
                for (int a = 1; a >= 0; a--) {
                    const int ex = ix - a;
                    if (ex < 0 || ex >= g.nelx) continue;
                    const TE Ee = E[((int64_t)ex * g.nely + ey) * g.nelz + ez];
                    if (!((double)Ee > thr)) continue;
                    const int L = local_node(a, bb, c);
                    const int64_t n0 = ((int64_t)ex * nny + ey) * nnz + ez;
                    T loc[24];
#pragma unroll
                    for (int m = 0; m < 8; m++) {
                        int ma, mb, mc;
                        local_offsets(m, ma, mb, mc);
                        const int64_t nm = n0 + ma * nyz + mb * nnz + mc;
                        loc[3 * m] = p[3 * nm];
                        loc[3 * m + 1] = p[3 * nm + 1];
                        loc[3 * m + 2] = p[3 * nm + 2];
                    }
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int row = (3 * L + d) * 24;
                        T s = T(0);
#pragma unroll
                        for (int j = 0; j < 24; j++) s += ke[row + j] * loc[j];
                        acc[d] += (T)Ee * s;
                    }
                }
            }
        }
        for (int d = 0; d < 3; d++) {
            const int64_t i = 3 * n + d;
            T v = acc[d];
            if (mode == 1) {
                if (fixed[i]) v = T(0);
            } else if (mode == 2) {
                v = fixed[i] ? T(0) : b[i] - v;
            } else if (mode == 3) {
                const T r = b[i] - v;
                v = fixed[i] ? p[i] : p[i] + omega * invD[i] * r;
            }
            out[i] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Deterministic reductions. Each block reduces its grid-stride share in a
// fixed tree, writes partial[block]; the last block to finish (ticket
// counter) adds the partials in block order and runs the epilogue. Results
// depend only on n and the launch shape, never on timing.
constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* sh /* NV*kRedThreads/64 */) {
    // wave64 tree with fixed shuffle pattern, then the 4 wave sums in order
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NV; k++) sh[k * (kRedThreads / 64) + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int w = 0; w < kRedThreads / 64; w++) s += sh[k * (kRedThreads / 64) + w];
            v[k] = s;
        }
}

// returns true in thread 0 of the last block, with tot[] = the grid sums.
// The last block adds the partials with all its threads: thread t takes
// blocks t, t+256, ... in order, then the fixed wave/block tree.
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double* partials, unsigned* ticket, double (&tot)[NV]) {
    __shared__ double sh[NV * (kRedThreads / 64)];
    __shared__ bool last;
    block_reduce<NV>(v, sh);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) partials[k * kRedBlocks + blockIdx.x] = v[k];
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double w[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0;
        for (unsigned bI = threadIdx.x; bI < gridDim.x; bI += blockDim.x)
            s += __hip_atomic_load(&partials[k * kRedBlocks + bI], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w[k] = s;
    }
    __syncthreads();   // sh is reused
    block_reduce<NV>(w, sh);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) tot[k] = w[k];
        *ticket = 0u;   // ready for the next launch on this stream
        return true;
    }
    return false;
}


// The CG kernels handle NR right-hand sides in one launch (NR = 1: a single
// solve): vector q starts at q * n, its scalars live in st[q], and a
// converged right-hand side (st[q].done) is skipped while the others go on.
// Every block reads the same st[] values (written by the last block of the
// previous launch), so all blocks take part in the reduction consistently.

// rn = ||r||^2 and rz = r.z; used at CG start
template <typename T, int NR = 1>
__global__ void __launch_bounds__(kRedThreads) k_cg_start(const T* __restrict__ r, const T* __restrict__ z, int64_t n,
                                                         CGState* st, double* partials, unsigned* ticket) {
    double v[2 * NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0, b = 0.0;
        const T* rq = r + q * n;
        const T* zq = z + q * n;
        for (int64_t i = gtid(); i < n; i += gstride()) {
            const double ri = (double)rq[i];
            a += ri * ri;
            b += ri * (double)zq[i];
        }
        v[2 * q] = a;
        v[2 * q + 1] = b;
    }
    double tot[2 * NR];
    if (grid_reduce<2 * NR>(v, partials, ticket, tot)) {
        for (int q = 0; q < NR; q++) {
            CGState& s = st[q];
            s.r0 = tot[2 * q];
            s.rn = tot[2 * q];
            s.rz = tot[2 * q + 1];
            s.tolSq = s.abs_tol ? s.tol * s.tol : s.tol * s.tol * (tot[2 * q] > 1e-30 ? tot[2 * q] : 1e-30);
            s.done = (s.maxit <= 0) || (tot[2 * q] < s.tolSq);
            s.iters = 0;
            s.nonfinite = 0;
        }
    }
}

__device__ __forceinline__ bool all_done(const CGState* st, int nr) {
    for (int q = 0; q < nr; q++)
        if (!st[q].done) return false;
    return true;
}

// pAp = p.Ap -> alpha = rz/(pAp + 1e-12) (ebePCG :695-699)
template <typename T, int NR = 1>
__global__ void __launch_bounds__(kRedThreads) k_cg_alpha(const T* __restrict__ p, const T* __restrict__ Ap, int64_t n,
                                                         CGState* st, double* partials, unsigned* ticket) {
    if (all_done(st, NR)) return;
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const T* pq = p + q * n;
            const T* aq = Ap + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) a += (double)pq[i] * (double)aq[i];
        }
        v[q] = a;
    }
    double tot[NR];
    if (grid_reduce<NR>(v, partials, ticket, tot)) {
        for (int q = 0; q < NR; q++) {
            CGState& s = st[q];
            if (s.done) continue;
            s.pAp = tot[q];
            s.alpha = s.rz / (tot[q] + s.guard);
            if (!isfinite(s.alpha)) s.nonfinite = 1;
        }
    }
}

// Jacobi variant: U += alpha p; r -= alpha Ap; z = invD r; rz_new, ||r||^2;
// then beta = rz_new/(rz + 1e-12) (ebePCG :701-719). The convergence test
// of the next iteration (:645-650) is folded in: done = ||r||^2 < tolSq.
template <typename T, typename TU, int NR = 1>
__global__ void __launch_bounds__(kRedThreads) k_cg_update_jacobi(TU* __restrict__ U, T* __restrict__ r, T* __restrict__ z,
                                                                 const T* __restrict__ p, const T* __restrict__ Ap,
                                                                 const T* __restrict__ invD, int64_t n, CGState* st,
                                                                 double* partials, unsigned* ticket) {
    if (all_done(st, NR)) return;
    double v[2 * NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0, b = 0.0;
        if (!st[q].done) {
            const double alpha = st[q].alpha;
            TU* Uq = U + q * n;
            T* rq = r + q * n;
            T* zq = z + q * n;
            const T* pq = p + q * n;
            const T* aq = Ap + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) {
                const T ri = (T)((double)rq[i] - alpha * (double)aq[i]);
                const T zi = invD[i] * ri;
                Uq[i] = (TU)((double)Uq[i] + alpha * (double)pq[i]);
                rq[i] = ri;
                zq[i] = zi;
                a += (double)ri * (double)zi;
                b += (double)ri * (double)ri;
            }
        }
        v[2 * q] = a;
        v[2 * q + 1] = b;
    }
    double tot[2 * NR];
    if (grid_reduce<2 * NR>(v, partials, ticket, tot)) {
        for (int q = 0; q < NR; q++) {
            CGState& s = st[q];
            if (s.done) continue;
            s.rz_new = tot[2 * q];
            s.rn = tot[2 * q + 1];
            s.beta = tot[2 * q] / (s.rz + s.guard);
            s.rz = tot[2 * q];
            s.iters += 1;
            if (!isfinite(tot[2 * q]) || !isfinite(tot[2 * q + 1])) s.nonfinite = 1;
            s.done = (tot[2 * q + 1] < s.tolSq) || (s.iters >= s.maxit) || s.nonfinite;
        }
    }
}

// MG variant part 1: U += alpha p; r -= alpha Ap (ebeMGPCG :1158-1166) and
// ||r||^2 for the next convergence test.
template <typename T, typename TU, int NR = 1>
__global__ void __launch_bounds__(kRedThreads) k_cg_update_ur(TU* __restrict__ U, T* __restrict__ r, const T* __restrict__ p,
                                                             const T* __restrict__ Ap, int64_t n, CGState* st,
                                                             double* partials, unsigned* ticket) {
    if (all_done(st, NR)) return;
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const double alpha = st[q].alpha;
            TU* Uq = U + q * n;
            T* rq = r + q * n;
            const T* pq = p + q * n;
            const T* aq = Ap + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) {
                Uq[i] = (TU)((double)Uq[i] + alpha * (double)pq[i]);
                const T ri = (T)((double)rq[i] - alpha * (double)aq[i]);
                rq[i] = ri;
                a += (double)ri * (double)ri;
            }
        }
        v[q] = a;
    }
    double tot[NR];
    if (grid_reduce<NR>(v, partials, ticket, tot)) {
        for (int q = 0; q < NR; q++) {
            CGState& s = st[q];
            if (s.done) continue;
            s.rn = tot[q];
            s.iters += 1;
            if (!isfinite(tot[q])) s.nonfinite = 1;
        }
    }
}

// MG variant part 2 (after z = M^-1 r with fixed DOFs zeroed):
// rz_new = r.z, beta = rz_new/(rz + 1e-12) (:1174-1187); the convergence
// flag is set here so the p-update of a converged iteration still runs,
// exactly like the reference (which tests at the top of the next iteration).
template <typename T, int NR = 1>
__global__ void __launch_bounds__(kRedThreads) k_cg_rz(const T* __restrict__ r, const T* __restrict__ z, int64_t n,
                                                      CGState* st, double* partials, unsigned* ticket) {
    if (all_done(st, NR)) return;
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const T* rq = r + q * n;
            const T* zq = z + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) a += (double)rq[i] * (double)zq[i];
        }
        v[q] = a;
    }
    double tot[NR];
    if (grid_reduce<NR>(v, partials, ticket, tot)) {
        for (int q = 0; q < NR; q++) {
            CGState& s = st[q];
            if (s.done) continue;
            s.rz_new = tot[q];
            s.beta = tot[q] / (s.rz + s.guard);
            s.rz = tot[q];
            if (!isfinite(tot[q])) s.nonfinite = 1;
        }
    }
}

// p = z + beta p (blockIdx.y = right-hand side); afterwards evaluate the
// convergence flag
template <typename T>
__global__ void k_cg_pupdate(T* __restrict__ p, const T* __restrict__ z, int64_t n, const CGState* st) {
    st += blockIdx.y;
    if (st->done) return;
    p += batch_off(n);
    z += batch_off(n);
    const T beta = (T)st->beta;
    for (int64_t i = gtid(); i < n; i += gstride()) p[i] = z[i] + beta * p[i];
}
__global__ void k_cg_check(CGState* st, int nr) {
    for (int q = 0; q < nr; q++) {
        CGState& s = st[q];
        if (s.done) continue;
        s.done = (s.rn < s.tolSq) || (s.iters >= s.maxit) || s.nonfinite;
    }
}

// ---------------------------------------------------------------------------
// Fence-free CG iteration. A producer kernel leaves one partial sum per
// block (no atomics, no fences); each kernel that needs the grid sum adds
// the partials itself, every block in the same fixed order, so all blocks
// compute the same bits, and k_cgp_finish keeps the books. The kernel
// boundary makes the partials visible. This replaces the last-block
// pattern in the iteration loops: there each block's device-scope release
// (an L2 write-back per block) serialises, ~20 ns per block.
// Quantities: 0 = p.Ap, 1 = ||r||^2, 2 = r.z; row (k, q) of the partials
// holds quantity k of right-hand side q for each block of its producer
// (kCgBlocks for the vector kernels; p.Ap and r.z may also come from the
// x-march kernels, one partial per (yz tile, node plane), up to kCgStride).
constexpr int kCgBlocks = 1024;
constexpr int kCgStride = 16384;
constexpr int kCgRows = 4;   // right-hand sides per quantity
__host__ __device__ __forceinline__ int64_t cg_row(int k, int q) { return (int64_t)(k * kCgRows + q) * kCgStride; }

// v[i] -> row (k0 + i / nr, i % nr), this block's column
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double* part, int k0, int nr) {
    __shared__ double sh[NV * (kRedThreads / 64)];
    block_reduce<NV>(v, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) part[cg_row(k0 + i / nr, i % nr) + blockIdx.x] = v[i];
}
// tot[i] = sum of row (k0 + i / nr, i % nr) over its nb columns, in a
// fixed order (thread t takes columns t + 256 m into 8 accumulators by
// m mod 8, which meet in a fixed tree; then the block tree), so the loads
// of a thread are in flight together; every thread of the block gets the
// result
template <int NV>
__device__ __forceinline__ void sum_partials(const double* part, int k0, int nr, double (&tot)[NV],
                                             int nb = kCgBlocks) {
    __shared__ double sh[NV * (kRedThreads / 64)];
    __shared__ double res[NV];
    double w[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const double* row = part + cg_row(k0 + i / nr, i % nr);
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int b = threadIdx.x;
        for (; b + 7 * kRedThreads < nb; b += 8 * kRedThreads) {
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] += row[b + u * kRedThreads];
        }
        // the tail's loads are issued together (a column past the end reads
        // entry 0 and adds an exact +0: the accumulators start at +0 and
        // never hold -0)
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = b + u * kRedThreads;
            v[u] = row[c < nb ? c : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] += (b + u * kRedThreads < nb) ? v[u] : 0.0;
        w[i] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    block_reduce<NV>(w, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) res[i] = w[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) tot[i] = res[i];
}

// sum_partials of two quantities at once (rows (ka + i / nr, i % nr) over
// nba columns into ta, rows (kb + i / nr, i % nr) over nbb into tb), each
// row summed exactly as sum_partials sums it; one pass, so the loads of all
// the rows are in flight together
template <int NR>
__device__ __forceinline__ void sum_partials2(const double* part, int nr, int ka, int nba, double (&ta)[NR],
                                              int kb, int nbb, double (&tb)[NR]) {
    constexpr int NV = 2 * NR;
    __shared__ double sh[NV * (kRedThreads / 64)];
    __shared__ double res[NV];
    double w[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const int nb = i < NR ? nba : nbb;
        const double* row = part + (i < NR ? cg_row(ka + i / nr, i % nr) : cg_row(kb + (i - NR) / nr, (i - NR) % nr));
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int b = threadIdx.x;
        for (; b + 7 * kRedThreads < nb; b += 8 * kRedThreads) {
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] += row[b + u * kRedThreads];
        }
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = b + u * kRedThreads;
            v[u] = row[c < nb ? c : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] += (b + u * kRedThreads < nb) ? v[u] : 0.0;
        w[i] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    block_reduce<NV>(w, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) res[i] = w[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NR; i++) {
        ta[i] = res[i];
        tb[i] = res[NR + i];
    }
}

// p.Ap partials (quantity 0)
template <typename T, int NR>
__global__ void __launch_bounds__(kRedThreads) k_cgp_pAp(const T* __restrict__ p, const T* __restrict__ Ap, int64_t n,
                                                        const CGState* st, double* part) {
    if (all_done(st, NR)) return;
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const T* pq = p + q * n;
            const T* aq = Ap + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) a += (double)pq[i] * (double)aq[i];
        }
        v[q] = a;
    }
    put_partials<NR>(v, part, 0, NR);
}

// alpha = rz / (p.Ap + guard); U += alpha p; r -= alpha Ap (ebeMGPCG
// :1158-1166); r_lo (may be null) receives r rounded to the V-cycle's
// precision; ||r||^2 partials (quantity 1). With U null the U update is
// left to k_cgp_pupdate (which reads p anyway).
template <typename T, typename TU, int NR, typename TL>
__global__ void __launch_bounds__(kRedThreads) k_cgp_update_ur(TU* __restrict__ U, T* __restrict__ r,
                                                              const T* __restrict__ p, const T* __restrict__ Ap,
                                                              int64_t n, const CGState* st, double* part,
                                                              TL* __restrict__ r_lo, int nb0) {
    if (all_done(st, NR)) return;
    // the first KP elements of r and Ap of right-hand side 0 are loaded
    // before the partial sums (the U update, when here, reads as it goes)
    constexpr int KP = 4;
    T rp[KP], ap[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int64_t i = gtid() + k * gstride();
        const int64_t ii = i < n ? i : 0;
        rp[k] = r[ii];
        ap[k] = Ap[ii];
    }
    double pAp[NR];
    sum_partials<NR>(part, 0, NR, pAp, nb0);
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const double alpha = st[q].rz / (pAp[q] + st[q].guard);
            TU* Uq = U + q * n;
            T* rq = r + q * n;
            const T* pq = p + q * n;
            const T* aq = Ap + q * n;
            TL* lq = r_lo ? r_lo + q * n : nullptr;
            int64_t i = gtid();
            if (q == 0) {
#pragma unroll
                for (int k = 0; k < KP; k++, i += gstride())
                    if (i < n) {
                        if (U) Uq[i] = (TU)((double)Uq[i] + alpha * (double)pq[i]);
                        const T ri = (T)((double)rp[k] - alpha * (double)ap[k]);
                        rq[i] = ri;
                        if (lq) lq[i] = (TL)ri;
                        a += (double)ri * (double)ri;
                    }
            }
            for (; i < n; i += gstride()) {
                if (U) Uq[i] = (TU)((double)Uq[i] + alpha * (double)pq[i]);
                const T ri = (T)((double)rq[i] - alpha * (double)aq[i]);
                rq[i] = ri;
                if (lq) lq[i] = (TL)ri;
                a += (double)ri * (double)ri;
            }
        }
        v[q] = a;
    }
    put_partials<NR>(v, part, 1, NR);
}


// Jacobi-PCG (ebePCG assembly/index.ts:701-719, one right-hand side):
// alpha = rz / (p.Ap + guard) (p.Ap from nb0 partials), r -= alpha Ap, and
// the ||r||^2 and r.z partials (quantities 1, 2) with z =
// invD r formed here and not stored (the p update forms it again from the
// same r and invD); U += alpha p is left to the p update, which reads p.
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_cgp_update_rj(T* __restrict__ r, const T* __restrict__ Ap,
                                                              const T* __restrict__ invD, int64_t n, const CGState* st,
                                                              double* part, int nb0) {
    if (st->done) return;
    double pAp[1];
    sum_partials<1>(part, 0, 1, pAp, nb0);
    const double alpha = st->rz / (pAp[0] + st->guard);
    double v[2] = {0.0, 0.0};
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const T ri = (T)((double)r[i] - alpha * (double)Ap[i]);
        const T zi = invD[i] * ri;
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)zi;
    }
    put_partials<2>(v, part, 1, 1);
}
// Jacobi-PCG p update: U += alpha p (this iteration's alpha), p = z + beta p
// with z = invD r
template <typename T, typename TU>
__global__ void __launch_bounds__(kRedThreads) k_cgp_pupdate_rj(T* __restrict__ p, const T* __restrict__ r,
                                                               const T* __restrict__ invD, TU* __restrict__ U,
                                                               int64_t n, const CGState* st, const double* part,
                                                               int nb0) {
    if (st->done) return;
    double rz[1], pAp[1];
    sum_partials<1>(part, 2, 1, rz);
    sum_partials<1>(part, 0, 1, pAp, nb0);
    const T beta = (T)(rz[0] / (st->rz + st->guard));
    const double alpha = st->rz / (pAp[0] + st->guard);
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const T pi = p[i];
        const T zi = invD[i] * r[i];
        U[i] = (TU)((double)U[i] + alpha * (double)pi);
        p[i] = zi + beta * pi;
    }
}

// r.z partials (quantity 2) with z = (MASK && fixed ? 0 : zz), zz the
// preconditioner's output (precision TZ). MASK = false when zz is zero at
// the fixed DOFs by construction (the V-cycle: invD = 0 there, sweeps keep x,
// prolongation adds to free DOFs only; Jacobi z = invD r).
template <typename T, typename TZ, int NR, bool MASK>
__global__ void __launch_bounds__(kRedThreads) k_cgp_rz(const T* __restrict__ r, const TZ* __restrict__ zz,
                                                       const uint8_t* __restrict__ fixed, int64_t n,
                                                       const CGState* st, double* part) {
    if (all_done(st, NR)) return;
    double v[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        double a = 0.0;
        if (!st[q].done) {
            const T* rq = r + q * n;
            const TZ* zq = zz + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) {
                const double zi = (MASK && fixed[i]) ? 0.0 : (double)zq[i];
                a += (double)rq[i] * zi;
            }
        }
        v[q] = a;
    }
    put_partials<NR>(v, part, 2, NR);
}

// beta = rz_new / (rz + guard) from the r.z partials; p = z + beta p with
// z = (MASK && fixed ? 0 : zz). With U given, first U += alpha p with the
// iteration's alpha = rz / (p.Ap + guard) (the p.Ap partials, nb0 of them):
// the update k_cgp_update_ur leaves out, on the p this kernel reads anyway.
template <typename T, typename TZ, int NR, bool MASK, typename TU = T>
__global__ void __launch_bounds__(kRedThreads) k_cgp_pupdate(T* __restrict__ p, const TZ* __restrict__ zz,
                                                            const uint8_t* __restrict__ fixed, int64_t n,
                                                            const CGState* st, const double* part,
                                                            TU* __restrict__ U = nullptr, int nb0 = kCgBlocks,
                                                            int nb2 = kCgBlocks) {
    if (all_done(st, NR)) return;
    double rz[NR];
    sum_partials<NR>(part, 2, NR, rz, nb2);
    double pAp[NR];
    if (U) sum_partials<NR>(part, 0, NR, pAp, nb0);
#pragma unroll
    for (int q = 0; q < NR; q++) {
        if (st[q].done) continue;
        const T beta = (T)(rz[q] / (st[q].rz + st[q].guard));
        T* pq = p + q * n;
        const TZ* zq = zz + q * n;
        if (U) {
            const double alpha = st[q].rz / (pAp[q] + st[q].guard);
            TU* Uq = U + q * n;
            for (int64_t i = gtid(); i < n; i += gstride()) {
                const T pi = pq[i];
                Uq[i] = (TU)((double)Uq[i] + alpha * (double)pi);
                pq[i] = ((MASK && fixed[i]) ? T(0) : (T)zq[i]) + beta * pi;
            }
        } else {
            for (int64_t i = gtid(); i < n; i += gstride())
                pq[i] = ((MASK && fixed[i]) ? T(0) : (T)zq[i]) + beta * pq[i];
        }
    }
}

// k_cgp_pupdate (with U, MASK false) that also keeps the books of the
// iteration, so no k_cgp_finish follows: the iteration's state st is only
// read, and block 0 writes the updated state into stn, which the next
// iteration's kernels read (the two state arrays trade roles every
// iteration; writing st in place would race with the blocks still reading
// it). The books are k_cgp_finish's, from the same partial sums in the same
// order, so the iterates are the same bits. A right-hand side that is done
// is copied through unchanged.
template <typename T, typename TZ, int NR, typename TU = T>
__global__ void __launch_bounds__(kRedThreads) k_cgp_pupdate_books(T* __restrict__ p, const TZ* __restrict__ zz,
                                                                  int64_t n, const CGState* st, CGState* stn,
                                                                  const double* part, TU* __restrict__ U, int nb0,
                                                                  int nb2) {
    const bool books = blockIdx.x == 0;
    if (all_done(st, NR)) {
        if (books && threadIdx.x < NR) stn[threadIdx.x] = st[threadIdx.x];
        return;
    }
    // the first KP elements of this thread's share of right-hand side 0 are
    // loaded before the partial sums, so the streaming starts under their
    // latency
    constexpr int KP = 4;
    T pp[KP];
    TZ zp[KP];
    TU up[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int64_t i = gtid() + k * gstride();
        const int64_t ii = i < n ? i : 0;
        pp[k] = p[ii];
        zp[k] = zz[ii];
        up[k] = U[ii];
    }
    double rz[NR], pAp[NR];
    sum_partials2<NR>(part, NR, 2, nb2, rz, 0, nb0, pAp);
#pragma unroll
    for (int q = 0; q < NR; q++) {
        if (st[q].done) continue;
        const T beta = (T)(rz[q] / (st[q].rz + st[q].guard));
        const double alpha = st[q].rz / (pAp[q] + st[q].guard);
        T* pq = p + q * n;
        const TZ* zq = zz + q * n;
        TU* Uq = U + q * n;
        int64_t i = gtid();
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < KP; k++, i += gstride())
                if (i < n) {
                    const T pi = pp[k];
                    Uq[i] = (TU)((double)up[k] + alpha * (double)pi);
                    pq[i] = (T)zp[k] + beta * pi;
                }
        }
        for (; i < n; i += gstride()) {
            const T pi = pq[i];
            Uq[i] = (TU)((double)Uq[i] + alpha * (double)pi);
            pq[i] = (T)zq[i] + beta * pi;
        }
    }
    if (!books) return;
    double rn[NR];
    sum_partials<NR>(part, 1, NR, rn);
    if (threadIdx.x != 0) return;
    for (int q = 0; q < NR; q++) {
        CGState s = st[q];
        if (!s.done) {
            s.pAp = pAp[q];
            s.alpha = s.rz / (pAp[q] + s.guard);
            s.rn = rn[q];
            s.rz_new = rz[q];
            s.beta = rz[q] / (s.rz + s.guard);
            s.rz = rz[q];
            s.iters += 1;
            if (!isfinite(s.alpha) || !isfinite(rn[q]) || !isfinite(rz[q])) s.nonfinite = 1;
            s.done = (rn[q] < s.tolSq) || (s.iters >= s.maxit) || s.nonfinite;
        }
        stn[q] = s;
    }
}

// the books of one iteration (one block of kRedThreads): alpha, ||r||^2,
// rz, beta, iteration count, non-finite flag, convergence
template <int NR>
__global__ void __launch_bounds__(kRedThreads) k_cgp_finish(CGState* st, const double* part, int nb0,
                                                           int nb2 = kCgBlocks) {
    if (all_done(st, NR)) return;
    double pAp[NR], tot[3 * NR];
    sum_partials<NR>(part, 0, NR, pAp, nb0);
    {
        double rn[NR], rz[NR];
        sum_partials<NR>(part, 1, NR, rn);
        sum_partials<NR>(part, 2, NR, rz, nb2);
#pragma unroll
        for (int q = 0; q < NR; q++) {
            tot[q] = pAp[q];
            tot[NR + q] = rn[q];
            tot[2 * NR + q] = rz[q];
        }
    }
    if (threadIdx.x != 0) return;
    for (int q = 0; q < NR; q++) {
        CGState& s = st[q];
        if (s.done) continue;
        const double pAp = tot[q], rn = tot[NR + q], rz = tot[2 * NR + q];
        s.pAp = pAp;
        s.alpha = s.rz / (pAp + s.guard);
        s.rn = rn;
        s.rz_new = rz;
        s.beta = rz / (s.rz + s.guard);
        s.rz = rz;
        s.iters += 1;
        if (!isfinite(s.alpha) || !isfinite(rn) || !isfinite(rz)) s.nonfinite = 1;
        s.done = (rn < s.tolSq) || (s.iters >= s.maxit) || s.nonfinite;
    }
}

// ---------------------------------------------------------------------------
// vector helpers
template <typename T>
__global__ void k_fill(T* __restrict__ v, T val, int64_t n, const int* gate) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    v += batch_off(n);
    for (int64_t i = gtid(); i < n; i += gstride()) v[i] = val;
}
template <typename TD, typename TS>
__global__ void k_convert(TD* __restrict__ d, const TS* __restrict__ s, int64_t n) {
    d += batch_off(n);
    s += batch_off(n);
    for (int64_t i = gtid(); i < n; i += gstride()) d[i] = (TD)s[i];
}
template <typename TD, typename TS>
__global__ void k_convert_gated(TD* __restrict__ d, const TS* __restrict__ s, int64_t n, const int* gate) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    d += batch_off(n);
    s += batch_off(n);
    for (int64_t i = gtid(); i < n; i += gstride()) d[i] = (TD)s[i];
}
template <typename T>
__global__ void k_mask(T* __restrict__ v, const uint8_t* __restrict__ fixed, int64_t n) {
    v += batch_off(n);
    for (int64_t i = gtid(); i < n; i += gstride())
        if (fixed[i]) v[i] = T(0);
}
// r = masked F (cold start of ebeMGPCG :1104-1112)
template <typename T>
__global__ void k_masked_copy(T* __restrict__ d, const double* __restrict__ s, const uint8_t* __restrict__ fixed, int64_t n) {
    for (int64_t i = gtid(); i < n; i += gstride()) d[i] = fixed[i] ? T(0) : (T)s[i];
}
// z = invD r
template <typename T>
__global__ void k_jacobi_apply(T* __restrict__ z, const T* __restrict__ r, const T* __restrict__ invD, int64_t n) {
    for (int64_t i = gtid(); i < n; i += gstride()) z[i] = invD[i] * r[i];
}
// U += s * d (refinement step, FP64 += FP32)
template <typename TS>
__global__ void k_axpy_refine(double* __restrict__ U, const TS* __restrict__ d, double scale, int64_t n) {
    for (int64_t i = gtid(); i < n; i += gstride()) U[i] += scale * (double)d[i];
}

__global__ void k_add_f(float* __restrict__ a, const float* __restrict__ b, int64_t n) {
    for (int64_t i = gtid(); i < n; i += gstride()) a[i] = a[i] + b[i];
}

// dot products into a device double (deterministic)
template <typename TA, typename TB>
__global__ void __launch_bounds__(kRedThreads) k_dot(const TA* __restrict__ a, const TB* __restrict__ b, int64_t n,
                                                    double* out, double* partials, unsigned* ticket) {
    double v[1] = {0.0};
    for (int64_t i = gtid(); i < n; i += gstride()) v[0] += (double)a[i] * (b ? (double)b[i] : 1.0);
    double tot[1];
    if (grid_reduce<1>(v, partials, ticket, tot)) *out = tot[0];
}

// (b == nullptr: sum of a)

// scaled residual for the mixed inner solve: rs = (float)(r / s), s = ||r||_inf-ish scale
template <typename T>
__global__ void k_scale_to(T* __restrict__ d, const double* __restrict__ s, const double* scale, int64_t n) {
    const double inv = 1.0 / *scale;
    for (int64_t i = gtid(); i < n; i += gstride()) d[i] = (T)(s[i] * inv);
}

// power iteration for lambda_max(D^-1 A): v0 hashed on free DOFs, nv = ||v0||^2
__global__ void k_power_init(double* __restrict__ v, const uint8_t* __restrict__ fixed, int64_t n, double* sc) {
    for (int64_t i = gtid(); i < n; i += gstride())
        v[i] = fixed[i] ? 0.0 : (double)(((uint64_t)i * 2654435761ULL) & 0xffffffffULL) / 4294967296.0 - 0.5;
    if (gtid() == 0) sc[0] = -1.0;   // nv computed by the first scale step
}
// w *= invD; nw = ||w||^2; rho = sqrt(nw/nv) (sc: [nv, nw, rho, 1/sqrt(nw)])
__global__ void __launch_bounds__(kRedThreads) k_power_scale(double* __restrict__ w, const double* __restrict__ invD,
                                                            int64_t n, double* sc, double* partials, unsigned* ticket) {
    double v[1] = {0.0};
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const double t = w[i] * invD[i];
        w[i] = t;
        v[0] += t * t;
    }
    double tot[1];
    if (grid_reduce<1>(v, partials, ticket, tot)) {
        sc[1] = tot[0];
        sc[3] = tot[0] > 0 ? 1.0 / sqrt(tot[0]) : 0.0;
    }
}
__global__ void k_power_norm(double* __restrict__ v, const double* __restrict__ w, int64_t n, const double* sc) {
    // v holds the previous iterate: nv = ||v||^2 is 1 after the first step
    const double s = sc[3];
    for (int64_t i = gtid(); i < n; i += gstride()) v[i] = w[i] * s;
}

// ---------------------------------------------------------------------------
// Multigrid transfer (trilinear, clamped at odd/boundary nodes like
// _buildProlongMap js/optimizer-worker-3d.js:591-603).
__device__ __forceinline__ void pw1d(int f, int nC, int& c0, int& c1, double& w0, double& w1) {
    c0 = min(f >> 1, nC - 1);
    c1 = min(c0 + 1, nC - 1);
    const double t = (c0 != c1 && (f & 1)) ? 0.5 : 0.0;
    w0 = 1 - t;
    w1 = t;
}

// Contributors of coarse index c along one axis under pw1d: fine 2c-1
// (weight 1/2, when c > 0), 2c (1) and 2c+1 (1/2, or 1 at the last coarse
// index; absent past the end of the fine axis). Absent slots get weight 0
// and an in-range index, so the gather below has a fixed trip count.
template <typename T>
__device__ __forceinline__ void rw1d(int c, int nF, int nC, int (&f)[3], T (&w)[3]) {
    f[0] = c > 0 ? 2 * c - 1 : 0;
    w[0] = c > 0 ? T(0.5) : T(0);
    f[1] = 2 * c;
    w[1] = T(1);
    const bool has = 2 * c + 1 < nF;
    f[2] = has ? 2 * c + 1 : 2 * c;
    w[2] = has ? (c < nC - 1 ? T(0.5) : T(1)) : T(0);
}

// bC = P^T res, gathered per coarse node; contributions are added in the
// reference's visiting order (fine fz, fy, fx ascending; one corner each,
// mgVcycle assembly/index.ts:868-961), then fixed coarse DOFs zeroed (:964-967).
// The weights are products of 1/2 and 1, exact in T; a zero-weight slot adds
// an exact zero, so the sums are those of the contributors alone. The 27
// fine node values are all loaded before the sums (the compiler otherwise
// waits for each load before issuing the next).
// With bf (the coarsest level's direct solve), the free coarse DOFs are
// written there instead, at their positions freeOf[] of the solve's order
// and rounded to TB as the solve's gather of bC would (bC is not written;
// the positions without a DOF are left as they are: zero).
template <typename T, typename TB = T>
__global__ void __launch_bounds__(256) k_restrict(const T* __restrict__ res, T* __restrict__ bC, Grid gf, Grid gc,
                                                  const uint8_t* __restrict__ fixedC, const int* gate,
                                                  TB* __restrict__ bf = nullptr,
                                                  const int32_t* __restrict__ freeOf = nullptr, int64_t nfp = 0) {
#pragma clang fp contract(off)
    gate = batch_gate(gate);
    if (gate && *gate) return;
    res += batch_off(gf.ndof());
    bC += batch_off(gc.ndof());
    if (bf) bf += batch_off(nfp);
    const int nxF = gf.nnx(), nyF = gf.nny(), nzF = gf.nnz();
    const int nxC = gc.nnx(), nyC = gc.nny(), nzC = gc.nnz();
    for (int64_t nc = xcd_gtid(); nc < gc.nnodes(); nc += gstride()) {
        // 32-bit index arithmetic (node counts below 2^32)
        const uint32_t u = (uint32_t)nc, t = u / (uint32_t)nzC, cx = t / (uint32_t)nyC;
        const int cz = (int)(u - t * (uint32_t)nzC), cy = (int)(t - cx * (uint32_t)nyC);
        int fx[3], fy[3], fz[3];
        T wx[3], wy[3], wz[3];
        rw1d<T>((int)cx, nxF, nxC, fx, wx);
        rw1d<T>(cy, nyF, nyC, fy, wy);
        rw1d<T>(cz, nzF, nzC, fz, wz);
        T v[27][3];
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int iy = 0; iy < 3; iy++)
#pragma unroll
                for (int ix = 0; ix < 3; ix++) {
                    const uint32_t nf = ((uint32_t)fx[ix] * (uint32_t)nyF + (uint32_t)fy[iy]) * (uint32_t)nzF +
                                        (uint32_t)fz[iz];
#pragma unroll
                    for (int d = 0; d < 3; d++) v[(iz * 3 + iy) * 3 + ix][d] = res[3 * (int64_t)nf + d];
                }
        uint8_t fc[3];
#pragma unroll
        for (int d = 0; d < 3; d++) fc[d] = fixedC[3 * nc + d];
        T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int iy = 0; iy < 3; iy++)
#pragma unroll
                for (int ix = 0; ix < 3; ix++) {
                    const T w = (wx[ix] * wy[iy]) * wz[iz];
#pragma unroll
                    for (int d = 0; d < 3; d++) acc[d] += w * v[(iz * 3 + iy) * 3 + ix][d];
                }
        if (bf) {
            for (int d = 0; d < 3; d++) {
                const int32_t pos = freeOf[3 * nc + d];
                if (pos >= 0) bf[pos] = (TB)(fc[d] ? T(0) : acc[d]);
            }
        } else {
            for (int d = 0; d < 3; d++) bC[3 * nc + d] = fc[d] ? T(0) : acc[d];
        }
    }
}

// x += P xC on free fine DOFs (mgVcycle :977-1042): the 8 corner terms are
// summed left to right in the order 000,100,010,110,001,101,011,111.
// Thread C, a coarse node: the fine nodes f = (2Cx + dx, 2Cy + dy, 2Cz + dz),
// d* in {0,1} (inside the grid), have the corners C and min(C+1, nC-1) per
// axis (pw1d of f gives c0 = C), so one load of the 8 coarse values serves
// all eight; consecutive threads update consecutive z-pairs of each of the
// four (dx, dy) rows.
template <typename T>
__global__ void __launch_bounds__(256) k_prolong_add(T* __restrict__ x, const T* __restrict__ xC, Grid gf, Grid gc,
                              const uint8_t* __restrict__ fixedF, const int* gate) {
#pragma clang fp contract(off)
    gate = batch_gate(gate);
    if (gate && *gate) return;
    x += batch_off(gf.ndof());
    xC += batch_off(gc.ndof());
    const int nxF = gf.nnx(), nyF = gf.nny(), nzF = gf.nnz();
    const int nxC = gc.nnx(), nyC = gc.nny(), nzC = gc.nnz();
    const int64_t ncn = gc.nnodes();
    for (int64_t nc = xcd_gtid(); nc < ncn; nc += gstride()) {
        // 32-bit index arithmetic (node counts below 2^32)
        const uint32_t u = (uint32_t)nc, t = u / (uint32_t)nzC, ux = t / (uint32_t)nyC;
        const int cz = (int)(u - t * (uint32_t)nzC), cy = (int)(t - ux * (uint32_t)nyC), cx = (int)ux;
        const int c1[3] = {min(cx + 1, nxC - 1), min(cy + 1, nyC - 1), min(cz + 1, nzC - 1)};
        const int c0[3] = {cx, cy, cz};
        // every load first, none behind a branch (a fine node outside the
        // grid reads entry 0 and is not written)
        T v8[8][3];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int64_t n = ((int64_t)(k & 1 ? c1[0] : c0[0]) * nyC + (k & 2 ? c1[1] : c0[1])) * nzC +
                              (k & 4 ? c1[2] : c0[2]);
#pragma unroll
            for (int d = 0; d < 3; d++) v8[k][d] = xC[3 * n + d];
        }
        bool ok[4][2];
        int64_t i0[4][2];
        T xv[4][2][3];
        uint8_t fv[4][2][3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int fx = 2 * cx + (q & 1), fy = 2 * cy + (q >> 1);
            const bool inxy = fx < nxF && fy < nyF;
#pragma unroll
            for (int dzf = 0; dzf < 2; dzf++) {
                const int fz = 2 * cz + dzf;
                ok[q][dzf] = inxy && fz < nzF;
                i0[q][dzf] = ok[q][dzf] ? 3 * (((int64_t)fx * nyF + fy) * nzF + fz) : 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    xv[q][dzf][d] = x[i0[q][dzf] + d];
                    fv[q][dzf][d] = fixedF[i0[q][dzf] + d];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int dxf = q & 1, dyf = q >> 1;
            const double tx = (c0[0] != c1[0] && dxf) ? 0.5 : 0.0;
            const double ty = (c0[1] != c1[1] && dyf) ? 0.5 : 0.0;
            const double wx[2] = {1 - tx, tx}, wy[2] = {1 - ty, ty};
#pragma unroll
            for (int dzf = 0; dzf < 2; dzf++) {
                const double tz = (c0[2] != c1[2] && dzf) ? 0.5 : 0.0;
                const double wz[2] = {1 - tz, tz};
                // no branch on the mask: a fixed DOF adds an exact zero
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    T v = v8[0][d] * (T)((wx[0] * wy[0]) * wz[0]);
#pragma unroll
                    for (int k = 1; k < 8; k++)
                        v = v + v8[k][d] * (T)((wx[k & 1] * wy[(k >> 1) & 1]) * wz[k >> 2]);
                    const T xi = xv[q][dzf][d];
                    if (ok[q][dzf]) x[i0[q][dzf] + d] = fv[q][dzf][d] ? xi : xi + v;
                }
            }
        }
    }
}

// coarse moduli: E_C = 2 * mean of the existing fine children (dz,dy,dx
// order), _restrictEvalAvg js/optimizer-worker-3d.js:702-724
__global__ void k_coarsen_E(const double* __restrict__ Ef, double* __restrict__ Ec, float* __restrict__ Ec32,
                            Grid gf, Grid gc) {
#pragma clang fp contract(off)
    for (int64_t i = gtid(); i < gc.nel(); i += gstride()) {
        const int cz = (int)(i % gc.nelz);
        const int64_t t = i / gc.nelz;
        const int cy = (int)(t % gc.nely), cx = (int)(t / gc.nely);
        double sum = 0;
        int cnt = 0;
        for (int dz = 0; dz < 2; dz++) {
            const int fz = 2 * cz + dz;
            if (fz >= gf.nelz) continue;
            for (int dy = 0; dy < 2; dy++) {
                const int fy = 2 * cy + dy;
                if (fy >= gf.nely) continue;
                for (int dx = 0; dx < 2; dx++) {
                    const int fx = 2 * cx + dx;
                    if (fx >= gf.nelx) continue;
                    sum += Ef[((int64_t)fx * gf.nely + fy) * gf.nelz + fz];
                    cnt++;
                }
            }
        }
        const double v = cnt > 0 ? (sum / cnt) * 2.0 : 0.0;
        Ec[i] = v;
        if (Ec32) Ec32[i] = (float)v;
    }
}

// ---------------------------------------------------------------------------
// dense levels (Galerkin operators, ndof <= galerkin_max_ndof)
// y = K x, one wave per row, or a damped-Jacobi sweep / residual with it
template <typename T>
__global__ void __launch_bounds__(256) k_dense_apply(const T* __restrict__ K, const T* __restrict__ xin, T* __restrict__ out,
                                                     int64_t n, const uint8_t* __restrict__ fixed, int mode,
                                                     const T* __restrict__ b, const T* __restrict__ invD,
                                                     const double* __restrict__ omp, const int* gate) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    {
        const int64_t bo = batch_off(n);
        xin += bo; out += bo;
        if (b) b += bo;
    }
    const T omega = omp ? (T)*omp : T(0);
    const int lane = threadIdx.x & 63;
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (row >= n) return;
    double s = 0;
    for (int64_t j = lane; j < n; j += 64) s += (double)K[row * n + j] * (double)xin[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) {
        T v = (T)s;
        if (mode == 1) {
            if (fixed[row]) v = T(0);
        } else if (mode == 2) {
            v = fixed[row] ? T(0) : b[row] - v;
        } else if (mode == 3) {
            const T r = b[row] - v;
            v = fixed[row] ? xin[row] : xin[row] + omega * invD[row] * r;
        }
        out[row] = v;
    }
}

__global__ void k_dense_invdiag(const double* __restrict__ K, int64_t n, const uint8_t* __restrict__ fixed,
                                double* __restrict__ invD, float* __restrict__ invD32) {
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const double d = K[i * n + i];
        const double v = (!fixed[i] && d > 1e-30) ? 1.0 / d : 0.0;
        invD[i] = v;
        if (invD32) invD32[i] = (float)v;
    }
}

// Galerkin P^T K_f P from fine elements (_assembleGalerkinFromElements
// js/optimizer-worker-3d.js:628-662); one thread per (element, i, j),
// FP64 atomics into the dense coarse matrix.
__global__ void k_galerkin_elems(const double* __restrict__ Ef, double thr, Grid gf, Grid gc, double* __restrict__ Kc,
                                 const double* __restrict__ ke) {
    const int nyF = gf.nny(), nzF = gf.nnz();
    const int nxC = gc.nnx(), nyC = gc.nny(), nzC = gc.nnz();
    const int64_t ndofC = gc.ndof();
    const int64_t total = gf.nel() * 576;
    for (int64_t w = gtid(); w < total; w += gstride()) {
        const int64_t e = w / 576;
        const int ij = (int)(w % 576), i = ij / 24, j = ij % 24;
        const double Ee = Ef[e];
        if (!(Ee > thr)) continue;
        const double kij = Ee * ke[i * 24 + j];
        if (kij == 0) continue;
        const int ez = (int)(e % gf.nelz);
        const int64_t t = e / gf.nelz;
        const int ey = (int)(t % gf.nely), ex = (int)(t / gf.nely);
        int a, b, c;
        local_offsets(i / 3, a, b, c);
        const int fi[3] = {ex + a, ey + b, ez + c};
        local_offsets(j / 3, a, b, c);
        const int fj[3] = {ex + a, ey + b, ez + c};
        const int di = i % 3, dj = j % 3;
        int ci[3][2], cj[3][2];
        double wi[3][2], wj[3][2];
        const int nC[3] = {nxC, nyC, nzC};
        for (int ax = 0; ax < 3; ax++) {
            pw1d(fi[ax], nC[ax], ci[ax][0], ci[ax][1], wi[ax][0], wi[ax][1]);
            pw1d(fj[ax], nC[ax], cj[ax][0], cj[ax][1], wj[ax][0], wj[ax][1]);
        }
        (void)nyF; (void)nzF;
        for (int qa = 0; qa < 8; qa++) {
            const double wa = wi[0][qa & 1] * wi[1][(qa >> 1) & 1] * wi[2][qa >> 2];
            if (!(wa > 0)) continue;
            const int64_t cr = 3 * (((int64_t)ci[0][qa & 1] * nyC + ci[1][(qa >> 1) & 1]) * nzC + ci[2][qa >> 2]) + di;
            for (int qb = 0; qb < 8; qb++) {
                const double wb = wj[0][qb & 1] * wj[1][(qb >> 1) & 1] * wj[2][qb >> 2];
                if (!(wb > 0)) continue;
                const int64_t cc = 3 * (((int64_t)cj[0][qb & 1] * nyC + cj[1][(qb >> 1) & 1]) * nzC + cj[2][qb >> 2]) + dj;
                atomicAdd(&Kc[cr * ndofC + cc], wa * kij * wb);
            }
        }
    }
}

// T = K_f P  (nF x nC), then Kc = P^T T (_assembleGalerkinFromMatrix :665-699)
__device__ __forceinline__ int prolong_entries(int64_t fdof, Grid gf, Grid gc, int64_t* cdof, double* w) {
    const int nyF = gf.nny(), nzF = gf.nnz();
    const int64_t nfn = fdof / 3;
    const int d = (int)(fdof % 3);
    const int fz = (int)(nfn % nzF);
    const int64_t t = nfn / nzF;
    const int fy = (int)(t % nyF), fx = (int)(t / nyF);
    int cx[2], cy[2], cz[2];
    double wx[2], wy[2], wz[2];
    pw1d(fx, gc.nnx(), cx[0], cx[1], wx[0], wx[1]);
    pw1d(fy, gc.nny(), cy[0], cy[1], wy[0], wy[1]);
    pw1d(fz, gc.nnz(), cz[0], cz[1], wz[0], wz[1]);
    int cnt = 0;
    for (int q = 0; q < 8; q++) {
        const double ww = wx[q & 1] * wy[(q >> 1) & 1] * wz[q >> 2];
        if (ww > 0) {
            cdof[cnt] = 3 * (((int64_t)cx[q & 1] * gc.nny() + cy[(q >> 1) & 1]) * gc.nnz() + cz[q >> 2]) + d;
            w[cnt] = ww;
            cnt++;
        }
    }
    return cnt;
}
// T (nF x nC) zeroed first; T[i][c] += K_f[i][k] * P[k][c]
__global__ void k_galerkin_KP(const double* __restrict__ Kf, int64_t nF, Grid gf, Grid gc, double* __restrict__ T) {
    const int64_t nC = gc.ndof();
    for (int64_t w = gtid(); w < nF * nF; w += gstride()) {
        const int64_t i = w / nF, k = w % nF;
        const double kik = Kf[i * nF + k];
        if (kik == 0) continue;
        int64_t cd[8];
        double ww[8];
        const int cnt = prolong_entries(k, gf, gc, cd, ww);
        for (int q = 0; q < cnt; q++) atomicAdd(&T[i * nC + cd[q]], kik * ww[q]);
    }
}
// Kc[c][j] += P[k][c] * T[k][j]
__global__ void k_galerkin_PtT(const double* __restrict__ T, int64_t nF, Grid gf, Grid gc, double* __restrict__ Kc) {
    const int64_t nC = gc.ndof();
    for (int64_t w = gtid(); w < nF * nC; w += gstride()) {
        const int64_t k = w / nC, j = w % nC;
        const double tkj = T[k * nC + j];
        if (tkj == 0) continue;
        int64_t cd[8];
        double ww[8];
        const int cnt = prolong_entries(k, gf, gc, cd, ww);
        for (int q = 0; q < cnt; q++) atomicAdd(&Kc[cd[q] * nC + j], ww[q] * tkj);
    }
}
// dense K of an EbE level (for the coarsest Cholesky): K[r][c] += E K0
__global__ void k_assemble_dense(const double* __restrict__ E, double thr, Grid g, double* __restrict__ K,
                                 const double* __restrict__ ke) {
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t ndof = g.ndof();
    for (int64_t w = gtid(); w < g.nel() * 576; w += gstride()) {
        const int64_t e = w / 576;
        const int ij = (int)(w % 576), i = ij / 24, j = ij % 24;
        const double Ee = E[e];
        if (!(Ee > thr)) continue;
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        int a, b, c;
        local_offsets(i / 3, a, b, c);
        const int64_t ni = ((int64_t)(ex + a) * nny + (ey + b)) * nnz + (ez + c);
        local_offsets(j / 3, a, b, c);
        const int64_t nj = ((int64_t)(ex + a) * nny + (ey + b)) * nnz + (ez + c);
        atomicAdd(&K[(3 * ni + i % 3) * ndof + 3 * nj + j % 3], Ee * ke[i * 24 + j]);
    }
}

// ---------------------------------------------------------------------------
// SIMP element kernels (zyx element order)

// Heaviside projection _heavisideProject js/optimizer-worker-3d.js:3500-3509
__global__ void k_heaviside(const float* __restrict__ x, float* __restrict__ xp, int64_t n, double beta, double eta) {
#pragma clang fp contract(off)
    const double tb = tanh(beta * eta), t1 = tanh(beta * (1 - eta)), den = tb + t1;
    for (int64_t i = gtid(); i < n; i += gstride()) xp[i] = (float)((tb + tanh(beta * ((double)x[i] - eta))) / den);
}

// energies w_e = u_e^T K0 u_e (f64 sum, Float32 store; :2163-2173 with
// _computeElementEnergyFlat :1561-1571), stress = E_e w_e, and the
// compliance sensitivity dc (:2175-2192):
//   dc = -p xPhys^(p-1) E0 w * (proj ? beta (1 - tanh^2(beta(xd - eta))) / (2 tanh(beta/2)) : 1)
// where xd is the projected field's argument (x, or the density-filtered x).
// clamp = 1 sets energies below 0 to 0 (u^T K0 u >= 0 for the positive
// semidefinite K0, so a negative sum is rounding of a near-rigid element
// motion); clamp = 0 keeps the reference's value.
__global__ void __launch_bounds__(256) k_energy_sens(const double* __restrict__ U, const float* __restrict__ xphys, const float* __restrict__ xd,
                              double penal, double E0, double Emin, int use_proj, double beta, double eta, Grid g,
                              float* __restrict__ w, float* __restrict__ stress, float* __restrict__ dc,
                              const double* __restrict__ ke, int clamp) {
#pragma clang fp contract(off)
    // K0 staged in LDS (rows read back as broadcasts), and each thread's
    // element vector (u_i for the row loop, which is not unrolled)
    __shared__ double s_ke[576];
    __shared__ double s_u[24][256];
    for (int k = threadIdx.x; k < 576; k += blockDim.x) s_ke[k] = ke[k];
    __syncthreads();
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t nyz = (int64_t)nny * nnz;
    const double denom = 2 * tanh(beta * 0.5);
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        const int64_t n0 = ((int64_t)ex * nny + ey) * nnz + ez;
        double u[24];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            int a, b, c;
            local_offsets(m, a, b, c);
            const int64_t nm = n0 + a * nyz + b * nnz + c;
            u[3 * m] = U[3 * nm];
            u[3 * m + 1] = U[3 * nm + 1];
            u[3 * m + 2] = U[3 * nm + 2];
        }
        // the reference's order, term by term: row i of K0 and u_i from LDS,
        // the u_j in registers
#pragma unroll
        for (int i = 0; i < 24; i++) s_u[i][threadIdx.x] = u[i];
        double en = 0;
#pragma unroll 1
        for (int i = 0; i < 24; i++) {
            const double ui = s_u[i][threadIdx.x];
            const double* kr = s_ke + i * 24;
#pragma unroll
            for (int j = 0; j < 24; j++) en += ui * kr[j] * u[j];
        }
        if (clamp && en < 0.0) en = 0.0;
        const float wf = (float)en;
        w[e] = wf;
        const double xp = (double)xphys[e];
        const double stiff = Emin + pow(xp, penal) * (E0 - Emin);
        if (stress) stress[e] = (float)(stiff * (double)wf);
        const double dphys = -penal * pow(xp, penal - 1) * E0 * (double)wf;
        double dd = 1.0;
        if (use_proj) {
            const double th = tanh(beta * ((double)xd[e] - eta));
            dd = beta * (1 - th * th) / denom;
        }
        dc[e] = (float)(dphys * dd);
    }
}

// Sensitivity filter (filterSensitivities js/optimizer-worker-3d.js:3052-3064
// over the prepareFilter neighbourhood :3006-3050): for each element the
// neighbours are visited x-major, then y, then z, accumulating in Float32
// with FP64 products; Hs is precomputed the same way.
constexpr int kFilterMaxR = 5;   // window half width up to 5 (filterRadius up to 5 in the reference UI)
constexpr int kFilterMaxCount = (2 * kFilterMaxR + 1) * (2 * kFilterMaxR + 1) * (2 * kFilterMaxR + 1);
struct FilterStencil {
    int r;            // window half width
    int count;        // number of (dx,dy,dz) with dist <= rmin
    int off[kFilterMaxCount][3];
    double w[kFilterMaxCount];    // rmin - dist
};

__global__ void k_filter_hs(float* __restrict__ Hs, Grid g, const FilterStencil* __restrict__ st) {
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        float hs = 0.f;
        for (int q = 0; q < st->count; q++) {
            const int kx = ex + st->off[q][0], ky = ey + st->off[q][1], kz = ez + st->off[q][2];
            if (kx < 0 || kx >= g.nelx || ky < 0 || ky >= g.nely || kz < 0 || kz >= g.nelz) continue;
            hs = (float)((double)hs + st->w[q]);
        }
        Hs[e] = hs;
    }
}

__global__ void k_sens_filter(const float* __restrict__ x, const float* __restrict__ dc, const float* __restrict__ Hs,
                              float* __restrict__ dcn, Grid g, const FilterStencil* __restrict__ st) {
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        float acc = 0.f;
        for (int q = 0; q < st->count; q++) {
            const int kx = ex + st->off[q][0], ky = ey + st->off[q][1], kz = ez + st->off[q][2];
            if (kx < 0 || kx >= g.nelx || ky < 0 || ky >= g.nely || kz < 0 || kz >= g.nelz) continue;
            const int64_t j = ((int64_t)kx * g.nely + ky) * g.nelz + kz;
            acc = (float)((double)acc + st->w[q] * (double)x[j] * (double)dc[j]);
        }
        const double xe = (double)x[e];
        dcn[e] = (float)((double)acc / ((double)Hs[e] * (xe > 1e-3 ? xe : 1e-3)));
    }
}

// Density filter (applyDensityFilter assembly/index.ts:224-273: window
// ceil(r), weight r - dist, normalised; neighbours z-major). FP64 sums.
// sw[e] = sum of weights. transpose=1: out[j] = sum_i w_ij in[i] / sw[i].
__global__ void k_density_sw(double* __restrict__ sw, Grid g, const FilterStencil* __restrict__ st) {
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        double s = 0;
        for (int q = 0; q < st->count; q++) {
            const int kx = ex + st->off[q][0], ky = ey + st->off[q][1], kz = ez + st->off[q][2];
            if (kx < 0 || kx >= g.nelx || ky < 0 || ky >= g.nely || kz < 0 || kz >= g.nelz) continue;
            s += st->w[q];
        }
        sw[e] = s;
    }
}
__global__ void k_density_filter(const float* __restrict__ in, float* __restrict__ out, const double* __restrict__ sw,
                                 Grid g, int transpose, const FilterStencil* __restrict__ st) {
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        double s = 0;
        for (int q = 0; q < st->count; q++) {
            const int kx = ex + st->off[q][0], ky = ey + st->off[q][1], kz = ez + st->off[q][2];
            if (kx < 0 || kx >= g.nelx || ky < 0 || ky >= g.nely || kz < 0 || kz >= g.nelz) continue;
            const int64_t j = ((int64_t)kx * g.nely + ky) * g.nelz + kz;
            s += transpose ? st->w[q] * (double)in[j] / sw[j] : st->w[q] * (double)in[j];
        }
        out[e] = transpose ? (float)s : (float)(s / sw[e]);
    }
}

// ---------------------------------------------------------------------------
// OC bisection (OC js/optimizer-worker-3d.js:3066-3101) evaluated for a
// whole subtree of bisection steps per pass: the K = 2^depth - 1 midpoints
// of the next `depth` steps are evaluated in one sweep over the elements,
// then the path through the tree is walked on the device. The sequence of
// lmid values is exactly the reference's.
constexpr int kOcDepth = 3;
constexpr int kOcCand = (1 << kOcDepth) - 1;

struct OCState {
    double l1, l2;
    double lam[kOcCand];   // candidate midpoints (tree in BFS order)
    double rsl[kOcCand];   // 1 / sqrt(lam) (the fast path's scale)
    int nlam;
    int passes;            // bisection steps taken
    int finished;
    double target;         // volfrac * nel
    double vol_last;
    double lmid_last;      // midpoint of the last executed step
};

__device__ __forceinline__ float oc_xnew(float xf, float dcf, double lmid, double move, uint8_t pres, uint8_t vd) {
#pragma clang fp contract(off)
    if (pres) return 1.0f;
    if (vd) return 0.0f;
    const double Be = -(double)dcf / lmid;
    const double xi = (double)xf;
    return (float)js_max(0.0, js_max(xi - move, js_min(1.0, js_min(xi + move, xi * sqrt(Be)))));
}

// The same xnew for an element with finite x > 0 and finite dc < 0 (not
// preserved, not void), from s = x sqrt(-dc) (per element) and
// rsl = 1 / sqrt(lmid) (per candidate): t = s rsl equals x sqrt(-dc / lmid)
// to within 4 FP64 ulps, so where t is clearly above min(1, x + move) the
// reference's clamp chain returns that bound, clearly below x - move (> 0)
// that bound, and in between t itself, whose float rounding is the exact
// value's unless t lies within 8 FP64 ulps of a float rounding midpoint.
// Returns false when it cannot decide (the caller runs oc_xnew).
__device__ __forceinline__ bool oc_xnew_fast(double s, double xi, double move, double rsl, float* out) {
#pragma clang fp contract(off)
    constexpr double eps = 1e-12;
    const double t = s * rsl;
    const double a = xi + move;
    const double hi = a < 1.0 ? a : 1.0;
    if (t > hi * (1.0 + eps)) { *out = (float)hi; return true; }
    if (t >= hi * (1.0 - eps)) return false;
    const double b = xi - move;
    if (b > 0.0) {
        if (t < b * (1.0 - eps)) { *out = (float)b; return true; }
        if (t <= b * (1.0 + eps)) return false;
    }
    // interior: the chain returns t (0 <= t < 1)
    const float f = (float)t;
    const double fd = (double)f;
    const unsigned fb = __float_as_uint(f);
    const double up = 0.5 * ((double)__uint_as_float(fb + 1u) - fd);          // f >= 0
    const double dn = fb ? 0.5 * (fd - (double)__uint_as_float(fb - 1u)) : up;
    const double d = t - fd;
    const double tol = 8.0 * 2.220446049250313e-16 * t;
    if (d >= 0.0 ? (d + tol < up) : (tol - d < dn)) { *out = f; return true; }
    return false;
}

// OC update of the compliant mechanism (the damped OC of the classic
// mechanism-synthesis codes; no reference counterpart): B_e = max(1e-10,
// -dc_e / lmid), x_new = clamp(x B^0.3, x -+ move, [0, 1]).
__device__ __forceinline__ float oc_xnew_mech(float xf, float dcf, double lmid, double move, uint8_t pres, uint8_t vd) {
    if (pres) return 1.0f;
    if (vd) return 0.0f;
    const double Be = fmax(1e-10, -(double)dcf / lmid);
    const double xi = (double)xf;
    return (float)fmax(0.0, fmax(xi - move, fmin(1.0, fmin(xi + move, xi * pow(Be, 0.3)))));
}
__device__ __forceinline__ float oc_xnew_any(int mech, float xf, float dcf, double lmid, double move, uint8_t pres,
                                             uint8_t vd) {
    return mech ? oc_xnew_mech(xf, dcf, lmid, move, pres, vd) : oc_xnew(xf, dcf, lmid, move, pres, vd);
}

// fill the candidate tree for the next steps from [l1, l2]
__global__ void k_oc_plan(OCState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (st->finished) { st->nlam = 0; return; }
    // node k (BFS, root 0) covers interval [lo, hi]; its midpoint is lam[k]
    double lo[kOcCand], hi[kOcCand];
    lo[0] = st->l1;
    hi[0] = st->l2;
    for (int k = 0; k < kOcCand; k++) {
        const double m = 0.5 * (hi[k] + lo[k]);
        st->lam[k] = m;
        st->rsl[k] = 1.0 / sqrt(m);
        const int lc = 2 * k + 1, rc = 2 * k + 2;
        if (lc < kOcCand) { lo[lc] = m; hi[lc] = hi[k]; }      // sum > target: l1 = lmid
        if (rc < kOcCand) { lo[rc] = lo[k]; hi[rc] = m; }      // else l2 = lmid
    }
    st->nlam = kOcCand;
}

__global__ void __launch_bounds__(kRedThreads) k_oc_eval(const float* __restrict__ x, const float* __restrict__ dc,
                                                        const uint8_t* __restrict__ pres, const uint8_t* __restrict__ vd,
                                                        double move, int64_t n, OCState* st, double* partials,
                                                        unsigned* ticket, int mech) {
    if (st->finished) return;
    double v[kOcCand];
    double lam[kOcCand], rsl[kOcCand];
#pragma unroll
    for (int k = 0; k < kOcCand; k++) { v[k] = 0.0; lam[k] = st->lam[k]; rsl[k] = st->rsl[k]; }
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const float xf = x[i], dcf = dc[i];
        const uint8_t pr = pres ? pres[i] : 0, vv = vd ? vd[i] : 0;
        const double xi = (double)xf, di = (double)dcf;
        // the fast path: finite x > 0, finite dc < 0, compliance OC
        const bool fast = !mech && !pr && !vv && xi > 0.0 && xi <= 1e30 && di < 0.0 && di > -1e300;
        const double se = fast ? xi * sqrt(-di) : 0.0;
#pragma unroll
        for (int k = 0; k < kOcCand; k++) {
            float val;
            if (!(fast && oc_xnew_fast(se, xi, move, rsl[k], &val)))
                val = oc_xnew_any(mech, xf, dcf, lam[k], move, pr, vv);
            v[k] += (double)val;
        }
    }
    double tot[kOcCand];
    if (grid_reduce<kOcCand>(v, partials, ticket, tot)) {
        // walk the tree with the reference's loop condition
        int k = 0;
        while (k < kOcCand) {
            if (!((st->l2 - st->l1) / (st->l2 + st->l1) > 1e-3)) { st->finished = 1; break; }
            st->vol_last = tot[k];
            st->lmid_last = st->lam[k];
            if (tot[k] > st->target) { st->l1 = st->lam[k]; k = 2 * k + 1; }
            else { st->l2 = st->lam[k]; k = 2 * k + 2; }
            st->passes++;
        }
        if (!((st->l2 - st->l1) / (st->l2 + st->l1) > 1e-3)) st->finished = 1;
    }
}

// final xnew with the last lmid of the reference loop (the xnew left in the
// array by the last executed pass)
__global__ void k_oc_final(const float* __restrict__ x, const float* __restrict__ dc, const uint8_t* __restrict__ pres,
                           const uint8_t* __restrict__ vd, double move, int64_t n, const double* lmid_last,
                           float* __restrict__ xnew, int mech) {
    const double lm = *lmid_last;
    for (int64_t i = gtid(); i < n; i += gstride())
        xnew[i] = oc_xnew_any(mech, x[i], dc[i], lm, move, pres ? pres[i] : 0, vd ? vd[i] : 0);
}

// mode 0: change = max |xnew - xold| with the NaN rule of JS Math.max (a NaN
// entry makes the result NaN, js/optimizer-worker-3d.js:2244-2247), volume =
// sum xnew; mode 1 (b unused): max(0, max_e a_e) with the comparison
// `a_e > m` of the max-stress loop (:2180), which passes NaN over.
__device__ __forceinline__ double chg_max(int mode, double m, double v) {
    if (mode == 0) return js_max(m, v);
    return v > m ? v : m;
}
__global__ void __launch_bounds__(kRedThreads) k_change(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                                                       double* out2, double* partials, unsigned* ticket, int mode) {
    __shared__ double sh[kRedThreads / 64];
    double m = 0.0, s = 0.0;
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const double v = mode == 0 ? (double)fabsf(a[i] - b[i]) : (double)a[i];
        m = chg_max(mode, m, v);
        s += (double)a[i];
    }
    for (int off = 32; off > 0; off >>= 1) m = chg_max(mode, m, __shfl_down(m, off, 64));
    double vs[1] = {s};
    double tot[1];
    // max via a second small reduction in shared memory
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0;
        for (int w = 0; w < kRedThreads / 64; w++) mm = chg_max(mode, mm, sh[w]);
        partials[2 * kRedBlocks + blockIdx.x] = mm;
    }
    __syncthreads();
    // the last block (grid_reduce returns true in its thread 0) also takes
    // the max over the block maxima, with all its threads
    __shared__ int is_last;
    __shared__ double smax[kRedThreads / 64];
    const bool fin = grid_reduce<1>(vs, partials, ticket, tot);
    if (threadIdx.x == 0) is_last = fin ? 1 : 0;
    __syncthreads();
    if (!is_last) return;
    double mm = 0.0;
    for (unsigned bI = threadIdx.x; bI < gridDim.x; bI += blockDim.x)
        mm = chg_max(mode, mm, __hip_atomic_load(&partials[2 * kRedBlocks + bI], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
    for (int off = 32; off > 0; off >>= 1) mm = chg_max(mode, mm, __shfl_down(mm, off, 64));
    if (lane == 0) smax[wid] = mm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m2 = 0.0;
        for (int w = 0; w < kRedThreads / 64; w++) m2 = chg_max(mode, m2, smax[w]);
        out2[0] = m2;
        out2[1] = tot[0];
    }
}

}  // namespace topo
