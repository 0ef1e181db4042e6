// kernels_chol.hpp -- FP64 banded Cholesky of the coarsest MG level and the
// explicit inverse of its factor (build-only: the reference smooths its
// coarsest level instead of solving it).
//
// The free-DOF block A (n x n, n a multiple of 64, row-major, symmetric) of a
// structured-grid operator is banded: with the reference's x-major node order
// every nonzero satisfies |i - j| <= w, w ~ 3 (nny nnz + nnz + 1). Cholesky
// keeps the band, so the blocked right-looking factorization only visits the
// 64 x 64 blocks within wb = ceil(w / 64) block-diagonals:
//   for K: L_KK = chol(A_KK) and L_KK^-1     (k_chol_diag, one wavefront)
//          A_IK <- A_IK L_KK^-T, I in (K, K+wb]         (k_chol_panel)
//          A_IJ -= A_IK A_JK^T,  K < J <= I <= K+wb      (k_chol_update)
// Z = L^-1 is dense lower triangular; its row blocks follow from
//   Z_II = L_II^-1,  Z_I,0:I = -L_II^-1 sum_{K=max(0,I-wb)}^{I-1} L_IK Z_K,0:I
// (two DGEMMs per row block on the host side). The coarse solve is then
// x = Z^T (Z b): two triangular matrix-vector products over Z (row sums by
// k_tri_part / k_tri_apply, column sums from the same rows by
// k_tri_colpart, for Z and for its FP32 and FP16 copies alike).
// All of it is O(n w^2 + n^2 w) work instead of the O(n^3) of a dense inverse.
#pragma once
#include <type_traits>
#include "kernels.hpp"

namespace topo {

constexpr int kCB = 64;           // block size
constexpr int kCS = kCB + 1;      // LDS row stride (bank spread)

// One front of the nested-dissection factorization: a dense row-major matrix
// whose rows and columns are the node's own DOFs (a multiple of 64) followed
// by those of its ancestors. The chain kernels below, given an array of
// these (blockIdx.y = front), factor the own blocks of several independent
// fronts in the same launches: NB = own size / 64 block columns, window =
// all of them (the own block is dense).
struct FrontDev {
    double* A;           // the front
    int64_t ld;          // its size
    double* Linv;        // NB inverses of the diagonal blocks
    const double* d0;    // original diagonal of the own DOFs (pivot test)
    int NB;
};

// load a 64 x 64 block (row-major, ld) into LDS with stride kCS by NT
// threads; all of a thread's global reads are issued before the LDS writes
template <int NT>
__device__ __forceinline__ void blk_load(double* s, const double* A, int64_t ld) {
    constexpr int IT = kCB * kCB / NT;
    double v[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = (int)threadIdx.x + it * NT;
        v[it] = A[(int64_t)(e / kCB) * ld + e % kCB];
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = (int)threadIdx.x + it * NT;
        s[(e / kCB) * kCS + e % kCB] = v[it];
    }
}


// Cholesky of the diagonal block K and the inverse of its factor, by one
// wavefront (lane i owns row i), blocked by 16 columns. Within a sub-block
// the lane's 16 entries stay in registers and the column steps need no LDS
// or barrier: the pivot and the column below it are read from the owning
// lanes (v_readlane), 1/l_jj comes from the hardware reciprocal square root
// refined by two Newton steps (l_jj = d / sqrt(d) to ~1 ulp), and each lane
// updates its own entries. After a sub-block the rows go to LDS and every
// lane updates its trailing columns k with the sub-block's panel,
// r[k] -= l_ij l_kj (the panel read as LDS broadcasts). The inverse, lane c
// solving L x = e_c by forward substitution, multiplies by the kept
// reciprocals and splits each dot product over four partial sums. The block
// load issues all of a lane's reads at once and the pivot thresholds are
// read once. Writes L_KK (upper part zeroed) back to A and L_KK^-1 to
// Linv[K]. A pivot <= tiny * (original diagonal of that DOF) sets *flag.
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__global__ void __launch_bounds__(64) k_chol_diag(double* __restrict__ A, int64_t ld, int K,
                                                  double* __restrict__ Linv, const double* __restrict__ d0,
                                                  double tiny, int* flag, const FrontDev* fr = nullptr) {
    constexpr int SB = 16;           // sub-block
    constexpr int ILP = 4;           // partial sums of the inverse's dot products
    if (fr) {
        const FrontDev d = fr[blockIdx.y];
        if (K >= d.NB) return;
        A = d.A;
        ld = d.ld;
        Linv = d.Linv;
        d0 = d.d0;
    }
    __shared__ double a[kCB * kCS];
    __shared__ double rdiag[kCB];
    double* Ab = A + (int64_t)K * kCB * ld + (int64_t)K * kCB;
    const int i = threadIdx.x;
    // the block: all 64 loads of a lane in flight together
#pragma unroll 16
    for (int rw = 0; rw < kCB; rw++) a[rw * kCS + i] = Ab[(int64_t)rw * ld + i];
    // this lane's pivot threshold, read once
    const double thr_i = tiny * d0[(int64_t)K * kCB + i];
    __syncthreads();
    double* ri = a + i * kCS;        // this lane's row
#pragma unroll
    for (int c0 = 0; c0 < kCB; c0 += SB) {
        double seg[SB];
#pragma unroll
        for (int m = 0; m < SB; m++) seg[m] = ri[c0 + m];
#pragma unroll
        for (int m = 0; m < SB; m++) {
            const int j = c0 + m;
            double colv[SB];         // a_kj of the lanes k below the pivot
#pragma unroll
            for (int m2 = m + 1; m2 < SB; m2++) colv[m2] = readlane_d(seg[m], c0 + m2);
            const double d = readlane_d(seg[m], j);
            if (i == j && !(d > thr_i)) *flag = 1;
            const double dd = d > 0 ? d : 1.0;
            double y = __builtin_amdgcn_rsq(dd);
            y = y * (1.5 - 0.5 * dd * y * y);
            y = y * (1.5 - 0.5 * dd * y * y);
            const double ljj = dd * y;
            const double lij = i == j ? ljj : (i > j ? seg[m] * y : 0.0);
            seg[m] = lij;
            if (i == 0) rdiag[j] = y;
#pragma unroll
            for (int m2 = m + 1; m2 < SB; m2++) seg[m2] -= lij * (colv[m2] * y);
        }
#pragma unroll
        for (int m = 0; m < SB; m++) ri[c0 + m] = seg[m];
        __syncthreads();
        if (c0 + SB < kCB) {
#pragma unroll 4
            for (int k = c0 + SB; k < kCB; k++) {
                const double* rk = a + k * kCS + c0;
                double v = ri[k];
#pragma unroll
                for (int m = 0; m < SB; m++) v -= seg[m] * rk[m];
                ri[k] = v;
            }
            __syncthreads();
        }
    }
    double r[kCB];
#pragma unroll
    for (int k = 0; k < kCB; k++) {
        const double v = ri[k];
        r[k] = k <= i ? v : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCB; k++) ri[k] = r[k];
    __syncthreads();
    // inverse: lane c solves L x = e_c (x_r = 0 for r < c falls out)
    const int c = i;
#pragma unroll
    for (int rr = 0; rr < kCB; rr++) {
        double s[ILP];
#pragma unroll
        for (int q = 0; q < ILP; q++) s[q] = 0.0;
#pragma unroll
        for (int k = 0; k < rr; k++) s[k % ILP] += a[rr * kCS + k] * r[k];
        double t = rr == c ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < ILP; q++) t -= s[q];
        r[rr] = t * rdiag[rr];
    }
    double* Li = Linv + (int64_t)K * kCB * kCB;
#pragma unroll
    for (int rr = 0; rr < kCB; rr++) Li[rr * kCB + c] = r[rr];
    for (int e = i; e < kCB * kCB; e += blockDim.x) {
        const int rw = e / kCB, cl = e % kCB;
        Ab[(int64_t)rw * ld + cl] = a[rw * kCS + cl];
    }
}

// C (64 x 64, 4 x 4 per thread) = sum_k X[r][k] Y[c][k] over LDS blocks
__device__ __forceinline__ void blk_mul_nt(const double* X, const double* Y, double (&acc)[4][4]) {
    const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
    for (int k = 0; k < kCB; k++) {
        double xv[4], yv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) { xv[i] = X[(r0 + i) * kCS + k]; yv[i] = Y[(c0 + i) * kCS + k]; }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++) acc[i][jj] += xv[i] * yv[jj];
    }
}

// A_IK <- A_IK L_KK^-T for the wb blocks below the diagonal block K
__global__ void __launch_bounds__(256) k_chol_panel(double* __restrict__ A, int64_t ld, int K, int NB, int wb,
                                                    const double* __restrict__ Linv, const FrontDev* fr = nullptr) {
    if (fr) {
        const FrontDev d = fr[blockIdx.y];
        A = d.A;
        ld = d.ld;
        Linv = d.Linv;
        NB = d.NB;
        wb = d.NB;
    }
    const int I = K + 1 + (int)blockIdx.x;
    if (I >= NB || I > K + wb) return;
    __shared__ double x[kCB * kCS];
    __shared__ double y[kCB * kCS];
    double* Ab = A + (int64_t)I * kCB * ld + (int64_t)K * kCB;
    blk_load<256>(x, Ab, ld);
    blk_load<256>(y, Linv + (int64_t)K * kCB * kCB, kCB);
    __syncthreads();
    double acc[4][4] = {};
    blk_mul_nt(x, y, acc);   // (A_IK L^-T)[r][c] = sum_k A_IK[r][k] Linv[c][k]
    const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
    for (int i = 0; i < 4; i++)
        for (int jj = 0; jj < 4; jj++) Ab[(int64_t)(r0 + i) * ld + c0 + jj] = acc[i][jj];
}

// A_IJ -= A_IK A_JK^T for K < J <= I <= min(K + wb, NB - 1). part 0: every
// pair of the window (blockIdx.x runs over its lower-triangular pairs);
// part 1: the next block column only (J = K + 1, blockIdx.x = I - K - 1),
// which is all the next diagonal and panel steps need; part 2: the other
// pairs (J >= K + 2), left to run beside those steps.
__global__ void __launch_bounds__(256) k_chol_update(double* __restrict__ A, int64_t ld, int K, int NB, int wb,
                                                     int part = 0, const FrontDev* fr = nullptr) {
    if (fr) {
        const FrontDev d = fr[blockIdx.y];
        A = d.A;
        ld = d.ld;
        NB = d.NB;
    }
    int I, J;
    if (part == 1) {
        I = K + 1 + (int)blockIdx.x;
        J = K + 1;
    } else {
        int p = (int)blockIdx.x, di = 0;
        while (p > di) { p -= di + 1; di++; }
        const int o = part == 2 ? 2 : 1;
        I = K + o + di;
        J = K + o + p;
    }
    if (I >= NB) return;
    __shared__ double x[kCB * kCS];
    __shared__ double y[kCB * kCS];
    blk_load<256>(x, A + (int64_t)I * kCB * ld + (int64_t)K * kCB, ld);
    blk_load<256>(y, A + (int64_t)J * kCB * ld + (int64_t)K * kCB, ld);
    __syncthreads();
    double acc[4][4] = {};
    blk_mul_nt(x, y, acc);
    double* Cb = A + (int64_t)I * kCB * ld + (int64_t)J * kCB;
    const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
    for (int i = 0; i < 4; i++)
        for (int jj = 0; jj < 4; jj++) Cb[(int64_t)(r0 + i) * ld + c0 + jj] -= acc[i][jj];
}

// copy a 64 x 64 block (row-major, leading dimensions lds / ldd)
__global__ void k_copy_block(const double* __restrict__ src, int64_t lds, double* __restrict__ dst, int64_t ldd) {
    for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
        const int r = e / kCB, c = e % kCB;
        dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
    }
}

// Front of a nested-dissection node (f x f, row-major): F[a][b] =
// A[g[a]][g[b]] where a or b is one of the node's own m DOFs, 0 between two
// ancestor DOFs, plus the update blocks of its (at most two) children, whose
// rows and columns are exactly this front's, added in child order. A: the
// coarse free block in the nested-dissection order (ld n); g: front index ->
// position; C0 / C1: a child's update block (ld0 / ld1; its lower triangle,
// read symmetrically) or null.
__global__ void k_front_assemble(double* __restrict__ F, int64_t f, int64_t m, const int32_t* __restrict__ g,
                                 const double* __restrict__ A, int64_t n, const double* __restrict__ C0, int64_t ld0,
                                 const double* __restrict__ C1, int64_t ld1) {
    for (int64_t w = gtid(); w < f * f; w += gstride()) {
        const int64_t a = w / f, b = w - a * f;
        double v = (a < m || b < m) ? A[(int64_t)g[a] * n + g[b]] : 0.0;
        const int64_t hi = a > b ? a : b, lo = a > b ? b : a;
        if (C0) v += C0[hi * ld0 + lo];
        if (C1) v += C1[hi * ld1 + lo];
        F[w] = v;
    }
}

// FP16 copy of a lower-triangular Z with one scale per row: Zq[i][j] =
// Z[i][j] / s_i, s_i = max_j |Z[i][j]| (1 for a zero row), s2[i] = s_i^2.
// Z^T (Z b) = Zq^T (S^2 (Zq b)) with the same Zq in both products: the
// operator applied stays symmetric positive definite whatever the rounding.
// One wavefront per row.
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
// (rfirst, optional: row i is zero before column rfirst[i], not read there.)
__global__ void __launch_bounds__(256) k_rowscale_f16(const double* __restrict__ Z, int64_t n,
                                                      _Float16* __restrict__ Zq, double* __restrict__ s2,
                                                      const int32_t* __restrict__ rfirst = nullptr) {
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= n) return;
    const double* row = Z + i * n;
    const int64_t rf = rfirst ? (int64_t)rfirst[i] : 0;
    double m = 0.0;
    for (int64_t j = rf + lane; j <= i; j += 64) m = fmax(m, fabs(row[j]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    const double sc = m > 0.0 ? m : 1.0;
    const double inv = 1.0 / sc;
    for (int64_t j = lane; j < n; j += 64) Zq[i * n + j] = (_Float16)(float)(j >= rf && j <= i ? row[j] * inv : 0.0);
    if (lane == 0) s2[i] = sc * sc;
}

// Triangular matrix-vector product for NR right-hand sides: rows of M (n x n,
// row-major, TK = double, float or _Float16) against bf (NR x n, TB = FP64,
// or FP32 with the FP16 matrix); the sums are FP64;
// lower: row i uses columns [0, i], upper: [i, n) (the other part of M is
// zero, so whole 16-byte vectors around the diagonal can be read). One wave
// per 4 rows. scatter = 0: y[q n + i] = result (TY), times ys[i] when ys
// is given; scatter = 1: x[q ndof + free[i]] = result in T, for i < nf
// (converged right-hand sides are not written).
template <typename T, int NR, typename TK, typename TB = double, typename TY = double>
__global__ void __launch_bounds__(256) k_tri_apply(const TK* __restrict__ M, int64_t n, int upper,
                                                   const TB* __restrict__ bf, TY* __restrict__ y,
                                                   T* __restrict__ x, const int32_t* __restrict__ freeIdx, int64_t nf,
                                                   int64_t ndof, const int* gate, const double* __restrict__ ys = nullptr) {
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
    if (row0 >= n) return;
    double acc[4][NR];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < NR; q++) acc[r][q] = 0.0;
    constexpr int VN = 16 / (int)sizeof(TK);
    // column range in vectors of VN entries
    const int64_t v0 = upper ? row0 / VN : 0;
    const int64_t v1 = upper ? n / VN : (row0 + 4 + VN - 1) / VN;
    const int64_t nvr = n / VN;
    for (int64_t j = v0 + lane; j < v1; j += 64) {
        double b[NR][VN];
        float b32[NR][VN];
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
            for (int u = 0; u < VN; u++) {
                if constexpr (VN == 8) b32[q][u] = (float)bf[q * n + j * VN + u];
                else b[q][u] = (double)bf[q * n + j * VN + u];
            }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if constexpr (VN == 2) {
                const double2 a = ((const double2*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++) acc[r][q] += a.x * b[q][0] + a.y * b[q][1];
            } else if constexpr (VN == 8) {
                // FP16 entries: the 8 products of a vector are summed in
                // FP32 (errors ~1e-7, far below the 2^-11 of the entries),
                // the vector sums in FP64
                const half8_t a = ((const half8_t*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    float t = 0.0f;
#pragma unroll
                    for (int u = 0; u < 8; u++) t = fmaf((float)a[u], b32[q][u], t);
                    acc[r][q] += (double)t;
                }
            } else {
                const float4 a = ((const float4*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++)
                    acc[r][q] += (double)a.x * b[q][0] + (double)a.y * b[q][1] + (double)a.z * b[q][2] +
                                 (double)a.w * b[q][3];
            }
        }
        (void)nvr;
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < NR; q++)
            for (int off = 32; off > 0; off >>= 1) acc[r][q] += __shfl_down(acc[r][q], off, 64);
    if (lane == 0)
        for (int q = 0; q < NR; q++) {
            if (!live[q]) continue;
            for (int r = 0; r < 4; r++) {
                const int64_t i = row0 + r;
                if (!x) y[q * n + i] = (TY)(ys ? ys[i] * acc[r][q] : acc[r][q]);
                else if (i < nf && freeIdx[i] >= 0) x[q * ndof + freeIdx[i]] = (T)acc[r][q];
            }
        }
}

// Column-chunked form of k_tri_apply for long rows: block (ch, g) covers RW
// rows per wave (4 waves, 4 RW rows) over the columns [ch cw, (ch+1) cw)
// of their triangular range and leaves one FP64 partial per row and chunk,
// part[(q nch + ch) n + i] (0 where the chunk misses the range). No wave
// walks more than cw columns, so the kernel is not held by the latency chain
// of the longest rows; k_tri_sum adds the chunks in ascending order.
// rfirst (lower only, optional): first column of each row's nonzeros (the
// same for the 4 RW rows of a wave, a multiple of 16); the columns before
// it are zero and not read.
template <int NR, typename TK, typename TB, int RW = 4>
__global__ void __launch_bounds__(256) k_tri_part(const TK* __restrict__ M, int64_t n, int upper,
                                                  const TB* __restrict__ bf, double* __restrict__ part, int cw, int nch,
                                                  const int* gate, const int32_t* __restrict__ rfirst = nullptr) {
    bool live[NR];
    bool any = false;
#pragma unroll
    for (int q = 0; q < NR; q++) {
        live[q] = !(gate && gate[(int64_t)q * kGateStride]);
        any = any || live[q];
    }
    if (!any) return;
    const int lane = threadIdx.x & 63;
    const int64_t row0 = (((int64_t)blockIdx.y * blockDim.x + threadIdx.x) >> 6) * RW;
    if (row0 >= n) return;
    const int ch = blockIdx.x;
    constexpr int VN = 16 / (int)sizeof(TK);
    const int64_t c0 = (int64_t)ch * cw;
    const int64_t c1 = c0 + cw < n ? c0 + cw : n;
    // triangular range in whole VN-vectors (the other triangle of M is 0)
    const int64_t rf = (!upper && rfirst) ? (int64_t)rfirst[row0] : 0;
    const int64_t lo = upper ? (c0 > row0 / VN * VN ? c0 : row0 / VN * VN) : (c0 > rf ? c0 : rf);
    const int64_t hv = (row0 + RW + VN - 1) / VN * VN;
    const int64_t hi = upper ? c1 : (c1 < hv ? c1 : hv);
    double acc[RW][NR];
#pragma unroll
    for (int r = 0; r < RW; r++)
#pragma unroll
        for (int q = 0; q < NR; q++) acc[r][q] = 0.0;
    for (int64_t j = lo / VN + lane; j < hi / VN; j += 64) {
        double b[NR][VN];
        float b32[NR][VN];
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
            for (int u = 0; u < VN; u++) {
                if constexpr (VN == 8) b32[q][u] = (float)bf[q * n + j * VN + u];
                else b[q][u] = (double)bf[q * n + j * VN + u];
            }
#pragma unroll
        for (int r = 0; r < RW; r++) {
            if constexpr (VN == 2) {
                const double2 a = ((const double2*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++) acc[r][q] += a.x * b[q][0] + a.y * b[q][1];
            } else if constexpr (VN == 8) {
                // FP16 entries: the 8 products of a vector are summed in
                // FP32 (errors ~1e-7, far below the 2^-11 of the entries),
                // the vector sums in FP64
                const half8_t a = ((const half8_t*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    float t = 0.0f;
#pragma unroll
                    for (int u = 0; u < 8; u++) t = fmaf((float)a[u], b32[q][u], t);
                    acc[r][q] += (double)t;
                }
            } else {
                const float4 a = ((const float4*)(M + (row0 + r) * n))[j];
#pragma unroll
                for (int q = 0; q < NR; q++)
                    acc[r][q] += (double)a.x * b[q][0] + (double)a.y * b[q][1] + (double)a.z * b[q][2] +
                                 (double)a.w * b[q][3];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RW; r++)
#pragma unroll
        for (int q = 0; q < NR; q++)
            for (int off = 32; off > 0; off >>= 1) acc[r][q] += __shfl_down(acc[r][q], off, 64);
    if (lane == 0)
        for (int q = 0; q < NR; q++)
            for (int r = 0; r < RW; r++) part[((int64_t)q * nch + ch) * n + row0 + r] = acc[r][q];
}

// Column sums x_j = sum_i M[i][j] y_i of a triangular FP16 or FP32 M
// (row-major, zeros in the other triangle; upper = 1 for an upper-triangular
// M): each
// lane streams 16-byte pieces of whole rows, the right-hand side is one
// uniform value per row, so these tiles read the matrix at close to the
// copy bandwidth (the row-oriented k_tri_part re-reads the right-hand side
// with every row piece). Z b is the column sums of Z^T, Z^T y those of Z.
// Block
// (cb, rb) covers the rows [rb 256, rb 256 + 256) and the columns
// [cb 512, cb 512 + 512) (FP32: 256 columns): lane l owns 8 (4) columns, wave w 64 of the rows
// (y_i read once per wave as a uniform value); the 4 waves' FP32 column sums
// meet in LDS and leave one FP64 partial per column and row block,
// part[(q nrb + rb) n + j] (0 for blocks above the diagonal), summed by
// k_tri_sum. rfirst (optional): first nonzero column of each row, the same
// for the rows of a wave; a wave whose rows are zero over the block's
// columns does not read them. With ypart, y is not read but formed by the
// block for its rows from the ynch chunk partials of k_tri_part, exactly as
// k_tri_sum would write it (y_i = (ys ? ys_i : 1) * sum, rounded to TB), so
// the row-chunk sums need no kernel of their own.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }
constexpr int kTcRows = 256;      // rows per block
constexpr int kTcCols = 512;      // FP16 columns per block (64 lanes x 8)
template <typename TK> constexpr int tc_cols() { return 64 * (16 / (int)sizeof(TK)); }
// NW waves per block (4 or 8), kTcRows / NW rows each.
template <int NR, typename TB, typename TK = _Float16, int NW = 4>
__global__ void __launch_bounds__(64 * NW) k_tri_colpart(const TK* __restrict__ M, int64_t n,
                                                         const TB* __restrict__ y, double* __restrict__ part, int nrb,
                                                         const int* gate, int upper = 0,
                                                         const int32_t* __restrict__ rfirst = nullptr,
                                                         const double* __restrict__ ypart = nullptr, int ynch = 0,
                                                         const double* __restrict__ ys = nullptr) {
    constexpr int VN = 16 / (int)sizeof(TK);   // entries per 16-byte piece
    constexpr int kTcCols = 64 * VN;
    static_assert(NW == 4 || NW == 8, "4 or 8 waves per block");
    // FP16 / FP32 entries are summed in FP32 within a tile, FP64 in FP64
    typedef typename std::conditional<std::is_same<TK, double>::value, double, float>::type TA;
    __shared__ TA sh[NW][kTcCols];   // the waves' column sums, one right-hand side at a time
    const int cb = blockIdx.x, rb = blockIdx.y;
    const int64_t J0 = (int64_t)cb * kTcCols, I0 = (int64_t)rb * kTcRows;
    bool live[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) live[q] = !(gate && gate[(int64_t)q * kGateStride]);
    if (upper ? (I0 > J0 + kTcCols - 1) : (J0 > I0 + kTcRows - 1)) {
        // outside the triangle: zero partials
        for (int q = 0; q < NR; q++)
            for (int t = threadIdx.x; t < kTcCols; t += blockDim.x)
                if (J0 + t < n) part[((int64_t)q * nrb + rb) * n + J0 + t] = 0.0;
        return;
    }
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t j = J0 + VN * lane;
    const bool colok = j < n;
    // y of the block's rows from the chunk partials: chunks sub, sub + 4, ...
    // in ascending order for sub = 0..3, then (s0 + s1) + (s2 + s3), the
    // order of k_tri_sum's four lanes
    __shared__ TB ysh[NR][kTcRows];
    // this wave's rows and the first nonzero column of its rows, read before
    // the row sums below so the two loads overlap
    const int64_t r0 = I0 + (int64_t)w * (kTcRows / NW);
    const int64_t r1 = r0 + kTcRows / NW < n ? r0 + kTcRows / NW : n;
    const int64_t rf0 = rfirst ? (int64_t)rfirst[r0 < n ? r0 : 0] : 0;
    if (ypart) {
        const int t = threadIdx.x;
        const int64_t i = I0 + t;
        if (t < kTcRows && i < n) {
            const double ysv = ys ? ys[i] : 1.0;
#pragma unroll
            for (int q = 0; q < NR; q++) {
                double ss[4] = {0.0, 0.0, 0.0, 0.0};
                const double* pq = ypart + (int64_t)q * ynch * n + i;
                for (int c0 = 0; c0 < ynch; c0 += 16) {
                    // 16 chunks' loads together, then added in ascending
                    // order per sub (past the end: entry 0, an exact +0)
                    double v[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const int c = c0 + u;
                        v[u] = pq[(int64_t)(c < ynch ? c : 0) * n];
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++) ss[u & 3] += (c0 + u < ynch) ? v[u] : 0.0;
                }
                const double sum = (ss[0] + ss[1]) + (ss[2] + ss[3]);
                ysh[q][t] = (TB)(ys ? ysv * sum : sum);
            }
        }
        __syncthreads();
    }
    TA acc[NR][VN];
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
        for (int u = 0; u < VN; u++) acc[q][u] = TA(0);
    const bool wz = rfirst && r0 < n && rf0 > J0 + kTcCols - 1;   // wave-uniform
    if (colok && !wz) {
#pragma unroll 8
        for (int64_t i = r0; i < r1; i++) {
            TA av[VN];
            if constexpr (VN == 8) {
                const half8_t a = *(const half8_t*)(M + i * n + j);
#pragma unroll
                for (int u = 0; u < 8; u++) av[u] = (float)a[u];
            } else if constexpr (VN == 4) {
                const float4 a = *(const float4*)(M + i * n + j);
                av[0] = a.x; av[1] = a.y; av[2] = a.z; av[3] = a.w;
            } else {
                const double2 a = *(const double2*)(M + i * n + j);
                av[0] = a.x; av[1] = a.y;
            }
#pragma unroll
            for (int q = 0; q < NR; q++) {
                const TA yv = ypart ? (TA)ysh[q][i - I0] : (TA)y[q * n + i];
#pragma unroll
                for (int u = 0; u < VN; u++) acc[q][u] = fma_t(av[u], yv, acc[q][u]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NR; q++) {
#pragma unroll
        for (int u = 0; u < VN; u++) sh[w][VN * lane + u] = acc[q][u];
        __syncthreads();
        if (live[q]) {
            for (int t = threadIdx.x; t < kTcCols; t += blockDim.x) {
                if (J0 + t >= n) continue;
                double v = ((double)sh[0][t] + (double)sh[1][t]) + ((double)sh[2][t] + (double)sh[3][t]);
                if constexpr (NW == 8) v += ((double)sh[4][t] + (double)sh[5][t]) + ((double)sh[6][t] + (double)sh[7][t]);
                part[((int64_t)q * nrb + rb) * n + J0 + t] = v;
            }
        }
        __syncthreads();
    }
}

// row sums of the chunk partials, written like k_tri_apply's epilogue:
// y[q n + i] = (ys ? ys[i] : 1) * sum (x == null), or x[q ndof + free[i]] =
// sum for i < nf. Four lanes per row, lane s adding the chunks s, s+4, ...;
// the four sums meet by two exchanges (the same bits in all four lanes).
template <typename T, int NR, typename TY>
__global__ void k_tri_sum(const double* __restrict__ part, int nch, int64_t n, TY* __restrict__ y, T* __restrict__ x,
                          const int32_t* __restrict__ freeIdx, int64_t nf, int64_t ndof, const int* gate,
                          const double* __restrict__ ys) {
    const int sub = threadIdx.x & 3;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    const bool ok = i < n;
#pragma unroll
    for (int q = 0; q < NR; q++) {
        if (gate && gate[(int64_t)q * kGateStride]) continue;
        double s = 0.0;
        if (ok) {
            // chunks sub, sub + 4, ... in ascending order, loaded four at a
            // time (past the end: entry 0, adding an exact +0)
            const double* pq = part + (int64_t)q * nch * n + i;
            for (int c0 = sub; c0 < nch; c0 += 16) {
                double v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int c = c0 + 4 * u;
                    v[u] = pq[(int64_t)(c < nch ? c : 0) * n];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) s += (c0 + 4 * u < nch) ? v[u] : 0.0;
            }
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (ok && sub == 0) {
            if (!x) y[q * n + i] = (TY)(ys ? ys[i] * s : s);
            else if (i < nf && freeIdx[i] >= 0) x[q * ndof + freeIdx[i]] = (T)s;
        }
    }
}

}  // namespace topo
