// kernels_xm.hpp -- fast matrix-free hex8 stiffness apply for gfx950.
//
// The element operator is applied in the Hadamard (parity) basis of the 8
// element nodes: with H the 2x2x2 sign transform per displacement component,
// K0 = H^T B H / 64 where B is block diagonal over the 8 reflection-parity
// classes (cube symmetry of the isotropic unit-cube element): two 3x3
// classes of the form (a-b) I + b 1 1^T, three 2x2 classes (a translation
// member is zero) and three classes made of a 2x2 shear pair plus a scalar.
// That is ~40 multiply-adds per element instead of 576, plus two butterfly
// transforms.
//
// Work decomposition ("x-march"): a workgroup owns NT consecutive nodes of
// the flattened (y,z) plane (j = iy*nnz + iz, consecutive in memory) and
// marches along x over CX node planes. Per step it loads one node plane
// (coalesced), stages it in LDS for the y/z neighbours, computes the element
// plane behind it, and completes one node plane of the output:
//   - the y/z butterflies of a node-plane face are computed once and reused
//     by the two element planes sharing it (carried in registers);
//   - the element's output for its trailing face is carried, in the face
//     basis, to the next step and added before the inverse y/z butterflies;
//   - node sums over the 4 faces around a node through one LDS exchange;
//   - the next node plane is prefetched into registers while the current
//     element plane is computed (2 barriers per plane).
// Every node value is produced by exactly one thread with a fixed order of
// additions: results are deterministic (no atomics).
#pragma once
#include "kernels.hpp"

namespace topo {

template <typename T>
struct HadCoef {
    T am0, b0;            // class (0,0,0): y_i = am0 v_i + b0 (v0+v1+v2)
    T am7, b7;            // class (1,1,1)
    T oa[3], oc[3], ob[3];  // class e_k: 2x2 [[oa, oc], [oc, ob]]
    T ta[3], tc[3], tb[3];  // class ~e_k: shear pair 2x2
    T tf[3];                // class ~e_k: scalar (1,1,1) member
};

// Floating-point contraction is off in these kernels and the fused
// multiply-adds are written out, so every inlined copy of the march step
// (it is unrolled) forms each value with the same operations: the results do
// not depend on the x-chunking or on which copy computed a plane.
__device__ __forceinline__ float xfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double xfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// face arrays f[qy][qz][d] flattened as f[(qy*2 + qz)*3 + d]
#define FI(qy, qz, d) (((qy) * 2 + (qz)) * 3 + (d))

template <typename T>
__device__ __forceinline__ void yz_forward(const T (&n00)[3], const T (&n01)[3], const T (&n10)[3], const T (&n11)[3],
                                           T (&f)[12]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T s0 = n00[d] + n01[d], d0 = n00[d] - n01[d];
        const T s1 = n10[d] + n11[d], d1 = n10[d] - n11[d];
        f[FI(0, 0, d)] = s0 + s1;
        f[FI(1, 0, d)] = s0 - s1;
        f[FI(0, 1, d)] = d0 + d1;
        f[FI(1, 1, d)] = d0 - d1;
    }
}

// A = u^[qx=0], Bq = u^[qx=1] (in: transformed displacements; out: the
// block-diagonal operator applied; translation members -> 0). The element
// modulus is applied by the caller after the inverse x butterfly.
template <typename T>
__device__ __forceinline__ void had_block(T (&A)[12], T (&Bq)[12], const HadCoef<T>& c) {
#pragma clang fp contract(off)
    // class (0,0,0)
    {
        const T v0 = Bq[FI(0, 0, 0)], v1 = A[FI(1, 0, 1)], v2 = A[FI(0, 1, 2)];
        const T s = c.b0 * ((v0 + v1) + v2);
        Bq[FI(0, 0, 0)] = xfma(c.am0, v0, s);
        A[FI(1, 0, 1)] = xfma(c.am0, v1, s);
        A[FI(0, 1, 2)] = xfma(c.am0, v2, s);
    }
    // class (1,1,1)
    {
        const T v0 = A[FI(1, 1, 0)], v1 = Bq[FI(0, 1, 1)], v2 = Bq[FI(1, 0, 2)];
        const T s = c.b7 * ((v0 + v1) + v2);
        A[FI(1, 1, 0)] = xfma(c.am7, v0, s);
        Bq[FI(0, 1, 1)] = xfma(c.am7, v1, s);
        Bq[FI(1, 0, 2)] = xfma(c.am7, v2, s);
    }
    // classes e_k: members (d ascending, d != k); translation -> 0
    {
        T v1 = Bq[FI(1, 0, 1)], v2 = Bq[FI(0, 1, 2)];
        Bq[FI(1, 0, 1)] = xfma(c.oa[0], v1, c.oc[0] * v2);
        Bq[FI(0, 1, 2)] = xfma(c.oc[0], v1, c.ob[0] * v2);
        v1 = Bq[FI(1, 0, 0)];
        v2 = A[FI(1, 1, 2)];
        Bq[FI(1, 0, 0)] = xfma(c.oa[1], v1, c.oc[1] * v2);
        A[FI(1, 1, 2)] = xfma(c.oc[1], v1, c.ob[1] * v2);
        v1 = Bq[FI(0, 1, 0)];
        v2 = A[FI(1, 1, 1)];
        Bq[FI(0, 1, 0)] = xfma(c.oa[2], v1, c.oc[2] * v2);
        A[FI(1, 1, 1)] = xfma(c.oc[2], v1, c.ob[2] * v2);
        A[FI(0, 0, 0)] = T(0);
        A[FI(0, 0, 1)] = T(0);
        A[FI(0, 0, 2)] = T(0);
    }
    // classes ~e_k: scalar (1,1,1) member d = k, shear pair (d ascending)
    {
        Bq[FI(1, 1, 0)] = c.tf[0] * Bq[FI(1, 1, 0)];
        T v1 = A[FI(0, 1, 1)], v2 = A[FI(1, 0, 2)];
        A[FI(0, 1, 1)] = xfma(c.ta[0], v1, c.tc[0] * v2);
        A[FI(1, 0, 2)] = xfma(c.tc[0], v1, c.tb[0] * v2);
        Bq[FI(1, 1, 1)] = c.tf[1] * Bq[FI(1, 1, 1)];
        v1 = A[FI(0, 1, 0)];
        v2 = Bq[FI(0, 0, 2)];
        A[FI(0, 1, 0)] = xfma(c.ta[1], v1, c.tc[1] * v2);
        Bq[FI(0, 0, 2)] = xfma(c.tc[1], v1, c.tb[1] * v2);
        Bq[FI(1, 1, 2)] = c.tf[2] * Bq[FI(1, 1, 2)];
        v1 = A[FI(1, 0, 0)];
        v2 = Bq[FI(0, 0, 1)];
        A[FI(1, 0, 0)] = xfma(c.ta[2], v1, c.tc[2] * v2);
        Bq[FI(0, 0, 1)] = xfma(c.tc[2], v1, c.tb[2] * v2);
    }
}

// Whether the kernel keeps the staged node planes double-buffered in LDS and
// recomputes the previous plane's face transform from there instead of
// carrying it in registers: the FP64 kernels, whose 24 fewer live registers
// let two 512-thread workgroups share a CU (and keep the 1024-thread one
// within its 128 registers). The FP32 ones measured faster with the carry.
template <typename T, int NT>
constexpr bool xm_recomp() { return sizeof(T) == 8; }

// dynamic LDS of an NT-thread workgroup: s_in[nbuf][(NT + H) * 3] | s_q[NT * 9]
inline size_t xm_lds_bytes(int nnz, size_t elt, int NT) {
    const size_t nbuf = elt == 8 ? 2 : 1;
    return ((size_t)(NT + nnz + 1) * 3 * nbuf + (size_t)NT * 9) * elt;
}

// MODE as k_applyA_exact: 0 Ap, 1 masked, 2 b - Ap (free) / 0 (fixed),
// 3 Jacobi sweep x + omega invD (b - Ap) (free) / x (fixed); 4 the first two
// damped-Jacobi sweeps from x = 0 in one pass: the operand x0 = omega invD b
// is formed from b and invD as the planes are loaded (p is not read), and the
// output is the mode-3 sweep from x0 (the staged own b, invD and x0 serve the
// epilogue). A template parameter so each variant holds only the registers
// its epilogue needs. With `part` (MODE 1), each workgroup also leaves, per
// output node plane ix of its yz tile bx, the plane's share of p.Ap (own p
// of each output node times its output, FP64, fixed order) at
// part[blockIdx.y * kCgStride + bx * nnx + ix] for the CG update (gx * nnx
// partials whatever the CX, so the CG scalars do not depend on the
// autotuned x-chunking); MODE 5 is the mode-3 sweep that leaves the shares
// of r.z the same way (r = rdot, the CG's FP64 residual; z = the sweep's
// output, the V-cycle's last). CX <= 32.
// (The FP64 512-thread variants up to MODE 2 -- the CG operator and the
// residual -- are held to four waves per SIMD, two workgroups per CU.)
template <typename T, typename TE, int NT, int MODE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && NT == 512 && MODE <= 2 ? 4 : 1))) k_applyA_xm(const T* __restrict__ p, T* __restrict__ out,
                                                  const TE* __restrict__ E, double thr, Grid g,
                                                  const uint8_t* __restrict__ fixed,
                                                  const T* __restrict__ bvec, const T* __restrict__ invD,
                                                  const double* __restrict__ omp, const int* gate, HadCoef<T> co,
                                                  int CX, int gx, int gy, double* __restrict__ part,
                                                  const double* __restrict__ rdot = nullptr) {
#pragma clang fp contract(off)
    gate = batch_gate(gate);
    if (gate && *gate) return;
    {
        const int64_t bo = batch_off(g.ndof());
        p += bo; out += bo;
        if (bvec) bvec += bo;
        if (rdot) rdot += bo;
    }
    // XCD-aware work order: hardware places workgroup L on XCD L % 8; work
    // item w = (L % 8) * ceil(N/8) + L / 8 gives each XCD a contiguous range
    // of (column, x-chunk) items with the x-chunks of a column adjacent, so
    // the node plane two chunks share is read once into that XCD's L2.
    const int nwork = gx * gy;
    const int per = (nwork + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (w >= nwork) return;   // padding workgroup
    const int bx = w / gy, by = w - bx * gy;
    __shared__ double s_red[32 * (NT / 64)];   // per-plane wave sums (MODE 1, 5 with part; CX <= 32)
    const T omega = omp ? (T)*omp : T(0);
    extern __shared__ __align__(16) unsigned char xm_smem[];
    const int nnz = g.nnz(), nny = g.nny(), nnx = g.nnx();
    const int nyz = nny * nnz;
    const int H = nnz + 1;
    constexpr bool RECOMP = xm_recomp<T, NT>();
    T* s_in = reinterpret_cast<T*>(xm_smem);
    T* s_q = s_in + (NT + H) * 3 * (RECOMP ? 2 : 1);   // per thread: q01[3] q10[3] q11[3]
    // LDS buffer of node plane ix
    auto sbuf = [&](int ix) -> T* { return RECOMP ? s_in + (ix & 1) * (NT + H) * 3 : s_in; };
    const int t = threadIdx.x;
    const int outs = NT - H;
    const int jbase = bx * outs - H;
    const int j = jbase + t;
    const bool jv = (j >= 0) && (j < nyz);
    const int iy = jv ? j / nnz : 0;
    const int iz = jv ? j - iy * nnz : 0;
    const bool elem_yz = jv && iy < g.nely && iz < g.nelz;
    const bool is_out = (t >= H) && jv;
    const int jh = jbase + NT + t;                 // halo node loaded by t < H
    const bool jhv = (t < H) && (jh >= 0) && (jh < nyz);
    const int X0 = by * CX;
    const int X1 = min(X0 + CX, nnx);
    const int64_t plane = (int64_t)nyz * 3;
    // element (ex, iy, iz) in zyx order: (ex*nely + iy)*nelz + iz
    const int64_t erow = (int64_t)iy * g.nelz + iz;
    const int64_t eplane = (int64_t)g.nely * g.nelz;

    // register prefetch of one node plane (own node + halo node). The loads
    // are unconditional (a lane with no node reads entry 0, and its value is
    // replaced by 0 when the plane is staged), so no branch surrounds them
    // and the compiler's wait counts track each load exactly: a load is only
    // waited for where its value is used, a plane step later.
    T pf[3], ph[3];
    T pfb[3], pfd[3];   // MODE 4: own b and invD of the prefetched plane
    T phb[3], phd[3];   // MODE 4: the halo node's
    auto fetch = [&](int ix) {
        const bool xin = ix >= 0 && ix < nnx;
        const int64_t aj = (xin && jv) ? (int64_t)ix * plane + 3 * j : 0;
        const int64_t ah = (xin && jhv) ? (int64_t)ix * plane + 3 * jh : 0;
        if constexpr (MODE == 4) {
            // b and invD only: x0 = omega invD b is formed when the plane is
            // staged
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pfb[d] = bvec[aj + d];
                pfd[d] = invD[aj + d];
                phb[d] = bvec[ah + d];
                phd[d] = invD[ah + d];
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pf[d] = p[aj + d];
                ph[d] = p[ah + d];
            }
        }
    };
    // own node values of the last staged plane (MODE 3 reads x at the
    // output node from here instead of from memory)
    T stg[3] = {T(0), T(0), T(0)};
    T stgb[3] = {T(0), T(0), T(0)}, stgd[3] = {T(0), T(0), T(0)};   // MODE 4
    // stage the prefetched plane in LDS, prefetch plane ixn, and return this
    // thread's face transform of the staged plane
    auto stage_face = [&](int ixn, T (&f)[12]) {
        T* sb = sbuf(ixn - 1);
        // the prefetched plane's values, 0 for a lane with no node there
        const bool xin = ixn - 1 >= 0 && ixn - 1 < nnx;
        const bool okj = xin && jv, okh = xin && jhv;
        if constexpr (MODE == 4) {
            // x0 = 0 + omega invD b, as the separate first sweep forms it
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pfb[d] = okj ? pfb[d] : T(0);
                pfd[d] = okj ? pfd[d] : T(0);
                pf[d] = xfma(omega * pfd[d], pfb[d], T(0));
                ph[d] = okh ? xfma(omega * phd[d], phb[d], T(0)) : T(0);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pf[d] = okj ? pf[d] : T(0);
                ph[d] = okh ? ph[d] : T(0);
            }
        }
        T own[3] = {pf[0], pf[1], pf[2]};
#pragma unroll
        for (int d = 0; d < 3; d++) stg[d] = own[d];
        if constexpr (MODE == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) { stgb[d] = pfb[d]; stgd[d] = pfd[d]; }
        }
#pragma unroll
        for (int d = 0; d < 3; d++) sb[3 * t + d] = own[d];
        if (t < H) {
#pragma unroll
            for (int d = 0; d < 3; d++) sb[3 * (NT + t) + d] = ph[d];
        }
        fetch(ixn);
        __syncthreads();
        T n01[3], n10[3], n11[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            n01[d] = sb[3 * (t + 1) + d];
            n10[d] = sb[3 * (t + nnz) + d];
            n11[d] = sb[3 * (t + nnz + 1) + d];
        }
        yz_forward<T>(own, n01, n10, n11, f);
    };
    // face transform of node plane ix from its LDS buffer (RECOMP)
    auto face_lds = [&](int ix, const T (&own)[3], T (&f)[12]) {
        const T* sb = sbuf(ix);
        T n01[3], n10[3], n11[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            n01[d] = sb[3 * (t + 1) + d];
            n10[d] = sb[3 * (t + nnz) + d];
            n11[d] = sb[3 * (t + nnz + 1) + d];
        }
        yz_forward<T>(own, n01, n10, n11, f);
    };

    // face transforms of the two node planes around the current element
    // plane and the carried trailing-face outputs, in two register sets that
    // trade roles every step (the march is unrolled by two below, so no
    // register copies at the loop edge)
    T fa[12], fb[12], ca[12], cb[12];
#pragma unroll
    for (int k = 0; k < 12; k++) ca[k] = T(0);
    // element modulus of element plane ex, read one step before it is used
    // (unconditionally, entry 0 for a lane with no element there; eok says
    // whether the value is the element's)
    auto emod = [&](int ex) -> TE {
        const bool ok = elem_yz && ex >= 0 && ex < g.nelx;
        return E[ok ? (int64_t)ex * eplane + erow : 0];
    };
    TE enext = emod(X0 - 1);
    fetch(X0 - 1);
    stage_face(X0, fa);
    // one x step: element plane ex between node planes ex (face fprev) and
    // ex+1 (staged here into fnew); completes output node plane ex
    auto step = [&](const int ex, const T (&fprev)[12], T (&fnew)[12], const T (&cin)[12], T (&cout)[12]) {
        const TE ev = enext;
        const bool eok = elem_yz && ex >= 0 && ex < g.nelx;
        enext = emod(ex + 1);
        // operands of this step's output epilogue (node plane ex), loaded
        // before the element work so their latency overlaps it
        const bool outstep = is_out && ex >= X0;
        const int64_t obase = (int64_t)ex * plane + 3 * (int64_t)j;
        T eb[3] = {T(0), T(0), T(0)}, ed[3] = {T(0), T(0), T(0)};
        const T ep[3] = {stg[0], stg[1], stg[2]};
        if constexpr (MODE == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) { eb[d] = stgb[d]; ed[d] = stgd[d]; }
        }
        // (unconditional, entry 0 for a lane with no output node: only the
        // output nodes' values are used)
        uint8_t ef[3] = {0, 0, 0};
        double er[3] = {0.0, 0.0, 0.0};
        if constexpr (MODE >= 1) {
            const int64_t ob = outstep ? obase : 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                ef[d] = fixed[ob + d];
                if (MODE == 2 || MODE == 3 || MODE == 5) eb[d] = bvec[ob + d];
                if (MODE == 3 || MODE == 5) ed[d] = invD[ob + d];
                if (MODE == 5) er[d] = rdot[ob + d];
            }
        }
        T A[12], Bq[12];
        // reads of s_in for the previous plane finished before the last
        // barrier of the previous step; only the first step follows the
        // pre-loop staging directly (with double-buffered planes it writes
        // the other buffer)
        if (!RECOMP && ex == X0 - 1) __syncthreads();
        stage_face(ex + 2, fnew);
        if constexpr (RECOMP) {
            // plane ex was staged one step earlier (a barrier ago)
            T fp[12];
            face_lds(ex, ep, fp);
#pragma unroll
            for (int k = 0; k < 12; k++) {
                A[k] = fp[k] + fnew[k];
                Bq[k] = fp[k] - fnew[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 12; k++) {
                A[k] = fprev[k] + fnew[k];
                Bq[k] = fprev[k] - fnew[k];
            }
        }
        const T Ee = (eok && (double)ev > thr) ? (T)ev : T(0);
        had_block<T>(A, Bq, co);
        // inverse x butterfly: face at x=0 (plane ex) and x=1 (plane ex+1),
        // scaled by the element modulus
        T nv[12];
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const T f0 = A[k] + Bq[k];
            const T f1 = A[k] - Bq[k];
            nv[k] = xfma(Ee, f0, cin[k]);
            cout[k] = Ee * f1;
        }
        // inverse y/z butterflies of the face at node plane ex -> 4 node values
        T q00[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T a00 = nv[FI(0, 0, d)] + nv[FI(1, 0, d)];   // y = 0, qz = 0
            const T a01 = nv[FI(0, 1, d)] + nv[FI(1, 1, d)];   // y = 0, qz = 1
            const T a10 = nv[FI(0, 0, d)] - nv[FI(1, 0, d)];   // y = 1, qz = 0
            const T a11 = nv[FI(0, 1, d)] - nv[FI(1, 1, d)];   // y = 1, qz = 1
            q00[d] = a00 + a01;
            s_q[9 * t + d] = a00 - a01;          // (y,z) = (0,1)
            s_q[9 * t + 3 + d] = a10 + a11;      // (1,0)
            s_q[9 * t + 6 + d] = a10 - a11;      // (1,1)
        }
        __syncthreads();
        // node (iy, iz) = face(iy,iz).q00 + face(iy,iz-1).q01 + face(iy-1,iz).q10
        //               + face(iy-1,iz-1).q11 (faces beyond the grid carry zeros)
        double cpl = 0.0;   // this thread's share of the plane's p.Ap (MODE 1) or r.z (MODE 5)
        if (outstep) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                T v = q00[d] + s_q[9 * (t - 1) + d];
                v = v + (s_q[9 * (t - nnz) + 3 + d] + s_q[9 * (t - nnz - 1) + 6 + d]);
                if (MODE == 1) {
                    if (ef[d]) v = T(0);
                } else if (MODE == 2) {
                    v = ef[d] ? T(0) : eb[d] - v;
                } else if (MODE == 3 || MODE == 4 || MODE == 5) {
                    const T r = eb[d] - v;
                    v = ef[d] ? ep[d] : xfma(omega * ed[d], r, ep[d]);
                }
                if (MODE == 1) cpl = xfma((double)ep[d], (double)v, cpl);
                if (MODE == 5) cpl = xfma(er[d], (double)v, cpl);
                out[obase + d] = v;
            }
        }
        // one partial per (yz tile, node plane): the same sums whatever the
        // x-chunking, so the CG scalars do not depend on the element tiles.
        // Each wave sums the plane (DPP, lane 63) into its own LDS slot of
        // the plane (no barrier here); the slots are added after the march.
        if ((MODE == 1 || MODE == 5) && part && ex >= X0) {
            const double ws = wave_sum_dpp(cpl);
            if ((t & 63) == 63) s_red[(ex - X0) * (NT / 64) + (t >> 6)] = ws;
        }
    };
    if constexpr (RECOMP) {
        // the previous face comes from LDS: one carry, updated in place
        for (int ex = X0 - 1; ex < X1; ex++) step(ex, fa, fb, ca, ca);
    } else {
        int ex = X0 - 1;
        for (; ex + 1 < X1; ex += 2) {
            step(ex, fa, fb, ca, cb);
            step(ex + 1, fb, fa, cb, ca);
        }
        if (ex < X1) step(ex, fa, fb, ca, cb);
    }
    if ((MODE == 1 || MODE == 5) && part) {
        __syncthreads();
        for (int k = t; k < X1 - X0; k += NT) {
            double sum = 0.0;
            for (int wv = 0; wv < NT / 64; wv++) sum += s_red[k * (NT / 64) + wv];
            part[(int64_t)blockIdx.y * kCgStride + (int64_t)bx * nnx + X0 + k] = sum;
        }
    }
}

#undef FI

}  // namespace topo
