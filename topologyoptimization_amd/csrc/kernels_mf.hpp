// kernels_mf.hpp -- hex8 stiffness apply on FP32 vectors with the element
// product on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// The element product y_e = K0 u_e is a dense contraction: per wave, 64
// elements (four column blocks of 16) against the 24x24 K0, permuted and
// padded to 32x32 so that each 8-wide k group is one (x, y) pair of element
// nodes along z: row/column index 16 dx + 8 dy + 4 dz + c (c = 3 is a zero
// pad). Operands are split into two bfloat16 parts, x = hi + lo with
// lo = bf16(x - hi) (about 16 significant bits together), and the product
// is K_hi u_hi + K_hi u_lo + K_lo u_hi accumulated in FP32 (the lo*lo term is
// below the split's own rounding). Node values are split once when a node
// plane is staged in LDS, so each node's split is shared by its 8 elements.
//
// Work decomposition and output order are those of the x-march kernel
// (kernels_xm.hpp): a workgroup owns NT consecutive (y,z) node positions and
// marches along x; per step it stages one node plane (hi/lo parts, double
// buffered), runs the element products of the element plane behind it, and
// completes one node plane of the output. The 24 element outputs go through
// LDS (slot-major, conflict-free) and each output node gathers its 8
// contributions -- the 4 elements of the current plane and, carried in
// registers from the previous step, the 4 behind it -- in a fixed order,
// scaled by the element moduli: deterministic, no atomics.
#pragma once
#include "kernels_xm.hpp"

namespace topo {

typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

// row stride of the element-output slots: NT + 48 makes the four 16-lane
// groups of one store land on distinct LDS banks
template <int NT>
constexpr int mf_sd() { return NT + 48; }

// dynamic LDS: s_hi[2][NT+H] | s_lo[2][NT+H] (uint2: 4 bf16 per node, the
// 4th zero) | s_d[24][SD] float (element outputs, slot-major) | s_e[2][NT]
inline size_t mf_lds_bytes(int nnz, int NT) {
    const size_t np = (size_t)NT + nnz + 1;
    return 2 * np * 8 * 2 + (size_t)24 * (NT + 48) * 4 + (size_t)2 * NT * 4;
}

// hi/lo bfloat16 parts of a node's 3 components, packed 4 per 8 bytes
__device__ __forceinline__ void mf_split3(const float (&v)[3], uint2& hi, uint2& lo) {
    unsigned short h[3], l[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const __bf16 bh = (__bf16)v[d];
        const __bf16 bl = (__bf16)(v[d] - (float)bh);
        h[d] = __builtin_bit_cast(unsigned short, bh);
        l[d] = __builtin_bit_cast(unsigned short, bl);
    }
    hi.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    hi.y = (uint32_t)h[2];
    lo.x = (uint32_t)l[0] | ((uint32_t)l[1] << 16);
    lo.y = (uint32_t)l[2];
}

// MODE as k_applyA_xm (0 Ap, 1 masked with p.Ap partials, 2 b - Ap, 3
// Jacobi sweep, 4 the first two sweeps from x = 0, 5 sweep with r.z
// partials); same arguments, plus kf: the K0 fragments, 4 x 64 lanes x 8
// bf16 ([rb*2 + part][lane], part 0 = hi, 1 = lo; A operand of row block rb:
// lane l holds row 16 rb + l % 16, columns 8 (l / 16) .. + 7). NT = 512.
template <int NT, int MODE>
__global__ void __launch_bounds__(NT) k_applyA_mf(const float* __restrict__ p, float* __restrict__ out,
                                                  const float* __restrict__ E, double thr, Grid g,
                                                  const uint8_t* __restrict__ fixed,
                                                  const float* __restrict__ bvec, const float* __restrict__ invD,
                                                  const double* __restrict__ omp, const int* gate,
                                                  const uint4* __restrict__ kf, int CX, int gx, int gy,
                                                  double* __restrict__ part, const double* __restrict__ rdot = nullptr) {
#pragma clang fp contract(off)
    gate = batch_gate(gate);
    if (gate && *gate) return;
    {
        const int64_t bo = batch_off(g.ndof());
        p += bo; out += bo;
        if (bvec) bvec += bo;
        if (rdot) rdot += bo;
    }
    // XCD-aware work order, as k_applyA_xm
    const int nwork = gx * gy;
    const int per = (nwork + 7) / 8;
    const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (w >= nwork) return;   // padding workgroup
    const int bx = w / gy, by = w - bx * gy;
    __shared__ double s_red[32 * (NT / 64)];
    const float omega = omp ? (float)*omp : 0.f;
    extern __shared__ __align__(16) unsigned char mf_smem[];
    constexpr int SD = mf_sd<NT>();
    const int nnz = g.nnz(), nny = g.nny(), nnx = g.nnx();
    const int nyz = nny * nnz;
    const int H = nnz + 1;
    const int NP = NT + H;
    uint2* s_hi = reinterpret_cast<uint2*>(mf_smem);
    uint2* s_lo = s_hi + 2 * NP;
    float* s_d = reinterpret_cast<float*>(s_lo + 2 * NP);
    float* s_e = s_d + 24 * SD;
    const int t = threadIdx.x;
    const int outs = NT - H;
    const int jbase = bx * outs - H;
    const int j = jbase + t;
    const bool jv = (j >= 0) && (j < nyz);
    const int iy = jv ? j / nnz : 0;
    const int iz = jv ? j - iy * nnz : 0;
    const bool elem_yz = jv && iy < g.nely && iz < g.nelz;
    const bool is_out = (t >= H) && jv;
    const int jh = jbase + NT + t;   // halo node staged by t < H
    const bool jhv = (t < H) && (jh >= 0) && (jh < nyz);
    const int X0 = by * CX;
    const int X1 = min(X0 + CX, nnx);
    const int64_t plane = (int64_t)nyz * 3;
    const int64_t erow = (int64_t)iy * g.nelz + iz;
    const int64_t eplane = (int64_t)g.nely * g.nelz;

    // matrix-core roles of this lane: B operand of column block b is the
    // element at position wbase + 16 b + lc, k group lg = 2 dx + dy (its
    // nodes dz = 0, 1 at (dx, dy)); D rows 4 lg .. + 3 of row block rb are
    // that element's outputs at node (dx = rb, dy = lg >> 1, dz = lg & 1)
    const int lane = t & 63, wbase = t & ~63;
    const int lg = lane >> 4, lc = lane & 15;
    const int bdx = lg >> 1;
    const int boff = wbase + lc + (lg & 1) * nnz;
    mf_bf16x8 Ah[2], Al[2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        Ah[rb] = __builtin_bit_cast(mf_bf16x8, kf[(rb * 2 + 0) * 64 + lane]);
        Al[rb] = __builtin_bit_cast(mf_bf16x8, kf[(rb * 2 + 1) * 64 + lane]);
    }

    // register prefetch of one node plane (own node + halo node)
    float pf[3], ph[3], pfb[3], pfd[3];
    float phb[3], phd[3];   // MODE 4: b and invD of the halo node
    auto fetch = [&](int ix) {
#pragma unroll
        for (int d = 0; d < 3; d++) { pf[d] = 0.f; ph[d] = 0.f; pfb[d] = 0.f; pfd[d] = 0.f; }
        if constexpr (MODE == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) { phb[d] = 0.f; phd[d] = 0.f; }
        }
        if (ix >= 0 && ix < nnx) {
            const int64_t o = (int64_t)ix * plane;
            if constexpr (MODE == 4) {
                // b and invD only: x0 = omega invD b is formed at staging
                if (jv) {
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        pfb[d] = bvec[o + 3 * j + d];
                        pfd[d] = invD[o + 3 * j + d];
                    }
                }
                if (jhv) {
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        phb[d] = bvec[o + 3 * jh + d];
                        phd[d] = invD[o + 3 * jh + d];
                    }
                }
            } else {
                if (jv) {
#pragma unroll
                    for (int d = 0; d < 3; d++) pf[d] = p[o + 3 * j + d];
                }
                if (jhv) {
#pragma unroll
                    for (int d = 0; d < 3; d++) ph[d] = p[o + 3 * jh + d];
                }
            }
        }
    };
    // own FP32 values of the last staged plane (the epilogue's x / p)
    float stg[3] = {0.f, 0.f, 0.f};
    float stgb[3] = {0.f, 0.f, 0.f}, stgd[3] = {0.f, 0.f, 0.f};   // MODE 4
    // split the prefetched plane ix into its LDS buffer, prefetch ix + 1
    auto stage = [&](int ix) {
        if constexpr (MODE == 4) {
            // x0 = 0 + omega invD b, as the separate first sweep forms it
#pragma unroll
            for (int d = 0; d < 3; d++) {
                pf[d] = xfma(omega * pfd[d], pfb[d], 0.f);
                ph[d] = xfma(omega * phd[d], phb[d], 0.f);
            }
        }
        uint2 hi, lo;
        mf_split3(pf, hi, lo);
        const int bo = (ix & 1) * NP;
        s_hi[bo + t] = hi;
        s_lo[bo + t] = lo;
        if (t < H) {
            mf_split3(ph, hi, lo);
            s_hi[bo + NT + t] = hi;
            s_lo[bo + NT + t] = lo;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) stg[d] = pf[d];
        if constexpr (MODE == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) { stgb[d] = pfb[d]; stgd[d] = pfd[d]; }
        }
        fetch(ix + 1);
    };

    float carry[3] = {0.f, 0.f, 0.f};   // x = 1 face sums of the previous element plane at this node
    fetch(X0 - 1);
    stage(X0 - 1);
    for (int ex = X0 - 1; ex < X1; ex++) {
        const bool outstep = is_out && ex >= X0;
        const int64_t obase = (int64_t)ex * plane + 3 * (int64_t)j;
        float eb[3] = {0.f, 0.f, 0.f}, ed[3] = {0.f, 0.f, 0.f};
        const float ep[3] = {stg[0], stg[1], stg[2]};
        if constexpr (MODE == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) { eb[d] = stgb[d]; ed[d] = stgd[d]; }
        }
        uint8_t ef[3] = {0, 0, 0};
        double er[3] = {0.0, 0.0, 0.0};
        if (outstep && MODE >= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                ef[d] = fixed[obase + d];
                if (MODE == 2 || MODE == 3 || MODE == 5) eb[d] = bvec[obase + d];
                if (MODE == 3 || MODE == 5) ed[d] = invD[obase + d];
                if (MODE == 5) er[d] = rdot[obase + d];
            }
        }
        float Ee = 0.f;
        if (elem_yz && ex >= 0 && ex < g.nelx) {
            const float ev = E[(int64_t)ex * eplane + erow];
            if ((double)ev > thr) Ee = ev;
        }
        s_e[(ex & 1) * NT + t] = Ee;
        stage(ex + 1);
        __syncthreads();   // node planes ex, ex + 1 and the moduli of element plane ex staged
        {
            const int bo = ((ex + bdx) & 1) * NP + boff;
            const uint2* bh = s_hi + bo;
            const uint2* bl = s_lo + bo;
            // all operands first, then the 24 products as 8 independent
            // accumulation chains (the matrix-core latency overlaps), then
            // the stores
            mf_bf16x8 Bh[4], Bl[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint2 h0 = bh[16 * b], h1 = bh[16 * b + 1];
                const uint2 l0 = bl[16 * b], l1 = bl[16 * b + 1];
                Bh[b] = __builtin_bit_cast(mf_bf16x8, make_uint4(h0.x, h0.y, h1.x, h1.y));
                Bl[b] = __builtin_bit_cast(mf_bf16x8, make_uint4(l0.x, l0.y, l1.x, l1.y));
            }
            mf_f32x4 acc[4][2];
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
                    acc[b][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[rb], Bh[b], mf_f32x4{0.f, 0.f, 0.f, 0.f},
                                                                         0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
                    acc[b][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[rb], Bl[b], acc[b][rb], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
                    acc[b][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[rb], Bh[b], acc[b][rb], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++) {
                    float* dst = s_d + (4 * rb + lg) * 3 * SD + wbase + 16 * b + lc;
                    dst[0] = acc[b][rb][0];
                    dst[SD] = acc[b][rb][1];
                    dst[2 * SD] = acc[b][rb][2];
                }
        }
        __syncthreads();   // element outputs of plane ex in LDS
        // node (iy, iz) of plane ex: elements (iy, iz), (iy, iz-1), (iy-1, iz),
        // (iy-1, iz-1) of plane ex at positions t, t-1, t-nnz, t-nnz-1, their
        // local nodes (0,0,0), (0,0,1), (0,1,0), (0,1,1) = slots 0..3 (x = 0
        // face) and slots 4..7 (x = 1 face, the next node plane)
        double cpl = 0.0;
        if (t >= H) {
            const float* se = s_e + (ex & 1) * NT;
            const int q0 = t, q1 = t - 1, q2 = t - nnz, q3 = t - nnz - 1;
            const float e0 = se[q0], e1 = se[q1], e2 = se[q2], e3 = se[q3];
            float cn[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                float v = carry[d];
                v = fmaf(e0, s_d[(0 * 3 + d) * SD + q0], v);
                v = fmaf(e1, s_d[(1 * 3 + d) * SD + q1], v);
                v = fmaf(e2, s_d[(2 * 3 + d) * SD + q2], v);
                v = fmaf(e3, s_d[(3 * 3 + d) * SD + q3], v);
                float c1 = e0 * s_d[(4 * 3 + d) * SD + q0];
                c1 = fmaf(e1, s_d[(5 * 3 + d) * SD + q1], c1);
                c1 = fmaf(e2, s_d[(6 * 3 + d) * SD + q2], c1);
                c1 = fmaf(e3, s_d[(7 * 3 + d) * SD + q3], c1);
                cn[d] = c1;
                if (outstep) {
                    if (MODE == 1) {
                        if (ef[d]) v = 0.f;
                    } else if (MODE == 2) {
                        v = ef[d] ? 0.f : eb[d] - v;
                    } else if (MODE == 3 || MODE == 4 || MODE == 5) {
                        const float r = eb[d] - v;
                        v = ef[d] ? ep[d] : xfma(omega * ed[d], r, ep[d]);
                    }
                    if (MODE == 1) cpl = xfma((double)ep[d], (double)v, cpl);
                    if (MODE == 5) cpl = xfma(er[d], (double)v, cpl);
                    out[obase + d] = v;
                }
            }
#pragma unroll
            for (int d = 0; d < 3; d++) carry[d] = cn[d];
        }
        // CG partials per (yz tile, node plane), as k_applyA_xm
        if ((MODE == 1 || MODE == 5) && part && ex >= X0) {
            const double ws = wave_sum_dpp(cpl);
            if ((t & 63) == 63) s_red[(ex - X0) * (NT / 64) + (t >> 6)] = ws;
        }
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

}  // namespace topo
