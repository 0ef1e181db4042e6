// kernels_2d.hpp -- the 2-D SIMP path (js/optimizer-worker.js): bilinear
// quad4 elements on a structured nelx x nely grid in the reference numbering
// (_precomputeEdofs2D :240-268):
//   node n = (nely+1) ix + iy, DOFs 2n (x) and 2n+1 (y);
//   element e = ey + ex nely, local nodes n1 = (nely+1) ex + ey, n2 = n1 +
//   nely + 1, n2 + 1, n1 + 1 (local 0..3 at offsets (0,0) (1,0) (1,1)
//   (0,1)), DOF order [n1x n1y n2x n2y (n2+1)x (n2+1)y (n1+1)x (n1+1)y].
// Element arrays are in that (column-major) order. Every output value is
// produced by one thread with the reference's order of additions (no
// atomics): the FP64 results are bit-identical to the reference's
// scatter loops.
#pragma once
#include "kernels.hpp"

namespace topo {

struct Grid2 {
    int nelx, nely;
    __host__ __device__ int nny() const { return nely + 1; }
    __host__ __device__ int64_t nnodes() const { return (int64_t)(nelx + 1) * (nely + 1); }
    __host__ __device__ int64_t ndof() const { return 2 * nnodes(); }
    __host__ __device__ int64_t nel() const { return (int64_t)nelx * nely; }
};

// local index of the node at offset (dx, dy) from the element's first node
__device__ __forceinline__ int q4_local(int dx, int dy) { return dx == 0 ? (dy == 0 ? 0 : 3) : (dy == 0 ? 1 : 2); }

// the 8 DOF values of element (ex, ey) in local order
__device__ __forceinline__ void q4_gather(const double* __restrict__ v, int64_t n1, int nny, double (&u)[8]) {
    const int64_t n2 = n1 + nny;
    u[0] = v[2 * n1];     u[1] = v[2 * n1 + 1];
    u[2] = v[2 * n2];     u[3] = v[2 * n2 + 1];
    u[4] = v[2 * n2 + 2]; u[5] = v[2 * n2 + 3];
    u[6] = v[2 * n1 + 2]; u[7] = v[2 * n1 + 3];
}

// Operator apply, one thread per node, the <= 4 elements around it visited in
// ascending e (the order of the reference's active list) and each element
// row summed j = 0..7 (_ebeMatVec js/optimizer-worker.js:287-315, the EbE
// block of ebePCG assembly/index.ts:654-692). E holds E_e for active
// elements and 0 for the others (skipped, like elements off the active list).
// MODE 0: out = K p; 1: out = K p, 0 at fixed DOFs; 2: out = b - K p at free
// DOFs, 0 at fixed (the warm-start residual :589-619). With part (MODE 1)
// each block also leaves its share of p.Ap at part[blockIdx.x] (quantity 0
// of the CG partials; each thread's nodes in order, then the block tree).
template <int MODE>
__global__ void k_q4_apply(const double* __restrict__ p, double* __restrict__ out, const double* __restrict__ E,
                           Grid2 g, const uint8_t* __restrict__ fixed, const double* __restrict__ b, const int* gate,
                           const double* __restrict__ ke, double* __restrict__ part = nullptr) {
#pragma clang fp contract(off)
    if (gate && *gate) return;
    double pap[1] = {0.0};
    const int nny = g.nny();
    for (int64_t nd = gtid(); nd < g.nnodes(); nd += gstride()) {
        const int ix = (int)(nd / nny), iy = (int)(nd % nny);
        double acc0 = 0.0, acc1 = 0.0;
        for (int ex = ix - 1; ex <= ix; ex++) {
            if (ex < 0 || ex >= g.nelx) continue;
            for (int ey = iy - 1; ey <= iy; ey++) {
                if (ey < 0 || ey >= g.nely) continue;
                const double Ee = E[(int64_t)ex * g.nely + ey];
                if (Ee == 0.0) continue;
                double u[8];
                q4_gather(p, (int64_t)nny * ex + ey, nny, u);
                const int L = q4_local(ix - ex, iy - ey);
                const double* k0 = ke + (2 * L) * 8;
                const double* k1 = k0 + 8;
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    s0 += k0[j] * u[j];
                    s1 += k1[j] * u[j];
                }
                acc0 += Ee * s0;
                acc1 += Ee * s1;
            }
        }
        const int64_t i0 = 2 * nd;
        if (MODE == 0) {
            out[i0] = acc0;
            out[i0 + 1] = acc1;
        } else if (MODE == 1) {
            const double o0 = fixed[i0] ? 0.0 : acc0, o1 = fixed[i0 + 1] ? 0.0 : acc1;
            out[i0] = o0;
            out[i0 + 1] = o1;
            pap[0] += p[i0] * o0 + p[i0 + 1] * o1;
        } else {
            out[i0] = fixed[i0] ? 0.0 : b[i0] - acc0;
            out[i0 + 1] = fixed[i0 + 1] ? 0.0 : b[i0 + 1] - acc1;
        }
    }
    if (MODE == 1 && part) put_partials<1>(pap, part, 0, 1);
}

// Jacobi diagonal: diag_i = sum over the active elements around the node,
// ascending e, of E_e K_ii; invD = 1/diag on free DOFs with diag > 1e-30,
// else 0 (_computeDiagonal js/optimizer-worker.js:317-333, ebePCG
// assembly/index.ts:553-573).
__global__ void k_q4_invdiag(const double* __restrict__ E, const uint8_t* __restrict__ fixed, double* __restrict__ invD,
                             Grid2 g, const double* __restrict__ ke) {
#pragma clang fp contract(off)
    const int nny = g.nny();
    for (int64_t nd = gtid(); nd < g.nnodes(); nd += gstride()) {
        const int ix = (int)(nd / nny), iy = (int)(nd % nny);
        double d0 = 0.0, d1 = 0.0;
        for (int ex = ix - 1; ex <= ix; ex++) {
            if (ex < 0 || ex >= g.nelx) continue;
            for (int ey = iy - 1; ey <= iy; ey++) {
                if (ey < 0 || ey >= g.nely) continue;
                const double Ee = E[(int64_t)ex * g.nely + ey];
                if (Ee == 0.0) continue;
                const int L = q4_local(ix - ex, iy - ey);
                d0 += Ee * ke[(2 * L) * 9];
                d1 += Ee * ke[(2 * L + 1) * 9];
            }
        }
        const int64_t i0 = 2 * nd;
        invD[i0] = (!fixed[i0] && d0 > 1e-30) ? 1.0 / d0 : 0.0;
        invD[i0 + 1] = (!fixed[i0 + 1] && d1 > 1e-30) ? 1.0 / d1 : 0.0;
    }
}

// E_e = Emin + x^p (E0 - Emin) with Math.pow (_precomputeStiffness
// js/optimizer-worker.js:270-285); Em = E_e if E_e > thr (active), else 0.
template <typename TX>
__global__ void k_q4_moduli(const TX* __restrict__ x, double penal, double E0, double Emin, double thr,
                            double* __restrict__ Em, int64_t nel) {
#pragma clang fp contract(off)
    const double dE = E0 - Emin;
    for (int64_t e = gtid(); e < nel; e += gstride()) {
        const double v = Emin + pow((double)x[e], penal) * dE;
        Em[e] = v > thr ? v : 0.0;
    }
}

// Element energies w_e = sum_i sum_j (u_i K_ij) u_j in FP64 (the loop of
// _computeElementEnergyFlat :335-345), stored as Float32, and
//   sens = 1: dc_e = (-p xPhys^(p-1) E0 w_e) * dPhys/dx (FP64 w_e), with
//             dPhys/dx = beta (1 - tanh^2(beta (x_e - 1/2))) / (2 tanh(beta/2))
//             when projecting, else 1 (:655-672);
//   stress != null: stress_e = (Emin + xs^p (E0 - Emin)) w_e in FP64, stored
//             as Float32 (the FEA-only output :427-438; xs = the densities
//             analysed).
__global__ void k_q4_energy_sens(const double* __restrict__ U, const float* __restrict__ xphys,
                                 const float* __restrict__ x, double penal, double E0, double Emin, int sens,
                                 int use_proj, double beta, Grid2 g, float* __restrict__ w, float* __restrict__ dc,
                                 float* __restrict__ stress, const double* __restrict__ ke) {
#pragma clang fp contract(off)
    const int nny = g.nny();
    const double denom = 2 * tanh(beta * 0.5);
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ex = (int)(e / g.nely), ey = (int)(e % g.nely);
        double u[8];
        q4_gather(U, (int64_t)nny * ex + ey, nny, u);
        double en = 0.0;
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) en += u[i] * ke[i * 8 + j] * u[j];
        w[e] = (float)en;
        const double xp = (double)xphys[e];
        if (sens) {
            const double dphys = -penal * pow(xp, penal - 1) * E0 * en;
            double dd = 1.0;
            if (use_proj) {
                const double th = tanh(beta * ((double)x[e] - 0.5));
                dd = beta * (1 - th * th) / denom;
            }
            dc[e] = (float)(dphys * dd);
        }
        if (stress) stress[e] = (float)((Emin + pow(xp, penal) * (E0 - Emin)) * en);
    }
}

// stress metric of the mesh builder: stress_e = (Emin + x_e^p (E0 - Emin))
// w_e with the Float32 energies and the new design (buildAdaptiveMesh
// js/optimizer-worker.js:1166-1173), stored as Float32
__global__ void k_q4_stress(const float* __restrict__ x, const float* __restrict__ w, double penal, double E0,
                            double Emin, float* __restrict__ stress, int64_t nel) {
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < nel; e += gstride())
        stress[e] = (float)((Emin + pow((double)x[e], penal) * (E0 - Emin)) * (double)w[e]);
}

}  // namespace topo
