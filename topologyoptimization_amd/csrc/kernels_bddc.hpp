// kernels_bddc.hpp -- the domain-decomposition preconditioner of the
// reference's 'petsc' solver (bddcApply assembly/index.ts:1272-1371; the
// partition and local diagonals of KSPSolver js/optimizer-worker-3d.js:
// 1155-1278). Per application:
//   z0 = local step: z0[d] = (sum over the subdomains s that contain DOF d,
//        ascending s, of linv_s[d] r[d]) / (their count)
//   z_{k+1} = z_k + invD (r - A z_k), k < sweeps - 1    (global Jacobi)
//   z = z_last + (corner ? invD (r - A z_last) : 0), z = 0 at fixed DOFs
// The sweeps and A z are applyA launches (modes 3 and 0); the two kernels
// here are the local step and the corner correction. FP64 only.
#pragma once
#include "kernels.hpp"

namespace topo {

constexpr int kBddcSlots = 8;   // subdomains that can share one node (2 per axis)

// local step. val holds linv_s[d] slot-major (slot k of DOF d at
// val[k * n + d], slots in ascending s), cnt the number of subdomains that
// contain d. The products and sums are rounded one by one like the
// reference's z[gi] += invD * ri.
__global__ void k_bddc_local(const double* __restrict__ r, double* __restrict__ z, const double* __restrict__ val,
                             const uint8_t* __restrict__ cnt, int64_t n, const int* gate) {
#pragma clang fp contract(off)
    if (gate && *gate) return;
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const int c = cnt[i];
        const double ri = r[i];
        double acc = 0.0;
        for (int k = 0; k < c; k++) acc = acc + val[(int64_t)k * n + i] * ri;
        z[i] = c > 0 ? acc / (double)c : acc;
    }
}

// corner correction and the final mask (bddcApply :1346-1371):
// z = fixed ? 0 : z + (corner ? invD (r - Az) : 0)
__global__ void k_bddc_corner(double* __restrict__ z, const double* __restrict__ r, const double* __restrict__ Az,
                              const double* __restrict__ invD, const uint8_t* __restrict__ corner,
                              const uint8_t* __restrict__ fixed, int64_t n, const int* gate) {
#pragma clang fp contract(off)
    if (gate && *gate) return;
    for (int64_t i = gtid(); i < n; i += gstride()) {
        double v = z[i];
        if (corner[i]) {
            const double cc = invD[i] * (r[i] - Az[i]);
            v = v + cc;
        }
        z[i] = fixed[i] ? 0.0 : v;
    }
}

// Local inverse diagonals of a box partition of the structured grid: per
// node, the <= 8 elements around it (ascending reference element order
// e = ex + ey nelx + ez nelx nely) are grouped by subdomain s =
// (subz[ez] nsy + suby[ey]) nsx + subx[ex]; each subdomain's diagonal adds
// E_e K0_ii of its elements with E_e >= 1e-30 (_updateBDDC :1256-1270), and
// linv_s = 1/diag on free DOFs with diag > 1e-30 (:1271-1276). Slots are
// written in ascending s. E in zyx order.
__global__ void k_bddc_build(const double* __restrict__ E, const uint8_t* __restrict__ fixed, Grid g,
                             const int32_t* __restrict__ subx, const int32_t* __restrict__ suby,
                             const int32_t* __restrict__ subz, int nsx, int nsy, const double* __restrict__ ke,
                             double* __restrict__ val, uint8_t* __restrict__ cnt) {
#pragma clang fp contract(off)
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t n = g.ndof();
    for (int64_t nd = gtid(); nd < g.nnodes(); nd += gstride()) {
        const int iz = (int)(nd % nnz);
        const int64_t t = nd / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        int ss[kBddcSlots];
        double ld[kBddcSlots][3];
        int ns = 0;
        for (int ez = iz - 1; ez <= iz; ez++) {
            if (ez < 0 || ez >= g.nelz) continue;
            for (int ey = iy - 1; ey <= iy; ey++) {
                if (ey < 0 || ey >= g.nely) continue;
                for (int ex = ix - 1; ex <= ix; ex++) {
                    if (ex < 0 || ex >= g.nelx) continue;
                    const int s = (subz[ez] * nsy + suby[ey]) * nsx + subx[ex];
                    int k = 0;
                    while (k < ns && ss[k] != s) k++;
                    if (k == ns) {
                        ss[ns] = s;
                        ld[ns][0] = ld[ns][1] = ld[ns][2] = 0.0;
                        ns++;
                    }
                    const double Ee = E[((int64_t)ex * g.nely + ey) * g.nelz + ez];
                    if (Ee < 1e-30) continue;
                    const int L = local_node(ix - ex, iy - ey, iz - ez);
                    for (int d = 0; d < 3; d++) {
                        const int li = 3 * L + d;
                        const double c = Ee * ke[li * 24 + li];
                        ld[k][d] = ld[k][d] + c;
                    }
                }
            }
        }
        // ascending subdomain order (insertion sort of <= 8 slots)
        for (int a = 1; a < ns; a++)
            for (int b = a; b > 0 && ss[b - 1] > ss[b]; b--) {
                const int ts = ss[b]; ss[b] = ss[b - 1]; ss[b - 1] = ts;
                for (int d = 0; d < 3; d++) {
                    const double tv = ld[b][d]; ld[b][d] = ld[b - 1][d]; ld[b - 1][d] = tv;
                }
            }
        for (int d = 0; d < 3; d++) {
            const int64_t i = 3 * nd + d;
            cnt[i] = (uint8_t)ns;
            for (int k = 0; k < ns; k++)
                val[(int64_t)k * n + i] = (!fixed[i] && ld[k][d] > 1e-30) ? 1.0 / ld[k][d] : 0.0;
        }
    }
}

// corner (primal) DOF flags of a box partition: every DOF of a node whose
// coordinates are partition boundaries on all three axes (:1230-1254)
__global__ void k_bddc_corner_flags(Grid g, const uint8_t* __restrict__ bx, const uint8_t* __restrict__ by,
                                    const uint8_t* __restrict__ bz, uint8_t* __restrict__ corner) {
    const int nny = g.nny(), nnz = g.nnz();
    for (int64_t nd = gtid(); nd < g.nnodes(); nd += gstride()) {
        const int iz = (int)(nd % nnz);
        const int64_t t = nd / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        const uint8_t f = (bx[ix] && by[iy] && bz[iz]) ? 1 : 0;
        for (int d = 0; d < 3; d++) corner[3 * nd + d] = f;
    }
}

}  // namespace topo
