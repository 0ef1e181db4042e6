// kernels_adj.hpp -- design-map and adjoint kernels: the Helmholtz (PDE)
// density filter, Heaviside projection and its derivative in FP64, the
// element sensitivities of compliance and of the compliant-mechanism output
// displacement (non-self-adjoint, with an adjoint field), spring terms of the
// mechanism's input/output ports, and the double-precision density filter
// used by the backward pass. None of these has a counterpart in the
// reference's 3-D worker, except the Heaviside map (_heavisideProject
// js/optimizer-worker-3d.js:3500-3509) and the compliance sensitivity
// (:2175-2192), which the FP64 versions here generalise to any eta.
#pragma once
#include "kernels.hpp"

namespace topo {

// ---------------------------------------------------------------------------
// Helmholtz filter on the nodes of the element grid:
//   (r^2 K_L + M) y = T x,   xf = T^T y
// K_L, M: trilinear unit-cube Laplacian and mass matrices; T: element ->
// node map with weight 1/8 (the integral of a nodal shape function over the
// unit cube). T^T is the element average of its 8 nodal values, so the map
// x -> xf is symmetric, maps constants to themselves and keeps sum(x).
struct HelmKe {
    double a[64];   // r^2 K_L + M in the local node order of local_node()
};

// b_n = sum over the elements holding node n of x_e / 8 (x in zyx order)
template <typename TX>
__global__ void k_helm_rhs(const TX* __restrict__ x, double* __restrict__ b, Grid g) {
    const int nny = g.nny(), nnz = g.nnz();
    for (int64_t n = gtid(); n < g.nnodes(); n += gstride()) {
        const int iz = (int)(n % nnz);
        const int64_t t = n / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        double s = 0.0;
        for (int ex = ix - 1; ex <= ix; ex++) {
            if (ex < 0 || ex >= g.nelx) continue;
            for (int ey = iy - 1; ey <= iy; ey++) {
                if (ey < 0 || ey >= g.nely) continue;
                for (int ez = iz - 1; ez <= iz; ez++) {
                    if (ez < 0 || ez >= g.nelz) continue;
                    s += (double)x[((int64_t)ex * g.nely + ey) * g.nelz + ez];
                }
            }
        }
        b[n] = 0.125 * s;
    }
}

// out = A v (mode 0) or out = b - A v (mode 2), A = sum_e (r^2 K_L + M)_e,
// gathered per node over its <= 8 elements in a fixed order. A non-null
// gate that is set turns the launch into a no-op (converged CG).
__global__ void k_helm_apply(const double* __restrict__ v, double* __restrict__ out, Grid g, HelmKe ke, int mode,
                             const double* __restrict__ b, const int* gate) {
    if (gate && *gate) return;
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t nyz = (int64_t)nny * nnz;
    for (int64_t n = gtid(); n < g.nnodes(); n += gstride()) {
        const int iz = (int)(n % nnz);
        const int64_t t = n / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        double s = 0.0;
        for (int ex = ix - 1; ex <= ix; ex++) {
            if (ex < 0 || ex >= g.nelx) continue;
            for (int ey = iy - 1; ey <= iy; ey++) {
                if (ey < 0 || ey >= g.nely) continue;
                for (int ez = iz - 1; ez <= iz; ez++) {
                    if (ez < 0 || ez >= g.nelz) continue;
                    const int li = local_node(ix - ex, iy - ey, iz - ez);
                    const int64_t n0 = ((int64_t)ex * nny + ey) * nnz + ez;
#pragma unroll
                    for (int m = 0; m < 8; m++) {
                        int a, bb, c;
                        local_offsets(m, a, bb, c);
                        s += ke.a[li * 8 + m] * v[n0 + a * nyz + bb * nnz + c];
                    }
                }
            }
        }
        out[n] = mode == 2 ? b[n] - s : s;
    }
}

// inverse diagonal of A: each node's diagonal is (#elements holding it) * a_ii
__global__ void k_helm_invdiag(double* __restrict__ invD, Grid g, double aii) {
    const int nny = g.nny(), nnz = g.nnz();
    for (int64_t n = gtid(); n < g.nnodes(); n += gstride()) {
        const int iz = (int)(n % nnz);
        const int64_t t = n / nnz;
        const int iy = (int)(t % nny), ix = (int)(t / nny);
        const int cx = (ix > 0) + (ix < g.nelx), cy = (iy > 0) + (iy < g.nely), cz = (iz > 0) + (iz < g.nelz);
        invD[n] = 1.0 / ((double)(cx * cy * cz) * aii);
    }
}

// xf_e = (1/8) sum of the element's nodal values (zyx element order)
__global__ void k_helm_out(const double* __restrict__ y, double* __restrict__ xf, Grid g) {
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t nyz = (int64_t)nny * nnz;
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        const int64_t n0 = ((int64_t)ex * nny + ey) * nnz + ez;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            int a, b, c;
            local_offsets(m, a, b, c);
            s += y[n0 + a * nyz + b * nnz + c];
        }
        xf[e] = 0.125 * s;
    }
}

// ---------------------------------------------------------------------------
// linear density filter in FP64 (same stencil and normalisation as
// k_density_filter): transpose = 0: out_e = sum_j w_ej in_j / sw_e;
// transpose = 1: out_e = sum_j w_je in_j / sw_j
__global__ void k_density_filter_d(const double* __restrict__ in, double* __restrict__ out,
                                   const double* __restrict__ sw, Grid g, int transpose,
                                   const FilterStencil* __restrict__ st) {
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        double s = 0;
        for (int q = 0; q < st->count; q++) {
            const int kx = ex + st->off[q][0], ky = ey + st->off[q][1], kz = ez + st->off[q][2];
            if (kx < 0 || kx >= g.nelx || ky < 0 || ky >= g.nely || kz < 0 || kz >= g.nelz) continue;
            const int64_t j = ((int64_t)kx * g.nely + ky) * g.nelz + kz;
            s += transpose ? st->w[q] * in[j] / sw[j] : st->w[q] * in[j];
        }
        out[e] = transpose ? s : s / sw[e];
    }
}

// Heaviside projection in FP64 and its derivative (any eta):
//   H(u) = (tanh(beta eta) + tanh(beta (u - eta))) / (tanh(beta eta) + tanh(beta (1 - eta)))
//   H'(u) = beta (1 - tanh^2(beta (u - eta))) / (tanh(beta eta) + tanh(beta (1 - eta)))
__global__ void k_heaviside_d(const double* __restrict__ u, double* __restrict__ xp, int64_t n, double beta,
                              double eta) {
    const double tb = tanh(beta * eta), den = tb + tanh(beta * (1 - eta));
    for (int64_t i = gtid(); i < n; i += gstride()) xp[i] = (tb + tanh(beta * (u[i] - eta))) / den;
}
// g *= H'(u)
__global__ void k_heaviside_grad(const double* __restrict__ u, double* __restrict__ g, int64_t n, double beta,
                                 double eta) {
    const double den = tanh(beta * eta) + tanh(beta * (1 - eta));
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const double th = tanh(beta * (u[i] - eta));
        g[i] *= beta * (1 - th * th) / den;
    }
}

// element sensitivity of the objective with respect to the physical density:
//   dJ/dxp_e = -(lam_e^T K0 u_e) * p xp^(p-1) (E0 - Emin)
// summed over load cases (lam = u for compliance, the adjoint field for the
// mechanism); accumulate = 0 overwrites grad, 1 adds to it.
__global__ void k_adj_sens(const double* __restrict__ U, const double* __restrict__ Lam, const double* __restrict__ xp,
                           double penal, double E0, double Emin, Grid g, double* __restrict__ grad, int accumulate,
                           const double* __restrict__ ke) {
    const int nny = g.nny(), nnz = g.nnz();
    const int64_t nyz = (int64_t)nny * nnz;
    for (int64_t e = gtid(); e < g.nel(); e += gstride()) {
        const int ez = (int)(e % g.nelz);
        const int64_t t = e / g.nelz;
        const int ey = (int)(t % g.nely), ex = (int)(t / g.nely);
        const int64_t n0 = ((int64_t)ex * nny + ey) * nnz + ez;
        double u[24], l[24];
        for (int m = 0; m < 8; m++) {
            int a, b, c;
            local_offsets(m, a, b, c);
            const int64_t nm = n0 + a * nyz + b * nnz + c;
            for (int d = 0; d < 3; d++) {
                u[3 * m + d] = U[3 * nm + d];
                l[3 * m + d] = Lam[3 * nm + d];
            }
        }
        double en = 0;
        for (int i = 0; i < 24; i++) {
            double r = 0;
            for (int j = 0; j < 24; j++) r += ke[i * 24 + j] * u[j];
            en += l[i] * r;
        }
        const double x = xp[e];
        const double ge = -en * penal * pow(x, penal - 1) * (E0 - Emin);
        grad[e] = accumulate ? grad[e] + ge : ge;
    }
}

// ---------------------------------------------------------------------------
// springs of the mechanism ports (at most two DOFs): the fine-level operator
// becomes K + sum_s k_s e_s e_s^T. Applied after the element operator with
// the same mode convention (0: out = A in, 1: masked, 2: b - A in masked,
// 3: Jacobi sweep out = in + omega invD (b - A in)).
struct Springs {
    int n;
    int64_t dof[2];
    double k[2];
};
template <typename T>
__global__ void k_spring(const T* __restrict__ in, T* __restrict__ out, Springs sp, int mode,
                         const uint8_t* __restrict__ fixed, const T* __restrict__ invD, const double* __restrict__ omp,
                         const int* gate, int64_t n) {
    gate = batch_gate(gate);
    if (gate && *gate) return;
    in += batch_off(n);
    out += batch_off(n);
    const int s = threadIdx.x;
    if (blockIdx.x != 0 || s >= sp.n) return;
    const int64_t d = sp.dof[s];
    const double kv = sp.k[s] * (double)in[d];
    if (mode == 0) out[d] = (T)((double)out[d] + kv);
    else if (mode == 1) { if (!fixed[d]) out[d] = (T)((double)out[d] + kv); }
    else if (mode == 2) { if (!fixed[d]) out[d] = (T)((double)out[d] - kv); }
    else out[d] = (T)((double)out[d] - (*omp) * (double)invD[d] * kv);
}
// Jacobi diagonal with the springs: invD = 1 / (1/invD + k) on free DOFs
__global__ void k_spring_diag(double* __restrict__ invD, float* __restrict__ invD32, Springs sp) {
    const int s = threadIdx.x;
    if (blockIdx.x != 0 || s >= sp.n) return;
    const int64_t d = sp.dof[s];
    const double v = invD[d];
    if (v > 0) {
        const double nv = 1.0 / (1.0 / v + sp.k[s]);
        invD[d] = nv;
        if (invD32) invD32[d] = (float)nv;
    }
}

}  // namespace topo
