// host_setup.cpp -- host-side problem setup of libtopo_hip: element
// stiffness, numbering, boundary-condition and load presets, defaults.
// No GPU is touched here; these give the inputs of the device path and are
// bit-exact with the reference's worker (js/optimizer-worker-3d.js).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <set>
#include <vector>
#include "common.hpp"

namespace topo {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// Unit-cube hex8 stiffness, E = 1, 2x2x2 Gauss, B rows
// [exx, eyy, ezz, gxy, gyz, gxz] (lk() js/optimizer-worker-3d.js:3404-3491).
// Local node a sits at natural coordinates (sa_x, sa_y, sa_z) in {-1,+1}^3;
// dN_a/dxi = sa_x (1 + sa_y eta)(1 + sa_z zeta) / 8 and cyclic. Summation
// order (Gauss point xi-major, then strain row p) matches the reference so
// the 576 entries are identical.
void k0_matrix(double nu, double* KE) {
    static const int S[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    const double f = 1.0 / ((1 + nu) * (1 - 2 * nu));
    const double d0 = f * (1 - nu), d1 = f * nu, g = f * (1 - 2 * nu) / 2;
    const double Cm[6][6] = {{d0, d1, d1, 0, 0, 0}, {d1, d0, d1, 0, 0, 0}, {d1, d1, d0, 0, 0, 0},
                             {0, 0, 0, g, 0, 0},    {0, 0, 0, 0, g, 0},    {0, 0, 0, 0, 0, g}};
    std::memset(KE, 0, 576 * sizeof(double));
    const double q = 1 / std::sqrt(3.0);
    const double gp[2] = {-q, q};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                const double xi[3] = {gp[i], gp[j], gp[k]};
                double B[6][24];
                std::memset(B, 0, sizeof B);
                for (int a = 0; a < 8; a++) {
                    double fac[3];
                    for (int c = 0; c < 3; c++) fac[c] = S[a][c] > 0 ? (1 + xi[c]) : (1 - xi[c]);
                    // derivative along c: product of the other two factors in
                    // increasing axis order, signed by S[a][c]
                    double dn[3];
                    const double prod[3] = {fac[1] * fac[2], fac[0] * fac[2], fac[0] * fac[1]};
                    for (int c = 0; c < 3; c++) dn[c] = (S[a][c] > 0 ? prod[c] : -prod[c]) * 0.125;
                    B[0][3 * a] = dn[0];
                    B[1][3 * a + 1] = dn[1];
                    B[2][3 * a + 2] = dn[2];
                    B[3][3 * a] = dn[1];
                    B[3][3 * a + 1] = dn[0];
                    B[4][3 * a + 1] = dn[2];
                    B[4][3 * a + 2] = dn[1];
                    B[5][3 * a] = dn[2];
                    B[5][3 * a + 2] = dn[0];
                }
                double CB[6][24];
                for (int m = 0; m < 6; m++)
                    for (int n = 0; n < 24; n++) {
                        double s = 0;
                        for (int p = 0; p < 6; p++) s += Cm[m][p] * B[p][n];
                        CB[m][n] = s;
                    }
                for (int m = 0; m < 24; m++)
                    for (int n = 0; n < 24; n++) {
                        double acc = KE[m * 24 + n];
                        for (int p = 0; p < 6; p++) acc += B[p][m] * CB[p][n];
                        KE[m * 24 + n] = acc;
                    }
            }
}

static inline int64_t nid(const Grid& g, int ix, int iy, int iz) {
    return ((int64_t)ix * g.nny() + iy) * g.nnz() + iz;
}

static void fill_edofs(const Grid& g, int32_t* ed) {
    static const int O[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int ez = 0; ez < g.nelz; ez++)
        for (int ey = 0; ey < g.nely; ey++)
            for (int ex = 0; ex < g.nelx; ex++) {
                int64_t e = ex + (int64_t)g.nelx * (ey + (int64_t)g.nely * ez);
                for (int a = 0; a < 8; a++) {
                    int64_t n = nid(g, ex + O[a][0], ey + O[a][1], ez + O[a][2]);
                    for (int d = 0; d < 3; d++) ed[e * 24 + 3 * a + d] = (int32_t)(3 * n + d);
                }
            }
}

// Recover (nelx,nely,nelz) from an edofs array and check every entry.
bool edofs_structured(const int32_t* ed, int64_t ndof, Grid* out) {
    if (!ed || ndof <= 0 || ndof % 3) return false;
    const int64_t nnz = ed[9] / 3;            // local node 3 = (0,1,0): n = nnz
    const int64_t nnyz = ed[3] / 3;           // local node 1 = (1,0,0): n = nny*nnz
    if (ed[0] != 0 || nnz < 2 || nnyz < 2 * nnz || nnyz % nnz) return false;
    const int64_t nny = nnyz / nnz;
    const int64_t nnodes = ndof / 3;
    if (nnodes % nnyz) return false;
    const int64_t nnx = nnodes / nnyz;
    if (nnx < 2) return false;
    Grid g;
    g.nelx = (int)(nnx - 1);
    g.nely = (int)(nny - 1);
    g.nelz = (int)(nnz - 1);
    static const int O[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int ez = 0; ez < g.nelz; ez++)
        for (int ey = 0; ey < g.nely; ey++)
            for (int ex = 0; ex < g.nelx; ex++) {
                int64_t e = ex + (int64_t)g.nelx * (ey + (int64_t)g.nely * ez);
                for (int a = 0; a < 8; a++) {
                    int64_t n = nid(g, ex + O[a][0], ey + O[a][1], ez + O[a][2]);
                    for (int d = 0; d < 3; d++)
                        if (ed[e * 24 + 3 * a + d] != (int32_t)(3 * n + d)) return false;
                }
            }
    *out = g;
    return true;
}

}  // namespace topo

using namespace topo;

extern "C" {

const char* topo_version(void) { return "libtopo_hip 0.1 (gfx950)"; }
const char* topo_last_error(void) { return g_err; }

int topo_device_info(int32_t* count, char* name, int32_t name_len) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    if (name && name_len > 0) {
        name[0] = 0;
        hipDeviceProp_t p;
        if (n > 0 && hipGetDeviceProperties(&p, 0) == hipSuccess) {
            snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        }
    }
    return n > 0 ? TOPO_OK : TOPO_E_NODEVICE;
}

// js/optimizer-worker-3d.js:20-64 (module constants), :1623-1647 and
// :2056-2072 (config reads with their defaults), js/main.js:61-110 (UI
// defaults volfrac/penal/rmin/maxIterations).
int topo_config_default(topo_config* c, int32_t nelx, int32_t nely, int32_t nelz) {
    if (!c || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_config_default: bad arguments");
        return TOPO_E_ARG;
    }
    std::memset(c, 0, sizeof *c);
    c->nelx = nelx; c->nely = nely; c->nelz = nelz;
    c->nload = 1;
    c->E0 = 1.0; c->Emin = 1e-9; c->nu = 0.3;
    c->volfrac = 0.3;
    c->penal = 3.0;
    c->penal_start = 1.5;            // min(penalTarget, 1.5)
    c->max_iterations = 100;
    c->continuation_iters = 33;      // max(20, floor(maxIterations/3))
    c->rmin = 1.5;
    c->filter_type = TOPO_FILTER_SENSITIVITY;
    c->use_projection = 1;
    c->beta_max = 64; c->beta_interval = 5; c->beta_fixed = 0; c->eta = 0.5;
    c->linear_solver = TOPO_SOLVER_MGPCG;   // DEFAULT_LINEAR_SOLVER
    c->precision = TOPO_PREC_FP64;
    c->mg_levels = 0;
    c->mg_nu1 = 2; c->mg_nu2 = 2; c->mg_omega = 0.5;
    c->coarse_solver = TOPO_COARSE_SWEEPS; c->coarse_sweeps = 30;
    c->galerkin_max_ndof = 3000;
    c->cg_max_iter = 400;
    c->cg_tol = 0.0;                 // schedule 1e-3 -> 1e-8
    c->warm_start = -1;
    c->move = 0.2;
    c->objective = TOPO_OBJ_COMPLIANCE;
    c->mech_in_dof = -1; c->mech_out_dof = -1;
    c->mech_k_in = 1.0; c->mech_k_out = 0.001;
    c->rhs_tile = 0;
    c->use_graphs = 1;
    c->skip_factor = 1000.0;          // skipThreshold = Emin * 1000
    c->mfma_apply = 0;               // the VALU kernel is the faster one at these sizes
    return TOPO_OK;
}

int topo_k0(double nu, double* ke) {
    if (!ke || !(nu > -1.0 && nu < 0.5)) {
        set_error("topo_k0: bad arguments");
        return TOPO_E_ARG;
    }
    k0_matrix(nu, ke);
    return TOPO_OK;
}

int topo_edofs(int32_t nelx, int32_t nely, int32_t nelz, int32_t* edofs) {
    if (!edofs || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_edofs: bad arguments");
        return TOPO_E_ARG;
    }
    Grid g;
    g.nelx = nelx; g.nely = nely; g.nelz = nelz;
    fill_edofs(g, edofs);
    return TOPO_OK;
}

int topo_fixed_dofs(int32_t nelx, int32_t nely, int32_t nelz, int32_t position, int32_t dof_mask,
                    uint8_t* mask) {
    if (!mask || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_fixed_dofs: bad arguments");
        return TOPO_E_ARG;
    }
    Grid g;
    g.nelx = nelx; g.nely = nely; g.nelz = nelz;
    int cnt = 0;
    auto add = [&](int64_t n) {
        for (int d = 0; d < 3; d++)
            if (dof_mask & (1 << d)) {
                mask[3 * n + d] = 1;
                cnt++;
            }
    };
    switch (position) {
    case TOPO_FIX_LEFT:
    case TOPO_FIX_RIGHT: {
        const int ix = position == TOPO_FIX_LEFT ? 0 : nelx;
        for (int j = 0; j <= nely; j++)
            for (int k = 0; k <= nelz; k++) add(nid(g, ix, j, k));
        break;
    }
    case TOPO_FIX_BOTTOM:
    case TOPO_FIX_TOP: {
        const int iy = position == TOPO_FIX_BOTTOM ? 0 : nely;
        for (int i = 0; i <= nelx; i++)
            for (int k = 0; k <= nelz; k++) add(nid(g, i, iy, k));
        break;
    }
    case TOPO_FIX_BOTTOM_CORNERS:
        add(nid(g, 0, 0, 0));
        add(nid(g, nelx, 0, 0));
        add(nid(g, 0, nely, 0));
        add(nid(g, nelx, nely, 0));
        break;
    case TOPO_FIX_NONE: break;
    default:
        set_error("topo_fixed_dofs: unknown position %d", position);
        return TOPO_E_ARG;
    }
    return cnt;
}

int topo_load_vector(int32_t nelx, int32_t nely, int32_t nelz, int32_t direction, double mag,
                     const double* fv, float* F) {
    if (!F || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_load_vector: bad arguments");
        return TOPO_E_ARG;
    }
    Grid g;
    g.nelx = nelx; g.nely = nely; g.nelz = nelz;
    std::memset(F, 0, g.ndof() * sizeof(float));
    const int64_t corner = nid(g, nelx, nely, nelz);
    if (fv) {
        const double len = std::sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
        if (len > 0)
            for (int d = 0; d < 3; d++) F[3 * corner + d] = (float)((fv[d] / len) * mag);
        return TOPO_OK;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN: F[3 * corner + 1] = (float)-mag; break;
    case TOPO_LOAD_UP: F[3 * corner + 1] = (float)mag; break;
    case TOPO_LOAD_LEFT: F[3 * nid(g, nelx, nely / 2, nelz / 2)] = (float)-mag; break;
    case TOPO_LOAD_RIGHT: F[3 * nid(g, 0, nely / 2, nelz / 2)] = (float)mag; break;
    case TOPO_LOAD_TOP_CENTER: F[3 * nid(g, nelx / 2, nely / 2, nelz) + 2] = (float)mag; break;
    case TOPO_LOAD_NONE: break;
    default:
        set_error("topo_load_vector: unknown direction %d", direction);
        return TOPO_E_ARG;
    }
    return TOPO_OK;
}

int topo_preserved_elements(int32_t nelx, int32_t nely, int32_t nelz, int32_t position,
                            int32_t direction, uint8_t* pres) {
    if (!pres || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_preserved_elements: bad arguments");
        return TOPO_E_ARG;
    }
    auto E = [&](int ex, int ey, int ez) {
        return (int64_t)ex + (int64_t)nelx * (ey + (int64_t)nely * ez);
    };
    switch (position) {
    case TOPO_FIX_LEFT:
    case TOPO_FIX_RIGHT:
        for (int ez = 0; ez < nelz; ez++)
            for (int ey = 0; ey < nely; ey++) pres[E(position == TOPO_FIX_LEFT ? 0 : nelx - 1, ey, ez)] = 1;
        break;
    case TOPO_FIX_BOTTOM:
    case TOPO_FIX_TOP:
        for (int ez = 0; ez < nelz; ez++)
            for (int ex = 0; ex < nelx; ex++) pres[E(ex, position == TOPO_FIX_BOTTOM ? 0 : nely - 1, ez)] = 1;
        break;
    case TOPO_FIX_BOTTOM_CORNERS:
        pres[E(0, 0, 0)] = 1;
        pres[E(nelx - 1, 0, 0)] = 1;
        pres[E(0, nely - 1, 0)] = 1;
        pres[E(nelx - 1, nely - 1, 0)] = 1;
        break;
    default: break;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN:
    case TOPO_LOAD_UP: pres[E(nelx - 1, nely - 1, nelz - 1)] = 1; break;
    case TOPO_LOAD_LEFT: pres[E(nelx - 1, nely / 2, nelz / 2)] = 1; break;
    case TOPO_LOAD_RIGHT: pres[E(0, nely / 2, nelz / 2)] = 1; break;
    case TOPO_LOAD_TOP_CENTER: pres[E(nelx / 2, nely / 2, nelz - 1)] = 1; break;
    default: break;
    }
    return TOPO_OK;
}

// face nodes (FACE_NODE_INDICES js/optimizer-worker-3d.js:49-56)
static const int kFaceNodes[6][4] = {{0, 4, 7, 3}, {5, 1, 2, 6}, {0, 1, 5, 4},
                                     {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};

static void painted_nodes(const Grid& g, const int32_t* faces, int32_t nfaces, std::vector<int64_t>& order) {
    static const int O[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    std::set<int64_t> seen;
    const int64_t maxNode = g.nnodes();
    for (int32_t f = 0; f < nfaces; f++) {
        const int vx = faces[4 * f], vy = faces[4 * f + 1], vz = faces[4 * f + 2], fi = faces[4 * f + 3];
        int64_t all[8];
        for (int a = 0; a < 8; a++) all[a] = nid(g, vx + O[a][0], vy + O[a][1], vz + O[a][2]);
        const int cnt = (fi >= 0 && fi < 6) ? 4 : 8;
        for (int i = 0; i < cnt; i++) {
            const int64_t n = cnt == 4 ? all[kFaceNodes[fi][i]] : all[i];
            if (n >= 0 && n < maxNode && seen.insert(n).second) order.push_back(n);
        }
    }
}

int topo_load_paint(int32_t nelx, int32_t nely, int32_t nelz, const int32_t* faces, int32_t nfaces,
                    const double* dir3, double magnitude, float* F) {
    if (!F || !dir3 || (nfaces > 0 && !faces) || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_load_paint: bad arguments");
        return TOPO_E_ARG;
    }
    Grid g;
    g.nelx = nelx; g.nely = nely; g.nelz = nelz;
    std::vector<int64_t> nodes;
    painted_nodes(g, faces, nfaces, nodes);
    if (nodes.empty()) return 0;
    const double per = magnitude / (double)nodes.size();
    for (int64_t n : nodes)
        for (int d = 0; d < 3; d++) F[3 * n + d] = (float)((double)F[3 * n + d] + dir3[d] * per);
    return (int)nodes.size();
}

int topo_fixed_paint(int32_t nelx, int32_t nely, int32_t nelz, const int32_t* faces, int32_t nfaces,
                     int32_t dof_mask, uint8_t* mask) {
    if (!mask || (nfaces > 0 && !faces) || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("topo_fixed_paint: bad arguments");
        return TOPO_E_ARG;
    }
    Grid g;
    g.nelx = nelx; g.nely = nely; g.nelz = nelz;
    std::vector<int64_t> nodes;
    painted_nodes(g, faces, nfaces, nodes);
    int cnt = 0;
    for (int64_t n : nodes)
        for (int d = 0; d < 3; d++)
            if (dof_mask & (1 << d)) {
                mask[3 * n + d] = 1;
                cnt++;
            }
    return cnt;
}

// ---------------------------------------------------------------------------
// 2-D path (js/optimizer-worker.js): node n = (nely+1) ix + iy, element
// e = ey + ex nely.

// lk() :1732-1767: the plane-stress bilinear quad stiffness of the classic
// 88-line code, K = k[P(i,j)] / (1 - nu^2) with the eight entries below.
// Rows 4..7 repeat rows 0..3 with the two node pairs exchanged
// (P(i+4, j) = P(i, (j+4) mod 8)).
int topo_k0_2d(double nu, double* ke) {
    if (!ke || !(nu > -1.0 && nu < 0.5)) {
        set_error("topo_k0_2d: bad arguments");
        return TOPO_E_ARG;
    }
    const double k[8] = {1.0 / 2 - nu / 6,  1.0 / 8 + nu / 8, -1.0 / 4 - nu / 12, -1.0 / 8 + 3 * nu / 8,
                         -1.0 / 4 + nu / 12, -1.0 / 8 - nu / 8, nu / 6,            1.0 / 8 - 3 * nu / 8};
    static const int P[4][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2},
                                {2, 7, 0, 5, 6, 3, 4, 1}, {3, 6, 5, 0, 7, 2, 1, 4}};
    const double den = 1 - nu * nu;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            const int pij = i < 4 ? P[i][j] : P[i - 4][(j + 4) % 8];
            ke[i * 8 + j] = k[pij] * 1 / den;
        }
    return TOPO_OK;
}

int topo_edofs_2d(int32_t nelx, int32_t nely, int32_t* ed) {
    if (!ed || nelx < 1 || nely < 1) {
        set_error("topo_edofs_2d: bad arguments");
        return TOPO_E_ARG;
    }
    for (int ex = 0; ex < nelx; ex++)
        for (int ey = 0; ey < nely; ey++) {
            const int64_t e = (int64_t)ey + (int64_t)ex * nely;
            const int32_t n1 = (nely + 1) * ex + ey, n2 = (nely + 1) * (ex + 1) + ey;
            const int32_t v[8] = {2 * n1, 2 * n1 + 1, 2 * n2, 2 * n2 + 1, 2 * n2 + 2, 2 * n2 + 3, 2 * n1 + 2, 2 * n1 + 3};
            for (int i = 0; i < 8; i++) ed[e * 8 + i] = v[i];
        }
    return TOPO_OK;
}

int topo_fixed_dofs_2d(int32_t nelx, int32_t nely, int32_t position, int32_t dof_mask, uint8_t* mask) {
    if (!mask || nelx < 1 || nely < 1) {
        set_error("topo_fixed_dofs_2d: bad arguments");
        return TOPO_E_ARG;
    }
    int cnt = 0;
    auto add = [&](int64_t n) {
        for (int d = 0; d < 2; d++)
            if (dof_mask & (1 << d)) {
                mask[2 * n + d] = 1;
                cnt++;
            }
    };
    const int64_t nny = nely + 1;
    switch (position) {
    case TOPO_FIX_LEFT: for (int j = 0; j <= nely; j++) add(j); break;
    case TOPO_FIX_RIGHT: for (int j = 0; j <= nely; j++) add(nny * nelx + j); break;
    case TOPO_FIX_BOTTOM: for (int i = 0; i <= nelx; i++) add(nny * i); break;
    case TOPO_FIX_TOP: for (int i = 0; i <= nelx; i++) add(nny * i + nely); break;
    case TOPO_FIX_NONE: break;
    default:
        set_error("topo_fixed_dofs_2d: position %d is not a 2-D edge", position);
        return TOPO_E_ARG;
    }
    return cnt;
}

int topo_load_vector_2d(int32_t nelx, int32_t nely, int32_t direction, double mag, const double* fv, float* F) {
    if (!F || nelx < 1 || nely < 1) {
        set_error("topo_load_vector_2d: bad arguments");
        return TOPO_E_ARG;
    }
    const int64_t nny = nely + 1;
    std::memset(F, 0, 2 * nny * (nelx + 1) * sizeof(float));
    if (fv) {
        const double len = std::sqrt(fv[0] * fv[0] + fv[1] * fv[1]);
        if (len > 0) {
            const int64_t n = nny * nelx + nely;
            F[2 * n] = (float)((fv[0] / len) * mag);
            F[2 * n + 1] = (float)((fv[1] / len) * mag);
        }
        return TOPO_OK;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN: F[2 * (nny * nelx + nely) + 1] = (float)-mag; break;
    case TOPO_LOAD_UP: F[2 * (nny * nelx) + 1] = (float)mag; break;
    case TOPO_LOAD_LEFT: F[2 * (nny * nelx + nely / 2)] = (float)-mag; break;
    case TOPO_LOAD_RIGHT: F[2 * (nely / 2)] = (float)mag; break;
    case TOPO_LOAD_NONE: break;
    default:
        set_error("topo_load_vector_2d: direction %d is not a 2-D direction", direction);
        return TOPO_E_ARG;
    }
    return TOPO_OK;
}

int topo_preserved_elements_2d(int32_t nelx, int32_t nely, int32_t position, int32_t direction, uint8_t* pres) {
    if (!pres || nelx < 1 || nely < 1) {
        set_error("topo_preserved_elements_2d: bad arguments");
        return TOPO_E_ARG;
    }
    auto E = [&](int ex, int ey) { return (int64_t)ey + (int64_t)ex * nely; };
    switch (position) {
    case TOPO_FIX_LEFT: for (int ey = 0; ey < nely; ey++) pres[E(0, ey)] = 1; break;
    case TOPO_FIX_RIGHT: for (int ey = 0; ey < nely; ey++) pres[E(nelx - 1, ey)] = 1; break;
    case TOPO_FIX_BOTTOM: for (int ex = 0; ex < nelx; ex++) pres[E(ex, 0)] = 1; break;
    case TOPO_FIX_TOP: for (int ex = 0; ex < nelx; ex++) pres[E(ex, nely - 1)] = 1; break;
    default: break;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN: pres[E(nelx - 1, nely - 1)] = 1; break;
    case TOPO_LOAD_UP: pres[E(nelx - 1, 0)] = 1; break;
    case TOPO_LOAD_LEFT: pres[E(nelx - 1, nely / 2)] = 1; break;
    case TOPO_LOAD_RIGHT: pres[E(0, nely / 2)] = 1; break;
    default: break;
    }
    return TOPO_OK;
}

// corner nodes n1, n2, n2 + 1, n1 + 1 of painted cells, first occurrence order
static void painted_nodes_2d(int nelx, int nely, const int32_t* cells, int32_t n, std::vector<int64_t>& order) {
    std::set<int64_t> seen;
    const int64_t nny = nely + 1, maxNode = (int64_t)(nelx + 1) * nny;
    for (int32_t c = 0; c < n; c++) {
        const int64_t vx = cells[2 * c], vy = cells[2 * c + 1];
        const int64_t n1 = nny * vx + vy, n2 = nny * (vx + 1) + vy;
        const int64_t nodes[4] = {n1, n2, n2 + 1, n1 + 1};
        for (int64_t q : nodes)
            if (q >= 0 && q < maxNode && seen.insert(q).second) order.push_back(q);
    }
}

int topo_fixed_paint_2d(int32_t nelx, int32_t nely, const int32_t* cells, int32_t ncells, int32_t dof_mask,
                        uint8_t* mask) {
    if (!mask || (ncells > 0 && !cells) || nelx < 1 || nely < 1) {
        set_error("topo_fixed_paint_2d: bad arguments");
        return TOPO_E_ARG;
    }
    std::vector<int64_t> nodes;
    painted_nodes_2d(nelx, nely, cells, ncells, nodes);
    int cnt = 0;
    for (int64_t q : nodes)
        for (int d = 0; d < 2; d++)
            if (dof_mask & (1 << d)) {
                mask[2 * q + d] = 1;
                cnt++;
            }
    return cnt;
}

int topo_load_paint_2d(int32_t nelx, int32_t nely, const int32_t* cells, int32_t ncells, const double* dir2,
                       double magnitude, float* F) {
    if (!F || !dir2 || (ncells > 0 && !cells) || nelx < 1 || nely < 1) {
        set_error("topo_load_paint_2d: bad arguments");
        return TOPO_E_ARG;
    }
    std::vector<int64_t> nodes;
    painted_nodes_2d(nelx, nely, cells, ncells, nodes);
    if (nodes.empty()) return 0;
    const double per = magnitude / (double)nodes.size();
    for (int64_t q : nodes)
        for (int d = 0; d < 2; d++) F[2 * q + d] = (float)((double)F[2 * q + d] + dir2[d] * per);
    return (int)nodes.size();
}

// js/optimizer-worker.js:15-18 (CG constants), :361-378 and :570-578
// (config reads with their defaults)
int topo_config_default_2d(topo_config* c, int32_t nelx, int32_t nely) {
    const int rc = topo_config_default(c, nelx, nely, 1);
    if (rc < 0) return rc;
    c->linear_solver = TOPO_SOLVER_PCG;
    c->cg_max_iter = 2000;            // MAX_CG_ITERATIONS
    c->cg_tol = 0.0;                  // schedule max(CG_TOLERANCE, change * CG_TOLERANCE_COARSE)
    return TOPO_OK;
}

}  // extern "C"
