// topo_oracle.cpp -- CPU ORACLE (test infrastructure only).
//
// A single-threaded FP64 C++ restatement of the reference's 3-D SIMP path
// (Jozo132/TopologyOptimization: js/optimizer-worker-3d.js and
// assembly/index.ts). It exists to check the HIP library and to time a CPU
// baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and never as the thing measured or shipped.
//
// Every function cites the reference file:line it follows. The reference
// keeps several quantities in Float32 (x, xPhys, F, dc, dcn, energies, Hs);
// the restatement rounds at the same points.
//
// Pinning (what checks each function, see tests/test_oracle.py):
//   or_k0            symmetry, 6 rigid-body null modes, diagonal 0.47008547,
//                    closed-form entries of the textbook hex8 stiffness
//   or_edofs         closed-form index formulas, counts
//   or_applyA        brute-force assembled dense K (tiny grids), rigid modes,
//                    linearity, zero/no-active cases, symmetry
//   or_ebePCG        dense Cholesky solve on tiny grids, U^T K U = F^T U
//   or_mg_*          V-cycle symmetry <Ma,b>=<a,Mb>, MGPCG vs dense solve
//   or_energy        sum_e E_e w_e = U^T K U
//   or_sens_filter   brute-force neighbour sums
//   or_density_filter brute force + central finite differences of compliance
//   or_oc            volume constraint, brute-force lambda scan
//   or_heaviside     closed form and its derivative
//   or_bddc_*        partition counts and corner sets (hand-computed known
//                    answers), local step closed form on a uniform grid,
//                    KSP solve vs dense solve
// Functions of the build-only features (Cholesky coarse solve, Helmholtz
// filter, mechanism objective, multi-load) have no reference counterpart;
// they are pinned by the invariants named at each.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <set>
#include "../include/topo.h"
#include <omp.h>

// Threads of the CPU-baseline timing mode (or_set_threads). With 1 (the
// default, and what every test uses) each loop runs in the order the
// restatement states; with more, the element scatters run colour by colour
// (elements of one colour share no node), the vector loops and reductions
// in parallel, and the coarsest Cholesky by parallel rows: the same
// algorithm with other summation orders, for the CPU baseline only.
static int g_threads = 1;
extern "C" int or_set_threads(int n) {
    g_threads = n > 0 ? n : omp_get_max_threads();
    return g_threads;
}

typedef std::vector<double> dvec;

static inline int64_t node_id(int ix, int iy, int iz, int nny, int nnz) {
    return (int64_t)ix * nny * nnz + (int64_t)iy * nnz + iz;
}

extern "C" {

// ---------------------------------------------------------------------------
// K0: lk() js/optimizer-worker-3d.js:3404-3491, same loop and product order.
int or_k0(double nu, double* KE) {
    const double E = 1.0;
    double C[6][6] = {{0}};
    const double fact = E / ((1 + nu) * (1 - 2 * nu));
    C[0][0] = fact * (1 - nu); C[0][1] = fact * nu; C[0][2] = fact * nu;
    C[1][0] = fact * nu; C[1][1] = fact * (1 - nu); C[1][2] = fact * nu;
    C[2][0] = fact * nu; C[2][1] = fact * nu; C[2][2] = fact * (1 - nu);
    C[3][3] = fact * (1 - 2 * nu) / 2;
    C[4][4] = fact * (1 - 2 * nu) / 2;
    C[5][5] = fact * (1 - 2 * nu) / 2;
    for (int i = 0; i < 576; i++) KE[i] = 0.0;
    const double g[2] = {-1 / std::sqrt(3.0), 1 / std::sqrt(3.0)};
    // natural coordinates of the 8 local nodes (:1448-1494 order)
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                const double xi = g[i], eta = g[j], zeta = g[k];
                double dN[8][3];
                for (int n = 0; n < 8; n++) {
                    // (1 +/- eta)(1 +/- zeta) etc., sign applied to the product
                    const double fx = sx[n] > 0 ? (1 + xi) : (1 - xi);
                    const double fy = sy[n] > 0 ? (1 + eta) : (1 - eta);
                    const double fz = sz[n] > 0 ? (1 + zeta) : (1 - zeta);
                    double a = fy * fz, b = fx * fz, c = fx * fy;
                    dN[n][0] = (sx[n] > 0 ? a : -a) * 0.125;
                    dN[n][1] = (sy[n] > 0 ? b : -b) * 0.125;
                    dN[n][2] = (sz[n] > 0 ? c : -c) * 0.125;
                }
                double B[6][24] = {{0}};
                for (int n = 0; n < 8; n++) {
                    int col = n * 3;
                    B[0][col] = dN[n][0];
                    B[1][col + 1] = dN[n][1];
                    B[2][col + 2] = dN[n][2];
                    B[3][col] = dN[n][1];
                    B[3][col + 1] = dN[n][0];
                    B[4][col + 1] = dN[n][2];
                    B[4][col + 2] = dN[n][1];
                    B[5][col] = dN[n][2];
                    B[5][col + 2] = dN[n][0];
                }
                double CB[6][24] = {{0}};
                for (int m = 0; m < 6; m++)
                    for (int n = 0; n < 24; n++)
                        for (int p = 0; p < 6; p++) CB[m][n] += C[m][p] * B[p][n];
                for (int m = 0; m < 24; m++)
                    for (int n = 0; n < 24; n++)
                        for (int p = 0; p < 6; p++) KE[m * 24 + n] += B[p][m] * CB[p][n] * 1.0 * 1.0;
            }
    return 0;
}

// _precomputeEdofs3D js/optimizer-worker-3d.js:1448-1494
int or_edofs(int nelx, int nely, int nelz, int32_t* edof) {
    const int nny = nely + 1, nnz = nelz + 1;
    for (int ez = 0; ez < nelz; ez++)
        for (int ey = 0; ey < nely; ey++)
            for (int ex = 0; ex < nelx; ex++) {
                int64_t e = ex + (int64_t)ey * nelx + (int64_t)ez * nelx * nely;
                int64_t n[8] = {node_id(ex, ey, ez, nny, nnz),         node_id(ex + 1, ey, ez, nny, nnz),
                                node_id(ex + 1, ey + 1, ez, nny, nnz), node_id(ex, ey + 1, ez, nny, nnz),
                                node_id(ex, ey, ez + 1, nny, nnz),     node_id(ex + 1, ey, ez + 1, nny, nnz),
                                node_id(ex + 1, ey + 1, ez + 1, nny, nnz), node_id(ex, ey + 1, ez + 1, nny, nnz)};
                for (int a = 0; a < 8; a++)
                    for (int d = 0; d < 3; d++) edof[e * 24 + 3 * a + d] = (int32_t)(3 * n[a] + d);
            }
    return 0;
}

// getFixedDOFs js/optimizer-worker-3d.js:4436-4493, _addNodeDOFs :4429-4434
int or_fixed_dofs(int nelx, int nely, int nelz, int position, int dofmask, uint8_t* mask) {
    const int nny = nely + 1, nnz = nelz + 1;
    int cnt = 0;
    auto add = [&](int64_t n) {
        for (int d = 0; d < 3; d++)
            if (dofmask & (1 << d)) { mask[3 * n + d] = 1; cnt++; }
    };
    switch (position) {
    case TOPO_FIX_LEFT:
        for (int j = 0; j <= nely; j++) for (int k = 0; k <= nelz; k++) add(node_id(0, j, k, nny, nnz));
        break;
    case TOPO_FIX_RIGHT:
        for (int j = 0; j <= nely; j++) for (int k = 0; k <= nelz; k++) add(node_id(nelx, j, k, nny, nnz));
        break;
    case TOPO_FIX_BOTTOM:
        for (int i = 0; i <= nelx; i++) for (int k = 0; k <= nelz; k++) add(node_id(i, 0, k, nny, nnz));
        break;
    case TOPO_FIX_TOP:
        for (int i = 0; i <= nelx; i++) for (int k = 0; k <= nelz; k++) add(node_id(i, nely, k, nny, nnz));
        break;
    case TOPO_FIX_BOTTOM_CORNERS:
        add(node_id(0, 0, 0, nny, nnz)); add(node_id(nelx, 0, 0, nny, nnz));
        add(node_id(0, nely, 0, nny, nnz)); add(node_id(nelx, nely, 0, nny, nnz));
        break;
    default: break;
    }
    return cnt;
}

// getLoadVector js/optimizer-worker-3d.js:4495-4547 (F is Float32)
int or_load_vector(int nelx, int nely, int nelz, int direction, double mag, const double* fv, float* F) {
    const int nny = nely + 1, nnz = nelz + 1;
    const int64_t ndof = 3LL * (nelx + 1) * nny * nnz;
    for (int64_t i = 0; i < ndof; i++) F[i] = 0.f;
    if (fv) {
        double len = std::sqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
        if (len > 0) {
            int64_t n = node_id(nelx, nely, nelz, nny, nnz);
            for (int d = 0; d < 3; d++) F[3 * n + d] = (float)((fv[d] / len) * mag);
        }
        return 0;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN: F[3 * node_id(nelx, nely, nelz, nny, nnz) + 1] = (float)(-mag); break;
    case TOPO_LOAD_UP: F[3 * node_id(nelx, nely, nelz, nny, nnz) + 1] = (float)mag; break;
    case TOPO_LOAD_LEFT: F[3 * node_id(nelx, nely / 2, nelz / 2, nny, nnz)] = (float)(-mag); break;
    case TOPO_LOAD_RIGHT: F[3 * node_id(0, nely / 2, nelz / 2, nny, nnz)] = (float)mag; break;
    case TOPO_LOAD_TOP_CENTER: F[3 * node_id(nelx / 2, nely / 2, nelz, nny, nnz) + 2] = (float)mag; break;
    default: break;
    }
    return 0;
}

// getConstraintElements :4552-4592 and getForceElements :4597-4619
int or_preserved(int nelx, int nely, int nelz, int position, int direction, uint8_t* pres) {
    auto E = [&](int ex, int ey, int ez) { return (int64_t)ex + (int64_t)ey * nelx + (int64_t)ez * nelx * nely; };
    switch (position) {
    case TOPO_FIX_LEFT: for (int ez = 0; ez < nelz; ez++) for (int ey = 0; ey < nely; ey++) pres[E(0, ey, ez)] = 1; break;
    case TOPO_FIX_RIGHT: for (int ez = 0; ez < nelz; ez++) for (int ey = 0; ey < nely; ey++) pres[E(nelx - 1, ey, ez)] = 1; break;
    case TOPO_FIX_BOTTOM: for (int ez = 0; ez < nelz; ez++) for (int ex = 0; ex < nelx; ex++) pres[E(ex, 0, ez)] = 1; break;
    case TOPO_FIX_TOP: for (int ez = 0; ez < nelz; ez++) for (int ex = 0; ex < nelx; ex++) pres[E(ex, nely - 1, ez)] = 1; break;
    case TOPO_FIX_BOTTOM_CORNERS:
        pres[E(0, 0, 0)] = 1; pres[E(nelx - 1, 0, 0)] = 1; pres[E(0, nely - 1, 0)] = 1; pres[E(nelx - 1, nely - 1, 0)] = 1;
        break;
    default: break;
    }
    switch (direction) {
    case TOPO_LOAD_DOWN: case TOPO_LOAD_UP: pres[E(nelx - 1, nely - 1, nelz - 1)] = 1; break;
    case TOPO_LOAD_LEFT: pres[E(nelx - 1, nely / 2, nelz / 2)] = 1; break;
    case TOPO_LOAD_RIGHT: pres[E(0, nely / 2, nelz / 2)] = 1; break;
    case TOPO_LOAD_TOP_CENTER: pres[E(nelx / 2, nely / 2, nelz - 1)] = 1; break;
    default: break;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// _powDensity js/optimizer-worker-3d.js:293-306 (integer fast path)
static double pow_density(double x, double p) {
    if (std::isfinite(p) && p == std::floor(p)) {
        switch ((int)p) {
        case 0: return 1.0;
        case 1: return x;
        case 2: return x * x;
        case 3: return x * x * x;
        case 4: { double x2 = x * x; return x2 * x2; }
        case 5: { double x2 = x * x; return x2 * x2 * x; }
        case 6: { double x2 = x * x; return x2 * x2 * x2; }
        default: break;
        }
    }
    return std::pow(x, p);
}

// applyAEbe3D assembly/index.ts:403-439: Ap = sum over active e of E_e K0 u_e
// the element scatter Ap += E_e K0 p_e over 8 colours (ex, ey, ez parities):
// elements of one colour share no node, so each colour runs in parallel
static void ebe_colored(int nelx, int nely, const double* ke, const double* evals, const int32_t* active,
                        int64_t activeCount, const int32_t* edof, const double* p, double* Ap) {
    for (int col = 0; col < 8; col++) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t ai = 0; ai < activeCount; ai++) {
            const int64_t e = active[ai];
            const int64_t ex = e % nelx, t = e / nelx, ey = t % nely, ez = t / nely;
            if ((int)((ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2)) != col) continue;
            const double E = evals[e];
            const int32_t* ed = &edof[e * 24];
            double loc[24];
            for (int j = 0; j < 24; j++) loc[j] = p[ed[j]];
            for (int i = 0; i < 24; i++) {
                double s = 0.0;
                for (int j = 0; j < 24; j++) s += ke[i * 24 + j] * loc[j];
                Ap[ed[i]] += E * s;
            }
        }
    }
}

int or_applyA(int nelx, int nely, int nelz, const double* ke, const double* evals,
              const int32_t* active, int activeCount, const double* p, double* Ap) {
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    const int64_t nel = (int64_t)nelx * nely * nelz;
    std::vector<int32_t> edof(nel * 24);
    or_edofs(nelx, nely, nelz, edof.data());
    for (int64_t i = 0; i < ndof; i++) Ap[i] = 0.0;
    if (g_threads > 1) {
        ebe_colored(nelx, nely, ke, evals, active, activeCount, edof.data(), p, Ap);
        return 0;
    }
    double loc[24];
    for (int ai = 0; ai < activeCount; ai++) {
        int64_t e = active[ai];
        double E = evals[e];
        const int32_t* ed = &edof[e * 24];
        for (int j = 0; j < 24; j++) loc[j] = p[ed[j]];
        for (int i = 0; i < 24; i++) {
            double s = 0.0;
            for (int j = 0; j < 24; j++) s += ke[i * 24 + j] * loc[j];
            Ap[ed[i]] += E * s;
        }
    }
    return 0;
}

// Dense global K (test helper: brute-force assembly of sum_e E_e P_e^T K0 P_e)
int or_assemble_dense(int nelx, int nely, int nelz, const double* ke, const double* evals, double* K) {
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    const int64_t nel = (int64_t)nelx * nely * nelz;
    std::vector<int32_t> edof(nel * 24);
    or_edofs(nelx, nely, nelz, edof.data());
    for (int64_t i = 0; i < ndof * ndof; i++) K[i] = 0.0;
    for (int64_t e = 0; e < nel; e++)
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) K[(int64_t)edof[e * 24 + i] * ndof + edof[e * 24 + j]] += evals[e] * ke[i * 24 + j];
    return 0;
}

// ebePCG assembly/index.ts:504-733, restated on host arrays (3-D, edofSize 24).
// E_vals via Math.pow (:544); reduced (free-DOF) space; warm start from U.
static double g_pcg_skip_factor = 1000.0;
int or_set_pcg_skip_factor(double f) { g_pcg_skip_factor = f; return 0; }

int or_ebePCG(const double* dens, const double* ke, int nelx, int nely, int nelz, const double* F,
              double* U, const int32_t* freedofs, int nfree, double Emin, double E0, double penal,
              int maxIter, double tolerance) {
    const double EPS = 1e-12;
    const int64_t nel = (int64_t)nelx * nely * nelz;
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    const double dE = E0 - Emin, skip = g_pcg_skip_factor > 0 ? Emin * g_pcg_skip_factor : -1.0;
    dvec ev(nel);
    std::vector<int32_t> act;
    for (int64_t e = 0; e < nel; e++) {
        ev[e] = Emin + std::pow(dens[e], penal) * dE;
        if (ev[e] > skip) act.push_back((int32_t)e);
    }
    std::vector<int32_t> edof(nel * 24);
    or_edofs(nelx, nely, nelz, edof.data());
    dvec diag(ndof, 0.0);
    for (int32_t e : act)
        for (int i = 0; i < 24; i++) diag[edof[(int64_t)e * 24 + i]] += ev[e] * ke[i * 24 + i];
    dvec invD(nfree), Uf(nfree), r(nfree), z(nfree), pp(nfree), Ap(nfree), pfull(ndof), apfull(ndof);
    for (int i = 0; i < nfree; i++) {
        double d = diag[freedofs[i]];
        invD[i] = d > 1e-30 ? 1.0 / d : 0.0;
    }
    for (int i = 0; i < nfree; i++) Uf[i] = U[freedofs[i]];
    std::fill(pfull.begin(), pfull.end(), 0.0);
    for (int i = 0; i < nfree; i++) pfull[freedofs[i]] = Uf[i];
    or_applyA(nelx, nely, nelz, ke, ev.data(), act.data(), (int)act.size(), pfull.data(), apfull.data());
    for (int i = 0; i < nfree; i++) r[i] = F[freedofs[i]] - apfull[freedofs[i]];
    double rz = 0, r0 = 0;
    for (int i = 0; i < nfree; i++) {
        double zi = invD[i] * r[i];
        z[i] = zi; pp[i] = zi; rz += r[i] * zi; r0 += r[i] * r[i];
    }
    const double tolSq = tolerance * tolerance * (r0 > 1e-30 ? r0 : 1e-30);
    const int cgMax = maxIter < nfree ? maxIter : nfree;
    int ret = cgMax;
    for (int it = 0; it < cgMax; it++) {
        double rn = 0;
        for (int i = 0; i < nfree; i++) rn += r[i] * r[i];
        if (rn < tolSq) { ret = it; break; }
        std::fill(pfull.begin(), pfull.end(), 0.0);
        for (int i = 0; i < nfree; i++) pfull[freedofs[i]] = pp[i];
        or_applyA(nelx, nely, nelz, ke, ev.data(), act.data(), (int)act.size(), pfull.data(), apfull.data());
        for (int i = 0; i < nfree; i++) Ap[i] = apfull[freedofs[i]];
        double pAp = 0;
        for (int i = 0; i < nfree; i++) pAp += pp[i] * Ap[i];
        double alpha = rz / (pAp + EPS);
        double rzn = 0;
        for (int i = 0; i < nfree; i++) {
            double ri = r[i] - alpha * Ap[i];
            double zi = invD[i] * ri;
            Uf[i] += alpha * pp[i];
            r[i] = ri; z[i] = zi; rzn += ri * zi;
        }
        double beta = rzn / (rz + EPS);
        for (int i = 0; i < nfree; i++) pp[i] = z[i] + beta * pp[i];
        rz = rzn;
    }
    for (int64_t i = 0; i < ndof; i++) U[i] = 0.0;
    for (int i = 0; i < nfree; i++) U[freedofs[i]] = Uf[i];
    return ret;
}

// ---------------------------------------------------------------------------
// Geometric multigrid: MGPrecond3D js/optimizer-worker-3d.js:317-1025 and
// mgApplyA/mgSmooth/mgVcycle/ebeMGPCG assembly/index.ts:759-1193.
struct OrLevel {
    int nelx, nely, nelz;
    int64_t nel, ndof;
    std::vector<int32_t> edof;
    std::vector<uint8_t> fixed;
    std::vector<int32_t> freed;
    dvec E;
    std::vector<int32_t> active;
    dvec diag, invDiag, Au, res, b, x;
    dvec denseK;   // ndof*ndof when assembled
    dvec chol;     // Cholesky factor (lower, row-major) of the free-DOF block
    std::vector<int32_t> cholMap;
};

struct OrMG {
    std::vector<OrLevel> L;
    dvec ke;
    double omega = 0.5;
    int nu1 = 2, nu2 = 2, coarseSweeps = 30;
    int galerkinMax = 3000;
    int coarseSolver = TOPO_COARSE_SWEEPS;
    double omegaCap = 0.0;        // build-only safeguard, 0 = reference
    double skip = 1e-6;           // skipThreshold (Emin * 1000 in the reference)
    std::vector<double> omegaL;   // per-level effective omega
};

static void level_init(OrLevel& l, int nx, int ny, int nz) {
    l.nelx = nx; l.nely = ny; l.nelz = nz;
    l.nel = (int64_t)nx * ny * nz;
    l.ndof = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
    l.edof.resize(l.nel * 24);
    or_edofs(nx, ny, nz, l.edof.data());
    l.E.assign(l.nel, 0.0);
    l.diag.assign(l.ndof, 0.0); l.invDiag.assign(l.ndof, 0.0);
    l.Au.assign(l.ndof, 0.0); l.res.assign(l.ndof, 0.0);
    l.b.assign(l.ndof, 0.0); l.x.assign(l.ndof, 0.0);
}

static void build_free(OrLevel& l) {  // _buildFreeDofsFromFixedMask :375-382
    l.freed.clear();
    for (int64_t i = 0; i < l.ndof; i++) if (!l.fixed[i]) l.freed.push_back((int32_t)i);
}

// _downsampleFixedMaskBy2 :384-406 (injection from fine node min(2c, nF-1))
static void downsample_fixed(const OrLevel& f, OrLevel& c) {
    const int nxF = f.nelx + 1, nyF = f.nely + 1, nzF = f.nelz + 1;
    const int nxC = c.nelx + 1, nyC = c.nely + 1, nzC = c.nelz + 1;
    c.fixed.assign(c.ndof, 0);
    for (int cz = 0; cz < nzC; cz++) {
        int fz = std::min(cz * 2, nzF - 1);
        for (int cy = 0; cy < nyC; cy++) {
            int fy = std::min(cy * 2, nyF - 1);
            for (int cx = 0; cx < nxC; cx++) {
                int fx = std::min(cx * 2, nxF - 1);
                int64_t nF = node_id(fx, fy, fz, nyF, nzF), nC = node_id(cx, cy, cz, nyC, nzC);
                for (int d = 0; d < 3; d++) c.fixed[3 * nC + d] = f.fixed[3 * nF + d];
            }
        }
    }
}

// 1-D prolongation weights (:591-603): fine f -> coarse (c0,c1,w0,w1)
static inline void pweights(int f, int nC, int& c0, int& c1, double& w0, double& w1) {
    c0 = std::min(f >> 1, nC - 1);
    c1 = std::min(c0 + 1, nC - 1);
    double t = (c0 != c1 && (f & 1)) ? 0.5 : 0.0;
    w0 = 1 - t; w1 = t;
}

static void level_applyA(const OrMG& mg, const OrLevel& l, const double* p, double* Ap) {
    if (!l.denseK.empty()) {  // mgApplyA dense branch assembly/index.ts:777-788
        for (int64_t i = 0; i < l.ndof; i++) {
            double s = 0;
            for (int64_t j = 0; j < l.ndof; j++) s += l.denseK[i * l.ndof + j] * p[j];
            Ap[i] = s;
        }
        return;
    }
    for (int64_t i = 0; i < l.ndof; i++) Ap[i] = 0.0;
    if (g_threads > 1) {
        ebe_colored(l.nelx, l.nely, mg.ke.data(), l.E.data(), l.active.data(), (int64_t)l.active.size(),
                    l.edof.data(), p, Ap);
        return;
    }
    double loc[24];
    for (int32_t e : l.active) {
        double E = l.E[e];
        const int32_t* ed = &l.edof[(int64_t)e * 24];
        for (int j = 0; j < 24; j++) loc[j] = p[ed[j]];
        for (int i = 0; i < 24; i++) {
            double s = 0;
            for (int j = 0; j < 24; j++) s += mg.ke[i * 24 + j] * loc[j];
            Ap[ed[i]] += E * s;
        }
    }
}

static void set_invdiag(OrLevel& l) {
    std::fill(l.invDiag.begin(), l.invDiag.end(), 0.0);
    for (int32_t d : l.freed) l.invDiag[d] = l.diag[d] > 1e-30 ? 1.0 / l.diag[d] : 0.0;
}

// _updateOperatorLevel :753-796
static void update_level0(OrMG& mg, const float* x, double penal, double E0, double Emin) {
    OrLevel& l = mg.L[0];
    const double dE = E0 - Emin, skip = mg.skip;
    l.active.clear();
    for (int64_t e = 0; e < l.nel; e++) {
        l.E[e] = Emin + pow_density((double)x[e], penal) * dE;
        if (l.E[e] > skip) l.active.push_back((int32_t)e);
    }
    std::fill(l.diag.begin(), l.diag.end(), 0.0);
    for (int32_t e : l.active)
        for (int i = 0; i < 24; i++) l.diag[l.edof[(int64_t)e * 24 + i]] += l.E[e] * mg.ke[i * 24 + i];
    set_invdiag(l);
}

// _restrictEvalAvg :702-739 (coarse E = 2 x mean of existing children)
static void restrict_eavg(OrMG& mg, const OrLevel& f, OrLevel& c, double Emin) {
    (void)Emin;
    const double skip = mg.skip;
    c.active.clear();
    int64_t idx = 0;
    for (int cz = 0; cz < c.nelz; cz++)
        for (int cy = 0; cy < c.nely; cy++)
            for (int cx = 0; cx < c.nelx; cx++, idx++) {
                double sum = 0; int cnt = 0;
                for (int dz = 0; dz < 2; dz++) { int fz = 2 * cz + dz; if (fz >= f.nelz) continue;
                    for (int dy = 0; dy < 2; dy++) { int fy = 2 * cy + dy; if (fy >= f.nely) continue;
                        for (int dx = 0; dx < 2; dx++) { int fx = 2 * cx + dx; if (fx >= f.nelx) continue;
                            sum += f.E[fx + (int64_t)fy * f.nelx + (int64_t)fz * f.nelx * f.nely]; cnt++; } } }
                double E = cnt > 0 ? (sum / cnt) * 2.0 : 0;
                c.E[idx] = E;
                if (E > skip) c.active.push_back((int32_t)idx);
            }
    std::fill(c.diag.begin(), c.diag.end(), 0.0);
    for (int32_t e : c.active)
        for (int i = 0; i < 24; i++) c.diag[c.edof[(int64_t)e * 24 + i]] += c.E[e] * mg.ke[i * 24 + i];
    set_invdiag(c);
    c.denseK.clear();
}

// Prolongation as explicit per-node lists (_buildProlongMap :585-625)
static void prolong_map(const OrLevel& f, const OrLevel& c, std::vector<int32_t>& pn, dvec& pw, std::vector<uint8_t>& pc) {
    const int nxF = f.nelx + 1, nyF = f.nely + 1, nzF = f.nelz + 1;
    const int nxC = c.nelx + 1, nyC = c.nely + 1, nzC = c.nelz + 1;
    int64_t nn = (int64_t)nxF * nyF * nzF;
    pn.assign(nn * 8, 0); pw.assign(nn * 8, 0.0); pc.assign(nn, 0);
    for (int fz = 0; fz < nzF; fz++) {
        int cz[2]; double wz[2]; pweights(fz, nzC, cz[0], cz[1], wz[0], wz[1]);
        for (int fy = 0; fy < nyF; fy++) {
            int cy[2]; double wy[2]; pweights(fy, nyC, cy[0], cy[1], wy[0], wy[1]);
            for (int fx = 0; fx < nxF; fx++) {
                int cx[2]; double wx[2]; pweights(fx, nxC, cx[0], cx[1], wx[0], wx[1]);
                int64_t nF = node_id(fx, fy, fz, nyF, nzF);
                int cnt = 0;
                for (int dz = 0; dz < 2; dz++) for (int dy = 0; dy < 2; dy++) for (int dx = 0; dx < 2; dx++) {
                    double w = wx[dx] * wy[dy] * wz[dz];
                    if (w > 0) { pn[nF * 8 + cnt] = (int32_t)node_id(cx[dx], cy[dy], cz[dz], nyC, nzC); pw[nF * 8 + cnt] = w; cnt++; }
                }
                pc[nF] = (uint8_t)cnt;
            }
        }
    }
}

static void diag_from_dense(OrLevel& l) {  // _diagFromDense :742-751
    for (int64_t i = 0; i < l.ndof; i++) l.diag[i] = l.denseK[i * l.ndof + i];
    set_invdiag(l);
}

// _assembleGalerkinFromElements :628-662
static void galerkin_from_elements(OrMG& mg, const OrLevel& f, OrLevel& c) {
    std::vector<int32_t> pn; dvec pw; std::vector<uint8_t> pc;
    prolong_map(f, c, pn, pw, pc);
    const int64_t n = c.ndof;
    c.denseK.assign(n * n, 0.0);
    for (int32_t e : f.active) {
        double E = f.E[e];
        for (int i = 0; i < 24; i++) {
            int32_t fi = f.edof[(int64_t)e * 24 + i];
            int32_t ni = fi / 3, di = fi % 3;
            for (int j = i; j < 24; j++) {
                double kij = E * mg.ke[i * 24 + j];
                if (kij == 0) continue;
                int32_t fj = f.edof[(int64_t)e * 24 + j];
                int32_t nj = fj / 3, dj = fj % 3;
                for (int a = 0; a < pc[ni]; a++) {
                    int64_t cr = 3LL * pn[ni * 8 + a] + di;
                    double wi = pw[ni * 8 + a];
                    for (int b = 0; b < pc[nj]; b++) {
                        int64_t cc = 3LL * pn[nj * 8 + b] + dj;
                        double v = wi * kij * pw[nj * 8 + b];
                        c.denseK[cr * n + cc] += v;
                        if (i != j) c.denseK[cc * n + cr] += v;
                    }
                }
            }
        }
    }
    diag_from_dense(c);
}

// _assembleGalerkinFromMatrix :665-699
static void galerkin_from_matrix(OrMG& mg, const OrLevel& f, OrLevel& c) {
    (void)mg;
    std::vector<int32_t> pn; dvec pw; std::vector<uint8_t> pc;
    prolong_map(f, c, pn, pw, pc);
    const int64_t nF = f.ndof, nC = c.ndof;
    dvec P(nF * nC, 0.0);
    int64_t nnF = nF / 3;
    for (int64_t k = 0; k < nnF; k++)
        for (int a = 0; a < pc[k]; a++)
            for (int d = 0; d < 3; d++) P[(3 * k + d) * nC + 3LL * pn[k * 8 + a] + d] = pw[k * 8 + a];
    dvec T(nF * nC, 0.0);
    for (int64_t i = 0; i < nF; i++)
        for (int64_t j = 0; j < nC; j++) {
            double s = 0;
            for (int64_t k = 0; k < nF; k++) s += f.denseK[i * nF + k] * P[k * nC + j];
            T[i * nC + j] = s;
        }
    c.denseK.assign(nC * nC, 0.0);
    for (int64_t i = 0; i < nC; i++)
        for (int64_t j = 0; j < nC; j++) {
            double s = 0;
            for (int64_t k = 0; k < nF; k++) s += P[k * nC + i] * T[k * nC + j];
            c.denseK[i * nC + j] = s;
        }
    diag_from_dense(c);
}

// Dense Cholesky of the coarsest free-DOF block (build-only feature; pinned
// by the factor residual ||L L^T - A|| and the solve residual).
static int coarse_cholesky(OrMG& mg, OrLevel& c) {
    const int64_t nf = (int64_t)c.freed.size();
    dvec A(nf * nf, 0.0);
    if (!c.denseK.empty()) {
        for (int64_t i = 0; i < nf; i++)
            for (int64_t j = 0; j < nf; j++) A[i * nf + j] = c.denseK[(int64_t)c.freed[i] * c.ndof + c.freed[j]];
    } else {
        dvec Kd(c.ndof * c.ndof, 0.0);
        for (int32_t e : c.active)
            for (int i = 0; i < 24; i++)
                for (int j = 0; j < 24; j++)
                    Kd[(int64_t)c.edof[(int64_t)e * 24 + i] * c.ndof + c.edof[(int64_t)e * 24 + j]] += c.E[e] * mg.ke[i * 24 + j];
        for (int64_t i = 0; i < nf; i++)
            for (int64_t j = 0; j < nf; j++) A[i * nf + j] = Kd[(int64_t)c.freed[i] * c.ndof + c.freed[j]];
    }
    // free DOFs with an empty row (no active element) get a unit diagonal
    for (int64_t i = 0; i < nf; i++)
        if (A[i * nf + i] <= 1e-30) { for (int64_t j = 0; j < nf; j++) A[i * nf + j] = A[j * nf + i] = 0; A[i * nf + i] = 1.0; }
    // a pivot below 1e-11 of the DOF's diagonal marks a DOF dependent on
    // the earlier ones (a mechanism after voided elements): it is dropped
    // (zero row/column of the factor's inverse), as in the HIP library
    dvec d0(nf);
    for (int64_t j = 0; j < nf; j++) d0[j] = A[j * nf + j];
    for (int64_t j = 0; j < nf; j++) {
        double s = A[j * nf + j];
        for (int64_t k = 0; k < j; k++) s -= A[j * nf + k] * A[j * nf + k];
        if (!(s > 1e-11 * d0[j])) {
            for (int64_t i = j; i < nf; i++) A[i * nf + j] = 0.0;
            continue;
        }
        double d = std::sqrt(s);
        A[j * nf + j] = d;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1 && nf - j > 256) schedule(static)
        for (int64_t i = j + 1; i < nf; i++) {
            double t = A[i * nf + j];
            for (int64_t k = 0; k < j; k++) t -= A[i * nf + k] * A[j * nf + k];
            A[i * nf + j] = t / d;
        }
    }
    c.chol.swap(A);
    return 0;
}

static void coarse_solve_chol(const OrLevel& c, const double* b, double* x) {
    const int64_t nf = (int64_t)c.freed.size();
    dvec y(nf);
    for (int64_t i = 0; i < nf; i++) {
        double s = b[c.freed[i]];
        for (int64_t k = 0; k < i; k++) s -= c.chol[i * nf + k] * y[k];
        y[i] = c.chol[i * nf + i] != 0.0 ? s / c.chol[i * nf + i] : 0.0;
    }
    for (int64_t i = nf - 1; i >= 0; i--) {
        double s = y[i];
        for (int64_t k = i + 1; k < nf; k++) s -= c.chol[k * nf + i] * y[k];
        y[i] = c.chol[i * nf + i] != 0.0 ? s / c.chol[i * nf + i] : 0.0;
    }
    for (int64_t i = 0; i < c.ndof; i++) x[i] = 0.0;
    for (int64_t i = 0; i < nf; i++) x[c.freed[i]] = y[i];
}

static void smooth(const OrMG& mg, OrLevel& l, const double* b, double* x, double omega) {  // mgSmooth :820-834
    level_applyA(mg, l, x, l.Au.data());
    const int64_t nf = (int64_t)l.freed.size();
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
    for (int64_t k = 0; k < nf; k++) {
        const int32_t d = l.freed[k];
        x[d] += omega * l.invDiag[d] * (b[d] - l.Au[d]);
    }
}

// Build-only smoother safeguard (no reference counterpart): estimate
// rho = lambda_max(D^-1 A) on a level by 20 power iterations from a hashed
// pseudo-random vector on free DOFs (v_i = frac(i * 2654435761 / 2^32) - 1/2);
// omega_eff = min(omega, cap / rho). Pinned by the eigenvalues of D^-1 A
// from numpy on small grids.
static double power_rho(const OrMG& mg, OrLevel& l) {
    dvec v(l.ndof, 0.0), w(l.ndof, 0.0);
    for (int32_t d : l.freed) v[d] = (double)(((uint64_t)d * 2654435761ULL) & 0xffffffffULL) / 4294967296.0 - 0.5;
    double rho = 0.0;
    for (int it = 0; it < 20; it++) {
        double nv = 0;
        for (int64_t i = 0; i < l.ndof; i++) nv += v[i] * v[i];
        level_applyA(mg, l, v.data(), w.data());
        double nw = 0;
        for (int64_t i = 0; i < l.ndof; i++) { w[i] *= l.invDiag[i]; nw += w[i] * w[i]; }
        if (!(nw > 0) || !(nv > 0)) return rho;
        rho = std::sqrt(nw / nv);
        const double s = 1.0 / std::sqrt(nw);
        for (int64_t i = 0; i < l.ndof; i++) v[i] = w[i] * s;
    }
    return rho;
}

// mgVcycle assembly/index.ts:837-1048
static void vcycle(OrMG& mg, int li, const double* b, double* x) {
    OrLevel& l = mg.L[li];
    const bool coarsest = li == (int)mg.L.size() - 1;
    if (coarsest && mg.coarseSolver == TOPO_COARSE_CHOLESKY) {
        coarse_solve_chol(l, b, x);
        return;
    }
    const double om = mg.omegaL.empty() ? mg.omega : mg.omegaL[li];
    for (int s = 0; s < mg.nu1; s++) smooth(mg, l, b, x, om);
    level_applyA(mg, l, x, l.Au.data());
    std::fill(l.res.begin(), l.res.end(), 0.0);
    for (int32_t d : l.freed) l.res[d] = b[d] - l.Au[d];
    if (coarsest) {
        for (int s = 0; s < mg.coarseSweeps; s++) smooth(mg, l, b, x, om);
        return;
    }
    OrLevel& c = mg.L[li + 1];
    const int nxF = l.nelx + 1, nyF = l.nely + 1, nzF = l.nelz + 1;
    const int nxC = c.nelx + 1, nyC = c.nely + 1, nzC = c.nelz + 1;
    std::fill(c.b.begin(), c.b.end(), 0.0);
    for (int fz = 0; fz < nzF; fz++) {
        int cz[2]; double wz[2]; pweights(fz, nzC, cz[0], cz[1], wz[0], wz[1]);
        for (int fy = 0; fy < nyF; fy++) {
            int cy[2]; double wy[2]; pweights(fy, nyC, cy[0], cy[1], wy[0], wy[1]);
            for (int fx = 0; fx < nxF; fx++) {
                int cx[2]; double wx[2]; pweights(fx, nxC, cx[0], cx[1], wx[0], wx[1]);
                int64_t nF = node_id(fx, fy, fz, nyF, nzF);
                const double r0 = l.res[3 * nF], r1 = l.res[3 * nF + 1], r2 = l.res[3 * nF + 2];
                // corner order 000,100,010,110,001,101,011,111 (:904-955)
                for (int dz = 0; dz < 2; dz++) for (int dy = 0; dy < 2; dy++) for (int dx = 0; dx < 2; dx++) {
                    double w = wx[dx] * wy[dy] * wz[dz];
                    if (w > 0) {
                        int64_t c3 = 3 * node_id(cx[dx], cy[dy], cz[dz], nyC, nzC);
                        c.b[c3] += w * r0; c.b[c3 + 1] += w * r1; c.b[c3 + 2] += w * r2;
                    }
                }
            }
        }
    }
    for (int64_t i = 0; i < c.ndof; i++) if (c.fixed[i]) c.b[i] = 0;
    std::fill(c.x.begin(), c.x.end(), 0.0);
    vcycle(mg, li + 1, c.b.data(), c.x.data());
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
    for (int fz = 0; fz < nzF; fz++) {
        int cz[2]; double wz[2]; pweights(fz, nzC, cz[0], cz[1], wz[0], wz[1]);
        for (int fy = 0; fy < nyF; fy++) {
            int cy[2]; double wy[2]; pweights(fy, nyC, cy[0], cy[1], wy[0], wy[1]);
            for (int fx = 0; fx < nxF; fx++) {
                int cx[2]; double wx[2]; pweights(fx, nxC, cx[0], cx[1], wx[0], wx[1]);
                int64_t nF = node_id(fx, fy, fz, nyF, nzF);
                for (int d = 0; d < 3; d++) {
                    if (l.fixed[3 * nF + d]) continue;
                    double v = 0;
                    for (int dz = 0; dz < 2; dz++) for (int dy = 0; dy < 2; dy++) for (int dx = 0; dx < 2; dx++)
                        v += c.x[3 * node_id(cx[dx], cy[dy], cz[dz], nyC, nzC) + d] * (wx[dx] * wy[dy] * wz[dz]);
                    x[3 * nF + d] += v;
                }
            }
        }
    }
    for (int s = 0; s < mg.nu2; s++) smooth(mg, l, b, x, om);
}

// MG hierarchy (MGPrecond3D.ensure :408-442). nlevels<=0: reference rule.
void* or_mg_create(int nelx, int nely, int nelz, const uint8_t* fixed0, const double* ke, int nlevels,
                   double omega, int nu1, int nu2, int coarseSolver, int coarseSweeps, int galerkinMax) {
    OrMG* mg = new OrMG();
    mg->ke.assign(ke, ke + 576);
    mg->omega = omega; mg->nu1 = nu1; mg->nu2 = nu2;
    mg->coarseSolver = coarseSolver; mg->coarseSweeps = coarseSweeps; mg->galerkinMax = galerkinMax;
    mg->L.emplace_back();
    level_init(mg->L[0], nelx, nely, nelz);
    mg->L[0].fixed.assign(fixed0, fixed0 + mg->L[0].ndof);
    build_free(mg->L[0]);
    const int maxL = nlevels > 0 ? nlevels : 6;
    int cx = nelx, cy = nely, cz = nelz;
    for (int li = 1; li < maxL; li++) {
        int nx = cx / 2, ny = cy / 2, nz = cz / 2;
        if (nx < 2 || ny < 2 || nz < 2) break;
        mg->L.emplace_back();
        OrLevel& c = mg->L.back();
        level_init(c, nx, ny, nz);
        downsample_fixed(mg->L[li - 1], c);
        build_free(c);
        cx = nx; cy = ny; cz = nz;
    }
    return mg;
}

int or_mg_nlevels(void* h) { return (int)((OrMG*)h)->L.size(); }
void or_mg_destroy(void* h) { delete (OrMG*)h; }

// updateFromFine :558-582 (+ Cholesky of the coarsest when requested)
int or_mg_set_skip(void* h, double skip) { ((OrMG*)h)->skip = skip; return 0; }

int or_mg_update(void* h, const float* x, double penal, double E0, double Emin) {
    OrMG& mg = *(OrMG*)h;
    update_level0(mg, x, penal, E0, Emin);
    for (size_t li = 1; li < mg.L.size(); li++) {
        OrLevel& c = mg.L[li];
        if (c.ndof <= mg.galerkinMax) {
            if (!mg.L[li - 1].denseK.empty()) galerkin_from_matrix(mg, mg.L[li - 1], c);
            else galerkin_from_elements(mg, mg.L[li - 1], c);
        } else {
            restrict_eavg(mg, mg.L[li - 1], c, Emin);
        }
    }
    mg.omegaL.assign(mg.L.size(), mg.omega);
    if (mg.omegaCap > 0)
        for (size_t li = 0; li < mg.L.size(); li++) {
            const double rho = power_rho(mg, mg.L[li]);
            if (rho > 0) mg.omegaL[li] = std::min(mg.omega, mg.omegaCap / rho);
        }
    if (mg.coarseSolver == TOPO_COARSE_CHOLESKY) return coarse_cholesky(mg, mg.L.back());
    return 0;
}

// per-level data as the reference's serializer hands it to ebeMGPCG
// (_solveWithWasmMGPCG js/optimizer-worker-3d.js:3690-3772)
int or_mg_level_data(void* h, int li, double* invDiag, double* denseK, int32_t* active, int32_t* nactive) {
    OrMG& mg = *(OrMG*)h;
    OrLevel& l = mg.L[li];
    if (invDiag) std::copy(l.invDiag.begin(), l.invDiag.end(), invDiag);
    if (denseK && !l.denseK.empty()) std::copy(l.denseK.begin(), l.denseK.end(), denseK);
    if (active) std::copy(l.active.begin(), l.active.end(), active);
    if (nactive) *nactive = (int32_t)l.active.size();
    return 0;
}

int or_mg_set_omega_cap(void* h, double cap) { ((OrMG*)h)->omegaCap = cap; return 0; }
double or_mg_level_omega(void* h, int li) { return ((OrMG*)h)->omegaL[li]; }

// one V-cycle z = M^-1 r from z = 0 (MGPrecond3D.apply :1021-1024)
int or_mg_vcycle(void* h, const double* r, double* z) {
    OrMG& mg = *(OrMG*)h;
    for (int64_t i = 0; i < mg.L[0].ndof; i++) z[i] = 0.0;
    vcycle(mg, 0, r, z);
    return 0;
}

int or_mg_level_applyA(void* h, int li, const double* p, double* Ap) {
    OrMG& mg = *(OrMG*)h;
    level_applyA(mg, mg.L[li], p, Ap);
    return 0;
}

int or_mg_level_info(void* h, int li, int32_t* dims4, double* E, int32_t* has_dense) {
    OrMG& mg = *(OrMG*)h;
    OrLevel& l = mg.L[li];
    dims4[0] = l.nelx; dims4[1] = l.nely; dims4[2] = l.nelz; dims4[3] = (int32_t)l.active.size();
    if (E) std::copy(l.E.begin(), l.E.end(), E);
    if (has_dense) *has_dense = !l.denseK.empty();
    return 0;
}

// ebeMGPCG assembly/index.ts:1076-1193 (cold start unless warm, full space)
int or_mgpcg(void* h, const double* F, double* U, int maxIter, double tol, int warm) {
    OrMG& mg = *(OrMG*)h;
    OrLevel& l = mg.L[0];
    const int64_t n = l.ndof;
    const double EPS = 1e-12;
    dvec r(n), z(n), p(n), Ap(n);
    if (!warm) for (int64_t i = 0; i < n; i++) U[i] = 0.0;
    else for (int64_t i = 0; i < n; i++) if (l.fixed[i]) U[i] = 0.0;
    if (warm) {
        level_applyA(mg, l, U, Ap.data());
        for (int64_t i = 0; i < n; i++) r[i] = l.fixed[i] ? 0.0 : F[i] - Ap[i];
    } else {
        for (int64_t i = 0; i < n; i++) r[i] = l.fixed[i] ? 0.0 : F[i];
    }
    or_mg_vcycle(h, r.data(), z.data());
    for (int64_t i = 0; i < n; i++) if (l.fixed[i]) z[i] = 0.0;
    p = z;
    double rz = 0, r0 = 0;
    for (int64_t i = 0; i < n; i++) { rz += r[i] * z[i]; r0 += r[i] * r[i]; }
    const double tolSq = tol * tol * (r0 > 1e-30 ? r0 : 1e-30);
    for (int it = 0; it < maxIter; it++) {
        double rn = 0;
        for (int64_t i = 0; i < n; i++) rn += r[i] * r[i];
        if (rn < tolSq) return it;
        level_applyA(mg, l, p.data(), Ap.data());
        for (int64_t i = 0; i < n; i++) if (l.fixed[i]) Ap[i] = 0.0;
        double pAp = 0;
        for (int64_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
        double alpha = rz / (pAp + EPS);
        for (int64_t i = 0; i < n; i++) { U[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        or_mg_vcycle(h, r.data(), z.data());
        for (int64_t i = 0; i < n; i++) if (l.fixed[i]) z[i] = 0.0;
        double rzn = 0;
        for (int64_t i = 0; i < n; i++) rzn += r[i] * z[i];
        double beta = rzn / (rz + EPS);
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    return maxIter;
}

// JS Math.max(a, b): NaN if either operand is NaN (the change loop
// js/optimizer-worker-3d.js:2244-2247 and js/optimizer-worker.js:704-706)
static double js_math_max(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return std::nan("");
    return a > b ? a : b;
}

// energy clamp (topo_config.energy_clamp, no reference counterpart): a
// negative sum of the positive semidefinite form is rounding
static int g_energy_clamp = 0;
void or_set_energy_clamp(int on) { g_energy_clamp = on; }

// ---------------------------------------------------------------------------
// element energies js/optimizer-worker-3d.js:2163-2173 and
// _computeElementEnergyFlat :1561-1571 (f64 sum, Float32 store)
int or_energy(int nelx, int nely, int nelz, const double* ke, const double* U, float* w) {
    const int64_t nel = (int64_t)nelx * nely * nelz;
    std::vector<int32_t> edof(nel * 24);
    or_edofs(nelx, nely, nelz, edof.data());
    double Ue[24];
    for (int64_t e = 0; e < nel; e++) {
        for (int i = 0; i < 24; i++) Ue[i] = U[edof[e * 24 + i]];
        double en = 0;
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) en += Ue[i] * ke[i * 24 + j] * Ue[j];
        if (g_energy_clamp && en < 0.0) en = 0.0;
        w[e] = (float)en;
    }
    return 0;
}

// _heavisideProject js/optimizer-worker-3d.js:3500-3509
int or_heaviside(const float* x, int64_t n, double beta, double eta, float* xp) {
    const double tb = std::tanh(beta * eta), t1 = std::tanh(beta * (1 - eta)), den = tb + t1;
    for (int64_t i = 0; i < n; i++) xp[i] = (float)((tb + std::tanh(beta * ((double)x[i] - eta))) / den);
    return 0;
}

// filter neighbourhood of prepareFilter js/optimizer-worker-3d.js:3006-3050
// (window +-floor(rmin), weight max(0, rmin - dist), order x, y, z)
static void filter_weights_ref(int nelx, int nely, int nelz, double rmin, int64_t e1i, int ix, int iy, int iz,
                               std::vector<std::pair<int64_t, double>>& out) {
    (void)e1i;
    out.clear();
    const int r = (int)std::floor(rmin);
    for (int k = std::max(ix - r, 0); k <= std::min(ix + r, nelx - 1); k++)
        for (int l = std::max(iy - r, 0); l <= std::min(iy + r, nely - 1); l++)
            for (int m = std::max(iz - r, 0); m <= std::min(iz + r, nelz - 1); m++) {
                double dist = std::sqrt((double)((ix - k) * (ix - k) + (iy - l) * (iy - l) + (iz - m) * (iz - m)));
                if (dist <= rmin) out.push_back({k + (int64_t)l * nelx + (int64_t)m * nelx * nely, std::max(0.0, rmin - dist)});
            }
}

// filterSensitivities :3052-3064 with Hs (Float32, :3043-3047)
int or_sens_filter(int nelx, int nely, int nelz, double rmin, const float* x, const float* dc, float* dcn) {
    const int64_t nel = (int64_t)nelx * nely * nelz;
    std::vector<float> Hs(nel, 0.f);
    for (int64_t e = 0; e < nel; e++) dcn[e] = 0.f;
    std::vector<std::pair<int64_t, double>> nb;
    // H entries are visited in prepareFilter's order: e1 by (i,j,m) x-outer
    for (int i = 0; i < nelx; i++)
        for (int j = 0; j < nely; j++)
            for (int m = 0; m < nelz; m++) {
                int64_t e1 = i + (int64_t)j * nelx + (int64_t)m * nelx * nely;
                filter_weights_ref(nelx, nely, nelz, rmin, e1, i, j, m, nb);
                for (auto& q : nb) {
                    Hs[e1] = (float)((double)Hs[e1] + q.second);
                    dcn[e1] = (float)((double)dcn[e1] + q.second * (double)x[q.first] * (double)dc[q.first]);
                }
            }
    for (int64_t e = 0; e < nel; e++)
        dcn[e] = (float)((double)dcn[e] / ((double)Hs[e] * std::max(1e-3, (double)x[e])));
    return 0;
}

// density filter, applyDensityFilter assembly/index.ts:224-273 (window
// ceil(r), weight r - dist, normalised), Float32 store.
// transpose=1 gives the chain rule dc_x[j] = sum_i w_ij in[i] / Sw_i.
int or_density_filter(int nelx, int nely, int nelz, double radius, const float* in, float* out, int transpose) {
    const int64_t nel = (int64_t)nelx * nely * nelz;
    const int r = (int)std::ceil(radius);
    const double r2 = radius * radius;
    std::vector<double> sw(nel, 0.0), acc(nel, 0.0);
    for (int ez = 0; ez < nelz; ez++)
        for (int ey = 0; ey < nely; ey++)
            for (int ex = 0; ex < nelx; ex++) {
                int64_t o = ex + (int64_t)ey * nelx + (int64_t)ez * nelx * nely;
                double s = 0, w = 0;
                for (int kz = std::max(0, ez - r); kz <= std::min(nelz - 1, ez + r); kz++)
                    for (int ky = std::max(0, ey - r); ky <= std::min(nely - 1, ey + r); ky++)
                        for (int kx = std::max(0, ex - r); kx <= std::min(nelx - 1, ex + r); kx++) {
                            double dx = ex - kx, dy = ey - ky, dz = ez - kz;
                            double d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 <= r2) {
                                double wt = radius - std::sqrt(d2);
                                s += wt * (double)in[kx + (int64_t)ky * nelx + (int64_t)kz * nelx * nely];
                                w += wt;
                            }
                        }
                sw[o] = w;
                acc[o] = s;
            }
    if (!transpose) {
        for (int64_t e = 0; e < nel; e++) out[e] = (float)(acc[e] / sw[e]);
        return 0;
    }
    // transpose: out[j] = sum_i w_ij in[i]/sw_i (w symmetric)
    for (int ez = 0; ez < nelz; ez++)
        for (int ey = 0; ey < nely; ey++)
            for (int ex = 0; ex < nelx; ex++) {
                int64_t o = ex + (int64_t)ey * nelx + (int64_t)ez * nelx * nely;
                double s = 0;
                for (int kz = std::max(0, ez - r); kz <= std::min(nelz - 1, ez + r); kz++)
                    for (int ky = std::max(0, ey - r); ky <= std::min(nely - 1, ey + r); ky++)
                        for (int kx = std::max(0, ex - r); kx <= std::min(nelx - 1, ex + r); kx++) {
                            double dx = ex - kx, dy = ey - ky, dz = ez - kz;
                            double d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 <= r2) {
                                int64_t i = kx + (int64_t)ky * nelx + (int64_t)kz * nelx * nely;
                                s += (radius - std::sqrt(d2)) * (double)in[i] / sw[i];
                            }
                        }
                out[o] = (float)s;
            }
    return 0;
}

// OC js/optimizer-worker-3d.js:3066-3101 (Float32 xnew, f64 volume sum)
int or_oc(int64_t nel, const float* x, double volfrac, const float* dc, const uint8_t* pres, const uint8_t* voidm,
          double move, float* xnew) {
    double l1 = 0, l2 = 1e9;
    int passes = 0;
    auto jsmax = [](double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (a > b ? a : b); };
    auto jsmin = [](double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (a < b ? a : b); };
    while ((l2 - l1) / (l2 + l1) > 1e-3) {
        double lmid = 0.5 * (l2 + l1);
        for (int64_t i = 0; i < nel; i++) {
            if (pres && pres[i]) xnew[i] = 1.0f;
            else if (voidm && voidm[i]) xnew[i] = 0.0f;
            else {
                double Be = -(double)dc[i] / lmid;
                double xi = x[i];
                xnew[i] = (float)jsmax(0.0, jsmax(xi - move, jsmin(1.0, jsmin(xi + move, xi * std::sqrt(Be)))));
            }
        }
        double s = 0;
        for (int64_t i = 0; i < nel; i++) s += xnew[i];
        if (s > volfrac * nel) l1 = lmid; else l2 = lmid;
        passes++;
    }
    return passes;
}

// KSP-CG with the domain-decomposition preconditioner (defined below)
void* or_bddc_create(int nelx, int nely, int nelz);
void or_bddc_destroy(void* h);
static int bddc_fe(void* h, const double* ke, const float* xPhys, double penal, double E0, double Emin, double skip,
                   const uint8_t* fixed, const double* F, double* U, int maxIter, double tol);


// ---------------------------------------------------------------------------
// SIMP loop: optimize() js/optimizer-worker-3d.js:1923-2352 restricted to the
// hot path (no AMR/manufacturing/UI). Runs `niter` iterations (or stops at
// change <= 0.01 like :2074). Outputs per-iteration compliance and the final
// design. Solver per cfg->linear_solver: PCG = ebePCG with warm start
// (:3572-3629); MGPCG = ebeMGPCG cold start (:3642-3807) with the MG options;
// KSP_BDDC = ebeKSP_BDDC cold start (:3814-3999).
// When set, or_simp_run copies the first iteration's physical densities and
// unfiltered sensitivities dc (dJ/dx through the projection, before the
// sensitivity filter) here: or_simp_sens reads them to pin the loop's own
// sensitivity code against finite differences of its compliance.
static float* g_probe_xphys = nullptr;
static float* g_probe_dc = nullptr;

int or_simp_run(const topo_config* cfg, const uint8_t* fixed, const float* F32, const uint8_t* pres,
                const uint8_t* voidm, int niter, float* x_io, double* comp_out, double* change_out,
                int32_t* cg_out, double* U_out) {
    const int nelx = cfg->nelx, nely = cfg->nely, nelz = cfg->nelz;
    const int64_t nel = (int64_t)nelx * nely * nelz;
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    dvec ke(576);
    or_k0(cfg->nu, ke.data());
    std::vector<int32_t> freed;
    for (int64_t i = 0; i < ndof; i++) if (!fixed[i]) freed.push_back((int32_t)i);
    dvec F(ndof);
    for (int64_t i = 0; i < ndof; i++) F[i] = F32[i];
    std::vector<float> x(x_io, x_io + nel), xnew(nel), xold(nel), xPhys(nel), w(nel), dc(nel), dcn(nel), xf(nel);
    dvec U(ndof, 0.0);
    void* mg = nullptr;
    void* bddc = cfg->linear_solver == TOPO_SOLVER_KSP_BDDC ? or_bddc_create(nelx, nely, nelz) : nullptr;
    if (cfg->linear_solver == TOPO_SOLVER_MGPCG)
    {
        mg = or_mg_create(nelx, nely, nelz, fixed, ke.data(), cfg->mg_levels, cfg->mg_omega, cfg->mg_nu1, cfg->mg_nu2,
                          cfg->coarse_solver, cfg->coarse_sweeps, cfg->galerkin_max_ndof);
        or_mg_set_omega_cap(mg, cfg->mg_omega_cap);
        or_mg_set_skip(mg, cfg->skip_factor > 0 ? cfg->Emin * cfg->skip_factor : -1.0);
    }
    or_set_energy_clamp(cfg->energy_clamp);
    const double penalTarget = cfg->penal, penalStart = cfg->penal_start;
    const int contIters = cfg->continuation_iters;
    const bool proj = cfg->use_projection != 0;
    const int maxIt = cfg->max_iterations;
    double change = 1.0;
    int loop = 0;
    std::vector<double> sw;
    while (change > 0.01 && loop < maxIt && loop < niter) {
        loop++;
        xold = x;
        const double penal = penalStart + (penalTarget - penalStart) * std::min(1.0, (double)(loop - 1) / contIters);
        double beta = 1;
        if (proj) beta = cfg->beta_fixed > 0 ? cfg->beta_fixed
                                             : std::min(cfg->beta_max, std::pow(2.0, std::floor((double)(loop - 1) / cfg->beta_interval)));
        // design -> physical density
        if (cfg->filter_type == TOPO_FILTER_DENSITY) {
            or_density_filter(nelx, nely, nelz, cfg->rmin, x.data(), xf.data(), 0);
        } else {
            xf = x;
        }
        if (proj) or_heaviside(xf.data(), nel, beta, cfg->eta, xPhys.data());
        else xPhys = xf;
        // FE (:4155)
        const double progress = maxIt > 1 ? (double)(loop - 1) / (maxIt - 1) : 1.0;
        const double tol = cfg->cg_tol > 0 ? cfg->cg_tol
                                           : std::exp(std::log(1e-3) + progress * (std::log(1e-8) - std::log(1e-3)));
        int its;
        if (cfg->linear_solver == TOPO_SOLVER_PCG) {
            or_set_pcg_skip_factor(cfg->skip_factor);
            dvec dens(xPhys.begin(), xPhys.end());
            if (!(loop > 1 && cfg->warm_start != 0)) std::fill(U.begin(), U.end(), 0.0);
            int maxI = std::min((int)freed.size(), cfg->cg_max_iter);
            its = or_ebePCG(dens.data(), ke.data(), nelx, nely, nelz, F.data(), U.data(), freed.data(), (int)freed.size(),
                            cfg->Emin, cfg->E0, penal, maxI, tol);
        } else if (bddc) {
            int maxI = std::min((int)freed.size(), cfg->cg_max_iter);
            its = bddc_fe(bddc, ke.data(), xPhys.data(), penal, cfg->E0, cfg->Emin, cfg->skip_factor, fixed, F.data(),
                          U.data(), maxI, tol);
        } else {
            or_mg_update(mg, xPhys.data(), penal, cfg->E0, cfg->Emin);
            int maxI = std::min((int)freed.size(), cfg->cg_max_iter);
            its = or_mgpcg(mg, F.data(), U.data(), maxI, tol, cfg->warm_start == 1 && loop > 1);
        }
        double c = 0;
        for (int64_t i = 0; i < ndof; i++) c += F[i] * U[i];
        or_energy(nelx, nely, nelz, ke.data(), U.data(), w.data());
        // sensitivities :2175-2192
        const double denom = 2 * std::tanh(beta * 0.5);
        for (int64_t e = 0; e < nel; e++) {
            double dphys = -penal * std::pow((double)xPhys[e], penal - 1) * cfg->E0 * (double)w[e];
            double d = 1.0;
            if (proj) {
                double th = std::tanh(beta * ((double)xf[e] - 0.5));
                d = beta * (1 - th * th) / denom;
            }
            dc[e] = (float)(dphys * d);
        }
        if (loop == 1 && g_probe_dc) {
            std::copy(dc.begin(), dc.end(), g_probe_dc);
            std::copy(xPhys.begin(), xPhys.end(), g_probe_xphys);
        }
        if (cfg->filter_type == TOPO_FILTER_DENSITY) {
            or_density_filter(nelx, nely, nelz, cfg->rmin, dc.data(), dcn.data(), 1);
        } else {
            or_sens_filter(nelx, nely, nelz, cfg->rmin, x.data(), dc.data(), dcn.data());
        }
        or_oc(nel, x.data(), cfg->volfrac, dcn.data(), pres, voidm, cfg->move, xnew.data());
        change = 0;
        for (int64_t i = 0; i < nel; i++) change = js_math_max(change, (double)std::fabs(xnew[i] - xold[i]));
        std::swap(x, xnew);
        if (comp_out) comp_out[loop - 1] = c;
        if (change_out) change_out[loop - 1] = change;
        if (cg_out) cg_out[loop - 1] = its;
    }
    or_set_pcg_skip_factor(1000.0);
    or_set_energy_clamp(0);
    std::copy(x.begin(), x.end(), x_io);
    if (U_out) std::copy(U.begin(), U.end(), U_out);
    if (mg) or_mg_destroy(mg);
    if (bddc) or_bddc_destroy(bddc);
    return loop;
}

// one iteration of or_simp_run from x: its compliance, physical densities
// and unfiltered sensitivities
int or_simp_sens(const topo_config* cfg, const uint8_t* fixed, const float* F32, const float* x, double* comp,
                 float* xphys_out, float* dc_out) {
    const int64_t nel = (int64_t)cfg->nelx * cfg->nely * cfg->nelz;
    std::vector<float> xi(x, x + nel);
    g_probe_xphys = xphys_out;
    g_probe_dc = dc_out;
    const int n = or_simp_run(cfg, fixed, F32, nullptr, nullptr, 1, xi.data(), comp, nullptr, nullptr, nullptr);
    g_probe_xphys = nullptr;
    g_probe_dc = nullptr;
    return n;
}

// ---------------------------------------------------------------------------
// KSP-CG with the domain-decomposition preconditioner of the 'petsc' solver
// (additive-Schwarz local Jacobi, two global Jacobi sweeps, a Jacobi
// correction on the subdomain-corner DOFs). Partition and local diagonals:
// KSPSolver._buildSubdomains / _updateBDDC js/optimizer-worker-3d.js:1155-1262
// (with the E_vals of _solveWithWasmKSP_BDDC :3824-3845); preconditioner
// bddcApply assembly/index.ts:1272-1371; solve ebeKSP_BDDC :1410-1517.
struct OrBddc {
    int nelx, nely, nelz, ns[3];
    int64_t ndof;
    std::vector<std::vector<int32_t>> elements, dofs;
    std::vector<dvec> linv;          // per subdomain, full ndof (0 off its DOFs)
    std::vector<int32_t> coarse;     // sorted corner DOFs
};

// subdomains per axis (:1162-1166)
static void bddc_counts(int nelx, int nely, int nelz, int ns[3]) {
    const double tot = (double)nelx * nely * nelz;
    const double target = std::max(8.0, std::min(64.0, std::ceil(std::cbrt(tot / 100))));
    auto one = [&](int n, int a, int b) {
        const double v = std::round(std::cbrt(target * ((double)n * n) / ((double)a * b)));
        return (int)std::max(2.0, std::min((double)n, v));
    };
    ns[0] = one(nelx, nely, nelz);
    ns[1] = one(nely, nelx, nelz);
    ns[2] = one(nelz, nelx, nely);
}

void* or_bddc_create(int nelx, int nely, int nelz) {
    OrBddc* b = new OrBddc();
    b->nelx = nelx; b->nely = nely; b->nelz = nelz;
    b->ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    bddc_counts(nelx, nely, nelz, b->ns);
    const int nsx = b->ns[0], nsy = b->ns[1], nsz = b->ns[2];
    const int nny = nely + 1, nnz = nelz + 1;
    std::set<int32_t> cset;
    for (int sz = 0; sz < nsz; sz++) {
        const int z0 = (int)std::floor((double)sz * nelz / nsz), z1 = (int)std::floor((double)(sz + 1) * nelz / nsz);
        for (int sy = 0; sy < nsy; sy++) {
            const int y0 = (int)std::floor((double)sy * nely / nsy), y1 = (int)std::floor((double)(sy + 1) * nely / nsy);
            for (int sx = 0; sx < nsx; sx++) {
                const int x0 = (int)std::floor((double)sx * nelx / nsx), x1 = (int)std::floor((double)(sx + 1) * nelx / nsx);
                std::vector<int32_t> el;
                for (int ez = z0; ez < z1; ez++)
                    for (int ey = y0; ey < y1; ey++)
                        for (int ex = x0; ex < x1; ex++) el.push_back(ex + ey * nelx + ez * nelx * nely);
                if (el.empty()) continue;
                b->elements.push_back(el);
                // corner nodes -> primal DOFs (:1236-1254)
                const int cx[2] = {x0, x1}, cy[2] = {y0, y1}, cz[2] = {z0, z1};
                for (int k = 0; k < 8; k++) {
                    const int X = std::min(cx[k & 1], nelx), Y = std::min(cy[(k >> 1) & 1], nely),
                              Z = std::min(cz[k >> 2], nelz);
                    const int64_t n = node_id(X, Y, Z, nny, nnz);
                    for (int d = 0; d < 3; d++) cset.insert((int32_t)(3 * n + d));
                }
            }
        }
    }
    b->coarse.assign(cset.begin(), cset.end());
    // DOFs touched by each subdomain, ascending (:1240-1248)
    std::vector<int32_t> edof((int64_t)nelx * nely * nelz * 24);
    or_edofs(nelx, nely, nelz, edof.data());
    for (auto& el : b->elements) {
        std::set<int32_t> ds;
        for (int32_t e : el)
            for (int i = 0; i < 24; i++) ds.insert(edof[(int64_t)e * 24 + i]);
        b->dofs.emplace_back(ds.begin(), ds.end());
    }
    b->linv.assign(b->elements.size(), dvec());
    return b;
}
void or_bddc_destroy(void* h) { delete (OrBddc*)h; }

// nsubs, ns[3], ncoarse
int or_bddc_info(void* h, int32_t* out5) {
    OrBddc* b = (OrBddc*)h;
    out5[0] = (int32_t)b->elements.size();
    out5[1] = b->ns[0]; out5[2] = b->ns[1]; out5[3] = b->ns[2];
    out5[4] = (int32_t)b->coarse.size();
    return 0;
}

// local inverse diagonals from the element moduli (_updateBDDC :1256-1278):
// per subdomain, its elements ascending, E >= 1e-30, K0 diagonal; 1/diag
// on free DOFs of the subdomain with diag > 1e-30
int or_bddc_update(void* h, const double* ke, const double* evals, const uint8_t* fixed) {
    OrBddc* b = (OrBddc*)h;
    std::vector<int32_t> edof((int64_t)b->nelx * b->nely * b->nelz * 24);
    or_edofs(b->nelx, b->nely, b->nelz, edof.data());
    for (size_t s = 0; s < b->elements.size(); s++) {
        dvec ld(b->ndof, 0.0);
        for (int32_t e : b->elements[s]) {
            const double E = evals[e];
            if (E < 1e-30) continue;
            for (int i = 0; i < 24; i++) ld[edof[(int64_t)e * 24 + i]] += E * ke[i * 24 + i];
        }
        dvec& li = b->linv[s];
        li.assign(b->ndof, 0.0);
        for (int32_t d : b->dofs[s])
            if (!fixed[d] && ld[d] > 1e-30) li[d] = 1.0 / ld[d];
    }
    return 0;
}

// subdomain si: element and DOF counts, and (when the pointers are given)
// its elements, DOFs and local inverse diagonal (ndof)
int or_bddc_sub(void* h, int si, int32_t* ne_nd, int32_t* elements, int32_t* dofs, double* linv) {
    OrBddc* b = (OrBddc*)h;
    ne_nd[0] = (int32_t)b->elements[si].size();
    ne_nd[1] = (int32_t)b->dofs[si].size();
    if (elements) std::copy(b->elements[si].begin(), b->elements[si].end(), elements);
    if (dofs) std::copy(b->dofs[si].begin(), b->dofs[si].end(), dofs);
    if (linv) std::copy(b->linv[si].begin(), b->linv[si].end(), linv);
    return 0;
}
int or_bddc_coarse(void* h, int32_t* coarse) {
    OrBddc* b = (OrBddc*)h;
    std::copy(b->coarse.begin(), b->coarse.end(), coarse);
    return 0;
}

// bddcApply assembly/index.ts:1272-1371 (z = M^-1 r); applyA over the active
// elements (bddcApplyA :1227-1262 = applyAEbe3D)
static void bddc_apply(const OrBddc* b, const double* ke, const double* evals, const int32_t* active, int nactive,
                       const uint8_t* fixed, const double* invD, const double* r, double* z, int sweeps) {
    const int64_t n = b->ndof;
    dvec w(n, 0.0), Az(n), cc(n, 0.0);
    for (int64_t i = 0; i < n; i++) z[i] = 0.0;
    for (size_t s = 0; s < b->elements.size(); s++)
        for (int32_t d : b->dofs[s]) {
            z[d] += b->linv[s][d] * r[d];
            w[d] += 1.0;
        }
    for (int64_t i = 0; i < n; i++)
        if (w[i] > 1e-12) z[i] /= w[i];
    for (int it = 1; it < sweeps; it++) {
        or_applyA(b->nelx, b->nely, b->nelz, ke, evals, active, nactive, z, Az.data());
        for (int64_t i = 0; i < n; i++) z[i] += invD[i] * (r[i] - Az[i]);
    }
    if (!b->coarse.empty()) {
        or_applyA(b->nelx, b->nely, b->nelz, ke, evals, active, nactive, z, Az.data());
        for (int32_t d : b->coarse) cc[d] = invD[d] * (r[d] - Az[d]);
    }
    for (int64_t i = 0; i < n; i++) z[i] += cc[i];
    for (int64_t i = 0; i < n; i++)
        if (fixed[i]) z[i] = 0.0;
}

int or_bddc_precond(void* h, const double* ke, const double* evals, const int32_t* active, int nactive,
                    const uint8_t* fixed, const double* invD, const double* r, double* z, int sweeps) {
    bddc_apply((OrBddc*)h, ke, evals, active, nactive, fixed, invD, r, z, sweeps);
    return 0;
}

// ebeKSP_BDDC assembly/index.ts:1410-1517: cold start, r = masked F, PCG
// with the +1e-12 guards; returns the iteration count (maxIter if not
// converged)
int or_bddc_solve(void* h, const double* ke, const double* evals, const int32_t* active, int nactive,
                  const uint8_t* fixed, const double* invD, const double* F, double* U, int maxIter, double tol,
                  int sweeps) {
    const OrBddc* b = (const OrBddc*)h;
    const int64_t n = b->ndof;
    const double EPS = 1e-12;
    dvec r(n), z(n), p(n), Ap(n);
    for (int64_t i = 0; i < n; i++) {
        U[i] = 0.0;
        r[i] = fixed[i] ? 0.0 : F[i];
    }
    bddc_apply(b, ke, evals, active, nactive, fixed, invD, r.data(), z.data(), sweeps);
    p = z;
    double rz = 0, r0 = 0;
    for (int64_t i = 0; i < n; i++) {
        rz += r[i] * z[i];
        r0 += r[i] * r[i];
    }
    const double tolSq = tol * tol * (r0 > 1e-30 ? r0 : 1e-30);
    for (int it = 0; it < maxIter; it++) {
        double rn = 0;
        for (int64_t i = 0; i < n; i++) rn += r[i] * r[i];
        if (rn < tolSq) return it;
        or_applyA(b->nelx, b->nely, b->nelz, ke, evals, active, nactive, p.data(), Ap.data());
        for (int64_t i = 0; i < n; i++)
            if (fixed[i]) Ap[i] = 0.0;
        double pAp = 0;
        for (int64_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
        const double alpha = rz / (pAp + EPS);
        for (int64_t i = 0; i < n; i++) {
            U[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        bddc_apply(b, ke, evals, active, nactive, fixed, invD, r.data(), z.data(), sweeps);
        double rzn = 0;
        for (int64_t i = 0; i < n; i++) rzn += r[i] * z[i];
        const double beta = rzn / (rz + EPS);
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    return maxIter;
}

// _solveWithWasmKSP_BDDC js/optimizer-worker-3d.js:3814-3999 around the
// above: E by Math.pow, active above skip * Emin, global inverse diagonal of
// the active elements; 3 smoother sweeps (KSP_BDDC_SMOOTHER_ITERS :1031)
static int bddc_fe(void* h, const double* ke, const float* xPhys, double penal, double E0, double Emin, double skip,
                   const uint8_t* fixed, const double* F, double* U, int maxIter, double tol) {
    OrBddc* b = (OrBddc*)h;
    const int64_t nel = (int64_t)b->nelx * b->nely * b->nelz;
    dvec E(nel);
    std::vector<int32_t> act;
    for (int64_t e = 0; e < nel; e++) {
        E[e] = Emin + std::pow((double)xPhys[e], penal) * (E0 - Emin);
        if (E[e] > skip * Emin) act.push_back((int32_t)e);
    }
    std::vector<int32_t> edof(nel * 24);
    or_edofs(b->nelx, b->nely, b->nelz, edof.data());
    dvec diag(b->ndof, 0.0), invD(b->ndof, 0.0);
    for (int32_t e : act)
        for (int i = 0; i < 24; i++) diag[edof[(int64_t)e * 24 + i]] += E[e] * ke[i * 24 + i];
    for (int64_t i = 0; i < b->ndof; i++)
        if (!fixed[i] && diag[i] > 1e-30) invD[i] = 1.0 / diag[i];
    or_bddc_update(h, ke, E.data(), fixed);
    return or_bddc_solve(h, ke, E.data(), act.data(), (int)act.size(), fixed, invD.data(), F, U, maxIter, tol, 3);
}

int or_bddc_fe(void* h, const double* ke, const float* xPhys, double penal, double E0, double Emin, double skip,
               const uint8_t* fixed, const double* F, double* U, int maxIter, double tol) {
    return bddc_fe(h, ke, xPhys, penal, E0, Emin, skip, fixed, F, U, maxIter, tol);
}

// ---------------------------------------------------------------------------
// Differentiation path (no reference counterpart; pinned by invariants in
// tests/test_oracle.py: the Helmholtz map keeps constants and sum(x) and is
// symmetric; it equals a dense numpy assembly of (r^2 K_L + M); central
// finite differences of J; the adjoint identity L^T u = lam^T f).

// Helmholtz filter: nodes of the grid, (r^2 K_L + M) y = T x, xf = T^T y with
// T the element -> node map of weight 1/8; r = rmin / (2 sqrt 3). Elements
// in the reference order (e = ex + ey nelx + ez nelx nely). CG (Jacobi) to
// 1e-14 relative.
int or_helmholtz(int nelx, int nely, int nelz, double rmin, const double* in, double* out) {
    const int nnx = nelx + 1, nny = nely + 1, nnz = nelz + 1;
    const int64_t nn = (int64_t)nnx * nny * nnz, nel = (int64_t)nelx * nely * nelz;
    const double rp = rmin / (2.0 * std::sqrt(3.0));
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const double m1[2][2] = {{1.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 3}}, k1[2][2] = {{1, -1}, {-1, 1}};
    double A[8][8];
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            const int* a = off[i];
            const int* b = off[j];
            const double mx = m1[a[0]][b[0]], my = m1[a[1]][b[1]], mz = m1[a[2]][b[2]];
            A[i][j] = rp * rp * (k1[a[0]][b[0]] * my * mz + mx * k1[a[1]][b[1]] * mz + mx * my * k1[a[2]][b[2]]) +
                      mx * my * mz;
        }
    auto nodes = [&](int64_t e, int64_t* n8) {
        const int ex = (int)(e % nelx), ey = (int)((e / nelx) % nely), ez = (int)(e / ((int64_t)nelx * nely));
        for (int m = 0; m < 8; m++) n8[m] = node_id(ex + off[m][0], ey + off[m][1], ez + off[m][2], nny, nnz);
    };
    auto apply = [&](const dvec& v, dvec& o) {
        std::fill(o.begin(), o.end(), 0.0);
        int64_t n8[8];
        for (int64_t e = 0; e < nel; e++) {
            nodes(e, n8);
            for (int i = 0; i < 8; i++) {
                double sacc = 0;
                for (int j = 0; j < 8; j++) sacc += A[i][j] * v[n8[j]];
                o[n8[i]] += sacc;
            }
        }
    };
    dvec b(nn, 0.0), d(nn, 0.0), y(nn, 0.0), r(nn), z(nn), pp(nn), Ap(nn);
    int64_t n8[8];
    for (int64_t e = 0; e < nel; e++) {
        nodes(e, n8);
        for (int m = 0; m < 8; m++) { b[n8[m]] += in[e] / 8.0; d[n8[m]] += A[m][m]; }
    }
    r = b;
    double rz = 0, rn0 = 0;
    for (int64_t i = 0; i < nn; i++) { z[i] = r[i] / d[i]; pp[i] = z[i]; rz += r[i] * z[i]; rn0 += r[i] * r[i]; }
    for (int it = 0; it < 10000 && rn0 > 0; it++) {
        apply(pp, Ap);
        double pAp = 0;
        for (int64_t i = 0; i < nn; i++) pAp += pp[i] * Ap[i];
        const double al = rz / pAp;
        double rn = 0, rz2 = 0;
        for (int64_t i = 0; i < nn; i++) { y[i] += al * pp[i]; r[i] -= al * Ap[i]; rn += r[i] * r[i]; }
        if (rn < 1e-28 * rn0) break;
        for (int64_t i = 0; i < nn; i++) { z[i] = r[i] / d[i]; rz2 += r[i] * z[i]; }
        const double be = rz2 / rz;
        rz = rz2;
        for (int64_t i = 0; i < nn; i++) pp[i] = z[i] + be * pp[i];
    }
    for (int64_t e = 0; e < nel; e++) {
        nodes(e, n8);
        double sacc = 0;
        for (int m = 0; m < 8; m++) sacc += y[n8[m]];
        out[e] = sacc / 8.0;
    }
    return 0;
}

// K(E) + springs, full space, fixed DOFs masked: Jacobi-PCG to 1e-14
// relative from zero (test solver of the forward/backward restatement)
static int fb_solve(int nelx, int nely, int nelz, const double* ke, const dvec& E, const uint8_t* fixed,
                    const int64_t* sdof, const double* sk, int ns, const dvec& rhs, dvec& u) {
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1), nel = (int64_t)nelx * nely * nelz;
    std::vector<int32_t> act(nel);
    for (int64_t e = 0; e < nel; e++) act[e] = (int32_t)e;
    auto apply = [&](const dvec& v, dvec& o) {
        or_applyA(nelx, nely, nelz, ke, E.data(), act.data(), (int)nel, v.data(), o.data());
        for (int s = 0; s < ns; s++) o[sdof[s]] += sk[s] * v[sdof[s]];
        for (int64_t i = 0; i < ndof; i++) if (fixed[i]) o[i] = 0;
    };
    dvec diag(ndof, 0.0);
    {   // diagonal: sum_e E_e K0_ii
        std::vector<int32_t> ed(nel * 24);
        or_edofs(nelx, nely, nelz, ed.data());
        for (int64_t e = 0; e < nel; e++)
            for (int i = 0; i < 24; i++) diag[ed[e * 24 + i]] += E[e] * ke[i * 24 + i];
        for (int s = 0; s < ns; s++) diag[sdof[s]] += sk[s];
    }
    u.assign(ndof, 0.0);
    dvec r(rhs), z(ndof), p(ndof), Ap(ndof);
    double rz = 0, rn0 = 0;
    for (int64_t i = 0; i < ndof; i++) {
        if (fixed[i]) r[i] = 0;
        z[i] = fixed[i] ? 0.0 : r[i] / diag[i];
        p[i] = z[i];
        rz += r[i] * z[i];
        rn0 += r[i] * r[i];
    }
    int it = 0;
    for (; it < 200000 && rn0 > 0; it++) {
        apply(p, Ap);
        double pAp = 0;
        for (int64_t i = 0; i < ndof; i++) pAp += p[i] * Ap[i];
        const double al = rz / pAp;
        double rn = 0, rz2 = 0;
        for (int64_t i = 0; i < ndof; i++) { u[i] += al * p[i]; r[i] -= al * Ap[i]; rn += r[i] * r[i]; }
        if (rn < 1e-28 * rn0) break;
        for (int64_t i = 0; i < ndof; i++) { z[i] = fixed[i] ? 0.0 : r[i] / diag[i]; rz2 += r[i] * z[i]; }
        const double be = rz2 / rz;
        rz = rz2;
        for (int64_t i = 0; i < ndof; i++) p[i] = z[i] + be * p[i];
    }
    return it;
}

// Forward/backward of the design map and objective (restates
// topo_solver_forward/backward): x (float, reference element order) ->
// density filter (filter_type DENSITY) or Helmholtz (HELMHOLTZ) -> Heaviside
// (use_projection, beta = beta_fixed or the schedule at iteration it) ->
// E = Emin + xp^p (E0 - Emin) (penal from the continuation at it) -> solve ->
// J = sum F^T U (compliance) or U[out] (mechanism, springs at the ports,
// adjoint K lam = e_out). Gradients dJ/dx, d mean(xp)/dx by the chain rule.
int or_fb(const topo_config* cfg, const uint8_t* fixed, const float* F32, const float* x, int it, double* J,
          double* vol, double* dJdx, double* dVdx, double* U_out, double* lam_out) {
    const int nelx = cfg->nelx, nely = cfg->nely, nelz = cfg->nelz;
    const int64_t nel = (int64_t)nelx * nely * nelz;
    const int64_t ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    const double penal = cfg->penal_start + (cfg->penal - cfg->penal_start) *
                                                std::min(1.0, (double)(it - 1) / std::max(1, cfg->continuation_iters));
    double beta = 1;
    if (cfg->use_projection)
        beta = cfg->beta_fixed > 0 ? cfg->beta_fixed
                                   : std::min(cfg->beta_max, std::pow(2.0, std::floor((double)(it - 1) / cfg->beta_interval)));
    const double eta = cfg->eta;
    dvec xd(x, x + nel), xf(nel), xp(nel);
    // density filter in FP64 (same weights as or_density_filter)
    auto dfilter = [&](const dvec& in, dvec& out, int transpose) {
        const double rmin = cfg->rmin;
        const int R = (int)std::ceil(rmin);
        dvec sw(nel, 0.0);
        for (int64_t e = 0; e < nel; e++) {
            const int ex = (int)(e % nelx), ey = (int)((e / nelx) % nely), ez = (int)(e / ((int64_t)nelx * nely));
            for (int kz = -R; kz <= R; kz++)
                for (int ky = -R; ky <= R; ky++)
                    for (int kx = -R; kx <= R; kx++) {
                        const double d2 = kx * kx + ky * ky + kz * kz;
                        const int jx = ex + kx, jy = ey + ky, jz = ez + kz;
                        if (d2 > rmin * rmin || jx < 0 || jx >= nelx || jy < 0 || jy >= nely || jz < 0 || jz >= nelz) continue;
                        sw[e] += rmin - std::sqrt(d2);
                    }
        }
        for (int64_t e = 0; e < nel; e++) {
            const int ex = (int)(e % nelx), ey = (int)((e / nelx) % nely), ez = (int)(e / ((int64_t)nelx * nely));
            double acc = 0;
            for (int kz = -R; kz <= R; kz++)
                for (int ky = -R; ky <= R; ky++)
                    for (int kx = -R; kx <= R; kx++) {
                        const double d2 = kx * kx + ky * ky + kz * kz;
                        const int jx = ex + kx, jy = ey + ky, jz = ez + kz;
                        if (d2 > rmin * rmin || jx < 0 || jx >= nelx || jy < 0 || jy >= nely || jz < 0 || jz >= nelz) continue;
                        const int64_t j = jx + (int64_t)nelx * (jy + (int64_t)nely * jz);
                        acc += transpose ? (rmin - std::sqrt(d2)) * in[j] / sw[j] : (rmin - std::sqrt(d2)) * in[j];
                    }
            out[e] = transpose ? acc : acc / sw[e];
        }
    };
    if (cfg->filter_type == TOPO_FILTER_DENSITY) dfilter(xd, xf, 0);
    else if (cfg->filter_type == TOPO_FILTER_HELMHOLTZ) or_helmholtz(nelx, nely, nelz, cfg->rmin, xd.data(), xf.data());
    else xf = xd;
    const double tb = std::tanh(beta * eta), den = tb + std::tanh(beta * (1 - eta));
    for (int64_t e = 0; e < nel; e++)
        xp[e] = cfg->use_projection ? (tb + std::tanh(beta * (xf[e] - eta))) / den : xf[e];
    dvec E(nel);
    for (int64_t e = 0; e < nel; e++) E[e] = cfg->Emin + std::pow(xp[e], penal) * (cfg->E0 - cfg->Emin);
    dvec ke(576);
    or_k0(cfg->nu, ke.data());
    const bool mech = cfg->objective == TOPO_OBJ_MECHANISM;
    int64_t sdof[2];
    double sk[2];
    int ns = 0;
    if (mech) {
        if (cfg->mech_k_in > 0) { sdof[ns] = cfg->mech_in_dof; sk[ns] = cfg->mech_k_in; ns++; }
        if (cfg->mech_k_out > 0) { sdof[ns] = cfg->mech_out_dof; sk[ns] = cfg->mech_k_out; ns++; }
    }
    dvec g(nel, 0.0), u, lam;
    double Jv = 0;
    std::vector<int32_t> ed(nel * 24);
    or_edofs(nelx, nely, nelz, ed.data());
    auto accumulate = [&](const dvec& uu, const dvec& ll) {
        for (int64_t e = 0; e < nel; e++) {
            double en = 0;
            for (int i = 0; i < 24; i++) {
                double r = 0;
                for (int j = 0; j < 24; j++) r += ke[i * 24 + j] * uu[ed[e * 24 + j]];
                en += ll[ed[e * 24 + i]] * r;
            }
            g[e] += -en * penal * std::pow(xp[e], penal - 1) * (cfg->E0 - cfg->Emin);
        }
    };
    for (int k = 0; k < cfg->nload; k++) {
        dvec f(ndof);
        for (int64_t i = 0; i < ndof; i++) f[i] = F32[(int64_t)k * ndof + i];
        fb_solve(nelx, nely, nelz, ke.data(), E, fixed, sdof, sk, ns, f, u);
        if (U_out) std::copy(u.begin(), u.end(), U_out + (int64_t)k * ndof);
        if (mech) {
            dvec L(ndof, 0.0);
            L[cfg->mech_out_dof] = 1.0;
            fb_solve(nelx, nely, nelz, ke.data(), E, fixed, sdof, sk, ns, L, lam);
            if (lam_out) std::copy(lam.begin(), lam.end(), lam_out);
            Jv = u[cfg->mech_out_dof];
            accumulate(u, lam);
        } else {
            for (int64_t i = 0; i < ndof; i++) Jv += f[i] * u[i];
            accumulate(u, u);
        }
    }
    dvec gv(nel, 1.0 / (double)nel);
    if (cfg->use_projection)
        for (int64_t e = 0; e < nel; e++) {
            const double th = std::tanh(beta * (xf[e] - eta));
            const double dH = beta * (1 - th * th) / den;
            g[e] *= dH;
            gv[e] *= dH;
        }
    dvec t(nel);
    if (cfg->filter_type == TOPO_FILTER_DENSITY) {
        dfilter(g, t, 1); g = t;
        dfilter(gv, t, 1); gv = t;
    } else if (cfg->filter_type == TOPO_FILTER_HELMHOLTZ) {
        or_helmholtz(nelx, nely, nelz, cfg->rmin, g.data(), t.data()); g = t;
        or_helmholtz(nelx, nely, nelz, cfg->rmin, gv.data(), t.data()); gv = t;
    }
    double vs = 0;
    for (int64_t e = 0; e < nel; e++) vs += xp[e];
    if (J) *J = Jv;
    if (vol) *vol = vs / (double)nel;
    if (dJdx) std::copy(g.begin(), g.end(), dJdx);
    if (dVdx) std::copy(gv.begin(), gv.end(), dVdx);
    return 0;
}

// ---------------------------------------------------------------------------
// 2-D SIMP path (js/optimizer-worker.js): quad4, node n = (nely+1) ix + iy,
// DOFs 2n, 2n+1, element e = ey + ex nely.
//
// Pinning (tests/test_oracle.py):
//   or2_k0        symmetry, the 3 rigid-body null modes, positive diagonal,
//                 equality with a 2x2 Gauss-quadrature plane-stress quad
//                 (an independent derivation in the test)
//   or2_edofs     closed-form index formulas
//   or2_applyA    brute-force assembled dense K, rigid modes, symmetry
//   or_ebePCG_ed  dense solve on small grids (the reference test's meshes),
//                 zero force -> zero U, U^T K U = F^T U
//   or2_fe        dense solve, patch test (uniform tension: u linear),
//                 cantilever compliance convergence and Timoshenko bound,
//                 Poisson contraction (tests/test-fea-benchmarks.js)
//   or2_sens_filter brute-force neighbour sums
//   or2_simp_run  volume constraint each iteration, compliance decrease,
//                 the shared OC / Heaviside restatements pinned above

// lk() js/optimizer-worker.js:1732-1767: the eight coefficients k_0..k_7 and
// the row pattern of the plane-stress bilinear quad, divided by (1 - nu^2)
int or2_k0(double nu, double* KE) {
    const double k[8] = {1.0 / 2 - nu / 6,  1.0 / 8 + nu / 8, -1.0 / 4 - nu / 12, -1.0 / 8 + 3 * nu / 8,
                         -1.0 / 4 + nu / 12, -1.0 / 8 - nu / 8, nu / 6,            1.0 / 8 - 3 * nu / 8};
    static const int P[8][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2}, {2, 7, 0, 5, 6, 3, 4, 1},
                                {3, 6, 5, 0, 7, 2, 1, 4}, {4, 5, 6, 7, 0, 1, 2, 3}, {5, 4, 3, 2, 1, 0, 7, 6},
                                {6, 3, 4, 1, 2, 7, 0, 5}, {7, 2, 1, 4, 3, 6, 5, 0}};
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) KE[i * 8 + j] = k[P[i][j]] * 1 / (1 - nu * nu);
    return 0;
}

// _precomputeEdofs2D :240-268
int or2_edofs(int nelx, int nely, int32_t* ed) {
    for (int ely = 0; ely < nely; ely++)
        for (int elx = 0; elx < nelx; elx++) {
            const int64_t o = ((int64_t)ely + (int64_t)elx * nely) * 8;
            const int32_t n1 = (nely + 1) * elx + ely, n2 = (nely + 1) * (elx + 1) + ely;
            ed[o] = 2 * n1; ed[o + 1] = 2 * n1 + 1; ed[o + 2] = 2 * n2; ed[o + 3] = 2 * n2 + 1;
            ed[o + 4] = 2 * n2 + 2; ed[o + 5] = 2 * n2 + 3; ed[o + 6] = 2 * n1 + 2; ed[o + 7] = 2 * n1 + 3;
        }
    return 0;
}

// _ebeMatVec :287-315 in full space: Ap = sum over active e (ascending) of
// E_e K u_e, scattered row by row
int or2_applyA(int nelx, int nely, const double* ke, const double* evals, const int32_t* active, int cnt,
               const double* p, double* Ap) {
    const int64_t nel = (int64_t)nelx * nely, ndof = 2LL * (nelx + 1) * (nely + 1);
    std::vector<int32_t> ed(nel * 8);
    or2_edofs(nelx, nely, ed.data());
    for (int64_t i = 0; i < ndof; i++) Ap[i] = 0.0;
    for (int a = 0; a < cnt; a++) {
        const int64_t e = active[a];
        const double E = evals[e];
        const int32_t* d = &ed[e * 8];
        for (int i = 0; i < 8; i++) {
            double sum = 0;
            for (int j = 0; j < 8; j++) sum += ke[i * 8 + j] * p[d[j]];
            Ap[d[i]] += E * sum;
        }
    }
    return 0;
}

// ebePCG assembly/index.ts:504-733 over an arbitrary connectivity (edofs,
// edofSize), the way the export runs it: E by pow, active above 1000 Emin,
// Jacobi diagonal, warm start from U at free DOFs, relative tolerance,
// cgMax = min(maxIter, nfree). Returns the iteration count.
int or_ebePCG_ed(const double* dens, const double* ke, const int32_t* edofs, int nel, int edofSize, int ndof,
                 const double* F, double* U, const int32_t* freedofs, int nfree, double Emin, double E0,
                 double penal, int maxIter, double tolerance) {
    const double EPS = 1e-12, dE = E0 - Emin, skip = Emin * 1000.0;
    dvec ev(nel);
    std::vector<int32_t> act;
    for (int e = 0; e < nel; e++) {
        ev[e] = Emin + std::pow(dens[e], penal) * dE;
        if (ev[e] > skip) act.push_back(e);
    }
    auto apply = [&](const dvec& pf, dvec& af) {
        std::fill(af.begin(), af.end(), 0.0);
        for (int32_t e : act) {
            const int32_t* d = &edofs[(int64_t)e * edofSize];
            for (int i = 0; i < edofSize; i++) {
                double s = 0;
                for (int j = 0; j < edofSize; j++) s += ke[i * edofSize + j] * pf[d[j]];
                af[d[i]] += ev[e] * s;
            }
        }
    };
    dvec diag(ndof, 0.0);
    for (int32_t e : act)
        for (int i = 0; i < edofSize; i++) diag[edofs[(int64_t)e * edofSize + i]] += ev[e] * ke[i * edofSize + i];
    dvec invD(nfree), Uf(nfree), r(nfree), z(nfree), pp(nfree), Ap(nfree), pfull(ndof, 0.0), apfull(ndof);
    for (int i = 0; i < nfree; i++) {
        const double d = diag[freedofs[i]];
        invD[i] = d > 1e-30 ? 1.0 / d : 0.0;
        Uf[i] = U[freedofs[i]];
        pfull[freedofs[i]] = Uf[i];
    }
    apply(pfull, apfull);
    for (int i = 0; i < nfree; i++) r[i] = F[freedofs[i]] - apfull[freedofs[i]];
    double rz = 0, r0 = 0;
    for (int i = 0; i < nfree; i++) {
        z[i] = invD[i] * r[i];
        pp[i] = z[i];
        rz += r[i] * z[i];
        r0 += r[i] * r[i];
    }
    const double tolSq = tolerance * tolerance * (r0 > 1e-30 ? r0 : 1e-30);
    const int cgMax = maxIter < nfree ? maxIter : nfree;
    int ret = cgMax;
    for (int it = 0; it < cgMax; it++) {
        double rn = 0;
        for (int i = 0; i < nfree; i++) rn += r[i] * r[i];
        if (rn < tolSq) { ret = it; break; }
        std::fill(pfull.begin(), pfull.end(), 0.0);
        for (int i = 0; i < nfree; i++) pfull[freedofs[i]] = pp[i];
        apply(pfull, apfull);
        for (int i = 0; i < nfree; i++) Ap[i] = apfull[freedofs[i]];
        double pAp = 0;
        for (int i = 0; i < nfree; i++) pAp += pp[i] * Ap[i];
        const double alpha = rz / (pAp + EPS);
        double rzn = 0;
        for (int i = 0; i < nfree; i++) {
            Uf[i] += alpha * pp[i];
            r[i] -= alpha * Ap[i];
            z[i] = invD[i] * r[i];
            rzn += r[i] * z[i];
        }
        const double beta = rzn / (rz + EPS);
        for (int i = 0; i < nfree; i++) pp[i] = z[i] + beta * pp[i];
        rz = rzn;
    }
    for (int i = 0; i < ndof; i++) U[i] = 0.0;
    for (int i = 0; i < nfree; i++) U[freedofs[i]] = Uf[i];
    return ret;
}

// FE :1780-1797 with solveCG :1799-1873: E by pow at xPhys, active above
// skip Emin, Jacobi diagonal on the free DOFs, PCG from U = 0 in the reduced
// space with the absolute test ||r||^2 < tol^2 at the top of each iteration,
// maxIter = min(nfree, maxit); c = sum F_i U_i (Float32 F). Returns the
// iteration count.
int or2_fe(int nelx, int nely, const double* ke, const float* x, double penal, double E0, double Emin,
           double skip, const uint8_t* fixed, const float* F, double tol, int maxit, double* U, double* c) {
    const int64_t nel = (int64_t)nelx * nely, ndof = 2LL * (nelx + 1) * (nely + 1);
    std::vector<int32_t> ed(nel * 8);
    or2_edofs(nelx, nely, ed.data());
    dvec ev(nel);
    std::vector<int32_t> act;
    const double thr = skip > 0 ? Emin * skip : -1.0;
    for (int64_t e = 0; e < nel; e++) {
        ev[e] = Emin + std::pow((double)x[e], penal) * (E0 - Emin);
        if (ev[e] > thr) act.push_back((int32_t)e);
    }
    std::vector<int32_t> fr;
    for (int64_t i = 0; i < ndof; i++) if (!fixed[i]) fr.push_back((int32_t)i);
    const int n = (int)fr.size();
    dvec diag(ndof, 0.0);
    for (int32_t e : act)
        for (int i = 0; i < 8; i++) diag[ed[(int64_t)e * 8 + i]] += ev[e] * ke[i * 8 + i];
    dvec invD(n), Uf(n, 0.0), r(n), z(n), p(n), Ap(n), pf(ndof), af(ndof);
    for (int i = 0; i < n; i++) {
        const double d = diag[fr[i]];
        invD[i] = d > 1e-30 ? 1.0 / d : 0.0;
        r[i] = F[fr[i]];
    }
    double rz = 0;
    for (int i = 0; i < n; i++) {
        z[i] = invD[i] * r[i];
        p[i] = z[i];
        rz += r[i] * z[i];
    }
    const int maxIter = std::min(n, maxit);
    const double tolSq = tol * tol;
    int it = 0;
    for (; it < maxIter; it++) {
        double rn = 0;
        for (int i = 0; i < n; i++) rn += r[i] * r[i];
        if (rn < tolSq) break;
        std::fill(pf.begin(), pf.end(), 0.0);
        for (int i = 0; i < n; i++) pf[fr[i]] = p[i];
        or2_applyA(nelx, nely, ke, ev.data(), act.data(), (int)act.size(), pf.data(), af.data());
        for (int i = 0; i < n; i++) Ap[i] = af[fr[i]];
        double pAp = 0;
        for (int i = 0; i < n; i++) pAp += p[i] * Ap[i];
        const double alpha = rz / (pAp + 1e-12);
        double rzn = 0;
        for (int i = 0; i < n; i++) {
            Uf[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            z[i] = invD[i] * r[i];
            rzn += r[i] * z[i];
        }
        const double beta = rzn / (rz + 1e-12);
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    for (int64_t i = 0; i < ndof; i++) U[i] = 0.0;
    for (int i = 0; i < n; i++) U[fr[i]] = Uf[i];
    double cc = 0;
    for (int64_t i = 0; i < ndof; i++) cc += F[i] * U[i];
    if (c) *c = cc;
    return it;
}

// prepareFilter :1391-1429 as the reference builds it (explicit iH, jH, sH
// lists, e1 x-major) and filterSensitivities :1431-1443 (Float32 dcn and Hs)
int or2_sens_filter(int nelx, int nely, double rmin, const float* x, const float* dc, float* dcn) {
    std::vector<int64_t> iH, jH;
    dvec sH;
    const int r = (int)std::floor(rmin);
    for (int i = 0; i < nelx; i++)
        for (int j = 0; j < nely; j++) {
            const int64_t e1 = (int64_t)i * nely + j;
            for (int k = std::max(i - r, 0); k <= std::min(i + r, nelx - 1); k++)
                for (int l = std::max(j - r, 0); l <= std::min(j + r, nely - 1); l++) {
                    const double dist = std::sqrt((double)((i - k) * (i - k) + (j - l) * (j - l)));
                    if (dist <= rmin) {
                        iH.push_back(e1);
                        jH.push_back((int64_t)k * nely + l);
                        sH.push_back(std::max(0.0, rmin - dist));
                    }
                }
        }
    const int64_t nel = (int64_t)nelx * nely;
    std::vector<float> Hs(nel, 0.f);
    for (size_t q = 0; q < iH.size(); q++) Hs[iH[q]] = (float)((double)Hs[iH[q]] + sH[q]);
    for (int64_t e = 0; e < nel; e++) dcn[e] = 0.f;
    for (size_t q = 0; q < iH.size(); q++)
        dcn[iH[q]] = (float)((double)dcn[iH[q]] + sH[q] * (double)x[jH[q]] * (double)dc[jH[q]]);
    for (int64_t e = 0; e < nel; e++) dcn[e] = (float)((double)dcn[e] / ((double)Hs[e] * std::max(1e-3, (double)x[e])));
    return 0;
}

// The loop of optimize() :583-700 without AMR, void filling and
// manufacturing constraints: niter iterations from x_io (or fewer, stopping
// at change <= 0.01 like :583). Per iteration: compliance, change, CG
// iterations and the mesh builder's max stress (:1166-1173, with the target
// penalty and the new design). Returns the iterations run.
int or2_simp_run(const topo_config* cfg, const uint8_t* fixed, const float* F, const uint8_t* pres,
                 const uint8_t* voidm, int niter, float* x_io, double* comp_out, double* change_out,
                 int32_t* cg_out, double* smax_out) {
    const int nelx = cfg->nelx, nely = cfg->nely;
    const int64_t nel = (int64_t)nelx * nely, ndof = 2LL * (nelx + 1) * (nely + 1);
    double ke[64];
    or2_k0(cfg->nu, ke);
    std::vector<int32_t> ed(nel * 8);
    or2_edofs(nelx, nely, ed.data());
    std::vector<float> x(x_io, x_io + nel), xold(nel), xp(nel), dc(nel), dcn(nel), xnew(nel), w(nel);
    dvec U(ndof);
    double change = 1.0;
    int loop = 0;
    const double pt = cfg->penal, ps = cfg->penal_start;
    while (change > 0.01 && loop < niter) {
        loop++;
        xold = x;
        const double tol = cfg->cg_tol > 0 ? cfg->cg_tol : std::max(1e-8, change * 1e-2);
        const double penal = ps + (pt - ps) * std::min(1.0, (double)(loop - 1) / cfg->continuation_iters);
        const int proj = cfg->use_projection != 0;
        const double beta = proj ? std::min(cfg->beta_max, std::pow(2.0, std::floor((double)(loop - 1) / cfg->beta_interval))) : 1.0;
        if (proj) or_heaviside(x.data(), nel, beta, 0.5, xp.data());
        else xp = x;
        double c = 0;
        const int it = or2_fe(nelx, nely, ke, xp.data(), penal, cfg->E0, cfg->Emin, cfg->skip_factor, fixed, F, tol,
                              cfg->cg_max_iter, U.data(), &c);
        const double denom = 2 * std::tanh(beta * 0.5);
        for (int64_t e = 0; e < nel; e++) {
            double Ue[8];
            for (int i = 0; i < 8; i++) Ue[i] = U[ed[e * 8 + i]];
            double en = 0;
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) en += Ue[i] * ke[i * 8 + j] * Ue[j];
            w[e] = (float)en;
            const double dphys = -penal * std::pow((double)xp[e], penal - 1) * cfg->E0 * en;
            const double th = std::tanh(beta * ((double)x[e] - 0.5));
            const double dd = proj ? beta * (1 - th * th) / denom : 1.0;
            dc[e] = (float)(dphys * dd);
        }
        or2_sens_filter(nelx, nely, cfg->rmin, x.data(), dc.data(), dcn.data());
        or_oc(nel, x.data(), cfg->volfrac, dcn.data(), pres, voidm, cfg->move, xnew.data());
        change = 0;
        for (int64_t e = 0; e < nel; e++) change = js_math_max(change, (double)std::fabs((double)xnew[e] - (double)xold[e]));
        x = xnew;
        double smax = 0;
        for (int64_t e = 0; e < nel; e++) {
            const float s = (float)((cfg->Emin + std::pow((double)x[e], cfg->penal) * (cfg->E0 - cfg->Emin)) * (double)w[e]);
            if (s > smax) smax = s;
        }
        if (comp_out) comp_out[loop - 1] = c;
        if (change_out) change_out[loop - 1] = change;
        if (cg_out) cg_out[loop - 1] = it;
        if (smax_out) smax_out[loop - 1] = smax;
    }
    std::copy(x.begin(), x.end(), x_io);
    return loop;
}

}  // extern "C"
