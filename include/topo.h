/*
 * topo.h -- C ABI of libtopo_hip.so, the MI355X (gfx950) implementation of the
 * 3-D SIMP inner loop of Jozo132/TopologyOptimization.
 *
 * Conventions shared by every entry point
 *   - plain C types only; every pointer is a HOST pointer unless the name says
 *     otherwise; arrays are caller-owned;
 *   - return value: >= 0 success (some entry points return an iteration
 *     count, exactly like the WASM export they replace), < 0 an error code
 *     (TOPO_E_*); topo_last_error() gives the message of the calling thread;
 *   - numbering is the reference's:  element e = ex + ey*nelx + ez*nelx*nely,
 *     node n = ix*(nely+1)*(nelz+1) + iy*(nelz+1) + iz, DOF = 3n + {0,1,2},
 *     element-local node order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),
 *     (1,0,1),(1,1,1),(0,1,1)  (js/optimizer-worker-3d.js:1448-1494);
 *   - every call is synchronous: results are in the caller's arrays when it
 *     returns (the semantics of the WASM exports it replaces).
 * Nothing here falls back to the CPU: with no usable GPU the compute entry
 * points return TOPO_E_NODEVICE.
 */
#ifndef TOPO_H
#define TOPO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOPO_ABI_VERSION 1

/* ---- error codes ---------------------------------------------------- */
#define TOPO_OK              0
#define TOPO_E_ARG          -1   /* bad argument / size                      */
#define TOPO_E_HIP          -2   /* HIP runtime or launch failure            */
#define TOPO_E_OOM          -3   /* device allocation failed                 */
#define TOPO_E_BREAKDOWN    -4   /* non-finite CG scalar (pAp, rz)           */
#define TOPO_E_UNSTRUCTURED -5   /* edofs is not a structured hex8/quad4 grid */
#define TOPO_E_NODEVICE     -6   /* no HIP device                            */
#define TOPO_E_STATE        -7   /* call out of order                        */
#define TOPO_E_UNSUPPORTED  -8   /* valid request this build does not do     */
#define TOPO_E_CANCELLED    -9   /* cancel flag observed                     */

/* ---- enums ---------------------------------------------------------- */
/* constraint positions of getFixedDOFs (js/optimizer-worker-3d.js:4436-4493) */
enum { TOPO_FIX_LEFT = 0, TOPO_FIX_RIGHT = 1, TOPO_FIX_BOTTOM = 2,
       TOPO_FIX_TOP = 3, TOPO_FIX_BOTTOM_CORNERS = 4, TOPO_FIX_NONE = 5 };
/* constraintDOFs string 'all' / contains 'x','y','z' (:4429-4434) */
enum { TOPO_DOF_X = 1, TOPO_DOF_Y = 2, TOPO_DOF_Z = 4, TOPO_DOF_ALL = 7 };
/* force directions of getLoadVector (js/optimizer-worker-3d.js:4495-4547) */
enum { TOPO_LOAD_DOWN = 0, TOPO_LOAD_UP = 1, TOPO_LOAD_LEFT = 2,
       TOPO_LOAD_RIGHT = 3, TOPO_LOAD_TOP_CENTER = 4, TOPO_LOAD_NONE = 5 };

enum { TOPO_SOLVER_PCG = 0,      /* Jacobi-PCG, ebePCG semantics              */
       TOPO_SOLVER_MGPCG = 1,    /* geometric MG-PCG, ebeMGPCG semantics      */
       TOPO_SOLVER_KSP_BDDC = 2 };/* 'petsc': PCG with the domain-decomposition
                                    preconditioner, ebeKSP_BDDC semantics (FP64) */
enum { TOPO_PREC_FP64 = 0,       /* everything in FP64                        */
       TOPO_PREC_MIXED = 1,      /* FP32 operator/V-cycle inside FP64 refinement */
       TOPO_PREC_MIXED_PC = 2 }; /* FP64 MG-PCG, FP32 V-cycle preconditioner (MG-PCG only;
                                    Jacobi-PCG runs in FP64)                   */
enum { TOPO_COARSE_SWEEPS = 0,   /* reference: 30 extra damped-Jacobi sweeps  */
       TOPO_COARSE_CHOLESKY = 1};/* FP64 dense Cholesky of the coarsest level */
enum { TOPO_FILTER_SENSITIVITY = 0,  /* reference filterSensitivities       */
       TOPO_FILTER_DENSITY = 1,      /* linear density filter (weights rmin-d) */
       TOPO_FILTER_HELMHOLTZ = 2 };  /* PDE filter -r^2 lap(xf) + xf = x on the
                                        nodes, r = rmin / (2 sqrt 3)         */
enum { TOPO_OBJ_COMPLIANCE = 0, TOPO_OBJ_MECHANISM = 1 };

/* ---- solver configuration (keys of the reference `config` object) ---- */
typedef struct topo_config {
    int32_t nelx, nely, nelz;
    int32_t nload;              /* load cases (columns of F); 1 in the reference      */
    double  E0, Emin, nu;       /* youngsModulus, Emin=1e-9, poissonsRatio            */
    double  volfrac;            /* volumeFraction                                      */
    double  penal;              /* penaltyFactor (target)                              */
    double  penal_start;        /* penalStart (reference default min(penal,1.5))       */
    int32_t continuation_iters; /* continuationIters (default max(20,floor(maxIt/3)))  */
    int32_t max_iterations;     /* maxIterations (SIMP)                                */
    double  rmin;               /* filterRadius                                        */
    int32_t filter_type;        /* TOPO_FILTER_*                                       */
    int32_t use_projection;     /* useProjection (default 1)                           */
    double  beta_max;           /* betaMax (64)                                        */
    int32_t beta_interval;      /* betaInterval (5)                                    */
    int32_t beta_fixed;         /* >0: constant beta (no continuation)                 */
    double  eta;                /* projection threshold (0.5)                          */
    int32_t linear_solver;      /* TOPO_SOLVER_*                                       */
    int32_t precision;          /* TOPO_PREC_*                                         */
    int32_t mg_levels;          /* 0: reference rule (floor(n/2), <=6 levels, dims>=2) */
    int32_t mg_nu1, mg_nu2;     /* V(nu1,nu2) (2,2)                                    */
    double  mg_omega;           /* damped-Jacobi weight (0.5)                          */
    int32_t coarse_solver;      /* TOPO_COARSE_*                                       */
    int32_t coarse_sweeps;      /* 30                                                  */
    int32_t galerkin_max_ndof;  /* dense Galerkin below this ndof (3000); 0 = never    */
    int32_t cg_max_iter;        /* 400 (capped by nfree)                               */
    double  cg_tol;             /* > 0 fixed relative tolerance; <= 0 reference schedule 1e-3 -> 1e-8 */
    int32_t warm_start;         /* -1 solver default (PCG warm, MGPCG cold), 0 cold, 1 warm */
    double  move;               /* OC move limit (0.2)                                 */
    int32_t objective;          /* TOPO_OBJ_* (mechanism: nload 1, F = input force)    */
    int32_t mech_in_dof;        /* mechanism: input DOF (load applied there)           */
    int32_t mech_out_dof;       /* mechanism: output DOF                               */
    double  mech_k_in;          /* input spring stiffness                              */
    double  mech_k_out;         /* output spring stiffness                             */
    int32_t rhs_tile;           /* load cases per batched MG-PCG (0 autotune, 1, 2, 4) */
    int32_t use_graphs;         /* capture MG-PCG iterations in hipGraphs (1 default)  */
    double  skip_factor;        /* elements with E <= skip_factor*Emin are left out of the
                                   operator (reference: 1000); 0 keeps every element so K
                                   stays SPD after voids disconnect parts of the design */
    int32_t exact_apply;        /* 1: operator applies in the reference's summation order
                                   (FP64 results bit-identical to applyAEbe3D), 0: fast
                                   Hadamard x-march kernel (default)                   */
    double  mg_omega_cap;       /* > 0: per level omega_eff = min(mg_omega, cap / rho),
                                   rho ~ lambda_max(D^-1 A) by 20 power iterations from
                                   a hashed vector on free DOFs; 0 = off (reference)  */
    int32_t coarse_fp32;        /* rounded copy of the coarsest level's FP64 factor
                                   inverse read by FP32 V-cycles: 0 none (FP64),
                                   1 FP32 (half the bytes), 2 FP16 with one scale per
                                   row (a quarter; the coarse operator stays SPD)  */
    int32_t energy_clamp;       /* 0 (reference): element energies u_e^T K0 u_e as
                                   summed, so a rounding-negative value gives dc > 0 and
                                   a NaN design entry in OC (sqrt of a negative), which
                                   makes `change` NaN and ends the loop; 1: energies
                                   below 0 are set to 0 (K0 is positive semidefinite, so
                                   a negative sum is rounding)                        */
    int32_t mfma_apply;         /* FP32 operator (V-cycle smoothers, FP32 CG) on the x-march
                                   levels: 0 (default) the FP32 VALU kernel (Hadamard
                                   form); 1 the element product on the matrix cores,
                                   operands split into two bfloat16 parts (about 16
                                   significant bits) with FP32 accumulation (slower at
                                   192x96x96: 28 against 21 us, bound by its LDS traffic) */
    int32_t reserved[5];
} topo_config;

/* Fill *cfg with the reference defaults (js/optimizer-worker-3d.js:20-64,
 * :2056-2072, js/main.js:61-110) for an nelx x nely x nelz grid. */
int topo_config_default(topo_config* cfg, int32_t nelx, int32_t nely, int32_t nelz);

/* ---- library / device ------------------------------------------------ */
const char* topo_version(void);
const char* topo_last_error(void);
/* number of HIP devices and the name of device 0 (no compute) */
int topo_device_info(int32_t* count, char* name, int32_t name_len);

/* ---- host-side problem setup (bit-exact reference numbering) ---------
 * These run on the host, need no GPU, and are what the reference's worker
 * computes before the solve. */

/* replaces lk() js/optimizer-worker-3d.js:3404-3491 + _flattenKE :1440-1446:
 * 24x24 row-major unit-cube hex8 stiffness, E=1, 2x2x2 Gauss. */
int topo_k0(double nu, double* ke576);

/* replaces _precomputeEdofs3D js/optimizer-worker-3d.js:1448-1494 */
int topo_edofs(int32_t nelx, int32_t nely, int32_t nelz, int32_t* edofs /* nel*24 */);

/* replaces getFixedDOFs js/optimizer-worker-3d.js:4436-4493 (+ mask build
 * :2001-2008). Sets fixedMask[dof]=1 (does not clear). Returns #DOFs listed. */
int topo_fixed_dofs(int32_t nelx, int32_t nely, int32_t nelz, int32_t position,
                    int32_t dof_mask, uint8_t* fixedMask /* ndof */);

/* replaces getLoadVector js/optimizer-worker-3d.js:4495-4547. forceVector may
 * be NULL (otherwise 3 doubles: custom direction at the top-right corner).
 * Writes F (float32, zeroed first). */
int topo_load_vector(int32_t nelx, int32_t nely, int32_t nelz, int32_t direction,
                     double magnitude, const double* forceVector, float* F /* ndof */);

/* replaces getConstraintElements (:4552-4592) and getForceElements
 * (:4597-4619): sets preserved[e]=1 for the elements pinned to x=1. */
int topo_preserved_elements(int32_t nelx, int32_t nely, int32_t nelz,
                            int32_t position, int32_t direction, uint8_t* preserved);

/* replaces getLoadVectorFromPaint (:4673-4745): faces is n x 4 int32
 * (vx,vy,vz,faceIndex; faceIndex<0 = all 8 nodes); dir is the unit force
 * direction; magnitude spread evenly over the unique nodes. F accumulated. */
int topo_load_paint(int32_t nelx, int32_t nely, int32_t nelz, const int32_t* faces,
                    int32_t nfaces, const double* dir3, double magnitude, float* F);

/* replaces getFixedDOFsFromPaint: same face encoding, marks fixedMask. */
int topo_fixed_paint(int32_t nelx, int32_t nely, int32_t nelz, const int32_t* faces,
                     int32_t nfaces, int32_t dof_mask, uint8_t* fixedMask);

/* ---- WASM-op twins (drop-in for assembly/index.ts exports) -----------
 * Same argument order and meaning as the AssemblyScript exports, with host
 * pointers in place of linear-memory offsets. The connectivity must be a
 * structured hex8 grid in the reference numbering (checked in full; any
 * other edofs -> TOPO_E_UNSTRUCTURED): the device computes it instead of
 * reading it. Each call copies its inputs to the device and its outputs
 * back; the device-resident path is topo_solver_* below. */

/* replaces applyAEbe3D assembly/index.ts:403-439 */
int topo_applyAEbe3D(const double* ke, const int32_t* edofs, const double* evals,
                     const int32_t* active, int32_t activeCount, const double* p,
                     double* Ap, int32_t ndof, double* scratch);

/* replaces ebePCG assembly/index.ts:504-733. Returns the CG iteration count
 * (cgMax if not converged), like the export; U is the warm start on input.
 * edofSize 24: the 3-D hex8 grid above; edofSize 8: the 2-D quad4 grid of
 * _precomputeEdofs2D (js/optimizer-worker.js:240-268, the use in
 * tests/test-wasm-ebePCG.js), checked in full like the 3-D one. `work` is
 * ignored (device workspace is owned by the library). */
int topo_ebePCG(const double* dens, const double* ke, const int32_t* edofs,
                const double* F, double* U, const int32_t* freedofs, int32_t nel,
                int32_t edofSize, int32_t ndof, int32_t nfree, double Emin, double E0,
                double penal, int32_t maxIter, double tolerance, double* work);

/* No counterpart in the reference: topo_ebePCG (edofSize 24) for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. Same arguments (less the
 * ignored `work`), semantics, errors and results as topo_ebePCG with
 * edofSize 24; every device buffer it needs (input copies, CG vectors,
 * reduction partials, CG state) lives in one file-scope device array of the
 * library (about 20 MB, held on each GPU it runs on until the process ends),
 * so it calls no hipMalloc / hipFree and no allocating library routine.
 * Limits: ndof <= 262144 and nel <= 131072 (TOPO_E_UNSUPPORTED beyond).
 * Runs on the null stream of the current device and returns after the
 * result is copied back. Calls are serialised within a process; calls from
 * several processes on one GPU each use their own copy of the array. */
int topo_ebePCG_noalloc(const double* dens, const double* ke, const int32_t* edofs,
                        const double* F, double* U, const int32_t* freedofs, int32_t nel,
                        int32_t edofSize, int32_t ndof, int32_t nfree, double Emin, double E0,
                        double penal, int32_t maxIter, double tolerance);

/* Level descriptor of ebeMGPCG (assembly/index.ts:743-757): the 20 x i32
 * record with its offsets widened to host pointers. */
typedef struct topo_mg_level {
    int32_t nelx, nely, nelz, ndof, nel, nfree, activeCount, hasDenseK;
    const int32_t* edofs;      const double* evals;   const int32_t* active;
    const uint8_t* fixedMask;  const int32_t* freeDofs; const double* invDiag;
    const double* denseK;      double* Au; double* res; double* b; double* x;
    double* scratch;
} topo_mg_level;

/* replaces ebeMGPCG assembly/index.ts:1076-1193 (cold start, V(2,2),
 * omega 0.5, 30 coarsest sweeps, dense levels where hasDenseK). Returns the
 * CG iteration count. */
int topo_ebeMGPCG(const topo_mg_level* levels, int32_t numLevels, const double* ke,
                  const double* F, double* U, const uint8_t* fixedMask0, int32_t ndof0,
                  int32_t maxIter, double tolerance, double* cgWork);

/* Subdomain descriptor of ebeKSP_BDDC (BDDC_SUB_STRIDE record,
 * assembly/index.ts:1204-1222): the 32-byte record with its offsets widened
 * to host pointers. localInvDiag has ndof entries. */
typedef struct topo_bddc_sub {
    int32_t numElements, numDofs;
    const int32_t* elements;   const int32_t* dofs;   const double* localInvDiag;
} topo_bddc_sub;

/* replaces ebeKSP_BDDC assembly/index.ts:1410-1517 (the 'petsc' solver's
 * PCG with the domain-decomposition preconditioner bddcApply :1272-1371:
 * averaged subdomain-local Jacobi, smootherIters - 1 global Jacobi sweeps,
 * Jacobi correction on the coarse (corner) DOFs); cold start. Returns the CG
 * iteration count (maxIter if not converged). A DOF may belong to at most 8
 * subdomains. `work` is ignored. */
int topo_ebeKSP_BDDC(const topo_bddc_sub* subs, int32_t numSubs, const int32_t* coarse,
                     int32_t numCoarse, const double* ke, const int32_t* edofs, const double* evals,
                     const int32_t* active, int32_t activeCount, const uint8_t* fixedMask,
                     const double* invDiag, const double* F, double* U, int32_t ndof,
                     int32_t maxIter, double tolerance, int32_t smootherIters, double* work);

/* ---- device-resident SIMP solver (the fast path) --------------------- */
typedef struct topo_solver topo_solver;

typedef struct topo_step_info {
    int32_t iteration;      /* loop counter after the step (1-based)          */
    int32_t cg_iters;       /* Krylov iterations of this step's solve(s)      */
    int32_t outer_iters;    /* FP64 refinement sweeps (mixed precision)       */
    int32_t applyA_fine;    /* fine-level operator applications               */
    int32_t vcycles;        /* V-cycles                                       */
    int32_t oc_passes;      /* OC bisection passes                            */
    double  compliance;     /* F^T U (summed over load cases); mechanism: u_out */
    double  change;         /* max |x_new - x_old|                            */
    double  penal, beta, tolerance, volume;
    double  max_stress;     /* max_e E_e * w_e                                */
    double  residual;       /* final relative residual of the solve           */
    double  t_solve_ms, t_total_ms;
    double  t_setup_ms;     /* operators of all levels from the densities   */
    double  t_post_ms;      /* energies, sensitivities, filter, OC, change  */
} topo_step_info;

/* fixedMask u8[ndof]; F float32[nload*ndof] (case-major); preserved/voidm
 * u8[nel] or NULL. x starts at volfrac (preserved -> 1, void -> 0) as in the
 * reference (js/optimizer-worker-3d.js:1928-2022). */
int topo_solver_create(const topo_config* cfg, const uint8_t* fixedMask, const float* F,
                       const uint8_t* preserved, const uint8_t* voidm, topo_solver** out);
int topo_solver_destroy(topo_solver* s);
/* design variables x (float32[nel], reference element order) */
int topo_solver_set_design(topo_solver* s, const float* x);
int topo_solver_get_design(topo_solver* s, float* x);
/* change the SIMP iteration counter (what loop the next step is) */
int topo_solver_set_iteration(topo_solver* s, int32_t loop);
/* live config patch (reference updateConfig keys, :2084-2112) */
int topo_solver_update(topo_solver* s, double penal, double rmin, double volfrac,
                       int32_t max_iterations);

/* one SIMP iteration: projection, FE solve, energies, sensitivities, filter,
 * OC, change (js/optimizer-worker-3d.js:2124-2252). */
int topo_solver_step(topo_solver* s, topo_step_info* info);

/* fields of the last step: */
enum { TOPO_FIELD_U = 0,        /* double[nload*ndof]                    */
       TOPO_FIELD_XPHYS = 1,    /* float[nel]                            */
       TOPO_FIELD_ENERGY = 2,   /* float[nel]  w_e = u_e^T k0 u_e        */
       TOPO_FIELD_DC = 3,       /* float[nel]                            */
       TOPO_FIELD_DCN = 4,      /* float[nel]  filtered                  */
       TOPO_FIELD_STRESS = 5,   /* float[nel]  E_e * w_e                 */
       TOPO_FIELD_ADJOINT = 6 };/* double[ndof] adjoint (mechanism)      */
int topo_solver_get_field(topo_solver* s, int32_t field, void* out);

/* ---- batching and tuning (no reference counterpart) ------------------
 * With nload > 1, MG-PCG and precision FP64 or MIXED_PC, the load cases are
 * solved by one batched MG-PCG per group of rhs_tile cases (one launch per
 * operation carries all of them, the coarsest-level inverse is read once per
 * group, each case keeps its own CG scalars and stops on its own). rhs_tile
 * 0 autotunes at the first solve: the x-march element-tile shape of the two
 * finest levels (x-planes per workgroup CX in {auto, 2, 4, 6, 8, 10, 12, 16,
 * 18, 20, 24, 32}; results are bit-identical for every CX), in the mixed
 * precisions also that of the FP64 CG operator on the finest level, and, with
 * several load cases, the group size R in {1, 2, 4} with the least time per
 * right-hand side for the fine operator plus one V-cycle. This reports the
 * choice (rhs_tile 1 when no batching applies, NT/CX of the finest level)
 * and the timings: log receives max_entries x {what, NT, CX, ms}, what = the
 * MG level of a tile timing (ms per launch), 1000 for a timing of the FP64
 * CG operator (ms per launch) or -R for a batch timing (ms per right-hand
 * side). Returns the number of tuning entries. */
int topo_solver_tuning(topo_solver* s, int32_t* rhs_tile, int32_t* xm_nt, int32_t* xm_cx, double* log,
                       int32_t max_entries);

/* ---- differentiation: forward/backward pair --------------------------
 * forward: at the current design x (set_design) and SIMP iteration counter,
 * maps x -> filter (density or Helmholtz) -> Heaviside projection (if on) ->
 * E = Emin + xPhys^p (E0 - Emin) -> FE solve(s) -> objective, all in FP64
 * (the solve in the configured precision). objective = sum_k F_k^T U_k
 * (compliance) or u_out = U[mech_out_dof] (mechanism, TOPO_OBJ_MECHANISM: the
 * input spring mech_k_in at mech_in_dof and the output spring mech_k_out at
 * mech_out_dof are part of K; the adjoint K lam = e_out is solved here too).
 * volume = mean(xPhys). Replaces, for compliance, the reference's FE +
 * compliance of js/optimizer-worker-3d.js:2124-2162 (no reference counterpart
 * for the mechanism or the Helmholtz filter).
 * backward: the hand-written adjoint of the last forward: dJdx and dVdx
 * (double[nel], reference element order; dVdx may be NULL). For compliance
 * with no filter it equals the reference's analytic sensitivity (:2175-2192)
 * up to its use of E0 in place of E0 - Emin. */
int topo_solver_forward(topo_solver* s, double* objective, double* volume, topo_step_info* info);
int topo_solver_backward(topo_solver* s, double* dJdx, double* dVdx);
/* Helmholtz filter of an element field (double[nel], reference order) for a
 * solver created with TOPO_FILTER_HELMHOLTZ; iters receives the CG count.
 * The map is symmetric, so it is also its own adjoint. */
int topo_op_helmholtz(topo_solver* s, const double* in, double* out, int32_t* iters);

/* FE() (js/optimizer-worker-3d.js:4155-4371): solve K(xPhys) U = F for the
 * given physical densities and penalty. U (double[nload*ndof]) is the warm
 * start when warm starting applies and receives the solution; c receives
 * F^T U. iteration/max_iterations drive the tolerance schedule. */
int topo_solver_fe(topo_solver* s, const float* xPhys, double penal, int32_t iteration,
                   int32_t max_iterations, double* U, double* c, topo_step_info* info);

/* ---- operator-level entry points (parity tests, profiling) ----------- */
/* Ap = K(E) p on MG level `level` with element moduli from the last
 * set_density/fe call, in precision prec (TOPO_PREC_FP64 or _MIXED=FP32). */
int topo_solver_set_density(topo_solver* s, const float* xPhys, double penal);
int topo_op_applyA(topo_solver* s, int32_t level, int32_t prec, const double* p, double* Ap);
/* z = M^-1 r, one V-cycle from z = 0 */
int topo_op_vcycle(topo_solver* s, int32_t prec, const double* r, double* z);
/* energies w_e (float[nel]) of displacement U (double[ndof]) */
int topo_op_energy(topo_solver* s, const double* U, float* w);
/* sensitivity / density filter forward: out = H(in) (float[nel]); for the
 * sensitivity filter x is the design (the max(1e-3,x) weighting), else NULL */
int topo_op_filter(topo_solver* s, int32_t filter_type, const float* x, const float* in,
                   float* out);
/* filter adjoint (transpose): out = H^T(in) */
int topo_op_filter_adjoint(topo_solver* s, int32_t filter_type, const float* in, float* out);
/* OC update (js/optimizer-worker-3d.js:3066-3101); returns passes */
int topo_op_oc(topo_solver* s, const float* x, const float* dcn, float* xnew,
               double volfrac);

/* ---- timing (HIP events on the stream the kernels run on) ------------ */
enum { TOPO_TIME_APPLYA = 0, TOPO_TIME_VCYCLE = 1, TOPO_TIME_ENERGY = 2,
       TOPO_TIME_FILTER = 3, TOPO_TIME_OC = 4, TOPO_TIME_COPY = 5,
       TOPO_TIME_APPLYA_MASKED = 6 /* the CG's operator: fixed DOFs masked, p.Ap partials */,
       TOPO_TIME_APPLYA_MFMA = 7   /* FP32 applyA on the matrix cores (as mfma_apply = 1) */ };
/* run `op` reps times back to back on device-resident data and return the
 * average duration of one launch in ms (and the algorithmic bytes of one). */
int topo_time_op(topo_solver* s, int32_t op, int32_t prec, int32_t level, int32_t reps,
                 double* avg_ms, double* bytes);

/* cancel flag polled between CG iterations and SIMP steps (thread-safe) */
int topo_solver_cancel(topo_solver* s);

/* ---- 2-D SIMP path (js/optimizer-worker.js) --------------------------
 * Bilinear quad4 elements, plane stress, on an nelx x nely grid in the
 * 2-D worker's numbering: node n = (nely+1)*ix + iy, DOFs 2n (x), 2n+1 (y),
 * element e = ey + ex*nely (column-major), element DOF order
 * [n1x n1y n2x n2y (n2+1)x (n2+1)y (n1+1)x (n1+1)y] with n1 = (nely+1)*ex +
 * ey, n2 = n1 + nely + 1 (_precomputeEdofs2D :240-268). Positions and
 * directions use the TOPO_FIX_* / TOPO_LOAD_* values above (left, right,
 * bottom, top; down, up, left, right), DOF masks TOPO_DOF_X / TOPO_DOF_Y. */

/* replaces lk() js/optimizer-worker.js:1732-1767 (+ _flattenKE :240-247):
 * 8x8 row-major quad4 stiffness, E = 1, unit thickness. */
int topo_k0_2d(double nu, double* ke64);

/* replaces _precomputeEdofs2D js/optimizer-worker.js:240-268 */
int topo_edofs_2d(int32_t nelx, int32_t nely, int32_t* edofs /* nel*8 */);

/* replaces getFixedDOFs js/optimizer-worker.js:1885-1921 (left, right,
 * bottom, top). Sets fixedMask[dof]=1 (does not clear). Returns #DOFs. */
int topo_fixed_dofs_2d(int32_t nelx, int32_t nely, int32_t position, int32_t dof_mask,
                       uint8_t* fixedMask /* ndof */);

/* replaces getLoadVector js/optimizer-worker.js:1923-1964. forceVector may be
 * NULL (otherwise 2 doubles: direction at node (nelx, nely)). Writes F
 * (float32, zeroed first). */
int topo_load_vector_2d(int32_t nelx, int32_t nely, int32_t direction, double magnitude,
                        const double* forceVector, float* F /* ndof */);

/* replaces getConstraintElements (:1966-1986) and getForceElements
 * (:1988-2014): sets preserved[e]=1 (position < 0 / direction < 0: none). */
int topo_preserved_elements_2d(int32_t nelx, int32_t nely, int32_t position, int32_t direction,
                               uint8_t* preserved /* nel */);

/* replaces getFixedDOFsFromPaint (:2016-2041): cells is n x 2 int32 (vx, vy)
 * of painted voxels; their 4 corner nodes are fixed in the masked DOFs. */
int topo_fixed_paint_2d(int32_t nelx, int32_t nely, const int32_t* cells, int32_t ncells,
                        int32_t dof_mask, uint8_t* fixedMask);

/* replaces getLoadVectorFromPaint (:2043-2098): magnitude spread evenly over
 * the unique corner nodes of the painted cells along dir2 (unit). F is
 * accumulated. Returns the node count. */
int topo_load_paint_2d(int32_t nelx, int32_t nely, const int32_t* cells, int32_t ncells,
                       const double* dir2, double magnitude, float* F);

/* 2-D defaults (the 2-D worker's constants :15-18 and config reads
 * :361-378, :570-578): E0 1, Emin 1e-9, nu 0.3, CG tolerance schedule
 * max(1e-8, change*1e-2) on ||r||, at most 2000 CG iterations; cfg->nelz = 1. */
int topo_config_default_2d(topo_config* cfg, int32_t nelx, int32_t nely);

/* Device-resident 2-D SIMP solver: the loop of optimize() :556-700 (Heaviside
 * projection with beta continuation, penalty continuation, Jacobi-PCG from a
 * cold start with the absolute tolerance schedule, energies, sensitivities,
 * sensitivity filter, OC, change) with every vector in HBM. Uses nelx, nely,
 * E0, Emin, nu, volfrac, penal, penal_start, continuation_iters,
 * max_iterations, rmin, use_projection, beta_max, beta_interval, move,
 * cg_max_iter, cg_tol (> 0: fixed absolute tolerance), skip_factor,
 * use_graphs of cfg; filter_type must be TOPO_FILTER_SENSITIVITY and nload 1.
 * fixedMask u8[ndof], F float32[ndof], preserved / voidm u8[nel] or NULL;
 * x starts at volfrac, preserved -> 1, void -> 0 (:517-553). The AMR
 * grouping, void filling and manufacturing constraints of the 2-D worker are
 * not part of this path. */
typedef struct topo_solver2d topo_solver2d;
int topo_solver2d_create(const topo_config* cfg, const uint8_t* fixedMask, const float* F,
                         const uint8_t* preserved, const uint8_t* voidm, topo_solver2d** out);
int topo_solver2d_destroy(topo_solver2d* s);
int topo_solver2d_set_design(topo_solver2d* s, const float* x);
int topo_solver2d_get_design(topo_solver2d* s, float* x);
/* live config patch (updateConfig :587-595): volfrac and max_iterations act
 * on the next step; penal changes only the stress metric (the loop keeps the
 * penalty schedule fixed at its start), rmin is recorded but the filter
 * keeps the weights prepared at the start, as in the 2-D worker. */
int topo_solver2d_update(topo_solver2d* s, double penal, double rmin, double volfrac,
                         int32_t max_iterations);
/* one SIMP iteration (:602-700); info->max_stress is the mesh builder's
 * max_e (Emin + x_e^penal (E0 - Emin)) w_e on the new design (:1166-1173) */
int topo_solver2d_step(topo_solver2d* s, topo_step_info* info);
/* FEA-only analysis (:384-485): FE at the densities x (float32[nel]) with
 * penalty penal and tolerance tol (absolute on ||r||; the worker uses 1e-8),
 * then energies and stress_e = (Emin + x_e^p (E0 - Emin)) w_e. compliance
 * and max_stress land in info; the fields U, ENERGY, STRESS are kept. */
int topo_solver2d_fea(topo_solver2d* s, const float* x, double penal, double tol, topo_step_info* info);
/* TOPO_FIELD_U (double[ndof]), XPHYS, ENERGY, DC, DCN, STRESS (float[nel]) */
int topo_solver2d_get_field(topo_solver2d* s, int32_t field, void* out);
int topo_solver2d_cancel(topo_solver2d* s);

#ifdef __cplusplus
}
#endif
#endif /* TOPO_H */
