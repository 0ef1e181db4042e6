"""Headline benchmark: SIMP iterations/s of the 192x96x96 cantilever
(1,769,472 elements, 5,447,811 DOF) with 4-level MG-PCG, mixed precision
(FP32 operator and V-cycle inside FP64 refinement) against the same
MG-PCG in FP64 and against the plain FP64 method (Jacobi-PCG), at matched
compliance, plus GDOF-matvec/s of the fine applyA.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 192,96,96]
                [--workload c3|c2|c4|c5|2d]

A step is one SIMP iteration (projection/filter, FE solve to tol 1e-6,
energies, sensitivities, filter, OC) on device-resident data. Prints one
JSON line. The default (c3) line also carries the 2-D leg and the c2, c4
and c5 workloads at their default step counts (`other_workloads`).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# smoothing-weight cap c of the cantilever workload: omega_eff = min(0.6, c / rho(D^-1 A)); measured
# at 192x96x96 over the timed steps: 1.2 / 1.4 / 1.6 / 1.8 / 1.9 take 1.45 / 1.18 / 1.06 / 0.94 / 0.95
# times the CG iterations of 1.6 ... (relative to 1.6), 2.0 stalls a solve at the 400-iteration cap
OMEGA_CAP = 1.8
# timed SIMP steps and untimed warm-up steps per workload
DEFAULT_STEPS = {"c3": (20, 5), "c2": (20, 1), "c4": (5, 2), "c5": (5, 15), "2d": (5, 1)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(T, P, workload, grid, prec, levels, **kw):
    nx, ny, nz = grid
    if workload == "c4":
        return P.bridge(nx, ny, nz, mg_levels=levels, mg_omega=0.6, coarse_solver=T.COARSE_CHOLESKY,
                        galerkin_max_ndof=0, precision=prec, **kw)
    if workload == "c5":
        return P.inverter(nx, ny, nz, precision=prec, mg_levels=levels, **kw)
    if (nx, ny, nz) == (192, 96, 96):
        cfg, fixed, F, pres, void, x0 = P.cantilever_c3(precision=prec, nlevels=levels)
    else:
        cfg, fixed, F, pres, void, x0 = P.cantilever(nx, ny, nz, linear_solver=T.SOLVER_MGPCG, precision=prec,
                                                     mg_levels=levels, mg_omega=0.6,
                                                     coarse_solver=T.COARSE_CHOLESKY, galerkin_max_ndof=0)
        nel, _ = T.dims(nx, ny, nz)
        x0 = P.splitmix64_uniform(nel)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg, fixed, F, pres, void, x0


def run_solver(T, P, grid, prec, steps, warmup, levels, plain=False, workload="c3", coarse_fp32=None,
               record=False, **kw):
    """Warm-up steps, then `steps` timed SIMP steps. Each step is timed on
    its own (it ends with a stream synchronisation: the host reads the
    step's scalars); with `record` the design before each timed step is
    copied out between the timed regions, for the matched-design legs.
    plain: the reference's plain method on the same problem -- Jacobi-PCG
    in FP64, with the 400-iteration cap lifted so it reaches the same
    tolerance (the compliance is then comparable)."""
    cfg, fixed, F, pres, void, x0 = make_problem(T, P, workload, grid, prec, levels, **kw)
    if workload == "c3":
        cfg.mg_omega_cap = OMEGA_CAP
    cfg.max_iterations = 1000
    # FP16 copy of the coarse factor inverse, FP32 for the inverter (whose
    # ill-conditioned coarse operator the FP16 copy does not represent well)
    cfg.coarse_fp32 = coarse_fp32 if coarse_fp32 is not None else (1 if workload == "c5" else 2)
    if plain:
        cfg.linear_solver = T.SOLVER_PCG
        cfg.precision = T.PREC_FP64
        cfg.cg_max_iter = 100000
    S = T.Solver(cfg, fixed, F)
    S.set_design(x0)
    infos, designs, times = [], [], []
    for i in range(warmup):
        t0 = time.perf_counter()
        info = S.step()
        log("warmup %d prec=%d: %.1f ms cg=%d outer=%d c=%.8g vol=%.6f change=%.4g ocp=%d" % (
            i, prec, (time.perf_counter() - t0) * 1e3, info.cg_iters, info.outer_iters, info.compliance,
            info.volume, info.change, info.oc_passes))
    for i in range(steps):
        if record:
            designs.append(S.design())
        t0 = time.perf_counter()
        info = S.step()
        times.append(time.perf_counter() - t0)
        infos.append(info.as_dict())
        log_step("step %d prec=%d" % (i, prec), info, times[-1])
    return S, infos, float(np.mean(times)) if times else 0.0, designs


def log_step(tag, info, dt):
    log("%s: %.1f ms (setup %.1f solve %.1f post %.1f) cg=%d outer=%d vcyc=%d c=%.8g vol=%.6f change=%.4g "
        "ocp=%d" % (tag, dt * 1e3, info.t_setup_ms, info.t_solve_ms, info.t_post_ms, info.cg_iters,
                    info.outer_iters, info.vcycles, info.compliance, info.volume, info.change, info.oc_passes))


def matched_leg(T, P, grid, prec, designs, first_iter, levels, plain=False, nsteps=None, workload="c3", **kw):
    """The same SIMP iterations as the timed mixed steps, from the same
    designs: each recorded design is handed over (set_design, iteration
    counter) and one step is timed on it. The FP64 MG-PCG is cold-started in
    every solve (as the mixed one); the plain Jacobi-PCG warm-starts from the
    previous handed design's displacements (the reference's PCG warm start),
    the first from zero. An untimed step on the first design first builds the
    hierarchy, tunes and captures the graphs."""
    cfg, fixed, F, _, _, _ = make_problem(T, P, workload, grid, prec, levels, **kw)
    if workload == "c3":
        cfg.mg_omega_cap = OMEGA_CAP
    cfg.max_iterations = 1000
    cfg.coarse_fp32 = 2
    if plain:
        cfg.linear_solver = T.SOLVER_PCG
        cfg.precision = T.PREC_FP64
        cfg.cg_max_iter = 100000
    S = T.Solver(cfg, fixed, F)
    n = len(designs) if nsteps is None else min(nsteps, len(designs))
    if not plain:
        S.set_design(designs[0])
        S.set_iteration(first_iter)
        S.step()
    infos, times = [], []
    for k in range(n):
        S.set_design(designs[k])
        S.set_iteration(first_iter + k)
        t0 = time.perf_counter()
        info = S.step()
        times.append(time.perf_counter() - t0)
        infos.append(info.as_dict())
        log_step("matched %s step %d" % ("plain Jacobi-PCG" if plain else "prec=%d" % prec, k), info, times[-1])
    S.close()
    return infos, float(np.mean(times))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # timed SIMP steps and untimed warm-up steps; defaults per workload:
    # c3 20 after 5 (the timed window reaches the late, harder designs), c2 20
    # after 1, c4 5 after 2, c5 5 after 15 (the inverting lever has formed),
    # 2d 5 after 1
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--grid", default="192,96,96")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed designs the plain FP64 Jacobi-PCG leg solves (each takes seconds at 5.4M DOF)")
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5", "2d"), default="c3",
                    help="c3: 192x96x96 cantilever (headline); c2: 96x48x48 cantilever, Jacobi-PCG FP64 and mixed; "
                         "c4: 128x64x64 bridge, 4 load cases, batched; c5: 160x80x80 force inverter, forward + "
                         "adjoint solves; 2d: the 2-D SIMP leg alone (1024x512 quad4 cantilever)")
    ap.add_argument("--coarse-fp32", type=int, default=None, choices=(0, 1, 2),
                    help="copy of the FP64 coarse factor inverse read by the FP32 V-cycle: 0 none (FP64), "
                         "1 FP32, 2 FP16 with one scale per row (default: the problem's own, FP16 for the "
                         "cantilever and the bridge, FP32 for the inverter)")
    ap.add_argument("--mixed", choices=("pc", "ir"), default="pc",
                    help="pc: FP64 MG-PCG with an FP32 V-cycle preconditioner; ir: FP64 iterative refinement "
                         "around an FP32 inner MG-PCG")
    ap.add_argument("--profile-kernel", action="store_true")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2-D SIMP leg")
    ap.add_argument("--no-others", action="store_true",
                    help="c3 only: skip the c2, c4 and c5 workloads reported beside the headline")
    ap.add_argument("--omega-cap", type=float, default=OMEGA_CAP,
                    help="cap c of the damped-Jacobi weight, omega_eff = min(0.6, c / rho(D^-1 A)) per level")
    args = ap.parse_args()
    dsteps, dwarm = DEFAULT_STEPS[args.workload]
    if args.steps is None:
        args.steps = dsteps
    if args.warmup is None:
        args.warmup = dwarm
    globals()["OMEGA_CAP"] = args.omega_cap
    if args.profile_kernel:
        return profile_kernels(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU: the solver has no multi-GPU path (one GPU holds the whole grid).",
              file=sys.stderr)
        sys.exit(2)
    import topologyoptimization_amd as T
    import topologyoptimization_amd.problems as P
    if args.workload == "2d":
        print(json.dumps(simp2d_leg(T, steps=args.steps, warmup=args.warmup, cpu=not args.no_cpu)))
        return
    if args.workload == "c2":
        print(json.dumps(c2_workload(T, P, args)))
        return
    if args.workload != "c3":
        print(json.dumps(other_workload(T, P, args)))
        return
    grid = tuple(int(v) for v in args.grid.split(","))
    n, name = T.device_info()
    assert n > 0, "no HIP device"
    log("device:", name)
    nel, ndof = T.dims(*grid)

    # mixed precision (the headline)
    mixed_prec = T.PREC_MIXED_PC if args.mixed == "pc" else T.PREC_MIXED
    S, infos, t_mixed, designs = run_solver(T, P, grid, mixed_prec, args.steps, args.warmup, args.levels,
                                            coarse_fp32=args.coarse_fp32, record=True)
    x_end = S.design()
    nonfinite = int((~np.isfinite(x_end)).sum())
    # dominant kernels, timed with HIP events on the solver's stream
    tR, tnt, tcx, tlog = S.tuning()
    ms32, by32 = S.time_op(T.TIME_APPLYA, T.PREC_MIXED, 0, 50)
    ms64, by64 = S.time_op(T.TIME_APPLYA, T.PREC_FP64, 0, 30)
    # the CG's FP64 operator (fixed DOFs masked, p.Ap partials in the epilogue)
    ms64m, by64m = S.time_op(T.TIME_APPLYA_MASKED, T.PREC_FP64, 0, 30)
    # the same FP32 applyA with the element product on the matrix cores
    msmf, bymf = S.time_op(T.TIME_APPLYA_MFMA, T.PREC_MIXED, 0, 50)
    msc, byc = S.time_op(T.TIME_COPY, T.PREC_FP64, 0, 30)
    msv, _ = S.time_op(T.TIME_VCYCLE, T.PREC_MIXED, 0, 5)
    S.close()
    # the same designs through FP64 MG-PCG and through the plain FP64 method
    infos64, t_fp64 = (None, None)
    if not args.no_fp64:
        infos64, t_fp64 = matched_leg(T, P, grid, T.PREC_FP64, designs, args.warmup, args.levels)
    infos_p, t_plain = (None, None)
    if not args.no_plain and args.plain_steps > 0:
        infos_p, t_plain = matched_leg(T, P, grid, T.PREC_FP64, designs, args.warmup, args.levels, plain=True,
                                       nsteps=args.plain_steps)
    value = 1.0 / t_mixed
    ach = by64m / (ms64m * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
            "traffic": measured_traffic(grid, "applyA_fp64_masked"),
            "kernel": "k_applyA_xm<double,double,512,1> (fine level, masked, p.Ap partials: the CG's operator)",
            "algorithmic_bytes": by64m, "avg_ms": ms64m}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(grid, infos, args.levels)
    fine_calls = np.mean([i["applyA_fine"] for i in infos])
    simp2d = None if args.no_2d else simp2d_leg(T, cpu=not args.no_cpu)
    # the other BASELINE workloads at their default step counts (the CPU
    # legs of the bridge and the inverter are left out: each would build a
    # full-size hierarchy on the host)
    others = None
    if not args.no_others:
        others = {}
        for wl in ("c2", "c4", "c5"):
            a = argparse.Namespace(**vars(args))
            a.workload = wl
            a.steps, a.warmup = DEFAULT_STEPS[wl]
            a.no_cpu = args.no_cpu or wl != "c2"
            try:
                others[wl] = c2_workload(T, P, a) if wl == "c2" else other_workload(T, P, a)
            except Exception as e:   # reported, never fatal to the headline
                others[wl] = {"error": repr(e)}
    out = {
        "metric": "SIMP iters/s + GDOF-matvec/s at 5.4M DOF, mixed-FP32 vs FP64 MGPCG",
        "value": value,
        "unit": "SIMP iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "cantilever %dx%dx%d hex8, %d DOF, %d-level MG-PCG V(2,2) omega 0.6 (capped), FP64 "
                               "direct coarsest, tol 1e-6, x0~U[0.05,1] seed %d" % (grid + (ndof, args.levels, P.SEED)),
                   "nel": nel, "ndof": ndof, "parallelism": "single GPU"},
        "trajectory": [{k: round(i[k], 9) if isinstance(i[k], float) else i[k]
                        for k in ("iteration", "compliance", "volume", "change", "oc_passes", "cg_iters",
                                  "t_setup_ms", "t_solve_ms", "t_post_ms")} for i in infos],
        "nonfinite_design_entries": nonfinite,
        "matched_designs": {
            "fp64_mgpcg_iters_per_s": (1.0 / t_fp64) if t_fp64 else None,
            "mixed_vs_fp64_mgpcg": (t_fp64 / t_mixed) if t_fp64 else None,
            "compliance_rel_vs_fp64_mgpcg": _rel(infos, infos64) if infos64 else None,
            "fp64_mgpcg_cg_iters": [i["cg_iters"] for i in infos64] if infos64 else None,
            "plain_fp64_pcg_iters_per_s": (1.0 / t_plain) if t_plain else None,
            "mixed_vs_plain_fp64_pcg": (t_plain / t_mixed) if t_plain else None,
            "compliance_rel_vs_plain_fp64_pcg": _rel(infos, infos_p) if infos_p else None,
            "plain_fp64_pcg_cg_iters": [i["cg_iters"] for i in infos_p] if infos_p else None,
            "plain_fp64_pcg_residual": [i["residual"] for i in infos_p] if infos_p else None,
        },
        "gdof_matvec_per_s": {"fp32": ndof / (ms32 * 1e-3) / 1e9, "fp64": ndof / (ms64 * 1e-3) / 1e9,
                              "fp64_masked_pap": ndof / (ms64m * 1e-3) / 1e9},
        "applyA_ms": {"fp32": ms32, "fp64": ms64, "fp64_masked_pap": ms64m, "fp32_mfma": msmf},
        "applyA_hbm_frac": {"fp32": by32 / (ms32 * 1e-3) / 8e12, "fp64": by64 / (ms64 * 1e-3) / 8e12,
                            "fp64_masked_pap": by64m / (ms64m * 1e-3) / 8e12, "fp32_mfma": bymf / (msmf * 1e-3) / 8e12},
        # matrix-core work of the split-bfloat16 element products: 3 products
        # of the 32x32 padded K0 with each element's 32-vector
        "mfma_apply": {"matrix_flops_per_launch": 6144.0 * nel,
                       "tflops": 6144.0 * nel / (msmf * 1e-3) / 1e12,
                       "frac_of_bf16_dense_peak": 6144.0 * nel / (msmf * 1e-3) / 2.5e15},
        "vcycle_ms_fp32": msv,
        "copy_GBps": byc / (msc * 1e-3) / 1e9,
        "fine_applyA_per_step": fine_calls,
        "compliance": infos[-1]["compliance"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotuned_fine_tile": {"xm_nt": tnt, "xm_cx": tcx},
        "autotune_candidates": [[int(a), int(b), int(c), float(d)] for a, b, c, d in tlog],
        "simp2d": simp2d,
        "other_workloads": others,
        "device": name,
    }
    print(json.dumps(out))


def c2_workload(T, P, args):
    """The 96x48x48 cantilever (698,691 DOF) with the reference's plain
    solver, Jacobi-PCG, tol 1e-6, in FP64 and in mixed precision (FP64
    iterative refinement around an FP32 Jacobi-PCG), from x = volfrac, 20
    SIMP iterations by default (--steps). The FP64 leg is timed on the mixed
    leg's designs (matched compliance)."""
    n, name = T.device_info()
    assert n > 0, "no HIP device"
    log("device:", name)
    grid = (96, 48, 48)
    nel, ndof = T.dims(*grid)
    steps = args.steps

    def make(prec):
        cfg, fixed, F, _, _, x0 = P.cantilever_c2(precision=prec)
        cfg.max_iterations = 1000
        cfg.cg_max_iter = 20000
        return cfg, fixed, F, x0

    cfg, fixed, F, x0 = make(T.PREC_MIXED)
    S = T.Solver(cfg, fixed, F)
    S.set_design(x0)
    designs, infos, times = [], [], []
    for i in range(args.warmup):
        S.step()
    for i in range(steps):
        designs.append(S.design())
        t0 = time.perf_counter()
        info = S.step()
        times.append(time.perf_counter() - t0)
        infos.append(info.as_dict())
        log_step("c2 mixed step %d" % i, info, times[-1])
    ms32, by32 = S.time_op(T.TIME_APPLYA, T.PREC_MIXED, 0, 50)
    S.close()
    cfg, fixed, F, x0 = make(T.PREC_FP64)
    S = T.Solver(cfg, fixed, F)
    infos64, times64 = [], []
    for k, x in enumerate(designs):
        S.set_design(x)
        S.set_iteration(args.warmup + k)
        t0 = time.perf_counter()
        info = S.step()
        times64.append(time.perf_counter() - t0)
        infos64.append(info.as_dict())
        log_step("c2 fp64 step %d" % k, info, times64[-1])
    S.close()
    t_m, t_64 = float(np.mean(times)), float(np.mean(times64))
    ach = by32 / (ms32 * 1e-3) / 1e9
    out = {
        "metric": "SIMP iters/s, 96x48x48 cantilever, Jacobi-PCG mixed vs FP64", "value": 1.0 / t_m,
        "unit": "SIMP iters/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": t_m * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": "cantilever 96x48x48 hex8, %d DOF, Jacobi-PCG tol 1e-6, x0 = volfrac 0.12" % ndof,
                   "nel": nel, "ndof": ndof, "parallelism": "single GPU"},
        "fp64_iters_per_s": 1.0 / t_64, "mixed_vs_fp64": t_64 / t_m,
        "compliance_rel_vs_fp64": _rel(infos, infos64),
        "cg_iters": {"mixed": [i["cg_iters"] for i in infos], "fp64": [i["cg_iters"] for i in infos64]},
        "trajectory": [{k: i[k] for k in ("iteration", "compliance", "volume", "change", "oc_passes")} for i in infos],
        "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                     "traffic": None, "kernel": "k_applyA_xm<float> (fine level)", "algorithmic_bytes": by32,
                     "avg_ms": ms32},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_c2(infos), "device": name,
    }
    return out


def other_workload(T, P, args):
    """c4: the 4-load bridge solved by the batched MG-PCG with autotuned batch
    and tile shapes, against the same solver one load case at a time; c5: the
    force inverter (forward and adjoint solves per step). One JSON line."""
    n, name = T.device_info()
    assert n > 0, "no HIP device"
    log("device:", name)
    mixed_prec = T.PREC_MIXED_PC if args.mixed == "pc" else T.PREC_MIXED
    if args.workload == "c4":
        grid = (128, 64, 64)
        S, infos, t_b, _ = run_solver(T, P, grid, mixed_prec, args.steps, args.warmup, args.levels, workload="c4",
                                      rhs_tile=0, coarse_fp32=args.coarse_fp32)
        R, nt, cx, tlog = S.tuning()
        ms32, by32 = S.time_op(T.TIME_APPLYA, T.PREC_MIXED, 0, 50)
        S.close()
        S1, infos1, t_s, _ = run_solver(T, P, grid, mixed_prec, args.steps, args.warmup, args.levels, workload="c4",
                                        rhs_tile=1, coarse_fp32=args.coarse_fp32)
        S1.close()
        extra = {"batched_vs_one_at_a_time": t_s / t_b, "one_at_a_time_iters_per_s": 1.0 / t_s,
                 "compliance_rel_vs_one_at_a_time": _rel(infos, infos1),
                 "autotuned": {"rhs_tile": R, "xm_nt": nt, "xm_cx": cx,
                               "candidates": [[int(a), int(b), int(c), float(d)] for a, b, c, d in tlog]}}
        workload = ("bridge 128x64x64 hex8, 1635075 DOF, 4 load cases (8x8-node patches, top face), batched "
                    "%d-level MG-PCG, FP32 V-cycle in FP64 CG, FP64 direct coarsest, tol 1e-6" % args.levels)
        metric = "SIMP iters/s, 4-load bridge, batched MG-PCG"
    else:
        grid = (160, 80, 80)
        S, infos, t_b, _ = run_solver(T, P, grid, mixed_prec, args.steps, args.warmup, args.levels, workload="c5",
                                      cg_max_iter=2000, coarse_fp32=args.coarse_fp32)
        ms32, by32 = S.time_op(T.TIME_APPLYA, T.PREC_MIXED, 0, 50)
        S.close()
        extra = {}
        workload = ("force inverter 160x80x80 hex8, 3168963 DOF, output displacement (adjoint solve), Helmholtz "
                    "filter rmin 2 + Heaviside beta 8, %d-level MG-PCG, FP32 V-cycle in FP64 CG, tol 1e-6" % args.levels)
        metric = "SIMP iters/s, compliant-mechanism (forward + adjoint) MG-PCG"
    nel, ndof = T.dims(*grid)
    ach = by32 / (ms32 * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline(grid, infos, args.levels)
    out = {
        "metric": metric, "value": 1.0 / t_b, "unit": "SIMP iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_b * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": workload, "nel": nel, "ndof": ndof, "parallelism": "single GPU"},
        "cg_iters_per_step": float(np.mean([i["cg_iters"] for i in infos])),
        "compliance": infos[-1]["compliance"],
        "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                     "traffic": measured_traffic(grid), "kernel": "k_applyA_xm<float> (fine level)",
                     "algorithmic_bytes": by32, "avg_ms": ms32},
        "cpu_baseline": cpu, "device": name,
    }
    out.update(extra)
    return out


def _rel(a, b):
    """max relative compliance difference over the steps both legs ran (both
    start from the same design)."""
    n = min(len(a), len(b))
    return max(abs(a[i]["compliance"] - b[i]["compliance"]) / abs(b[i]["compliance"]) for i in range(n))


def measured_traffic(grid, kernel="applyA_fp32"):
    """HBM bytes per launch of a fine-level applyA variant from the committed
    rocprofv3 counter runs (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), or None when no summary exists for it."""
    path = os.path.join(ROOT, "profiles", "applyA_pmc_%dx%dx%d.json" % grid)
    try:
        with open(path) as f:
            return json.load(f)[kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(grid, infos, levels):
    """CPU baseline: the oracle (the C++ restatement of the reference
    algorithm, -O3) on the host, on a bounded sample of the same workload:
    the 4-level hierarchy setup for the first timed design (E-average levels
    and the coarsest Cholesky), one V-cycle and one fine applyA at the same
    grid, all with OpenMP on the host's cores (at most 16), and the V-cycle
    and applyA again on one core. SIMP iters/s = 1 / (setup + n_cg x
    (applyA + V-cycle)), n_cg = the mean CG iteration count of the GPU steps
    (same algorithm and tolerance); energies, filter and OC are not counted."""
    try:
        import ctypes
        from oracle import oracle as O
        import topologyoptimization_amd.problems as P
        L = O.lib()
        L.or_set_threads.argtypes = [ctypes.c_int]
        nx, ny, nz = grid
        nel, ndof = O.dims(nx, ny, nz)
        cfg, fixed, F, _, _, _ = P.cantilever(nx, ny, nz)
        x = P.splitmix64_uniform(nel)
        ke = O.k0(0.3)
        E = 1e-9 + x.astype(np.float64) ** 3 * (1 - 1e-9)
        act = np.arange(nel, dtype=np.int32)
        p = np.random.default_rng(0).normal(size=ndof)
        r = F.astype(np.float64).copy()
        r[fixed > 0] = 0
        cores = min(16, os.cpu_count() or 1)
        out = {}
        L.or_set_threads(cores)
        mg = O.MG(nx, ny, nz, fixed, nlevels=levels, omega=0.6, coarse=1, galerkin_max=0, omega_cap=OMEGA_CAP,
                  skip=-1.0)
        t0 = time.perf_counter()
        mg.update(x, 3.0)
        t_setup = time.perf_counter() - t0
        for th in (cores, 1):
            L.or_set_threads(th)
            t0 = time.perf_counter()
            O.applyA(nx, ny, nz, ke, E, act, p)
            t_ap = time.perf_counter() - t0
            t0 = time.perf_counter()
            mg.vcycle(r)
            t_vc = time.perf_counter() - t0
            out[th] = (t_ap, t_vc)
        L.or_set_threads(1)
        n_cg = float(np.mean([i["cg_iters"] for i in infos]))
        t_ap, t_vc = out[cores]
        per_step = t_setup + n_cg * (t_ap + t_vc)
        t_ap1, t_vc1 = out[1]
        return {"value": 1.0 / per_step, "unit": "SIMP iters/s", "cores": cores, "kind": "port",
                "sample": "oracle (-O3, OpenMP, %d threads) at %dx%dx%d, %d-level hierarchy: setup %.2f s + 1 V-cycle "
                          "%.3f s + 1 fine applyA %.3f s (%.3f GDOF-matvec/s), scaled by %.1f CG iterations per step" %
                          (cores, nx, ny, nz, levels, t_setup, t_vc, t_ap, ndof / t_ap / 1e9, n_cg),
                "gdof_matvec_per_s": ndof / t_ap / 1e9,
                "single_core": {"applyA_s": t_ap1, "vcycle_s": t_vc1, "gdof_matvec_per_s": ndof / t_ap1 / 1e9,
                                "cg_iteration_s": t_ap1 + t_vc1},
                "host": platform.processor() or platform.machine(), "nproc": os.cpu_count()}
    except Exception as e:  # reported, never fatal
        return {"value": None, "error": repr(e)}


def cpu_baseline_c2(infos):
    """CPU baseline of the C2 workload: the oracle's ebePCG (the reference's
    Jacobi-PCG, FP64, one thread) at 96x48x48 for 20 CG iterations from the
    first design, scaled by the GPU steps' mean CG iteration count."""
    try:
        from oracle import oracle as O
        import topologyoptimization_amd.problems as P
        cfg, fixed, F, _, _, x0 = P.cantilever_c2()
        nx, ny, nz = 96, 48, 48
        nel, ndof = O.dims(nx, ny, nz)
        free = np.where(fixed == 0)[0].astype(np.int32)
        U = np.zeros(ndof)
        t0 = time.perf_counter()
        it = O.ebePCG(x0.astype(np.float64), O.k0(0.3), nx, ny, nz, F.astype(np.float64), U, free, 1e-9, 1.0, 3.0, 20,
                      1e-30)
        t_it = (time.perf_counter() - t0) / max(1, it)
        n_cg = float(np.mean([i["cg_iters"] for i in infos]))
        return {"value": 1.0 / (n_cg * t_it), "unit": "SIMP iters/s", "cores": 1, "kind": "port",
                "sample": "oracle ebePCG at 96x48x48, %d CG iterations (%.1f ms each), scaled by %.0f CG iterations "
                          "per step (energies, filter and OC not counted)" % (it, t_it * 1e3, n_cg)}
    except Exception as e:  # reported, never fatal
        return {"value": None, "error": repr(e)}


def simp2d_leg(T, grid=(1024, 512), steps=5, warmup=1, cpu=True):
    """The 2-D path (js/optimizer-worker.js loop) on a quad4 cantilever:
    left edge clamped, unit load down at node (nelx, nely), volfrac 0.4,
    rmin 1.5, penalty 1.5 -> 3 over 20 iterations, Heaviside projection,
    Jacobi-PCG (FP64) to max(1e-8, change/100) on ||r||, at most 2000
    iterations. SIMP iters/s over `steps` steps after `warmup`; the CPU
    baseline is the oracle's solveCG on 20 CG iterations at the same grid,
    scaled by the GPU steps' CG iteration counts."""
    nelx, nely = grid
    nel, ndof = T.dims2(nelx, nely)
    cfg = T.default_config_2d(nelx, nely, volfrac=0.4, max_iterations=100, penal_start=1.5, continuation_iters=33)
    fixed = T.fixed_dofs_2d(nelx, nely, T.FIX_LEFT)
    F = T.load_vector_2d(nelx, nely, T.LOAD_DOWN, 1.0)
    pres = T.preserved_elements_2d(nelx, nely, T.FIX_LEFT, T.LOAD_DOWN)
    with T.Solver2D(cfg, fixed, F, pres) as S:
        for _ in range(warmup):
            S.step()
        t0 = time.perf_counter()
        infos = [S.step() for _ in range(steps)]
        dt = time.perf_counter() - t0
    cg = [i.cg_iters for i in infos]
    solve_ms = sum(i.t_solve_ms for i in infos)
    out = {"workload": "2-D cantilever %dx%d quad4, %d DOF, Jacobi-PCG FP64, steps %d..%d" %
                       (nelx, nely, ndof, warmup + 1, warmup + steps),
           "value": steps / dt, "unit": "SIMP iters/s", "ms_per_step": dt / steps * 1e3, "cg_iters": cg,
           # the 2-D worker's own cap (js/optimizer-worker.js solveCG): at
           # this size its absolute tolerance is not reached within it
           "cg_max_iter": cfg.cg_max_iter, "solves_at_cap": int(sum(c >= cfg.cg_max_iter for c in cg)),
           "us_per_cg_iteration": solve_ms * 1e3 / max(1, sum(cg)),
           "compliance": infos[-1].compliance, "cpu_baseline": None}
    if cpu:
        try:
            from oracle import oracle as O
            x = np.full(nel, 0.4, np.float32)
            t0 = time.perf_counter()
            _, _, it = O.fe2(nelx, nely, O.k0_2d(0.3), x, 3.0, fixed, F, tol=1e-30, maxit=20)
            t_it = (time.perf_counter() - t0) / it
            n_cg = float(np.mean(cg))
            out["cpu_baseline"] = {"value": 1.0 / (n_cg * t_it), "unit": "SIMP iters/s", "cores": 1, "kind": "port",
                                   "sample": "oracle solveCG, %d CG iterations at %dx%d (%.1f ms each), scaled by %.0f "
                                             "CG iterations per step (energies, filter and OC not counted)" %
                                             (it, nelx, nely, t_it * 1e3, n_cg)}
        except Exception as e:  # reported, never fatal
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    return out


def profile_kernels(args):
    """Fixed kernel workload for rocprofv3 counter runs at the benchmark
    grid, 20 launches each: the fine applyA in FP32 (mode 0, VALU and
    matrix-core kernels), in FP64 (mode 0), the CG's FP64 operator (masked,
    p.Ap partials: mode 1), and a plain FP64 copy of one vector (the
    calibration kernel)."""
    import topologyoptimization_amd as T
    import topologyoptimization_amd.problems as P
    grid = tuple(int(v) for v in args.grid.split(","))
    if grid == (192, 96, 96):
        cfg, fixed, F, pres, void, x0 = P.cantilever_c3(precision=T.PREC_MIXED, nlevels=args.levels)
    else:
        cfg, fixed, F, pres, void, x0 = P.cantilever(*grid, precision=T.PREC_MIXED)
        x0 = P.splitmix64_uniform(T.dims(*grid)[0])
    S = T.Solver(cfg, fixed, F)
    # one FE solve at the design: the autotuner picks the element tiles the
    # solves use
    S.fe(x0, 3.0)
    out = {}
    for name, op, prec in (("applyA_fp32", T.TIME_APPLYA, T.PREC_MIXED), ("applyA_fp32_mfma", T.TIME_APPLYA_MFMA, T.PREC_MIXED),
                           ("applyA_fp64", T.TIME_APPLYA, T.PREC_FP64),
                           ("applyA_fp64_masked", T.TIME_APPLYA_MASKED, T.PREC_FP64),
                           ("copy_fp64", T.TIME_COPY, T.PREC_FP64)):
        ms, by = S.time_op(op, prec, 0, 20)
        out[name] = {"avg_ms": ms, "algorithmic_bytes": by}
    S.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
