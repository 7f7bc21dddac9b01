#!/usr/bin/env python3
"""Equilibrium steps/s of the shell arc-length path on one MI355X.

Default workload ("C2", the batched imperfection sweep): 256 shallow cylindrical panels (R = 2.5,
half-angle 0.1 rad, length 0.5, t = 0.0127, E = 3.1e3, nu = 0.3, degree 3), mesh n x n with n
drawn from {16,24,32,40,48}, each mid-surface perturbed by a seeded Gaussian random field
(correlation length 0.1, amplitude U[0, 0.5] t, PCG64(20261202 + i)); straight edges hinged,
central point load; Crisfield arc length, Scaling = 0 (benchmarks/benchmark_Roof.cpp:365),
TolF = 1e-8 (the metric's ||R||/||F|| bar), TolU = 1e-6, MaxIter 20, 60 arc-length steps per
instance, determinant stability (negative pivots of every corrector's tangent).  Arc length per
instance: the reference's 0.5 mm on its 45-dof mesh in metres, scaled with sqrt(n_free) (the same
physical step on every mesh) and by kappa = 2.6, so that each size class meets its limit point
near step 15 and leaves the unstable branch near step 46 of its 60.

A "step" in the JSON line is one arc-length step of the whole batch (each of the 256 instances one
converged step); the C2 path is 60 such steps whatever --steps asks.  `value` = converged instance
steps / wall seconds of the headline form, inputs resident on the GPU:
  "mixed+learned"   (default; the metric's form) FP32 band LDL^T with FP64 refinement, FP64 K, R,
                    constraint and pivot count, and the learned arc-length predictor;
  "mixed"           the same with the tangent predictor;
  "fp32k+learned"   K itself assembled in FP32 (element GEMMs on FP32 MFMA, FP32 band), FP32 LDL^T,
                    FP64 refinement against the FP32 K, FP64 R, constraint and pivot count (an FP64
                    re-assembly and factor where the count is not certified), learned predictor;
  "batched+learned" FP64 throughout with the learned predictor;
  "batched"         FP64 throughout with the tangent predictor.
The headline run is timed with profiling off; a second run of the same form with per-launch HIP
events gives the kernel table and the roofline.  The learned predictor (264-256-256-65 BF16 MLP) is
trained before any timing, on the GPU, on the FP64 paths of 256 other roofs (seeds 50261202 + i);
its data and training seconds are reported and amortised.  The held-out evaluation (64 roofs,
seeds 30261202 + i) has three arms: tangent predictor, quadratic extrapolation alone, and the
network.  The plain FP64 GPU form (one instance at a time, tangent predictor) is timed in full on
--plain-sample instances per size class (or on all with --plain-all) and scaled by the batch's size
histogram.  cpu_baseline: the FP64 CPU oracle port of the same loop (reference re-factorisations
kept), a bounded prefix of one instance per size class, scaled by the size histogram.

--workload c1: the single 32x32 perfect panel, plain FP64 vs mixed, 80 steps, the reference's arc
length (kappa = 1).
"""
from __future__ import annotations

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

METRIC = "equilibrium steps/s at ‖R‖/‖F‖≤1e-8: mixed+batched+learned vs plain FP64 GPU"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}
KERNELS = ["tangent_gp", "tangent_elem", "band_gather", "residual_elem", "vector_gather", "band_ldlt", "band_solve",
           "band_symv", "band_ldlt_batch_fp64", "band_ldlt_batch_fp32", "band_solve_batch_fp64", "band_solve_batch_fp32",
           "refine_residual_batch", "band_ldlt_early_fp64", "fp64_refactor_pivot", "fp64_refactor_stall",
           "fp64_refactor_early", "fp64_refactor_early_used", "fp64_refactor_count", "crisfield_complex_roots",
           "band_ldlt_tiled", "band_solve_tiled", "band_ldlt_recount_fp64", "fp64_recount_deferred", "engine_rounds",
           "engine_round_us"]
C2_STEPS, C1_STEPS = 60, 80
C2_KAPPA = 2.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_stats(lib):
    """(launches, total ms, algorithmic work) per profiled kernel, from the library's HIP-event
    brackets on the launching stream and the work each launcher counts for its matrices."""
    import ctypes as C
    out = {}
    for nm in KERNELS:
        cnt = C.c_int64()
        ms = C.c_double()
        work = C.c_double()
        lib.ksh_profile_query(nm.encode(), C.byref(cnt), C.byref(ms))
        lib.ksh_profile_query_work(nm.encode(), C.byref(work))
        out[nm] = (cnt.value, ms.value, work.value)
    return out


# bound of each kernel: MFMA (work in flop) or HBM (work in bytes)
BOUND = {"band_ldlt": "mfma", "band_ldlt_batch_fp64": "mfma", "band_ldlt_batch_fp32": "mfma", "band_ldlt_tiled": "mfma",
         "band_ldlt_early_fp64": "mfma", "band_ldlt_recount_fp64": "mfma",
         "tangent_elem": "mfma", "tangent_gp": "hbm", "band_solve": "hbm", "band_solve_batch_fp64": "hbm",
         "band_solve_batch_fp32": "hbm", "band_solve_tiled": "hbm", "band_gather": "hbm", "band_symv": "hbm",
         "refine_residual_batch": "hbm"}
FP32_KERNELS = {"band_ldlt_batch_fp32"}
# launched on the engine's second and third streams, beside the main stream's work: reported, but not a
# candidate for the dominant kernel (their time overlaps the main stream's)
OVERLAPPED = {"band_ldlt_early_fp64", "band_ldlt_recount_fp64"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (scaled by a device copy of known size), or None when no such measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if rec is None else float(rec["traffic"])


def roofline(stats, kernel=None):
    """The dominant kernel (largest total time on the main stream) against its roofline:
    achieved = algorithmic work per launch (counted by the launcher over the matrices of each
    launch) / average launch duration (HIP events on the launch stream)."""
    cands = [k for k in stats if stats[k][0] > 0 and stats[k][2] > 0 and k in BOUND and k not in OVERLAPPED]
    best = kernel or max(cands, key=lambda k: stats[k][1])
    cnt, ms, work = stats[best]
    per_launch = work / cnt
    achieved = per_launch / (ms / cnt / 1e3)
    if BOUND[best] == "mfma":
        peak = PEAK["fp32"] if best in FP32_KERNELS else PEAK["fp64"]
        return {"kernel": best, "bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": measured_traffic(best), "avg_launch_ms": ms / cnt, "launches": cnt,
                "work_per_launch": per_launch, "work_unit": "flop"}
    return {"kernel": best, "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK["hbm"], "traffic": measured_traffic(best), "avg_launch_ms": ms / cnt, "launches": cnt,
            "work_per_launch": per_launch, "work_unit": "byte"}


def host_cpu():
    info = {"nproc": os.cpu_count(), "machine": platform.machine()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)", "CPU max MHz")
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keep:
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_port(sps, by_size, steps, lengths, budget):
    """FP64 CPU oracle port (reference re-factorisations kept), a bounded prefix of one instance
    per size class, timed on this host; scaled by the batch's size histogram."""
    from oracle.solvers import run_crisfield
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    per_class = budget / max(1, len(by_size))
    per = {}
    for n, members in sorted(by_size.items()):
        i = members[0]
        t0 = time.perf_counter()
        r = run_crisfield(sps[i].patch, 1, lengths[i], max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=0.0, faithful=True)
        first = time.perf_counter() - t0
        k = int(max(1, min(steps - 1, (per_class - first) / max(first, 1e-3))))
        t0 = time.perf_counter()
        r2 = run_crisfield(sps[i].patch, 1 + k, lengths[i], max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=0.0,
                           faithful=True)
        dt = time.perf_counter() - t0
        per[n] = (r2["steps_done"], dt, i)
        log(f"[bench] CPU port n={n} (instance {i}): {r2['steps_done']} steps in {dt:.2f} s")
    t_total = sum(len(by_size[n]) * steps * per[n][1] / max(1, per[n][0]) for n in per)
    return {"value": round(steps * sum(len(v) for v in by_size.values()) / t_total, 4), "unit": "steps/s",
            "cores": cores, "kind": "port",
            "sample": "first k arc-length steps of one instance per size class ("
                      + ", ".join(f"n={n}: {per[n][0]} steps in {per[n][1]:.1f} s" for n in sorted(per))
                      + "), FP64 oracle, one instance at a time, reference re-factorisations kept (3 per corrector "
                        "iteration); per-class time per step scaled by the batch's size histogram",
            "host": host_cpu()}


def limit_points(sps, res):
    """Per size class: instances whose path crosses a limit point (a negative pivot appears), the
    median step where it first appears, and instances that return to a stable branch."""
    out = {}
    for sp, r in zip(sps, res):
        neg = np.asarray(r["negatives"])
        first = int(np.argmax(neg > 0)) if (neg > 0).any() else -1
        back = first >= 0 and bool((neg[first:] == 0).any())
        d = out.setdefault(str(sp.patch.nx), {"instances": 0, "crossed": 0, "first_negative_steps": [], "restabilised": 0})
        d["instances"] += 1
        if first >= 0:
            d["crossed"] += 1
            d["first_negative_steps"].append(first + 1)
            d["restabilised"] += int(back)
    for d in out.values():
        f = d.pop("first_negative_steps")
        d["median_first_negative_step"] = float(np.median(f)) if f else None
    return out


def c4_leg():
    """C4: the 10 lowest modal and buckling pairs of the 128 x 128 degree-3 plate (49,411 free
    dofs) in FP64 and with the FP32 factor + FP64 refinement: seconds to 10 converged pairs, the
    solves, and the largest relative eigenvalue error against ARPACK on the oracle's matrices
    (tests/golden/c4_eigen.json)."""
    from gsstructuralanalysis_amd import _lib as L
    from gsstructuralanalysis_amd import geometry as g
    from gsstructuralanalysis_amd.problem import ShellProblem
    with open(os.path.join(ROOT, "tests", "golden", "c4_eigen.json")) as f:
        fx = json.load(f)

    def z0(x, y):
        return 0.05 * np.sin(np.pi * x) * np.sin(np.pi * y)

    sp = ShellProblem(g.clamped_plate_problem(p=3, n=128, z0=z0, rho=1e3))
    out = {"workload": "C4: clamped plate 1 x 1, t 0.01, E 1e7, nu 0.3, rho 1e3, initial deflection "
                       "0.05 sin(pi x) sin(pi y), degree 3, 128 x 128 elements",
           "nfree": sp.nfree, "band": sp.bw}
    for name, run, key in (("modal", sp.run_modal, "modal_omega2"), ("buckling", sp.run_buckling, "buckling_mu")):
        ref = np.sort(np.asarray(fx[key]))
        for form, kw in (("fp64", dict(precision=L.FP64)), ("fp32_ir", dict(precision=L.FP32_IR))):
            t0 = time.perf_counter()
            r = run(10, tol=1e-10, **kw)
            dt = time.perf_counter() - t0
            err = float(np.max(np.abs(np.sort(r["values"]) - ref) / np.abs(ref)))
            out[f"{name}_{form}"] = {"seconds": round(dt, 3), "converged": int(r["converged"]), "solves": int(r["solves"]),
                                     "max_rel_error_vs_arpack": err}
            log(f"[bench] C4 {name} {form}: {dt:.2f} s, {r['solves']} solves, max rel. error vs ARPACK {err:.1e}")
    sp.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="requested steps (recorded; the C2 path is 60 arc-length steps, C1 80)")
    ap.add_argument("--warmup", type=int, default=2,
                    help="> 0: each timed form first runs once over the whole batch and path, untimed (0: no warm-up)")
    ap.add_argument("--workload", choices=["c2", "c1"], default="c2")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=24.0, help="seconds of CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 reference timing")
    ap.add_argument("--no-c4", action="store_true", help="skip the C4 eigen-analysis leg")
    ap.add_argument("--c4-only", action="store_true", help="run the C4 eigen-analysis leg alone and print its JSON")
    ap.add_argument("--plain-sample", type=int, default=2, help="plain FP64: instances timed per size class")
    ap.add_argument("--plain-all", action="store_true", help="plain FP64: time every instance")
    ap.add_argument("--form", choices=["mixed+learned", "mixed", "fp32k+learned", "batched+learned", "batched"],
                    default="mixed+learned",
                    help="headline form")
    ap.add_argument("--epochs", type=int, default=400, help="training epochs of the learned predictor")
    ap.add_argument("--only", action="store_true",
                    help="time the headline form only (no other forms, no held-out predictor evaluation)")
    ap.add_argument("--also", nargs="*", default=None, metavar="FORM",
                    help="the other forms to time (default: all); compared with the FP64 batched form")
    ap.add_argument("--max-refine", type=int, default=None,
                    help="FP32-factor forms: Krylov (FGMRES) steps per solve with the FP32 factor as preconditioner "
                         "(default per form: 4 for mixed+learned, 3 for mixed and fp32k+learned)")
    ap.add_argument("--refine-tol", type=float, default=None,
                    help="FP32-factor forms: solve stop ||b - K x|| <= tol ||b|| (the corrector's FP64 residual decides "
                         "convergence; default per form: 1e-6 for mixed+learned, 1e-4 for mixed and fp32k+learned)")
    ap.add_argument("--refine-accept", type=float, default=1e-2,
                    help="mixed form: a solve still above accept ||b|| after the sweeps is redone with an FP64 factor")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (device-to-device), for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()

    if args.c4_only:
        print(json.dumps(c4_leg()), flush=True)
        return

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} MiB x 5 (reads {args.copy_calibration * 1024} KiB and writes "
            f"{args.copy_calibration * 1024} KiB per copy)")
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; instances of a batch would be sharded over GPUs "
              "one process per GPU, which is not run here.", file=sys.stderr)
        sys.exit(2)

    from gsstructuralanalysis_amd import _lib as L
    from gsstructuralanalysis_amd import geometry as g
    from gsstructuralanalysis_amd.problem import ShellProblem, run_alm_batch

    lib = L.load()
    kw = dict(method=0, max_iter=20, tol=1e-6, tol_f=1e-8, tol_u=1e-6, scaling=0.0, bifurcation_method=0,
              max_halvings=10)
    # refinement of the FP32-factor forms, per form the faster of (3 steps, 1e-4) and (4 steps, 1e-6) on
    # this workload (warm runs): the learned proposals start the corrector close to the solution,
    # where inexact solves cost it iterations (47,283 against 43,757 on C2)
    def refine(m, tol):
        return dict(max_refine=args.max_refine if args.max_refine is not None else m,
                    refine_tol=args.refine_tol if args.refine_tol is not None else tol, refine_accept=args.refine_accept)
    forms = {"batched": dict(precision=L.FP64), "mixed": dict(precision=L.FP32_IR, **refine(3, 1e-4)),
             "batched+learned": dict(precision=L.FP64), "mixed+learned": dict(precision=L.FP32_IR, **refine(4, 1e-6)),
             "fp32k+learned": dict(precision=L.FP32, **refine(3, 1e-4))}

    t_build = time.perf_counter()
    if args.workload == "c1":
        steps = C1_STEPS
        pt = g.roof_problem(p=3, n=32)
        sps = [ShellProblem(pt)]
        lengths = np.array([g.roof_arc_length(sps[0].nfree, 1.0)])
        workload = (f"C1 shallow roof Crisfield, p=3 32x32 ({sps[0].nfree} free dofs, bw {sps[0].bw}), "
                    f"ds={lengths[0]:.4g} (the reference's 0.5 mm on 45 dofs, scaled by sqrt(n_free)), Scaling=0, "
                    f"TolF=1e-8, {steps} steps, single instance")
    else:
        steps = C2_STEPS
        pts = [g.imperfect_roof_instance(i) for i in range(args.instances)]
        sps = [ShellProblem(pt) for pt in pts]
        lengths = np.array([g.roof_arc_length(s.nfree, C2_KAPPA) for s in sps])
        workload = (f"C2 batch of {len(sps)} imperfect shallow roofs, p=3, n in {{16,24,32,40,48}}, GRF imperfection "
                    f"(l=0.1, a~U[0,0.5t]), Crisfield, ds = 2.6 x 0.5e-3 sqrt(n_free/45) per instance "
                    f"({lengths.min():.4g}..{lengths.max():.4g}), Scaling=0, TolF=1e-8, {steps} steps each")
    log(f"[bench] built {len(sps)} instances in {time.perf_counter() - t_build:.1f} s; dofs "
        f"{min(s.nfree for s in sps)}..{max(s.nfree for s in sps)}, mean {np.mean([s.nfree for s in sps]):.0f}")
    by_size = {}
    for i, sp in enumerate(sps):
        by_size.setdefault(sp.patch.nx, []).append(i)

    if args.only:
        timed = [args.form]
    else:
        timed = [args.form] + [f for f in (args.also if args.also is not None else forms) if f != args.form]

    def batch(problems, ls, **extra):
        t0 = time.perf_counter()
        if len(problems) == 1:
            r = [problems[0].run_alm(steps, float(ls[0]), **kw, **extra)]
        else:
            r, _ = run_alm_batch(problems, steps, ls, **kw, **extra)
        return r, time.perf_counter() - t0

    # learned predictor: trained (untimed) on the FP64 paths of other roofs of the family
    training = None
    net = None
    if any(f.endswith("+learned") for f in timed) or not args.only:
        from gsstructuralanalysis_amd.learned import LearnedPredictor, dataset
        t0 = time.perf_counter()
        ntrain = len(sps) if len(sps) > 1 else 64
        tr_sps = [ShellProblem(g.imperfect_roof_instance(i, seed0=50261202)) for i in range(ntrain)]
        tr_len = np.array([g.roof_arc_length(s.nfree, C2_KAPPA) for s in tr_sps])
        tr, _ = run_alm_batch(tr_sps, C2_STEPS, tr_len, features=True, **kw, **forms["batched"])
        X, Y, grp = dataset([r["features"][:r["steps_done"]] for r in tr], with_groups=True)
        t_data = time.perf_counter() - t0
        net = LearnedPredictor(seed=1)
        t0 = time.perf_counter()
        losses = net.fit(X, Y, epochs=args.epochs, batch=256, lr=1e-3, groups=grp, val_frac=0.125)
        t_fit = time.perf_counter() - t0
        del tr_sps
        training = {"instances": ntrain, "seeds": "50261202 + i", "samples": int(len(X)), "epochs": args.epochs,
                    "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
                    "validation": net.validation,
                    "data_seconds": round(t_data, 2), "train_seconds": round(t_fit, 2)}
        log(f"[bench] learned predictor: {len(X)} samples from {ntrain} roofs in {t_data:.1f} s; "
            f"{args.epochs} epochs in {t_fit:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
        for f in forms:
            if f.endswith("+learned"):
                forms[f]["learned"] = net

    # warm-up: every timed form once over the whole batch and path, untimed.  Code objects and
    # allocations, and the launch autotuner: each (precision, band width, band count, order) bucket a
    # form meets along its path times its three factor schedules on its first calls, so a cold run
    # carries that exploration (its time is reported as first_run_seconds).
    cold = {}
    if args.warmup > 0:
        for f in timed:
            cold[f] = round(batch(sps, lengths, **forms[f])[1], 3)
            log(f"[bench] warm-up {f}: {cold[f]:.3f} s")

    # timed: headline form, profiling off
    res, t_head = batch(sps, lengths, **forms[args.form])
    total_steps = int(sum(r["steps_done"] for r in res))
    value = total_steps / t_head
    log(f"[bench] {args.form}: {total_steps} steps in {t_head:.3f} s -> {value:.2f} steps/s; corrector iterations "
        f"{int(sum(r['iterations'].sum() for r in res))}")

    # kernel table and roofline: a second run of the headline form with per-launch HIP events
    lib.ksh_profile_reset()
    lib.ksh_profile_enable(1)
    res_p, t_prof = batch(sps, lengths, **forms[args.form])
    lib.ksh_profile_enable(0)
    stats = kernel_stats(lib)
    kernels = {}
    for k, (c, ms, work) in stats.items():
        if c:
            log(f"[bench]   {k:26s} launches {c:6d}  total {ms:10.2f} ms  avg {ms / c:9.4f} ms")
            if k in BOUND and work > 0:
                r = roofline(stats, kernel=k)
                kernels[k] = {"launches": c, "total_ms": round(ms, 2), "avg_ms": round(ms / c, 4),
                              "achieved": round(r["achieved"], 4), "unit": r["unit"], "frac": round(r["frac"], 5)}
            else:
                kernels[k] = {"count": c}
    same_profiled = all(np.array_equal(a["negatives"], b["negatives"]) and a["steps_done"] == b["steps_done"]
                        for a, b in zip(res, res_p))

    import ctypes as C
    tuned = {}
    shapes = {"tangent_tiles": lambda v: f"{1 << v} elements per workgroup",
              "band_gather": lambda v: f"{128 << v} threads per block",
              "band_factor": lambda v: ("plain", "lookahead 8 waves x 1 block", "lookahead 8 waves x 2 blocks")[v]}
    for k, shape in shapes.items():
        v, nv = C.c_int64(), C.c_int64()
        lib.ksh_tune_choice(k.encode(), C.byref(v), C.byref(nv))
        tuned[k] = {"variant": int(v.value), "variants": int(nv.value),
                    "shape": shape(v.value) if v.value >= 0 else None,
                    "of": "last call" if k == "band_factor" else "last launch"}

    def summary(rs, secs):
        st = int(sum(r["steps_done"] for r in rs))
        return {"steps_per_s": round(st / secs, 3), "seconds": round(secs, 3), "converged_steps": st,
                "corrector_iterations": int(sum(r["iterations"].sum() for r in rs)),
                "factorizations": int(sum(r["jacobians"] for r in rs)),
                "halvings": int(sum(r["halvings"] for r in rs)),
                "learned_steps": int(sum(r["learned_steps"] for r in rs)),
                "fallbacks": int(sum(r["fallbacks"] for r in rs))}

    # the other forms, each compared instance by instance with the FP64 batched form
    results = {args.form: (res, t_head)}
    for f in timed[1:]:
        results[f] = batch(sps, lengths, **forms[f])
        log(f"[bench] {f}: {summary(*results[f])}")
    others = {}
    r64 = results.get("batched", (None,))[0]
    for f, (rs, secs) in results.items():
        o = summary(rs, secs)
        if r64 is not None and f != "batched":
            o["instances_with_identical_negatives_vs_fp64_batch"] = sum(
                int(a["steps_done"] == b["steps_done"] and np.array_equal(a["negatives"], b["negatives"]))
                for a, b in zip(r64, rs))
            o["max_rel_lambda_diff_vs_fp64_batch"] = max(
                float(np.abs(a["lambda"][:a["steps_done"]] - b["lambda"][:a["steps_done"]]).max()
                      / max(1e-300, np.abs(a["lambda"][:a["steps_done"]]).max()))
                for a, b in zip(r64, rs) if a["steps_done"] == b["steps_done"] and a["steps_done"] > 0)
        if f in cold:
            o["first_run_seconds"] = cold[f]
        if "max_refine" in forms[f]:
            o["refinement"] = {"krylov_steps": forms[f]["max_refine"], "tol": forms[f]["refine_tol"],
                               "accept": forms[f]["refine_accept"]}
        others[f] = o
    same_all = None
    if r64 is not None and args.form != "batched":
        same_all = others[args.form]["instances_with_identical_negatives_vs_fp64_batch"] == len(sps)

    # held-out evaluation of the predictor: tangent, extrapolation only, network -- 64 roofs never trained on
    heldout = None
    if net is not None and len(sps) > 1 and not args.only:
        prec = forms[args.form] if args.form.endswith("+learned") else forms["mixed+learned"]
        base_kw = {k: v for k, v in prec.items() if k != "learned"}
        ho = [ShellProblem(g.imperfect_roof_instance(i, seed0=30261202)) for i in range(64)]
        ho_len = np.array([g.roof_arc_length(s.nfree, C2_KAPPA) for s in ho])
        if args.warmup > 0:  # the 64-instance batch meets buckets of its own
            for extra in (base_kw, dict(prec, learned_mode=1), prec):
                batch(ho, ho_len, **extra)
        rb, tb = batch(ho, ho_len, **base_kw)
        re_, te = batch(ho, ho_len, **prec, learned_mode=1)
        rl, tl = batch(ho, ho_len, **prec)
        sb, se, sl = summary(rb, tb), summary(re_, te), summary(rl, tl)
        heldout = {"instances": 64, "seeds": "30261202 + i", "tangent": sb, "extrapolation_only": se, "learned": sl,
                   "fallback_rate_learned": round(sl["fallbacks"] / max(1, sl["learned_steps"]), 4),
                   "fallback_rate_extrapolation": round(se["fallbacks"] / max(1, se["learned_steps"]), 4),
                   "identical_negatives": {
                       "extrapolation_only": sum(int(np.array_equal(a["negatives"], b["negatives"])) for a, b in zip(rb, re_)),
                       "learned": sum(int(np.array_equal(a["negatives"], b["negatives"])) for a, b in zip(rb, rl))},
                   "max_rel_lambda_diff_learned": max(
                       float(np.abs(a["lambda"] - b["lambda"]).max() / np.abs(a["lambda"]).max())
                       for a, b in zip(rb, rl) if a["steps_done"] == b["steps_done"])}
        log(f"[bench] held-out predictor evaluation: {heldout}")

    # plain FP64 GPU form: one instance at a time, tangent predictor
    plain = None
    plain_info = None
    if not args.no_plain:
        if args.plain_all or len(sps) == 1:
            idx = list(range(len(sps)))
        else:
            idx = [i for _, v in sorted(by_size.items()) for i in v[:args.plain_sample]]
        t_per = {}
        same = True
        for i in idx:
            t0 = time.perf_counter()
            r = sps[i].run_alm(steps, float(lengths[i]), precision=L.FP64, **kw)
            t_per[i] = (time.perf_counter() - t0, r["steps_done"])
            same &= bool(r["steps_done"] == res[i]["steps_done"] and np.array_equal(r["negatives"], res[i]["negatives"]))
        if len(idx) == len(sps):
            t_plain_total = sum(t for t, _ in t_per.values())
            plain_steps = sum(s for _, s in t_per.values())
            how = "measured: every instance one at a time"
        else:
            t_plain_total = 0.0
            for sz, members in sorted(by_size.items()):
                sample = [i for i in idx if sps[i].patch.nx == sz]
                per_step = sum(t_per[i][0] for i in sample) / max(1, sum(t_per[i][1] for i in sample))
                t_plain_total += len(members) * steps * per_step
            plain_steps = steps * len(sps)
            how = (f"estimated: {len(idx)} instances ({args.plain_sample} per size class) each timed over its full "
                   f"{steps}-step path one at a time; per-class seconds per step scaled by the batch's size histogram")
        plain = plain_steps / t_plain_total
        plain_info = {"steps_per_s": round(plain, 3), "how": how, "instances_timed": len(idx),
                      "seconds_timed": round(sum(t for t, _ in t_per.values()), 2),
                      "negatives_identical_to_headline": same}
        log(f"[bench] plain FP64 (one instance at a time): {plain:.2f} steps/s ({how})")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_port(sps, by_size, steps, lengths, args.cpu_budget)
        log(f"[bench] CPU port: {cpu['value']} steps/s")

    c4 = None
    if not args.no_c4 and not args.only:
        c4 = c4_leg()

    amortised = None
    if training is not None and args.form.endswith("+learned"):
        amortised = round(total_steps / (t_head + training["data_seconds"] + training["train_seconds"]), 3)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_head / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if args.form.startswith("batched") else
                  "f32 (K assembly, band LDL^T) + f64 (R, refinement, constraint, count)" if args.form.startswith("fp32k")
                  else "f64 (K, R, refinement, constraint, count) + f32 (band LDL^T)"),
        "data": "synthetic",
        "config": {"workload": workload,
                   "form": {"batched": "FP64 + batched", "mixed": "mixed + batched",
                            "batched+learned": "FP64 + batched + learned predictor",
                            "mixed+learned": "mixed + batched + learned predictor",
                            "fp32k+learned": "FP32 K + FP32 factor + FP64 refinement + batched + learned predictor"}[args.form]
                           + ("" if len(sps) > 1 else ", single instance"),
                   "instances": len(sps), "parallelism": "single GPU"},
        "steps_meaning": (f"arc-length steps of the whole batch (every instance one converged step); "
                          f"{total_steps} converged instance steps in {t_head:.3f} s"),
        "requested_steps": args.steps,
        "instance_steps": total_steps,
        "timed_seconds": round(t_head, 4),
        "plain_fp64_steps_per_s": round(plain, 3) if plain else None,
        "plain_fp64": plain_info,
        "speedup_vs_plain_fp64": round(value / plain, 3) if plain else None,
        "negatives_identical_vs_fp64_batch": same_all,
        "limit_points": limit_points(sps, res),
        "converged_steps": total_steps,
        "corrector_iterations": int(sum(r["iterations"].sum() for r in res)),
        "forms": others,
        "predictor_training": training,
        "amortised_steps_per_s_incl_training": amortised,
        "predictor_heldout": heldout,
        "roofline": roofline(stats),
        "roofline_measured_in": "a second run of the headline form with per-launch HIP events on the launch stream "
                                f"({t_prof:.2f} s; same pivot counts as the timed run: {same_profiled})",
        "autotuner": tuned,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "c4": c4,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
