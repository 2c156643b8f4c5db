"""Benchmark: n=16384 time-to-FP64-accuracy of the low-precision LU + FGMRES-IR solve vs the plain FP64 LU.

Workload (the single-large-system configuration): randsvd A (n=16384, kappa in {1e2, 1e6}, seed 9200/9201),
one right-hand side b = A x_true; low-precision blocked LU (FP16 trailing update, per-tile power-of-two
scaling, FP32 accumulation on MFMA) + FGMRES-IR (restart 30) with double-double outer residuals, stopping at
normwise backward error eta <= 1e-15; comparator: the plain FP64 LU (rocSOLVER dgetrf + dgetrs) on the same
GPU.  A "step" is one complete solve from resident inputs to the stopped iterate.  `value` is the median
time-to-accuracy of the headline case (kappa=1e6) in ms; every case's timings, iteration counts and errors
are reported alongside.

python bench.py [--steps K] [--warmup W] [--n N] [--kappas 1e2,1e6] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_method(P, method, params, steps, warmup, require_eta=True):
    """Warmup + timed solves of one method.  With require_eta, a method whose first solve does not reach
    eta <= 1e-15 is not repeated (its single record is reported, marked not converged); the FP64 LU
    comparator is always timed in full whatever accuracy it reaches."""
    recs = []
    for i in range(warmup + steps):
        X, rec = P.solve(method, **params)
        ok = (not require_eta) or (rec["terminal_reason"] in ("converged", "direct_solve") and rec["eta_max"] <= 1e-15)
        if i >= warmup or not ok:
            recs.append(rec)
        log(f"  {method} {'warm' if i < warmup else 'step'} {i}: {rec['device_ready_wall_s']*1e3:.2f} ms "
            f"{rec['terminal_reason']} eta={rec['eta_max']:.2e} outer={rec['outer_iters']} inner={rec['inner_iters_total']} "
            f"stages={rec['stages']}")
        if not ok:
            break
    times = [r["device_ready_wall_s"] * 1e3 for r in recs]
    last = recs[-1]
    return dict(method=method, ms_median=statistics.median(times), ms_all=times, terminal=last["terminal_reason"],
                eta=last["eta_max"], omega=last["omega_max"], outer=last["outer_iters"], inner=last["inner_iters_total"],
                factor_s=last["factor_total_s"], stages=last["stages"], einf=last.get("einf"),
                converged=all(r["terminal_reason"] in ("converged", "direct_solve") for r in recs) and last["eta_max"] <= 1e-15)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappas", default="1e2,1e6")
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--methods", default="mx16_gmresir_w64_rdd,mx16x3_gmresir_w64_rdd",
                    help="low-precision candidates (FP16 MFMA trailing update); the fastest converged one is reported")
    ap.add_argument("--fallback", default="mx32_gmresir_w64_rdd")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this solver build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from cpsc402_midterm_project_amd import Context, Problem

    n = args.n
    kappas = [float(k) for k in args.kappas.split(",")]
    common = dict(nb=args.nb, gmres_m=30, gmres_tol=1e-4, tau=1e-15, stop_on="eta", max_outer=50)
    ctx = Context(0)
    caps = ctx.caps()
    log(f"device {caps['device']} {caps['gcn_arch']} CUs={caps['multiprocessors']}")
    cases = []
    headline = None
    prof = None
    cpu = None
    for kappa in kappas:
        seed = 9200 if kappa < 1e4 else 9201
        t0 = time.time()
        P = Problem.generate(ctx, family="randsvd", n=n, nrhs=1, kappa=kappa, seed=seed, rhs="xtrue_uniform",
                             case_id=f"randsvd_n{n}_k{kappa:g}")
        crec = P.case_record()
        log(f"[setup] device randsvd n={n} kappa={kappa:g} seed={seed} in {time.time()-t0:.1f}s "
            f"(gen {crec['gen_seconds']:.1f}s, hash {crec['hash_seconds']:.1f}s)")
        case = dict(kappa=kappa, seed=seed, a_sha256=crec["a_sha256"], generator=crec["generator"])
        log("[fp64 lu] rocSOLVER dgetrf + dgetrs")
        case["fp64_lu"] = time_method(P, "native_lu64", dict(nb=args.nb), args.steps, args.warmup, require_eta=False)
        case["low_precision"] = []
        for method in [m for m in args.methods.split(",") if m]:
            log(f"[low precision] {method}")
            case["low_precision"].append(time_method(P, method, common, args.steps, args.warmup))
        conv = [r for r in case["low_precision"] if r["converged"]]
        if not conv and args.fallback:
            log(f"[fallback] {args.fallback}")
            case["fallback"] = time_method(P, args.fallback, common, args.steps, args.warmup)
            conv = [case["fallback"]] if case["fallback"]["converged"] else []
        best = min(conv, key=lambda r: r["ms_median"]) if conv else None
        case["best"] = best["method"] if best else None
        case["best_ms"] = best["ms_median"] if best else None
        case["speedup_vs_fp64_lu"] = (case["fp64_lu"]["ms_median"] / best["ms_median"]) if best else None
        cases.append(case)
        if kappa == max(kappas):
            headline = case
            # kernel-level timing of one more solve of the headline method (HIP events on the library stream)
            ctx.profile(True)
            P.solve(best["method"] if best else case["low_precision"][0]["method"], **common)
            prof = ctx.profile_report()
            ctx.profile(False)
            if not args.no_cpu:
                import threadpoolctl
                A, b, _ = P.host_copy()
                cores = len(os.sched_getaffinity(0))
                with threadpoolctl.threadpool_limits(cores):
                    t = time.time(); xc = np.linalg.solve(A, b); tc = time.time() - t
                r = b - A @ xc
                eta = np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(xc).max() + np.abs(b).max())
                cpu = dict(value=tc * 1e3, unit="ms", cores=cores, kind="port",
                           sample=f"numpy.linalg.solve (LAPACK dgesv, {np.__version__} OpenBLAS) on the same n={n} "
                                  f"kappa={kappa:g} system, one run, eta={eta:.2e}")
                log(f"[cpu] dgesv {tc*1e3:.0f} ms on {cores} threads")
        P.close()
    ctx.close()

    # roofline of the dominant kernel (largest total time among the profiled launches)
    roof = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
        name, v = dom
        per_ms = v["ms"] / v["count"]
        gbs = v["bytes"] / v["ms"] / 1e6
        roof = dict(kernel=name, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, launches=v["count"], avg_launch_ms=round(per_ms, 4),
                    all_kernels={k: dict(ms=round(x["ms"], 3), count=x["count"],
                                         GBps=round(x["bytes"] / max(x["ms"], 1e-9) / 1e6, 1),
                                         TFps=round(x["flops"] / max(x["ms"], 1e-9) / 1e9, 2)) for k, x in prof.items()})
    hb = headline or cases[-1]
    cands = hb["low_precision"] + ([hb["fallback"]] if hb.get("fallback") else [])
    best = next((r for r in cands if r["method"] == hb.get("best")), None) or cands[0]
    ms = best["ms_median"]
    out = {
        "metric": "n=16384 time-to-FP64-accuracy: low-prec LU-IR vs FP64 LU; serve p50/p99 ms",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp16", "precisions": "fp16 MFMA trailing update (two-term hi+lo operands when needed) with fp32 accumulation; fp32 panel, TRSM and getrs; fp64 and double-double residuals",
        "data": "synthetic randsvd (seeded), generated on the GPU",
        "config": {"workload": f"randsvd n={n} nrhs=1 kappa={hb['kappa']:g}: {best['method']} (nb={args.nb}, FGMRES restart 30, "
                               f"DD residual, eta<=1e-15) vs rocSOLVER FP64 LU", "n": n, "kappas": kappas,
                   "method": best["method"]},
        "fp64_lu_ms": round(hb["fp64_lu"]["ms_median"], 3),
        "speedup_vs_fp64_lu": None if hb["speedup_vs_fp64_lu"] is None else round(hb["speedup_vs_fp64_lu"], 3),
        "converged": bool(hb.get("best")),
        "serve_p50_ms": None, "serve_p99_ms": None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cases": cases,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
