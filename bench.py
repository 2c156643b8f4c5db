"""Benchmark: n=16384 time-to-FP64-accuracy of the low-precision LU + FGMRES-IR solve vs the plain FP64 LU.

Workload (the single-large-system configuration): randsvd A (n=16384, kappa in {1e2, 1e6}, seed 9200/9201),
one right-hand side b = A x_true; low-precision blocked LU (FP16 trailing update, per-tile power-of-two
scaling, FP32 accumulation on MFMA) + FGMRES-IR (restart 30) with double-double outer residuals, stopping at
normwise backward error eta <= 1e-15; comparators: the plain FP64 LU (rocSOLVER dgetrf + dgetrs) on the same
GPU, and this library's blocked FP64 LU (FP64 MFMA trailing update, vendor panel).  A "step" is one complete solve from resident inputs to the stopped iterate.  `value` is the median
time-to-accuracy of the headline case (kappa=1e6) in ms; every case's timings, iteration counts and errors
are reported alongside.

Further legs report the rest of the metric: the larger system (randsvd n=32768, 4 right-hand sides, kappa in
{1e4, 1e8, 1e10}, FP8 -> FP16 -> FP32 -> FP64 ladder, timed against rocSOLVER's FP64 LU), the uneven batch (4096
systems n <= 512 in one persistent launch) and the serving stream (p50/p99 arrival-to-completion latency at 0.5x/0.8x/0.95x of the measured
saturation rate; serve_p50_ms / serve_p99_ms are the 0.8x figures).

python bench.py [--steps K] [--warmup W] [--n N] [--kappas 1e2,1e6] [--batch-count C] [--serve-requests R]
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


def bench_batch(ctx, count, steps, warmup, cpu):
    """Uneven batch: `count` randsvd systems, n ~ U{32..512}, log10 kappa ~ U[1, 5], seeds 9300 + i, one right-hand
    side each, generated on the GPU into one packed buffer; one mp_batch_solve launch per step (FP32 LU +
    stationary IR with FP64 residuals, eta <= 1e-15 per system)."""
    import concurrent.futures as cf
    import threadpoolctl
    from cpsc402_midterm_project_amd.problems import randsvd
    rng = np.random.default_rng(9300)
    n = rng.integers(32, 513, size=count).astype(np.int32)
    kap = 10.0 ** rng.uniform(1, 5, size=count)
    a_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)[:-1]])
    b_off = np.concatenate([[0], np.cumsum(n.astype(np.int64))[:-1]])
    tot_a, tot_b = int((n.astype(np.int64) ** 2).sum()), int(n.sum())
    Ah, Bh = np.empty(tot_a), np.empty(tot_b)
    t0 = time.time()

    def gen(i):   # host randsvd (numpy QR of seeded Gaussians), written into the packed column-major buffers
        k = int(n[i])
        A, b, _, _ = randsvd(k, float(kap[i]), 9300 + i, nrhs=1, gpu=False)
        Ah[a_off[i]:a_off[i] + k * k] = A.ravel(order="F"); Bh[b_off[i]:b_off[i] + k] = b.ravel(order="F")

    workers = min(16, len(os.sched_getaffinity(0)))
    with threadpoolctl.threadpool_limits(1), cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(gen, range(count)))
    dA, dB, dX = ctx.buffer(Ah), ctx.buffer(Bh), ctx.empty(8 * tot_b)
    log(f"[batch] generated {count} systems (sum n^2 = {tot_a:.3e}) on {workers} host threads in {time.time()-t0:.1f}s")
    def timed(tune):
        recs = []
        for i in range(max(warmup, 1) + steps):      # the first call of the tuned shape runs the autotuner
            res, rec = ctx.batch_solve_packed(n, np.ones(count, np.int32), dA, a_off, dB, b_off, dX, tau=1e-15, stop_on="eta",
                                              tune=tune)
            log(f"  batch tune={tune} {'warm' if i < max(warmup, 1) else 'step'} {i}: {rec['device_s']*1e3:.2f} ms, "
                f"converged {rec['converged']}/{count}, eta_worst {rec['eta_worst']:.2e}, outer total {rec['outer_iters_total']}, "
                f"wg_per_cu {rec['wg_per_cu']} split {rec['split_by_n_bucket']} tuning {rec['tuning'].get('source')}")
            if i == 0:
                first = rec
            if i >= max(warmup, 1):
                recs.append(rec)
        return recs, first
    fixed, _ = timed("off")
    recs, tuned_first = timed("auto")
    ms = statistics.median(r["device_s"] * 1e3 for r in recs)
    ms_fixed = statistics.median(r["device_s"] * 1e3 for r in fixed)
    flops = float((2.0 / 3.0) * (n.astype(np.float64) ** 3).sum())
    out = dict(workload=f"{count} randsvd systems, n~U{{32..512}}, log10 kappa~U[1,5], seeds 9300+i, nrhs=1",
               method="FP32 LU (nb=32, FP32 trailing update) + stationary IR, FP64 residuals, one persistent launch",
               ms_median=round(ms, 3), ms_all=[round(r["device_s"] * 1e3, 3) for r in recs],
               systems_per_s=round(count / (ms * 1e-3), 1), converged=recs[-1]["converged"], systems=count,
               eta_worst=recs[-1]["eta_worst"], outer_iters_total=recs[-1]["outer_iters_total"],
               getrf_flops=flops, getrf_tflops_equiv=round(flops / (ms * 1e-3) / 1e12, 3),
               launch_shape=dict(wg_per_cu=recs[-1]["wg_per_cu"], split_by_n_bucket=recs[-1]["split_by_n_bucket"],
                                 panel_width=recs[-1]["panel_width"], threads=recs[-1]["threads"],
                                 tuning=tuned_first["tuning"]),
               fixed_shape_ms_median=round(ms_fixed, 3), tuned_over_fixed=round(ms_fixed / ms, 3))
    if cpu:
        cores = len(os.sched_getaffinity(0))
        t = time.time()
        with threadpoolctl.threadpool_limits(cores):
            for i in range(count):
                k = int(n[i]); A = Ah[a_off[i]:a_off[i] + k * k].reshape(k, k, order="F")
                np.linalg.solve(A, Bh[b_off[i]:b_off[i] + k])
        tc = time.time() - t
        out["cpu_dgesv_ms"] = round(tc * 1e3, 1)
        out["cpu_cores"] = cores
        log(f"[batch cpu] numpy dgesv over {count} systems: {tc*1e3:.0f} ms on {cores} threads")
    for b in (dA, dB, dX):
        b.free()
    return out


def bench_serve(ctx, nreq, pool_size, loads=(0.5, 0.8, 0.95)):
    """Serving stream (scaled): a pool of `pool_size` device-resident randsvd matrices, n log-uniform in
    [256, 16384], kappa log-uniform in [1e1, 1e8]; each request has 1-16 right-hand sides; 20% of requests repeat
    an earlier request's matrix id (cache hit if its factors are still resident), the others take a new id
    (factorized afresh).  Saturation is measured by a burst of `nreq` requests; then Poisson streams of `nreq`
    requests at 0.5x, 0.8x and 0.95x of it report p50/p99 arrival-to-completion latency and solves/s."""
    from cpsc402_midterm_project_amd import Server
    rng = np.random.default_rng(9400)
    ns = np.exp(rng.uniform(np.log(256), np.log(16384), size=pool_size)).astype(np.int64)
    ks = 10.0 ** rng.uniform(1, 8, size=pool_size)
    t0 = time.time()
    pool = []
    for j in range(pool_size):
        k = int(ns[j])
        dA, dB = ctx.empty(8 * k * k), ctx.empty(8 * k * 16)
        ctx.generate_into(dA, dB, None, family="randsvd", n=k, nrhs=16, kappa=float(ks[j]), seed=9400 + j, rhs="xtrue_uniform")
        pool.append((k, dA, dB))
    log(f"[serve] pool of {pool_size} matrices (n {ns.min()}..{ns.max()}) generated in {time.time()-t0:.1f}s")

    def stream(rate, seed):
        r = np.random.default_rng(seed)
        reqs, ids = [], []
        next_id = 1
        for i in range(nreq):
            if ids and r.random() < 0.2:
                mid, j = ids[int(r.integers(len(ids)))]
            else:
                mid, j = next_id, int(r.integers(pool_size)); next_id += 1
                ids.append((mid, j))
            reqs.append((mid, j, int(r.integers(1, 17))))
        gaps = r.exponential(1.0 / rate, size=nreq) if rate else np.zeros(nreq)
        return reqs, np.cumsum(gaps)

    def run(rate, seed, tag):
        reqs, arr = stream(rate, seed)
        with Server(tau=1e-15, hbm_budget_bytes=200 << 30, solve={"nb": 256}) as S:
            outs = {}
            t_start = Server.clock()
            done = []
            for (mid, j, nrhs), a in zip(reqs, arr):
                k, dA, dB = pool[j]
                while Server.clock() < t_start + a:
                    done += S.poll(wait=min(0.002, max(0.0, t_start + a - Server.clock())))
                X = np.zeros((k, nrhs), order="F")
                rid = S.submit(mid, k, nrhs, dA, dB, X, arrival=t_start + a)
                outs[rid] = X
            while len(done) < len(reqs):
                done += S.poll(wait=0.05)
            t_end = Server.clock()
            st = S.stats()
        lat = np.array([c["latency_s"] for c in done]) * 1e3
        ok = sum(1 for c in done if c["terminal"] == "converged" and c["eta_max"] <= 1e-15)
        res = dict(offered_per_s=None if not rate else round(rate, 2), requests=len(done),
                   solves_per_s=round(len(done) / (t_end - t_start), 2), p50_ms=round(float(np.percentile(lat, 50)), 2),
                   p99_ms=round(float(np.percentile(lat, 99)), 2), mean_ms=round(float(lat.mean()), 2),
                   converged=ok, cache_hits=st["cache_hits"], batches=st["batches"], small=st["small_requests"],
                   large=st["large_requests"], errors=st["errors"])
        log(f"[serve {tag}] {res}")
        return res

    burst = run(None, 9401, "burst")
    sat = burst["solves_per_s"]
    out = dict(workload=f"{nreq} requests per stream; pool of {pool_size} randsvd matrices n log-uniform [256,16384], "
                        f"kappa log-uniform [1e1,1e8], 1-16 rhs, 20% repeated ids, eta<=1e-15, HBM budget 200 GiB",
               saturation_solves_per_s=sat, burst=burst, loads={})
    for f in loads:
        out["loads"][str(f)] = run(f * sat, 9402 + int(f * 100), f"{f}x")
    for _, dA, dB in pool:
        dA.free(); dB.free()
    return out


def bench_c2(ctx, n, kappas, nrhs=4, seed=9202):
    """Larger-system leg: randsvd n=32768 (seed 9202, the context keeps U and V so the kappas share one pair of QR
    factorizations), 4 right-hand sides, kappa in {1e4, 1e8, 1e10}; policy_mx8_ladder (FP8 -> FP16 -> FP32 -> FP64
    factors, FGMRES-IR with double-double residuals, a stage left after one GMRES cycle that gains less than 10x),
    eta <= 1e-15.  One warm-up solve, then one timed solve per kappa; rocSOLVER's FP64 LU + solve on the same system
    timed once beside it (its time does not depend on kappa)."""
    from cpsc402_midterm_project_amd import Problem
    out = dict(workload=f"randsvd n={n} nrhs={nrhs} seed={seed}, kappa in {kappas}: policy_mx8_ladder (nb=256, FGMRES "
                        f"restart 30, DD residual, eta<=1e-15) vs rocSOLVER FP64 LU", kappas={})
    params = dict(nb=256, gmres_m=30, tau=1e-15, stop_on="eta")
    for i, kappa in enumerate(kappas):
        t0 = time.time()
        P = Problem.generate(ctx, family="randsvd", n=n, nrhs=nrhs, kappa=kappa, seed=seed, rhs="xtrue_uniform", keep_q=True)
        P.reference()
        log(f"[c2] kappa={kappa:g} generated + FP64 reference in {time.time() - t0:.1f} s")
        if i == 0:
            _, w = P.solve("policy_mx8_ladder", **params)
            log(f"[c2] warm-up {w['device_ready_wall_s']*1e3:.0f} ms {w['terminal_reason']} stages={w['stages']}")
            recs = []
            for _ in range(2):
                _, r64 = P.solve("native_lu64")
                recs.append(r64)
            out["fp64_lu_ms"] = recs[-1]["device_ready_wall_s"] * 1e3
            out["fp64_lu_eta"] = recs[-1]["eta_max"]
            log(f"[c2] rocSOLVER FP64 LU {out['fp64_lu_ms']:.0f} ms eta={out['fp64_lu_eta']:.2e}")
        _, r = P.solve("policy_mx8_ladder", **params)
        ms = r["device_ready_wall_s"] * 1e3
        out["kappas"][f"{kappa:g}"] = dict(ms=round(ms, 2), terminal=r["terminal_reason"], stages=r["stages"],
                                           stage_times_s=[round(t, 4) for t in r["stage_times_s"]], outer=r["outer_iters"],
                                           inner=r["inner_iters_total"], eta=r["eta_max"], omega=r["omega_max"],
                                           einf_vs_xtrue=max(r["einf"]) if r.get("einf") else None,
                                           finf_vs_fp64_reference=max(r["finf"]) if r.get("finf") else None,
                                           speedup_vs_fp64_lu=round(out["fp64_lu_ms"] / ms, 3))
        log(f"[c2] kappa={kappa:g}: {ms:.0f} ms {r['terminal_reason']} stages={r['stages']} outer={r['outer_iters']} "
            f"inner={r['inner_iters_total']} eta={r['eta_max']:.2e}")
        P.close()
    return out


def pmc_mfma(kernel):
    """Counter evidence for `kernel` from the committed rocprofv3 PMC summaries: MFMA FLOPs per factorization,
    MFMA-busy fraction of SIMD cycles and the rate when the kernel runs alone; None when no summary covers it."""
    for f in sorted((Path(__file__).resolve().parent / "profiles").glob("*_counters.md")):
        for line in f.read_text().splitlines():
            w = line.split()
            if len(w) == 5 and w[0] == "mfma_counters" and w[1] == kernel:
                return dict(flops_per_factorization=float(w[2]), mfma_busy_frac=float(w[3]), alone_tflops=float(w[4]),
                            source=f"profiles/{f.name}")
    return None


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size) from
    the committed rocprofv3 PMC summaries; None when no summary covers it."""
    for f in sorted((Path(__file__).resolve().parent / "profiles").glob("*_traffic.md")):
        for line in f.read_text().splitlines():
            w = line.split()
            if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel:
                return int(w[2])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappas", default="1e2,1e6")
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--methods", default="mx16_gmresir_w64_rdd,mxbf16_gmresir_w64_rdd,mx16x3_gmresir_w64_rdd",
                    help="low-precision candidates (FP16 and, separately, BF16 MFMA trailing updates); the fastest "
                         "converged one is reported")
    ap.add_argument("--fallback", default="mx32_gmresir_w64_rdd")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096, help="uneven-batch leg: systems per launch (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=240, help="serving leg: requests per stream (0: skip)")
    ap.add_argument("--serve-pool", type=int, default=40)
    ap.add_argument("--c2-n", type=int, default=32768, help="larger-system leg: n")
    ap.add_argument("--c2-kappas", default="1e4,1e8,1e10")
    ap.add_argument("--legs", default="headline,batch,serve,c2",
                    help="legs to run (a profiling run may select one; the headline fields are null without it)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this solver build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from cpsc402_midterm_project_amd import Context, Problem

    legs = set(args.legs.split(","))
    n = args.n
    kappas = [float(k) for k in args.kappas.split(",")] if "headline" in legs else []
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
        log("[fp64 lu, own] blocked LU with the FP64 MFMA trailing update (vendor panel)")
        case["fp64_lu_own"] = time_method(P, "mx64_lu_direct", dict(nb=512), 1, 1, require_eta=False)
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
                blas = [f"{i.get('internal_api')} {i.get('version')}" for i in threadpoolctl.threadpool_info()
                        if i.get("user_api") == "blas"]
                times = []
                with threadpoolctl.threadpool_limits(cores):
                    for rep in range(6):                  # one warm-up, then the median of 5
                        t = time.time(); xc = np.linalg.solve(A, b); tc = time.time() - t
                        if rep > 0:
                            times.append(tc)
                r = b - A @ xc
                eta = np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(xc).max() + np.abs(b).max())
                cpu = dict(value=statistics.median(times) * 1e3, unit="ms", cores=cores, kind="port",
                           sample=f"numpy.linalg.solve (LAPACK dgesv, numpy {np.__version__}, {', '.join(blas) or 'BLAS'}) "
                                  f"on the same n={n} kappa={kappa:g} system, median of 5 runs after 1 warm-up "
                                  f"({', '.join(f'{x * 1e3:.0f}' for x in times)} ms), eta={eta:.2e}")
                log(f"[cpu] dgesv median {cpu['value']:.0f} ms on {cores} threads ({blas})")
                # the CPU restatement of the configuration-0 mixed method (FP16-trailing blocked LU, nb=128, FP64-residual
                # stationary IR, n=1024, kappa=1e4, at most 50 steps), single-threaded C + numpy
                from cpsc402_midterm_project_amd.problems import randsvd
                from oracle import oracle as O
                A0, b0, _, _ = randsvd(1024, 1e4, 9200, gpu=False)
                with threadpoolctl.threadpool_limits(1):
                    t = time.time()
                    _, reason0, hist0 = O.ir_stationary(np.asfortranarray(A0), b0, fmt="fp16", nb=128, tau=1e-15, max_outer=50)
                    t0c = time.time() - t
                cpu["c0_mixed_restatement"] = dict(ms=t0c * 1e3, cores=1, terminal=reason0, steps=len(hist0) - 1,
                                                   eta_final=hist0[-1])
                log(f"[cpu] C0 mixed restatement {t0c*1e3:.0f} ms, {reason0} after {len(hist0) - 1} steps")
        P.close()
    batch = None
    if args.batch_count > 0 and "batch" in legs:
        batch = bench_batch(ctx, args.batch_count, args.steps, args.warmup, not args.no_cpu)
    serve = None
    if args.serve_requests > 0 and "serve" in legs:
        serve = bench_serve(ctx, args.serve_requests, args.serve_pool)
    c2 = None
    if "c2" in legs and args.c2_n > 0:
        c2 = bench_c2(ctx, args.c2_n, [float(k) for k in args.c2_kappas.split(",")])
    ctx.close()

    # roofline of the dominant kernel (largest total time among the profiled launches)
    roof = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
        name, v = dom
        per_ms = v["ms"] / v["count"]
        gbs = v["bytes"] / v["ms"] / 1e6
        roof = dict(kernel=name, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(gbs / HBM_PEAK_GBS, 4), traffic=pmc_traffic(name), launches=v["count"], avg_launch_ms=round(per_ms, 4),
                    all_kernels={k: dict(ms=round(x["ms"], 3), count=x["count"],
                                         GBps=round(x["bytes"] / max(x["ms"], 1e-9) / 1e6, 1),
                                         TFps=round(x["flops"] / max(x["ms"], 1e-9) / 1e9, 2)) for k, x in prof.items()})
    # the trailing update (MFMA-bound only at large nb): live rate from the HIP events of its launches during the
    # solve (where it overlaps the next panel), against the dense FP16 peak, with the PMC counter evidence
    mfma = None
    gemm = {k: v for k, v in (prof or {}).items() if "gemm" in k and "64" not in k}
    if gemm:
        gms = sum(v["ms"] for v in gemm.values()); gfl = sum(v["flops"] for v in gemm.values())
        tfs = gfl / max(gms, 1e-9) / 1e9
        mfma = dict(kernel="trailing_gemm", bound="mfma", achieved=round(tfs, 1), peak=2500.0, unit="TFLOP/s",
                    frac=round(tfs / 2500.0, 4), launches=sum(v["count"] for v in gemm.values()), ms=round(gms, 3),
                    counters=pmc_mfma("trailing_gemm"))
    if not cases:
        print(json.dumps({"metric": "n=16384 time-to-FP64-accuracy: low-prec LU-IR vs FP64 LU; serve p50/p99 ms", "value": None,
                          "unit": "ms", "legs": sorted(legs), "batch": batch, "serve": serve, "c2": c2}), flush=True)
        return
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
        "fp64_lu_own_ms": round(hb["fp64_lu_own"]["ms_median"], 3),
        "speedup_vs_fp64_lu_own": None if not hb.get("best") else round(hb["fp64_lu_own"]["ms_median"] / ms, 3),
        "speedup_vs_fp64_lu": None if hb["speedup_vs_fp64_lu"] is None else round(hb["speedup_vs_fp64_lu"], 3),
        "converged": bool(hb.get("best")),
        "serve_p50_ms": serve["loads"]["0.8"]["p50_ms"] if serve else None,
        "serve_p99_ms": serve["loads"]["0.8"]["p99_ms"] if serve else None,
        "serve": serve,
        "batch": batch,
        "c2": c2,
        "roofline": roof,
        "mfma": mfma,
        "cpu_baseline": cpu,
        "cases": cases,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
