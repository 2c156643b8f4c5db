"""Serving path: device-resident systems, factor reuse across solves, and the continuous-batching request loop.

The bar for the server is that every request's result is the one the same system gets when it is solved
alone: bitwise for batched small systems (mp_batch_solve on that system only) and for large systems (a fresh
problem solved by the same method and parameters; reused factors are the factors a fresh factorization
produces), whatever else was in flight and whatever was cached or evicted.
"""
import numpy as np
import pytest

from cpsc402_midterm_project_amd import Problem, Server

pytestmark = pytest.mark.gpu
TAU = 1e-15


def _gen(ctx, n, nrhs, kappa, seed):
    dA, dB = ctx.empty(8 * n * n), ctx.empty(8 * n * nrhs)
    ctx.generate_into(dA, dB, None, family="randsvd", n=n, nrhs=nrhs, kappa=kappa, seed=seed, rhs="xtrue_uniform")
    return dA, dB


def test_generate_into_matches_problem_generate(ctx):
    case = dict(family="randsvd", n=300, nrhs=2, kappa=1e3, seed=42, rhs="xtrue_uniform")
    print("[generate_into] n=300", flush=True)
    P = Problem.generate(ctx, **case)
    A0, B0, X0 = P.host_copy()
    dA, dB, dX = ctx.empty(300 * 300 * 8), ctx.empty(300 * 2 * 8), ctx.empty(300 * 2 * 8)
    exact = ctx.generate_into(dA, dB, dX, **case)
    assert exact == (X0 is not None)
    np.testing.assert_array_equal(dA.to_array((300, 300), np.float64), A0)
    np.testing.assert_array_equal(dB.to_array((300, 2), np.float64), B0)
    P.close()


def test_reuse_keeps_factors_and_results(ctx):
    n = 1200
    print(f"[reuse] n={n}", flush=True)
    dA, dB1 = _gen(ctx, n, 1, 1e4, 5)
    _, dB2 = _gen(ctx, n, 2, 1e4, 6)
    params = dict(nb=128, tau=TAU, stop_on="eta")
    P = ctx.problem_from_device(n, 1, dA, dB1)
    X1, r1 = P.solve("mx16x3_gmresir_w64_rdd", reuse=True, **params)
    assert r1["refactorizations"] == 1 and r1["terminal_reason"] == "converged"
    P.set_rhs(dB2, 2)
    X2, r2 = P.solve("mx16x3_gmresir_w64_rdd", reuse=True, **params)
    assert r2["refactorizations"] == 0 and any("(reused)" in s for s in r2["stages"]), r2["stages"]
    assert r2["terminal_reason"] == "converged" and r2["eta_max"] <= TAU
    F = ctx.problem_from_device(n, 2, dA, dB2)
    Xf, rf = F.solve("mx16x3_gmresir_w64_rdd", **params)
    np.testing.assert_array_equal(X2, Xf)
    assert r2["outer_iters"] == rf["outer_iters"]
    P.close(); F.close()


def _alone_small(ctx, dA, dB, n, nrhs):
    A = dA.to_array((n, n), np.float64); B = dB.to_array((n, nrhs), np.float64)
    X, res, _ = ctx.batch_solve([(A, B)], tau=TAU, stop_on="eta")
    return X[0], res[0]


def _alone_large(ctx, dA, dB, n, nrhs, nb):
    P = ctx.problem_from_device(n, nrhs, dA, dB)
    X, rec = P.solve("mx16x3_gmresir_w64_rdd", nb=nb, tau=TAU, stop_on="eta")
    P.close()
    return X, rec


@pytest.mark.parametrize("budget", [200 << 30, 6 << 20])
def test_server_results_equal_alone(ctx, budget):
    """Mixed small and large requests, repeated matrix ids with new right-hand sides; with the small budget
    the cached factors are evicted and rebuilt on the way."""
    sizes = {1: 64, 2: 200, 3: 480, 4: 700, 5: 1100}
    mats = {mid: _gen(ctx, n, 1, 10.0 ** (1 + mid), 100 + mid)[0] for mid, n in sizes.items()}
    plan = [(4, 1), (1, 1), (2, 3), (5, 2), (3, 1), (4, 2), (1, 4), (5, 1), (2, 1), (4, 1), (3, 16), (5, 3)]
    print(f"[server] budget={budget} requests={len(plan)}", flush=True)
    reqs = []
    for i, (mid, nrhs) in enumerate(plan):
        n = sizes[mid]
        _, dB = _gen(ctx, n, nrhs, 10.0, 1000 + i)
        reqs.append((mid, n, nrhs, dB, np.zeros((n, nrhs), order="F")))
    with Server(tau=TAU, hbm_budget_bytes=budget, solve={"nb": 128}) as S:
        ids = {}
        for mid, n, nrhs, dB, X in reqs:
            ids[S.submit(mid, n, nrhs, mats[mid], dB, X)] = (mid, n, nrhs, dB, X)
        done = {}
        for _ in range(600):
            for c in S.poll(wait=0.1):
                done[c["id"]] = c
            if len(done) == len(ids):
                break
        stats = S.stats()
    assert len(done) == len(ids), stats
    assert stats["errors"] == 0, stats
    if budget < (1 << 30):
        assert stats["evictions"] > 0, stats
    else:
        assert stats["cache_hits"] >= 3, stats
    for rid, (mid, n, nrhs, dB, X) in ids.items():
        c = done[rid]
        assert c["terminal"] == "converged" and c["eta_max"] <= TAU, c
        assert c["latency_s"] >= c["service_s"] > 0
        if n <= 512:
            assert c["path"] == "batch"
            Xa, ra = _alone_small(ctx, mats[mid], dB, n, nrhs)
            assert ra["outer_iters"] == c["outer_iters"]
        else:
            assert c["path"] == "large"
            Xa, ra = _alone_large(ctx, mats[mid], dB, n, nrhs, 128)
            assert ra["outer_iters"] == c["outer_iters"]
        np.testing.assert_array_equal(X, Xa)


def test_ladder_reaches_tau_and_resumes_at_final_stage(ctx):
    """kappa = 1e8: the two-term FP16 stage stops short of 1e-15 and the ladder escalates; with reuse and
    resume_ladder a second solve of the same matrix starts at the stage that converged, with its factors."""
    n = 1000
    print(f"[ladder] n={n} kappa=1e8", flush=True)
    dA, dB = _gen(ctx, n, 1, 1e8, 77)
    P = ctx.problem_from_device(n, 1, dA, dB)
    params = dict(nb=128, tau=TAU, stop_on="eta", reuse=True, resume_ladder=True)
    X1, r1 = P.solve("policy_mx16x3_ladder", **params)
    assert r1["terminal_reason"] == "converged" and r1["eta_max"] <= TAU, (r1["terminal_reason"], r1["stages"], r1["error"])
    X2, r2 = P.solve("policy_mx16x3_ladder", **params)
    assert r2["terminal_reason"] == "converged" and r2["refactorizations"] == 0
    assert r2["stages"][0] == r1["final_stage"] + "(reused)", (r1["stages"], r2["stages"])
    if r1["stages"] == [r1["final_stage"]]:       # no escalation: the same stage from the same start
        np.testing.assert_array_equal(X1, X2)
    P.close()


def test_factor_cache_put_get_evict(ctx):
    n = 900
    print(f"[factor cache] n={n}", flush=True)
    dA, dB = _gen(ctx, n, 1, 1e3, 44)
    P = ctx.problem_from_device(n, 1, dA, dB)
    ctx.factor_cache_put(7, P)
    assert not ctx.factor_cache_has(7)
    params = dict(nb=128, tau=TAU, stop_on="eta", reuse=True)
    _, r1 = P.solve("mx16x3_gmresir_w64_rdd", **params)
    assert ctx.factor_cache_has(7) and r1["refactorizations"] == 1
    _, r2 = P.solve("mx16x3_gmresir_w64_rdd", **params)
    assert r2["refactorizations"] == 0
    ctx.factor_cache_evict(7)
    with pytest.raises(Exception, match="not in the factor cache"):
        ctx.factor_cache_has(7)
    _, r3 = P.solve("mx16x3_gmresir_w64_rdd", **params)
    assert r3["refactorizations"] == 1 and r3["terminal_reason"] == "converged"
    ctx.factor_cache_put(8, P)
    P.close()
    with pytest.raises(Exception, match="not in the factor cache"):
        ctx.factor_cache_has(8)


def test_c4_shaped_stream_each_result_equals_its_replay(ctx):
    """A stream shaped like the serving configuration: matrices with n log-uniform in [256, 16384] and kappa
    log-uniform in [1e1, 1e8], 1-16 right-hand sides, 20% of the requests repeating an earlier matrix id, a
    200 GB budget, all submitted at once so small batches and large solves overlap.  Every request converges to
    eta <= 1e-15, and every result equals, bit for bit, the same system solved outside the server: a batch of
    that system alone (n <= 512), or its matrix's requests replayed in arrival order on a fresh problem with the
    server's parameters (reuse and resume_ladder: a repeat starts at the stage that served the matrix last)."""
    rng = np.random.default_rng(9404)
    nreq, npool = 40, 32
    ns = np.exp(rng.uniform(np.log(256), np.log(16384), size=npool)).astype(np.int64)
    ks = 10.0 ** rng.uniform(1, 8, size=npool)
    mats, plan, used = {}, [], []
    for i in range(nreq):
        if used and rng.uniform() < 0.2:
            mid = int(rng.choice(used))
        else:
            mid = len(used); used.append(mid)
        plan.append((mid, int(rng.integers(1, 17))))
    for mid in used:
        mats[mid] = _gen(ctx, int(ns[mid]), 1, float(ks[mid]), 9500 + mid)[0]
    print(f"[c4 stream] {nreq} requests over {len(used)} matrices, n {ns[used].min()}..{ns[used].max()}", flush=True)
    reqs = []
    for i, (mid, nrhs) in enumerate(plan):
        n = int(ns[mid])
        _, dB = _gen(ctx, n, nrhs, 10.0, 2000 + i)
        reqs.append((mid, n, nrhs, dB, np.zeros((n, nrhs), order="F")))
    nb = 256
    with Server(tau=TAU, hbm_budget_bytes=200 << 30, solve={"nb": nb}) as S:
        ids = {}
        for r in reqs:
            ids[S.submit(r[0], r[1], r[2], mats[r[0]], r[3], r[4])] = r
        done = {}
        for _ in range(3000):
            for c in S.poll(wait=0.1):
                done[c["id"]] = c
            if len(done) == len(ids):
                break
        stats = S.stats()
    print(f"[c4 stream] {stats}", flush=True)
    assert len(done) == len(ids) and stats["errors"] == 0, stats
    assert stats["cache_hits"] >= 1
    order = sorted(ids)                                        # request ids follow submission order
    replay = {}
    for rid in order:
        mid, n, nrhs, dB, X = ids[rid]
        c = done[rid]
        assert c["terminal"] == "converged" and c["eta_max"] <= TAU, (rid, n, c)
        if n <= 512:
            Xa, ra = _alone_small(ctx, mats[mid], dB, n, nrhs)
        else:
            if mid not in replay:
                replay[mid] = ctx.problem_from_device(n, nrhs, mats[mid], dB)
            else:
                replay[mid].set_rhs(dB, nrhs)
            Xa, ra = replay[mid].solve("policy_mx16x3_ladder", nb=nb, tau=TAU, stop_on="eta", reuse=True, resume_ladder=True)
        np.testing.assert_array_equal(X, Xa)
    for P in replay.values():
        P.close()
