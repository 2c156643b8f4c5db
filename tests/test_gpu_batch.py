"""Batched small systems (mp_batch_solve): parity with the oracle's FP32-LU stationary refinement, system by system.

The kernel runs, for every system, the reference's custom_blocked_lu with nb = 32 and the FP32 trailing format
(src/mp_core.cuh:129-150) followed by Engine::refine_stage with FP64 residuals (src/mp_core.cuh:281-296).  The
oracle runs the same method on the CPU (oracle.ir_stationary(fmt="fp32", nb=32)) with the same operation
order, so the bar is bitwise: the returned X, the number of refinement steps and the terminal reason must equal
the oracle's for each system.  At batch scale the pass criterion is the one of the uneven-batch workload:
every system's normwise backward error, recomputed by the oracle's FP64 metric, is <= tau.
"""
import numpy as np
import pytest

from cpsc402_midterm_project_amd import MPSolveError
from cpsc402_midterm_project_amd.problems import randsvd

pytestmark = pytest.mark.gpu


def _systems(specs, seed0):
    out = []
    for i, (n, nrhs, kappa) in enumerate(specs):
        A, b, _, _ = randsvd(n, kappa, seed0 + i, nrhs=nrhs, gpu=False)
        out.append((A, b))
    return out


def _oracle_solve(oracle, A, B, tau, max_outer, eta_only):
    x, reason, hist = oracle.ir_stationary(np.asfortranarray(A), B, fmt="fp32", nb=32, tau=tau, max_outer=max_outer,
                                           stop_eta_only=eta_only)
    return x, reason, len(hist) - 1


EDGE = [(1, 1, 10.0), (2, 3, 10.0), (31, 1, 1e3), (32, 2, 1e3), (33, 1, 1e4), (63, 4, 1e2), (64, 1, 1e5),
        (65, 5, 1e3), (100, 16, 1e4), (127, 1, 1e5), (200, 2, 1e3), (257, 1, 1e4), (300, 7, 1e5), (511, 1, 1e3),
        (512, 3, 1e5)]


@pytest.mark.parametrize("stop_on", ["eta", "eta_omega"])
def test_batch_matches_oracle_bitwise(ctx, oracle, stop_on):
    print(f"[batch parity] {len(EDGE)} systems stop_on={stop_on}", flush=True)
    systems = _systems(EDGE, 100)
    tau = 1e-15 if stop_on == "eta" else 100 * 2.0 ** -53
    X, res, rec = ctx.batch_solve(systems, tau=tau, stop_on=stop_on, max_outer=30)
    assert rec["systems"] == len(systems) and rec["device_s"] > 0
    for (A, B), x, r, spec in zip(systems, X, res, EDGE):
        xo, reason, outer = _oracle_solve(oracle, A, B, tau, 30, stop_on == "eta")
        assert (r["terminal"], r["outer_iters"]) == (reason, outer), (spec, r, reason, outer)
        np.testing.assert_array_equal(x, xo, err_msg=str(spec))
        m = oracle.metrics(np.asfortranarray(A), B, x)
        assert r["eta_max"] == m["eta_max"] and r["omega_max"] == m["omega_max"], (spec, r, m["eta_max"], m["omega_max"])
        assert r["info"] == 0


def test_batch_alone_equals_in_mix(ctx):
    """A system's result does not depend on what else is in the batch or on the launch shape."""
    specs = [(n, 1 + n % 5, 10.0 ** (1 + n % 5)) for n in (40, 77, 150, 333, 480, 96, 250)]
    systems = _systems(specs, 500)
    print(f"[batch alone-vs-mix] {len(systems)} systems", flush=True)
    Xmix, rmix, _ = ctx.batch_solve(systems, tau=1e-15, stop_on="eta")
    Xone, rone, _ = ctx.batch_solve(systems, tau=1e-15, stop_on="eta", wg_per_cu=1)
    for i, s in enumerate(systems):
        Xa, ra, _ = ctx.batch_solve([s], tau=1e-15, stop_on="eta")
        np.testing.assert_array_equal(Xa[0], Xmix[i])
        np.testing.assert_array_equal(Xone[i], Xmix[i])
        assert ra[0] == rmix[i] == rone[i]


def test_batch_uneven_workload_reaches_tau(ctx, oracle):
    """Uneven batch (n ~ U{32..512}, log10 kappa ~ U[1,5], seeds base + i): every system converges with
    eta <= 1e-15 by the oracle's FP64 metric."""
    rng = np.random.default_rng(7)
    count = 256
    specs = [(int(rng.integers(32, 513)), 1, float(10.0 ** rng.uniform(1, 5))) for _ in range(count)]
    systems = _systems(specs, 9000)
    print(f"[batch uneven] {count} systems", flush=True)
    X, res, rec = ctx.batch_solve(systems, tau=1e-15, stop_on="eta")
    assert rec["converged"] == count, rec
    for (A, B), x, r in zip(systems, X, res):
        m = oracle.metrics(np.asfortranarray(A), B, x)
        assert r["terminal"] == "converged" and m["eta_max"] <= 1e-15, (r, m["eta_max"])


def test_batch_singular_and_zero_rhs(ctx, oracle):
    n = 70
    Z = np.zeros((n, n), order="F")
    A, b, _, _ = randsvd(n, 1e2, 3, gpu=False)
    systems = [(Z, np.ones((n, 1))), (A, np.zeros((n, 2))), (A, b)]
    X, res, _ = ctx.batch_solve(systems, tau=1e-15, stop_on="eta")
    assert res[0]["terminal"] == "singular_factor" and res[0]["info"] == 1
    assert not X[0].any()
    assert res[1]["terminal"] == "converged" and res[1]["outer_iters"] == 0 and not X[1].any()
    assert res[2]["terminal"] == "converged"


def test_batch_rejects_out_of_range(ctx):
    A, b, _, _ = randsvd(8, 10.0, 1, gpu=False)
    with pytest.raises(MPSolveError, match="nrhs"):
        ctx.batch_solve([(A, np.ones((8, 17)))])
    big = np.eye(513, order="F")
    with pytest.raises(MPSolveError, match="512"):
        ctx.batch_solve([(big, np.ones((513, 1)))])
    X, res, rec = ctx.batch_solve([])
    assert X == [] and res == [] and rec["systems"] == 0


def test_batch_full_uneven_workload_every_system_by_oracle(ctx, oracle):
    """The uneven-batch configuration at its full size: 4096 randsvd systems, n ~ U{32..512}, log10 kappa ~
    U[1, 5], seeds 9300 + i (the bench's workload), one packed launch; every system's eta, recomputed by the
    oracle's FP64 metric from the returned X, is <= 1e-15."""
    import concurrent.futures as cf
    import os
    import threadpoolctl
    count = 4096
    rng = np.random.default_rng(9300)
    n = rng.integers(32, 513, size=count).astype(np.int32)
    kap = 10.0 ** rng.uniform(1, 5, size=count)
    a_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)[:-1]])
    b_off = np.concatenate([[0], np.cumsum(n.astype(np.int64))[:-1]])
    Ah, Bh = np.empty(int((n.astype(np.int64) ** 2).sum())), np.empty(int(n.sum()))
    systems = [None] * count

    def gen(i):
        k = int(n[i])
        A, b, _, _ = randsvd(k, float(kap[i]), 9300 + i, nrhs=1, gpu=False)
        Ah[a_off[i]:a_off[i] + k * k] = A.ravel(order="F"); Bh[b_off[i]:b_off[i] + k] = b.ravel(order="F")
        systems[i] = (A, b)

    print(f"[batch full] generating {count} systems", flush=True)
    with threadpoolctl.threadpool_limits(1), cf.ThreadPoolExecutor(min(16, len(os.sched_getaffinity(0)))) as ex:
        list(ex.map(gen, range(count)))
    dA, dB, dX = ctx.buffer(Ah), ctx.buffer(Bh), ctx.empty(8 * Bh.size)
    print(f"[batch full] solving {count} systems", flush=True)
    res, rec = ctx.batch_solve_packed(n, np.ones(count, np.int32), dA, a_off, dB, b_off, dX, tau=1e-15, stop_on="eta")
    assert rec["converged"] == count, rec
    X = dX.to_array((Bh.size,), np.float64)
    worst = 0.0
    for i, (A, b) in enumerate(systems):
        k = int(n[i])
        eta = oracle.metrics(A, b, X[b_off[i]:b_off[i] + k].reshape(k, 1))["eta_max"]
        assert eta <= 1e-15, (i, k, kap[i], eta)
        worst = max(worst, eta)
    print(f"[batch full] every system converged; worst eta by the oracle {worst:.2e}", flush=True)


def test_batch_autotuned_shape_gives_the_same_bits(ctx):
    """The autotuner (launch shape: workgroups per CU, row-split thresholds per n bucket, timed on a sample of
    the batch and cached per batch signature) changes only which thread or workgroup computes what: X and every
    per-system outcome equal the untuned launch bit for bit, the first tuned call reports the candidates it
    timed, the second one the cached choice."""
    rng = np.random.default_rng(41)
    n = rng.integers(20, 513, size=300).astype(np.int32)
    nrhs = rng.integers(1, 4, size=300).astype(np.int32)
    systems = [randsvd(int(k), 10.0 ** rng.uniform(1, 4), 5000 + i, nrhs=int(r), gpu=False)[:2] for i, (k, r) in enumerate(zip(n, nrhs))]
    print(f"[batch tune] {len(systems)} systems", flush=True)
    X0, r0, rec0 = ctx.batch_solve(systems, tau=1e-15, stop_on="eta", tune="off")
    assert rec0["tuning"]["source"] == "off" and rec0["split_by_n_bucket"] == [[64, 128]] * 4
    X1, r1, rec1 = ctx.batch_solve(systems, tau=1e-15, stop_on="eta", tune="auto")
    X2, r2, rec2 = ctx.batch_solve(systems, tau=1e-15, stop_on="eta")          # default: auto, now from the cache
    print(f"[batch tune] off {rec0['device_s']*1e3:.2f} ms; tuned {rec1['device_s']*1e3:.2f} ms wg_per_cu={rec1['wg_per_cu']} "
          f"split={rec1['split_by_n_bucket']} tuning={rec1['tuning']}", flush=True)
    assert rec1["tuning"]["source"] == "tuned" and rec2["tuning"]["source"] == "cache"
    assert rec1["wg_per_cu"] in (1, 2, 3) and rec2["wg_per_cu"] == rec1["wg_per_cu"]
    assert rec2["split_by_n_bucket"] == rec1["split_by_n_bucket"]
    assert set(rec1["tuning"]["buckets"]) == {"1-128", "129-256", "257-384", "385-512"}
    for (a, b, c) in zip(X0, X1, X2):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert r0 == r1 == r2
