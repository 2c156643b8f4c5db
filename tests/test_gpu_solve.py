"""End-to-end solves through mp_solve: stopping semantics and terminal reasons of the reference engine
(src/mp_core.cuh:281-314), final errors checked by the CPU oracle (double-double residual), determinism."""
import numpy as np
import pytest

from cpsc402_midterm_project_amd import Problem
from cpsc402_midterm_project_amd.problems import randsvd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys_k1e4():
    return randsvd(1024, 1e4, 9200, gpu=False)


def test_fgmres_ir_dd_reaches_fp64_accuracy(ctx, oracle, sys_k1e4):
    A, b, xt, _ = sys_k1e4
    print("[solve] mx16_gmresir_w64_rdd n=1024 kappa=1e4", flush=True)
    P = ctx.problem(A, b, x_exact=xt)
    X, rec = P.solve("mx16_gmresir_w64_rdd", nb=128, gmres_m=30, tau=1e-15, stop_on="eta")
    assert rec["ok"] and rec["terminal_reason"] == "converged", rec
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= 1e-15
    # the record's eta is evaluated with the FP64 residual (as the reference's eval_metrics): equal to the
    # oracle's FP64 evaluation up to the two residuals' rounding, each within (n+2) u T (T = |A||x| + |b|)
    m64 = oracle.metrics(A, b, X, dd=False)
    _, T = oracle.resid_fp64(A, X, b)
    n = A.shape[0]
    den = np.abs(A).sum(axis=1).max() * np.abs(X).max() + np.abs(b).max()
    assert abs(m64["eta_max"] - rec["eta_max"]) <= 2 * (n + 2) * 2.0 ** -53 * T.max() / den
    # forward error against x_true is at the kappa * u level
    assert rec["einf"][0] < 1e4 * 2.0 ** -53 * 50
    hist = P.history()
    assert hist and hist[0]["kind"] == "outer" and hist[-1]["eta_max"] == rec["solver_final_eta_max"]


def test_solve_is_deterministic(ctx, sys_k1e4):
    A, b, xt, _ = sys_k1e4
    P = ctx.problem(A, b)
    X1, r1 = P.solve("mx16_gmresir_w64_rdd", nb=128, gmres_m=30, tau=1e-15, stop_on="eta")
    X2, r2 = P.solve("mx16_gmresir_w64_rdd", nb=128, gmres_m=30, tau=1e-15, stop_on="eta")
    assert r1["x_sha256"] == r2["x_sha256"]
    np.testing.assert_array_equal(X1, X2)


@pytest.mark.parametrize("method", ["hetero16_ir_w64_r64", "custom32_ir_w64_r64", "native_ir32_w64_r64", "native_ir64_w64_r64",
                                    "mxbf16_gmresir_w64_rdd", "mx16_ir_w64_rdd", "native_gmresir32_w64_r64"])
def test_reference_methods_converge_kappa1e2(ctx, oracle, method):
    A, b, xt, _ = randsvd(768, 1e2, 9201, gpu=False)
    print(f"[solve] {method} n=768 kappa=1e2", flush=True)
    P = ctx.problem(A, b, x_exact=xt)
    X, rec = P.solve(method, nb=128)
    assert rec["ok"] and rec["terminal_reason"] == "converged", (rec["terminal_reason"], rec["error"], rec["eta_max"])
    tau = 100 * 2.0 ** -53
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= tau and m["omega_max"] <= tau


@pytest.mark.parametrize("method", ["native_leftgmres32_app32_w64_r64", "native_leftgmres32_app64_w64_r64",
                                    "native_gmresir32_app64_w64_r64", "hetero16_gmresir_app64_w64_r64",
                                    "hetero16_leftgmres_app64_w64_r64"])
def test_left_gmres_and_promoted_application(ctx, oracle, method):
    """Left-preconditioned GMRES (Krylov space of F^-1 A from F^-1 r, correction d += V y) and the FP64
    application of promoted FP32 factors (src/mp_core.cuh:178-186, 236-278) reach the FP64 stopping test."""
    A, b, xt, _ = randsvd(640, 1e4, 9210, gpu=False)
    print(f"[solve] {method} n=640 kappa=1e4", flush=True)
    P = ctx.problem(A, b, x_exact=xt)
    X, rec = P.solve(method, nb=128)
    assert rec["ok"] and rec["terminal_reason"] == "converged", (rec["terminal_reason"], rec["error"], rec["eta_max"])
    assert rec["precond_applies"] > 0 and rec["matvecs"] > 0
    tau = 100 * 2.0 ** -53
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= tau and m["omega_max"] <= tau


def _exact_lu_system(n, seed):
    """A = L U with dyadic factors that one FP32 panel (nb >= n) reproduces exactly: unit L with |l| <= 1/4,
    U with diagonal +-{1, 2, 4} and off-diagonals in {0, +-1/8}, so partial pivoting keeps the natural order
    and every quotient, product and sum is exact.  (Blocked factorizations are not exact here: the triangular
    solve for U12 goes through the inverse of L11, whose entries need more than 24 bits.)"""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.choice([-0.25, 0.0, 0.25], size=(n, n)), -1) + np.eye(n)
    U = np.triu(rng.choice([-0.125, 0.0, 0.125], size=(n, n)), 1)
    U += np.diag(rng.choice([-4.0, -2.0, -1.0, 1.0, 2.0, 4.0], size=n))
    return np.asfortranarray(L @ U), rng.standard_normal((n, 1))


def test_promoted_application_is_fp64(ctx, oracle):
    """The initial solve x0 = F^-1 b with FP32 factors that are exactly A's: applied in FP64 (apply64) x0 has an
    FP64-level backward error and the stage stops converged before any refinement step; applied in FP32 it
    carries FP32 rounding."""
    A, b = _exact_lu_system(64, 12)
    P = ctx.problem(A, b)
    f = oracle.factor(A, fmt="fp16", nb=64, scale_mode=1, tile_scale=False)
    L, U = oracle.unpack_lu(f["LU"])
    assert f["info"] == 0 and np.array_equal(L @ U, f["g"] * A)       # the factors are exactly those of A
    X64, r64 = P.solve("hetero16_leftgmres_app64_w64_r64", nb=64, max_outer=0)
    X32, r32 = P.solve("hetero16_gmresir_w64_r64", nb=64, max_outer=0)
    e64 = oracle.metrics(A, b, X64, dd=True)["eta_max"]
    e32 = oracle.metrics(A, b, X32, dd=True)["eta_max"]
    assert r64["terminal_reason"] == "converged" and r64["outer_iters"] == 0, (r64["terminal_reason"], e64)
    assert e64 <= 64 * 2.0 ** -53, e64
    assert e32 > 1e-10 and r32["terminal_reason"] == "max_outer", (e32, r32["terminal_reason"])
    np.testing.assert_allclose(X64, np.linalg.solve(A, b), rtol=0, atol=1e-12 * np.abs(X64).max())


def test_direct_fp64_lu(ctx, oracle):
    A, b, xt, _ = randsvd(1000, 1e3, 9202, gpu=False)
    P = ctx.problem(A, b)
    X, rec = P.solve("native_lu64")
    assert rec["terminal_reason"] == "direct_solve" and rec["outer_iters"] == 0
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] < 1e-15
    ref = np.linalg.solve(A, b)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e3 * 1e-13


def test_singular_and_unsupported_outcomes(ctx):
    n = 256
    A = np.ones((n, n)); b = np.ones((n, 1))
    P = ctx.problem(A, b)
    X, rec = P.solve("mx16_ir_w64_r64", nb=64)
    assert rec["terminal_reason"] == "singular_factor" and rec["factor_info"] == 2
    X, rec = P.solve("cusolver_irs_m64_l16_classical_fb_default")
    assert not rec["ok"] and rec["terminal_reason"] == "unsupported_on_platform"
    X, rec = P.solve("magma_ht_ir_w64_r64")
    assert rec["terminal_reason"] == "unsupported_on_platform"


def _rowcol_scaled(n, span, seed):
    """A = Dr0 B Dc0 with B randsvd (kappa 10) and power-of-two Dr0, Dc0 drawn from 2^[-span, span]."""
    rng = np.random.default_rng(seed)
    B0, _, _, _ = randsvd(n, 10.0, seed, gpu=False)
    A = np.ldexp(1.0, rng.integers(-span, span + 1, n))[:, None] * B0 * np.ldexp(1.0, rng.integers(-span, span + 1, n))[None, :]
    return np.asfortranarray(A), rng.standard_normal((n, 2))


@pytest.mark.parametrize("method,fmt", [("hetero16_ir_w64_r64_twosided", "fp16"), ("native_ir32_w64_r64_twosided", None)])
def test_two_sided_scaling(ctx, oracle, method, fmt):
    """Scale mode 2 on a system whose rows and columns span 2^+-24: the equilibrated factors solve it to the
    FP64 stopping test.  The initial solve x0 = g Dc F^-1 (Dr b) is checked against the oracle's equilibration:
    with the oracle's scales and (for the own factorization, bitwise equal to the oracle's) factors,
    y = x0 / (g dc) solves L U y = P (dr b) to FP32 backward accuracy, which fails for any wrong scale."""
    A, b = _rowcol_scaled(384, 24, 31)
    print(f"[solve] {method} n=384 span=2^24", flush=True)
    P = ctx.problem(A, b)
    X, rec = P.solve(method, nb=64)
    assert rec["ok"] and rec["terminal_reason"] == "converged", (rec["terminal_reason"], rec["error"], rec["eta_max"])
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= 100 * 2.0 ** -53 and m["omega_max"] <= 100 * 2.0 ** -53
    if fmt is None:
        return
    X0, _ = P.solve(method, nb=64, max_outer=0)
    F = oracle.factor(A, fmt, 64, scale_mode=2, tile_scale=False)
    L, U = oracle.unpack_lu(F["LU"])
    y = X0 / (F["g"] * F["dc"][:, None])
    rhs = (F["dr"][:, None] * b)[F["perm"]]
    res = np.abs(L @ (U @ y) - rhs).max(axis=0)
    scale = (np.abs(L) @ (np.abs(U) @ np.abs(y))).max(axis=0) + np.abs(rhs).max(axis=0)
    assert np.all(res / scale <= 4 * 384 * 2.0 ** -24), res / scale


def test_unattainable_tau_stagnates(ctx):
    A, b, xt, _ = randsvd(512, 1e2, 3, gpu=False)
    P = ctx.problem(A, b)
    X, rec = P.solve("mx16_ir_w64_r64", nb=64, tau_mult=0.001)
    assert rec["terminal_reason"] in ("stagnation", "max_outer")


def test_precision_ladder_escalates(ctx, oracle):
    A, b, xt, _ = randsvd(512, 1e6, 4, gpu=False)
    print("[solve] policy_mx8_ladder n=512 kappa=1e6", flush=True)
    P = ctx.problem(A, b)
    X, rec = P.solve("policy_mx8_ladder", nb=64, gmres_m=30, gmres_budget=120, tau=1e-15, stop_on="eta")
    assert rec["used_fallback"] and rec["refactorizations"] >= 2, rec["stages"]
    assert rec["terminal_reason"] == "converged", (rec["stages"], rec["terminal_reason"])
    assert oracle.metrics(A, b, X, dd=True)["eta_max"] <= 1e-15


@pytest.mark.parametrize("method", ["mx16x3_gmresir_w64_rdd", "policy_mx16_ladder"])
def test_two_term_fp16_kappa1e6(ctx, oracle, method):
    """Two-term FP16 operands (hi + lo) give FP32-level factors on FP16 MFMA, which FGMRES-IR takes to
    FP64 accuracy at kappa=1e6 where single-term FP16 stagnates."""
    A, b, xt, _ = randsvd(1024, 1e6, 9201, gpu=False)
    print(f"[solve] {method} n=1024 kappa=1e6", flush=True)
    P = ctx.problem(A, b, x_exact=xt)
    X, rec = P.solve(method, nb=128, gmres_m=30, tau=1e-15, stop_on="eta", max_outer=50)
    assert rec["terminal_reason"] == "converged", (rec["stages"], rec["terminal_reason"], rec["eta_max"])
    assert oracle.metrics(A, b, X, dd=True)["eta_max"] <= 1e-15
    if method == "mx16x3_gmresir_w64_rdd":
        assert rec["refactorizations"] == 1 and rec["outer_iters"] <= 6


def test_multiple_rhs_stationary(ctx, oracle):
    A, b, xt, _ = randsvd(640, 1e3, 6, nrhs=5, gpu=False)
    P = ctx.problem(A, b, x_exact=xt)
    X, rec = P.solve("mx16_ir_w64_rdd", nb=64)
    assert rec["terminal_reason"] == "converged" and len(rec["eta"]) == 5
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= 100 * 2.0 ** -53


def test_problem_outliving_its_context(oracle):
    """A problem freed after its context is destroyed (the order a garbage collector may choose) must not
    touch the destroyed context; using it with another context is an error, not a crash."""
    from cpsc402_midterm_project_amd import Context, MPSolveError
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((64, 64)))
    b = rng.standard_normal(64)
    c1 = Context(0)
    P = c1.problem(A, b)
    c1.close()
    c2 = Context(0)
    try:
        with pytest.raises(MPSolveError, match="another context"):
            P.ctx = c2
            P.solve("mx16_gmresir_w64_rdd", nb=32)
    finally:
        P.close()
        c2.close()


def test_left_gmres_ir_against_oracle_restatement(ctx, oracle):
    """Left-preconditioned GMRES-IR with FP32 factors and FP64 residuals against the oracle's left variant of
    Engine::gmres_solve (src/mp_core.cuh:236-278) on the same system: both converge, with the same number of
    outer steps (within 1) and inner iterations within 25% (the factorizations round differently)."""
    A, b, xt, _ = randsvd(512, 1e4, 9212, gpu=False)
    print("[solve] native_leftgmres32_app32_w64_r64 vs oracle n=512 kappa=1e4", flush=True)
    P = ctx.problem(A, b)
    X, rec = P.solve("native_leftgmres32_app32_w64_r64", nb=64, tau=1e-15, stop_on="eta")
    xo, reason, hist, inner = oracle.fgmres_ir(np.asfortranarray(A), b, fmt="fp32", nb=64, tau=1e-15, m=40, dd=False,
                                               tile_scale=False, left=True, cycles=1)
    print(f"[solve] gpu {rec['terminal_reason']} outer={rec['outer_iters']} inner={rec['inner_iters_total']}; "
          f"oracle {reason} outer={len(hist) - 1} inner={inner}", flush=True)
    assert rec["terminal_reason"] == reason == "converged"
    assert abs(rec["outer_iters"] - (len(hist) - 1)) <= 1
    assert abs(rec["inner_iters_total"] - inner) <= max(2, 0.25 * inner)
    assert oracle.metrics(A, b, X)["eta_max"] <= 1e-15


@pytest.mark.parametrize("method,fmt,kappa", [("mx16_gmresir_w64_rdd", "fp16", 1e3), ("mx32_gmresir_w64_rdd", "fp32", 1e6)])
def test_fgmres_ir_against_oracle_restatement(ctx, oracle, method, fmt, kappa):
    """Right-preconditioned FGMRES-IR with double-double residuals against the oracle's restatement
    (src/mp_core.cuh:237-278 over the blocked LU with the same trailing format): both converge, outer steps
    within 1, inner iterations within 30% (the MFMA accumulation rounds differently from the oracle's
    double-precision sums, so the factors are not bitwise equal on random matrices)."""
    A, b, xt, _ = randsvd(512, kappa, 9213, gpu=False)
    print(f"[solve] {method} vs oracle n=512 kappa={kappa:g}", flush=True)
    P = ctx.problem(A, b)
    X, rec = P.solve(method, nb=64, tau=1e-15, stop_on="eta")
    xo, reason, hist, inner = oracle.fgmres_ir(np.asfortranarray(A), b, fmt=fmt, nb=64, tau=1e-15, m=40, dd=True)
    print(f"[solve] gpu {rec['terminal_reason']} outer={rec['outer_iters']} inner={rec['inner_iters_total']}; "
          f"oracle {reason} outer={len(hist) - 1} inner={inner}", flush=True)
    assert rec["terminal_reason"] == reason == "converged"
    assert abs(rec["outer_iters"] - (len(hist) - 1)) <= 1
    assert abs(rec["inner_iters_total"] - inner) <= max(3, 0.3 * inner)
    assert oracle.metrics(A, b, X, dd=True)["eta_max"] <= 1e-15


REFERENCE_IDS = [  # src/mpsolve.cu:20-60 (the vendor IDs are covered by test_singular_and_unsupported_outcomes)
    "native_lu32", "native_lu64", "native_lu64_promoted32", "native_ir32_w32_r64", "native_ir32_w32_r32",
    "native_ir32_w64_r64", "native_ir64_w64_r64", "native_ir32_w64_r64_twosided", "native_ir32_w64_r64_global",
    "native_gmresir32_w64_r64", "native_gmresir32_w32_r64", "native_leftgmres32_app32_w64_r64",
    "native_leftgmres32_app64_w64_r64", "native_gmresir32_app64_w64_r64", "hetero16_lu_direct", "hetero16_ir_w64_r64",
    "hetero16_ir_w32_r64", "hetero16_ir_w64_r64_noscale", "hetero16_ir_w64_r64_twosided", "hetero16_gmresir_w64_r64",
    "hetero16_gmresir_w32_r64", "hetero16_gmresir_app64_w64_r64", "hetero16_leftgmres_app64_w64_r64",
    "custom32_lu_direct", "custom32_ir_w64_r64", "custom32_ir_w32_r64", "policy_hetero16_stationary_only",
    "policy_hetero16_gmres_only", "policy_hetero16_escalate_immediate", "policy_hetero16_cost_aware",
    "policy_native32_escalate_immediate"]


def test_every_reference_method_on_an_fp64_system(ctx, oracle):
    """Every method ID of the reference on one FP64-stored system (n=384, kappa=1e2), default parameters
    (tau = 100 u64 on eta and omega): direct methods report direct_solve at their factor precision; FP64-working
    methods converge by the oracle's metric; FP32-working methods stop at the FP32 level (their iterate cannot
    hold FP64 accuracy), and FP32 residual control needs FP32-stored data (runtime_error, as the reference's
    MPError would give)."""
    A, b, xt, _ = randsvd(384, 1e2, 9214, gpu=False)
    P = ctx.problem(A, b)
    tau = 100 * 2.0 ** -53
    for mid in REFERENCE_IDS:
        X, rec = P.solve(mid, nb=64)
        t = rec["terminal_reason"]
        if mid == "native_ir32_w32_r32":
            assert t == "runtime_error" and not rec["ok"] and "fp32 stored" in rec["error"], (mid, t, rec["error"])
            continue
        m = oracle.metrics(A, b, X, dd=True)
        if "lu" in mid and "direct" in mid or mid.startswith("native_lu"):
            assert t == "direct_solve", (mid, t)
            bound = 1e-15 if mid.startswith("native_lu64") else (1e-2 if mid.startswith("hetero16") else 1e-5)
            assert m["eta_max"] <= bound, (mid, m["eta_max"])      # FP64 / FP32 / FP16-operand factors
        elif "_w32_" in mid:
            assert t in ("stagnation", "max_outer", "converged"), (mid, t)
            assert m["eta_max"] <= 1e-6, (mid, m["eta_max"])
        else:
            assert t == "converged", (mid, t, rec["stages"], rec["error"])
            assert m["eta_max"] <= tau and m["omega_max"] <= tau, (mid, m["eta_max"], m["omega_max"])


def test_fp32_residual_control_on_an_fp32_system(ctx, oracle):
    """native_ir32_w32_r32 on an FP32-stored system (target FP32): FP32 residuals, omega not evaluated (-1 in
    the history), converged to 100 u32 by the oracle's FP64 metric of the stored system."""
    P = Problem.generate(ctx, family="randsvd", n=300, nrhs=2, kappa=1e2, seed=9215, rhs="random", dtype="fp32")
    X, rec = P.solve("native_ir32_w32_r32", nb=64)
    assert rec["terminal_reason"] == "converged", (rec["terminal_reason"], rec["error"])
    hist = [h for h in P.history() if h.get("kind") == "outer"]
    assert hist and all(h["omega_max"] == -1 or h["omega_max"] is None for h in hist), hist[:2]
    A, B, _ = P.host_copy()
    m = oracle.metrics(A.astype(np.float64), B, X)
    assert m["eta_max"] <= 100 * 2.0 ** -24, m["eta_max"]
    P.close()


@pytest.mark.parametrize("mid,ladder", [
    ("policy_hetero16_escalate_immediate", ["het16", "cs32", "cs64"]),
    ("policy_hetero16_budget120", ["het16", "cs32", "cs64"]),
    ("policy_hetero16_cost_aware", ["het16", "cs32", "cs64"]),
    ("policy_native32_escalate_immediate", ["cs32", "cs64"])])
def test_reference_policies_escalate(ctx, oracle, mid, ladder):
    """The reference's policies at kappa=1e6, where FP16-operand factors cannot reach the FP64 target:
    they escalate along het16 -> cs32 -> cs64 (cs32 -> cs64), src/mp_core.cuh:299-314, stop at the first stage
    that converges, and the final iterate meets tau = 100 u64 by the oracle's metric."""
    A, b, xt, _ = randsvd(384, 1e6, 9216, gpu=False)
    print(f"[policy] {mid} n=384 kappa=1e6", flush=True)
    P = ctx.problem(A, b)
    X, rec = P.solve(mid, nb=64)
    assert rec["terminal_reason"] == "converged", (mid, rec["terminal_reason"], rec["stages"])
    st = rec["stages"]
    assert st == ladder[:len(st)] and rec["final_stage"] == st[-1], (mid, st)
    assert rec["used_fallback"] == (len(st) > 1)
    tau = 100 * 2.0 ** -53
    m = oracle.metrics(A, b, X, dd=True)
    assert m["eta_max"] <= tau and m["omega_max"] <= tau, (mid, m["eta_max"], m["omega_max"])


def test_instrumented_history(ctx):
    """--instrument (src/mp_core.cuh:264,270): one inner record per Arnoldi step with its recursive relative
    residual, and a cycle_end record with the true relative residual of the correction equation."""
    A, b, xt, _ = randsvd(512, 1e4, 9217, gpu=False)
    P = ctx.problem(A, b)
    X, rec = P.solve("mx16_gmresir_w64_rdd", nb=64, tau=1e-15, stop_on="eta", instrument=True)
    assert rec["terminal_reason"] == "converged"
    hist = P.history()
    inner = [h for h in hist if h["kind"] == "inner" and h["note"] != "cycle_end"]
    ends = [h for h in hist if h["kind"] == "inner" and h["note"] == "cycle_end"]
    outer = [h for h in hist if h["kind"] == "outer"]
    assert len(inner) == rec["inner_iters_total"] > 0
    assert len(ends) >= rec["outer_iters"] and len(outer) == rec["outer_iters"] + 1
    for e in ends:
        assert e["true_rel"] is not None and 0 <= e["true_rel"] < 1
    by_cycle = {}
    for h in inner:
        by_cycle.setdefault((h["outer"], h["rhs"], h["cycle"]), []).append(h["recursive_rel"])
    for seq in by_cycle.values():
        assert all(b <= a * (1 + 1e-12) for a, b in zip(seq, seq[1:])), seq     # GMRES residuals never grow
    plain = P.solve("mx16_gmresir_w64_rdd", nb=64, tau=1e-15, stop_on="eta")
    assert not [h for h in P.history() if h["kind"] == "inner"]
    np.testing.assert_array_equal(plain[0], X)                                  # instrumentation changes nothing


def test_memory_admission_skips_trial(ctx):
    """Admission (src/mpsolve.cu:182-190): a trial whose estimated extra device memory exceeds the budget is
    recorded as memory_skip with its reason and no solve; within budget it runs."""
    A, b, xt, _ = randsvd(256, 1e2, 9218, gpu=False)
    P = ctx.problem(A, b)
    X, rec = P.solve("mx16_gmresir_w64_rdd", nb=64, hbm_budget_bytes=1 << 10)
    assert rec["terminal_reason"] == "memory_skip" and rec["skipped"] and "memory_skip" in rec["skip_reason"], rec
    assert rec["outer_iters"] == 0 and rec["refactorizations"] == 0
    X, rec = P.solve("mx16_gmresir_w64_rdd", nb=64, hbm_budget_bytes=64 << 30)
    assert rec["terminal_reason"] == "converged" and not rec["skipped"]


@pytest.mark.parametrize("n,dtype,method", [(1000, "fp64", "mx16x3_gmresir_w64_rdd"), (6144, "fp64", "mx16_gmresir_w64_rdd"),
                                            (777, "fp32", "native_gmresir32_w64_r64")])
def test_fgmres_product_beside_upper_solve(ctx, n, dtype, method, monkeypatch):
    """Each FGMRES step's product w = A z can be formed by extra workgroups of the upper substitution as its
    blocks appear (column chunks of 64 or 152 columns here, so some straddle two 256-row blocks, and a partial
    last block at n=777).  Its chunk partials are the residual kernel's, in the same order, so the iterate, the
    iteration counts and the residual history must be the same bits as with the separate product
    (MPSOLVE_GETRS_MV=0)."""
    print(f"[fgmres product] {method} n={n} {dtype}", flush=True)
    P = Problem.generate(ctx, family="randsvd", n=n, nrhs=1, kappa=1e4, seed=9230 + n, rhs="random", dtype=dtype)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_GETRS_MV", mode)
        X, rec = P.solve(method, nb=128, tau=1e-15, stop_on="eta")
        assert rec["ok"] and rec["terminal_reason"] == "converged", (mode, rec["terminal_reason"])
        assert rec["inner_iters_total"] > 0
        out[mode] = (X, rec, [h.get("eta_max") for h in P.history() if h.get("kind") == "outer"])
    P.close()
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    for k in ("inner_iters_total", "outer_iters", "matvecs", "x_sha256"):
        assert out["0"][1][k] == out["1"][1][k], k
    assert out["0"][2] == out["1"][2]


@pytest.mark.parametrize("n,dtype,method", [(1000, "fp64", "mx16x3_gmresir_w64_rdd"), (1500, "fp64", "hetero16_gmresir_w64_r64"),
                                            (777, "fp32", "native_gmresir32_w64_r64"),
                                            (900, "fp64", "native_leftgmres32_app32_w64_r64")])
def test_fgmres_check_and_next_residual_share_a_pass(ctx, n, dtype, method, monkeypatch):
    """The FP64 check r - A d at the end of an FGMRES cycle and the residual of x + d (double-double or FP64),
    which the next outer step evaluates after x += d, can be formed in one pass over A.  Each keeps its own
    kernel's sums in order, so the iterate, the counts and the history must be the same bits as with the two
    passes apart (MPSOLVE_RESID_PAIR=0), for the right- and left-preconditioned solves and FP32 storage."""
    print(f"[resid pair] {method} n={n} {dtype}", flush=True)
    P = Problem.generate(ctx, family="randsvd", n=n, nrhs=1, kappa=1e4, seed=9240 + n, rhs="random", dtype=dtype)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_RESID_PAIR", mode)
        X, rec = P.solve(method, nb=128, tau=1e-15, stop_on="eta")
        assert rec["ok"] and rec["terminal_reason"] == "converged", (mode, rec["terminal_reason"])
        assert rec["outer_iters"] >= 2 and rec["inner_iters_total"] > 0, (rec["outer_iters"], rec["inner_iters_total"])
        hist = [(h.get("eta_max"), h.get("omega_max"), h.get("note")) for h in P.history() if h.get("kind") == "outer"]
        out[mode] = (X, rec, hist)
    P.close()
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    for k in ("inner_iters_total", "outer_iters", "matvecs", "true_residual_checks", "x_sha256", "eta", "omega", "rinf",
              "worst_row"):
        assert out["0"][1][k] == out["1"][1][k], k
    assert out["0"][2] == out["1"][2]


def test_cgs_update_and_next_dots_fused(ctx, monkeypatch):
    """The first CGS pass's update of w and the second pass's dot products in one launch per 256-row block: the
    iterate, counts and history must be the same bits as with the two launches apart (MPSOLVE_CGS_FUSE=0)."""
    print("[cgs fuse] mx16x3_gmresir_w64_rdd n=1200", flush=True)
    P = Problem.generate(ctx, family="randsvd", n=1200, nrhs=1, kappa=1e5, seed=9260, rhs="random", dtype="fp64")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_CGS_FUSE", mode)
        X, rec = P.solve("mx16x3_gmresir_w64_rdd", nb=128, tau=1e-15, stop_on="eta")
        assert rec["ok"] and rec["inner_iters_total"] > 3, (mode, rec["terminal_reason"], rec["inner_iters_total"])
        hist = [(h.get("eta_max"), h.get("note")) for h in P.history() if h.get("kind") == "outer"]
        out[mode] = (X, rec, hist)
    P.close()
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    for k in ("terminal_reason", "inner_iters_total", "outer_iters", "x_sha256", "eta"):
        assert out["0"][1][k] == out["1"][1][k], k
    assert out["0"][2] == out["1"][2]
