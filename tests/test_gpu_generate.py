"""Seeded test-system generation (mp_problem_generate) against the oracle's restatement of the reference's
generators (src/mp_core.cuh:29-103, src/mp_kernels.cuh:39-49): the prescribed family's rotation stages, the
host streams for random right-hand sides, row/column scales and unit columns, and the dyadic counter hash.
The bar is bitwise, signed zeros included."""
import numpy as np
import pytest

from cpsc402_midterm_project_amd import Problem

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64); b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.view(np.uint64).tobytes() == b.view(np.uint64).tobytes()


@pytest.mark.parametrize("n,spectrum", [(1, "geometric"), (2, "geometric"), (17, "clustered"), (64, "geometric"),
                                        (65, "one_small")])
def test_prescribed_and_random_rhs_match_oracle(ctx, oracle, n, spectrum):
    seed = 9200 + n
    print(f"[generate] prescribed n={n} spectrum={spectrum}", flush=True)
    P = Problem.generate(ctx, family="prescribed", spectrum=spectrum, n=n, nrhs=2, kappa=1e3, seed=seed, rhs="random")
    A, B, X = P.host_copy()
    Ao = oracle.prescribed(n, 1e3, seed, spectrum)
    assert _same_bits(A, Ao), np.abs(A - Ao).max()
    assert _same_bits(B, oracle.rhs_random(n, 2, seed))
    assert X is None
    P.close()


def test_scaled_rhs_rowcol_and_unit_columns_match_oracle(ctx, oracle):
    n, seed = 40, 77
    P = Problem.generate(ctx, family="prescribed", n=n, nrhs=3, kappa=10.0, seed=seed, rhs="random_scaled", rowcol_span=10)
    A, B, _ = P.host_copy()
    r, c = oracle.rowcol_factors(n, 10, seed)
    assert _same_bits(A, oracle.prescribed(n, 10.0, seed) * r[:, None] * c[None, :])
    assert _same_bits(B, oracle.rhs_random(n, 3, seed, scaled=True))
    P.close()
    P = Problem.generate(ctx, family="grcar", n=n, nrhs=4, seed=seed, rhs="unit_column", global_exp=-3)
    A, B, X = P.host_copy()
    for k, j in enumerate(oracle.unit_columns(n, 4, seed)):
        assert _same_bits(B[:, k], A[:, j])
        assert X[j, k] == 1.0 and np.count_nonzero(X[:, k]) == 1
    P.close()


def test_dyadic_matches_oracle(ctx, oracle):
    n, seed = 50, 17
    P = Problem.generate(ctx, family="dyadic", n=n, nrhs=2, seed=seed, rhs="exact_sign", intmax=255, dyadic_exp=8)
    A, B, X = P.host_copy()
    Ao, Bo, Xo = oracle.dyadic(n, seed, 255, 8, nrhs=2)
    assert _same_bits(A, Ao) and _same_bits(B, Bo) and _same_bits(X, Xo)
    P.close()


def test_prescribed_large_keeps_spectrum(ctx):
    """n = 1024: the stages are orthogonal, so the singular values are the prescribed ones to rounding."""
    n = 1024
    P = Problem.generate(ctx, family="prescribed", n=n, nrhs=1, kappa=1e6, seed=3, rhs="random")
    A, _, _ = P.host_copy()
    s = np.linalg.svd(A, compute_uv=False)
    sig = 1e6 ** (-np.arange(n) / (n - 1))
    assert np.abs(s - sig).max() < 1e-13
    P.close()


def test_randsvd_kept_q_factors_give_the_same_matrix(ctx):
    """keep_q: the context keeps U and V of a randsvd generation and reuses them for the same (n, seed); the
    generated systems (A, B, x_true) are bitwise those of a generation from scratch, for the same and for
    another kappa, and a different seed replaces the kept factors."""
    n = 384
    print(f"[generate] randsvd keep_q n={n}", flush=True)
    ref = {}
    for kappa in (1e2, 1e6):
        P = Problem.generate(ctx, family="randsvd", n=n, nrhs=2, kappa=kappa, seed=77, rhs="xtrue_uniform")
        ref[kappa] = P.host_copy(); P.close()
    for kappa in (1e2, 1e6, 1e2):                                 # first generation fills the cache, the rest reuse it
        P = Problem.generate(ctx, family="randsvd", n=n, nrhs=2, kappa=kappa, seed=77, rhs="xtrue_uniform", keep_q=True)
        got = P.host_copy(); P.close()
        for a, b in zip(got, ref[kappa]):
            assert _same_bits(a, b)
    P = Problem.generate(ctx, family="randsvd", n=n, nrhs=2, kappa=1e2, seed=78, rhs="xtrue_uniform", keep_q=True)
    A78 = P.host_copy()[0]; P.close()
    P = Problem.generate(ctx, family="randsvd", n=n, nrhs=2, kappa=1e2, seed=78, rhs="xtrue_uniform")
    assert _same_bits(A78, P.host_copy()[0]) and not _same_bits(A78, ref[1e2][0])
    P.close()
