"""Python side of the CPU oracle.  TEST INFRASTRUCTURE ONLY (used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product library never imports it).

Wraps liboracle.so (mp_oracle.c) and restates, in numpy, the reference's metric definitions and refinement
loops:
  metrics         eval_metrics_T        src/mp_core.cuh:193-205 (eta, omega, relres, worst row)
  ir_stationary   Engine::refine_stage  src/mp_core.cuh:281-296 (stopping tests and terminal reasons)
  fgmres          Engine::gmres_solve   src/mp_core.cuh:237-278 (right-preconditioned flexible GMRES, and
                                        the left-preconditioned variant)
  factor / apply_inverse   factorize + apply_inverse, src/mp_core.cuh:158-190 (scale modes 0, 1, 2)
  Xoshiro, hash3, prescribed, rhs_random, rowcol_factors, unit_columns, dyadic
                  the reference's generator streams and families, src/mp_common.cuh:34-43,
                  src/mp_core.cuh:29-103, src/mp_kernels.cuh:39-49 (pinned by published splitmix64 /
                  xoshiro256** outputs, the prescribed spectrum and exact dyadic products)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, i64 = C.c_void_p, C.c_int64
    lib.or_resid_fp64.argtypes = [i64, C.c_int, C.c_int, vp, i64, vp, i64, vp, i64, vp, vp, i64]
    lib.or_resid_dd.argtypes = [i64, C.c_int, C.c_int, vp, i64, vp, i64, vp, i64, vp, vp, i64]
    lib.or_colstats.argtypes = [i64, C.c_int, vp, vp, vp, i64, vp, vp]
    lib.or_getrf.argtypes = [i64, C.c_int, vp, i64, vp, vp, C.c_int, C.c_int, vp]
    lib.or_getrf.restype = C.c_int
    lib.or_getrs.argtypes = [i64, C.c_int, vp, i64, vp, vp, i64, vp, i64, C.c_double]
    lib.or_pack.argtypes = [C.c_int, C.c_int, i64, C.c_int, vp, vp, vp, vp, vp]
    for f in ("or_round_fp16", "or_round_bf16", "or_round_e4m3"):
        getattr(lib, f).argtypes = [C.c_float]
        getattr(lib, f).restype = C.c_float
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


FMT = {"fp16": 0, "bf16": 1, "fp8": 2, "fp32": 3, "fp16x3": 4}


def resid_fp64(A, X, B=None):
    A = np.asfortranarray(A); X = np.asfortranarray(X, dtype=np.float64).reshape(A.shape[0], -1, order="F")
    n, nrhs = X.shape
    R = np.empty((n, nrhs), order="F"); T = np.empty((n, nrhs), order="F")
    Bc = None if B is None else np.asfortranarray(B, dtype=np.float64).reshape(n, nrhs, order="F")
    lib().or_resid_fp64(n, nrhs, 32 if A.dtype == np.float32 else 64, _p(A), n, _p(X), n, _p(Bc), n, _p(R), _p(T), n)
    return R, T


def resid_dd(A, X, B=None):
    A = np.asfortranarray(A); X = np.asfortranarray(X, dtype=np.float64).reshape(A.shape[0], -1, order="F")
    n, nrhs = X.shape
    Rh = np.empty((n, nrhs), order="F"); Rl = np.empty((n, nrhs), order="F")
    Bc = None if B is None else np.asfortranarray(B, dtype=np.float64).reshape(n, nrhs, order="F")
    lib().or_resid_dd(n, nrhs, 32 if A.dtype == np.float32 else 64, _p(A), n, _p(X), n, _p(Bc), n, _p(Rh), _p(Rl), n)
    return Rh, Rl


def colstats(R, X, T):
    R = np.asfortranarray(R); X = np.asfortranarray(X); T = np.asfortranarray(T)
    n, nrhs = R.shape
    out = np.empty(4 * nrhs); w = np.empty(nrhs, dtype=np.int32)
    lib().or_colstats(n, nrhs, _p(R), _p(X), _p(T), n, _p(out), _p(w))
    return out.reshape(nrhs, 4), w


def getrf(A32, nb=256, fmt="fp16", tile_scale=True):
    """Blocked LU of a copy of A32 (float32).  Returns (LU, ipiv (1-based), perm, info, counts)."""
    LU = np.array(A32, dtype=np.float32, order="F", copy=True)
    n = LU.shape[0]
    ipiv = np.zeros(n, dtype=np.int64); perm = np.zeros(n, dtype=np.int32); cnt = np.zeros(2, dtype=np.uint64)
    info = lib().or_getrf(n, nb, _p(LU), n, _p(ipiv), _p(perm), FMT[fmt], 1 if tile_scale else 0, _p(cnt))
    return LU, ipiv, perm, info, cnt


def getrs(LU, perm, R, scale=1.0):
    LU = np.asfortranarray(LU, dtype=np.float32)
    n = LU.shape[0]
    R = np.asfortranarray(R, dtype=np.float64).reshape(n, -1, order="F")
    D = np.empty_like(R)
    lib().or_getrs(n, R.shape[1], _p(LU), n, _p(perm), _p(R), n, _p(D), n, float(scale))
    return D


def round_fmt(x, fmt):
    f = {"fp16": lib().or_round_fp16, "bf16": lib().or_round_bf16, "fp8": lib().or_round_e4m3}[fmt]
    x = np.asarray(x, dtype=np.float32)
    return np.array([f(float(v)) for v in x.ravel()], dtype=np.float32).reshape(x.shape)


def pack(S, fmt="fp16", tile_scale=True):
    """One operand's conversion (or_pack; src/mp_kernels.cuh:14-18 with the per-tile power-of-two scale): S is
    nidx x K (index-major).  Returns the rounded scaled values (nidx x K, FP32), the low terms of the two-term
    format (zeros otherwise), the tile exponents and the [overflow, subnormal] counts."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    nidx, K = S.shape
    q = np.zeros_like(S); qlo = np.zeros_like(S)
    ex = np.zeros((nidx + 31) // 32, dtype=np.int32)
    cnt = np.zeros(2, dtype=np.uint64)
    lib().or_pack(FMT[fmt], 1 if tile_scale else 0, nidx, K, _p(S), _p(q), _p(qlo), _p(ex), _p(cnt))
    return q, qlo, ex, cnt


def unpack_lu(LU):
    L = np.tril(LU.astype(np.float64), -1) + np.eye(LU.shape[0])
    U = np.triu(LU.astype(np.float64))
    return L, U


# ------------------------------------------------------------------ metrics (src/mp_core.cuh:193-205)
def norms(A, B):
    A64 = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64).reshape(A64.shape[0], -1)
    return np.abs(A64).sum(axis=1).max(), np.abs(B).max(axis=0), np.sqrt((B * B).sum(axis=0))


def metrics(A, B, X, dd=False):
    """eta, omega per RHS from the oracle residual (FP64 sequential FMA or double-double)."""
    n = A.shape[0]
    X = np.asarray(X, dtype=np.float64).reshape(n, -1)
    B = np.asarray(B, dtype=np.float64).reshape(n, -1)
    if dd:
        Rh, Rl = resid_dd(A, X, B)
        R = Rh + Rl
        _, T = resid_fp64(A, X, B)
    else:
        R, T = resid_fp64(A, X, B)
    st, w = colstats(R, X, T)
    anorm, binf, b2 = norms(A, B)
    eta = np.array([0.0 if s[0] == 0 else (np.inf if anorm * s[1] + binf[c] == 0 else s[0] / (anorm * s[1] + binf[c]))
                    for c, s in enumerate(st)])
    return dict(eta=eta, omega=st[:, 2], rinf=st[:, 0], xinf=st[:, 1], worst=w, eta_max=eta.max(), omega_max=st[:, 2].max())


def exact_residual(A, x, b):
    """Exact rational residual b - A x for tiny systems (Fraction arithmetic)."""
    n = A.shape[0]
    return [Fraction(float(b[i])) - sum(Fraction(float(A[i, j])) * Fraction(float(x[j])) for j in range(n)) for i in range(n)]


# ------------------------------------------------------------------ refinement restatements
def _pow2_recip(m):
    """k_pow2_recip (src/mp_kernels.cuh:108): 2^-(e-1) for m in [2^(e-1), 2^e); 1 for m <= 0 or non-finite."""
    m = np.asarray(m, dtype=np.float64)
    out = np.ones_like(m)
    ok = (m > 0) & np.isfinite(m)
    out[ok] = np.ldexp(1.0, 1 - np.frexp(m[ok])[1])
    return out


def _absmax_rows(A, axis):
    a = np.abs(A.astype(np.float64))
    return np.where(np.isnan(a), 0.0, a).max(axis=axis)      # NaN never exceeds the running maximum


def factor(A, fmt="fp16", nb=128, scale_mode=1, tile_scale=True):
    """factorize (src/mp_core.cuh:158-167) for the own blocked kinds: the FP32 copy of A, for scale mode 2 the
    power-of-two row then column equilibration (src/mp_core.cuh:164-166), for modes 1 and 2 the global scale
    g = 2^-e from the copy's maximum f 2^e, applied in FP32; then the blocked LU."""
    A32 = np.asarray(A, dtype=np.float64).astype(np.float32)
    dr = dc = None
    if scale_mode == 2:
        dr = _pow2_recip(_absmax_rows(A32, 1))
        A32 = (A32.astype(np.float64) * dr[:, None]).astype(np.float32)
        dc = _pow2_recip(_absmax_rows(A32, 0))
        A32 = (A32.astype(np.float64) * dc[None, :]).astype(np.float32)
    g = 1.0
    amax = float(_absmax_rows(A32, None)) if A32.size else 0.0
    if scale_mode >= 1 and amax > 0 and np.isfinite(amax):
        g = float(np.ldexp(1.0, -int(np.frexp(amax)[1])))
        A32 = A32 * np.float32(g)
    LU, ipiv, perm, info, cnt = getrf(A32, nb=nb, fmt=fmt, tile_scale=tile_scale)
    return dict(LU=LU, ipiv=ipiv, perm=perm, info=info, g=g, dr=dr, dc=dc, counts=cnt)


def apply_inverse(F, R):
    """apply_inverse with FP32 factors (src/mp_core.cuh:182-190): (Dr r) -> FP32 substitution -> Dc, g."""
    R = np.asarray(R, dtype=np.float64).reshape(F["LU"].shape[0], -1)
    if F.get("dr") is not None:
        R = R * F["dr"][:, None]
    D = getrs(F["LU"], F["perm"], R, F["g"])
    if F.get("dc") is not None:
        D = D * F["dc"][:, None]
    return D


def ir_stationary(A, B, fmt="fp16", nb=128, tau=1e-15, max_outer=50, dd=False, stop_eta_only=True, tile_scale=True,
                  scale_mode=1):
    """Stationary IR as in Engine::refine_stage (src/mp_core.cuh:281-296): x0 = A^-1 b, then residual+metrics,
    stopping tests (converged / nonfinite / max_outer / stagnation), correction, x += d."""
    F = factor(A, fmt, nb, scale_mode, tile_scale)
    return ir_loop(A, B, F, tau=tau, max_outer=max_outer, dd=dd, stop_eta_only=stop_eta_only)


def ir_loop(A, B, F, tau=1e-15, max_outer=50, dd=False, stop_eta_only=True, metric=None, apply=None):
    """The loop of Engine::refine_stage (src/mp_core.cuh:283-296) over given factors F (a factor() dict, e.g.
    built around another factorization's LU): x0 = F^-1 b; per step the metrics of x, then in this order
    converged (finite, eta <= tau and, unless stop_eta_only, omega <= tau), nonfinite, max_outer, stagnation
    (eta > 0.9 * previous eta for 3 consecutive steps, src/mp_core.cuh:290), then x += F^-1 r.
    metric(x) -> dict(eta_max, omega_max, rinf, xinf) and apply(R) -> F^-1 R may be replaced (scripted
    sequences that pin the stopping rules).  Returns (x, terminal reason, eta history)."""
    n = A.shape[0]
    B = np.asarray(B, dtype=np.float64).reshape(n, -1)
    metric = metric or (lambda x: metrics(A, B, x, dd=dd))
    apply = apply or (lambda R: apply_inverse(F, R))
    x = apply(B)
    hist = []
    prev, stall, outer = np.inf, 0, 0
    while True:
        m = metric(x)
        hist.append(m["eta_max"])
        finite = np.isfinite(m["rinf"]).all() and np.isfinite(m["xinf"]).all()
        if finite and m["eta_max"] <= tau and (stop_eta_only or m["omega_max"] <= tau):
            return x, "converged", hist
        if not finite:
            return x, "nonfinite", hist
        if outer >= max_outer:
            return x, "max_outer", hist
        stall = stall + 1 if m["eta_max"] > 0.9 * prev else 0
        if stall >= 3:
            return x, "stagnation", hist
        prev = m["eta_max"]
        if dd:
            Rh, Rl = resid_dd(A, x, B); R = Rh + Rl
        else:
            R, _ = resid_fp64(A, x, B)
        x = x + apply(R)
        outer += 1


def fgmres(A, r, precond, m=30, tol=1e-4, cycles=4, budget=480, left=False):
    """Right-preconditioned flexible GMRES on A d = r (Engine::gmres_solve, src/mp_core.cuh:237-278), with
    classical Gram-Schmidt applied twice; left=True: GMRES on M^-1 A d = M^-1 r (src/mp_core.cuh:240-241,
    249-250, 266, 271): the Krylov space of M^-1 A from M^-1 r, d += V y, and the restart residual
    preconditioned.  Returns (d, status, iterations)."""
    n = A.shape[0]
    d = np.zeros(n)
    rcur = precond(r) if left else r.copy()
    beta0 = np.linalg.norm(rcur)
    if beta0 == 0 or not np.isfinite(beta0):
        return d, ("zero_rhs" if beta0 == 0 else "nonfinite"), 0
    its = 0; status = "budget"
    for cyc in range(cycles):
        if budget <= 0:
            break
        betac = np.linalg.norm(rcur)
        if betac == 0:
            return d, "converged_true", its
        V = np.zeros((n, m + 1)); Z = np.zeros((n, m)); H = np.zeros((m + 1, m)); g = np.zeros(m + 1); g[0] = betac
        cs = np.zeros(m); sn = np.zeros(m)
        V[:, 0] = rcur / betac
        j = 0; happy = False
        while j < m and budget > 0:
            if left:
                w = precond(A @ V[:, j])
            else:
                Z[:, j] = precond(V[:, j])
                w = A @ Z[:, j]
            h = V[:, :j + 1].T @ w; w = w - V[:, :j + 1] @ h
            h2 = V[:, :j + 1].T @ w; w = w - V[:, :j + 1] @ h2; h = h + h2
            hn = np.linalg.norm(w)
            H[:j + 1, j] = h; H[j + 1, j] = hn
            if hn > 0:
                V[:, j + 1] = w / hn
            for i in range(j):
                a, b = H[i, j], H[i + 1, j]
                H[i, j] = cs[i] * a + sn[i] * b; H[i + 1, j] = -sn[i] * a + cs[i] * b
            a, b = H[j, j], H[j + 1, j]; rho = np.hypot(a, b)
            cs[j], sn[j] = (1.0, 0.0) if rho == 0 else (a / rho, b / rho)
            H[j, j] = rho; H[j + 1, j] = 0.0; g0 = g[j]; g[j] = cs[j] * g0; g[j + 1] = -sn[j] * g0
            its += 1; budget -= 1; j += 1
            if hn == 0:
                happy = True; break
            if abs(g[j]) / beta0 <= tol:
                break
        k = j
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k]) if k else np.zeros(0)
        d = d + (V[:, :k] if left else Z[:, :k]) @ y
        sr = r - A @ d
        true_rel = np.linalg.norm(sr) / np.linalg.norm(r)
        rcur = precond(sr) if left else sr
        if true_rel <= tol:
            return d, ("happy_breakdown_verified" if happy else "converged_true"), its
        if happy:
            return d, "happy_breakdown_unverified", its
    return d, status, its


def fgmres_ir(A, B, fmt="fp16", nb=128, tau=1e-15, max_outer=30, m=30, tol=1e-4, dd=True, tile_scale=True,
              left=False, cycles=4):
    F = factor(A, fmt, nb, 1, tile_scale)
    n = A.shape[0]
    b = np.asarray(B, dtype=np.float64).reshape(n)
    pre = lambda v: apply_inverse(F, v)[:, 0]
    x = pre(b)
    hist = []; inner = 0
    for outer in range(max_outer + 1):
        m_ = metrics(A, b, x, dd=dd)
        hist.append(m_["eta_max"])
        if m_["eta_max"] <= tau:
            return x, "converged", hist, inner
        if dd:
            Rh, Rl = resid_dd(A, x, b); r = (Rh + Rl)[:, 0]
        else:
            r = resid_fp64(A, x, b)[0][:, 0]
        d, st, its = fgmres(A, r, pre, m=m, tol=tol, left=left, cycles=cycles)
        inner += its
        x = x + d
    return x, "max_outer", hist, inner


# ------------------------------------------------------------------ generator streams (src/mp_common.cuh:34-43)
# Python integers and exact rational arithmetic restate the reference's host streams and the prescribed family
# (src/mp_core.cuh:29-58, src/mp_kernels.cuh:39-49) for small n.
_M64 = (1 << 64) - 1


def splitmix_step(state):
    """One splitmix64 step: returns (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def hash3(a, b, c):
    x = ((a * 0x9E3779B97F4A7C15) & _M64) ^ (((b + 0x7F4A7C15) * 0xBF58476D1CE4E5B9) & _M64) \
        ^ (((c + 0x1234567) * 0x94D049BB133111EB) & _M64)
    x, _ = splitmix_step(x)
    _, out = splitmix_step(x)
    return out


class Xoshiro:
    """xoshiro256** with its state filled by four splitmix64 steps from the seed."""

    def __init__(self, seed=None, state=None):
        if state is None:
            x, state = seed & _M64, []
            for _ in range(4):
                x, v = splitmix_step(x)
                state.append(v)
        self.s = list(state)

    @staticmethod
    def _rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & _M64

    def next(self):
        s = self.s
        out = (self._rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = self._rotl(s[3], 45)
        return out

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def normal(self):
        import math
        u1 = self.uniform()
        while u1 <= 0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)

    def below(self, n):
        return self.next() % n


def fma(x, y, z):
    """Correctly rounded x*y + z (exact rational sum, one rounding), with IEEE signed zeros."""
    import math
    r = Fraction(x) * Fraction(y) + Fraction(z)
    if r != 0:
        return float(r)
    if (x == 0 or y == 0) and z == 0 and math.copysign(1, x) * math.copysign(1, y) < 0 and math.copysign(1, z) < 0:
        return -0.0
    return 0.0


def spectrum(sp, n, kappa):
    if sp == "geometric":
        return [1.0 if n == 1 else kappa ** (-i / (n - 1)) for i in range(n)]
    if sp == "clustered":
        k = max(2, n // 64)
        return [1.0 / kappa if i >= n - k else 1.0 for i in range(n)]
    if sp == "one_small":
        return [1.0 / kappa if i == n - 1 else 1.0 for i in range(n)]
    if sp == "identity":
        return [1.0] * n
    raise ValueError(sp)


def _rotation_stages(A, n, seed, side, stages):
    import math
    rng = Xoshiro((seed * 1000003 + 17 * side + 0xC0FFEE) & _M64)
    for _ in range(stages):
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for p in range(n // 2):
            c1, c2 = perm[2 * p], perm[2 * p + 1]
            th = rng.uniform() * 6.283185307179586
            c, s = math.cos(th), math.sin(th)
            for r in range(n):
                a, b = A[r][c1], A[r][c2]
                A[r][c1] = fma(c, a, s * b)
                A[r][c2] = fma(c, b, -(s * a))


def prescribed(n, kappa, seed, sp="geometric"):
    """The prescribed family: diag(sigma), rotation stages on the columns, transpose, stages again."""
    import math
    sig = spectrum(sp, n, kappa)
    A = [[sig[i] if i == j else 0.0 for j in range(n)] for i in range(n)]
    stages = int(math.ceil(math.log2(n))) + 4
    _rotation_stages(A, n, seed, 1, stages)
    A = [list(col) for col in zip(*A)]
    _rotation_stages(A, n, seed, 2, stages)
    return np.asfortranarray(np.array(A, dtype=np.float64))


def rhs_random(n, nrhs, seed, scaled=False):
    rng = Xoshiro((seed * 104729 + 0xBEEF + nrhs) & _M64)
    B = np.empty((n, nrhs), order="F")
    for c in range(nrhs):
        for i in range(n):
            v = rng.normal()
            B[i, c] = np.ldexp(v, 4 * c - 2 * nrhs) if scaled else v
    return B


def rowcol_factors(n, span, seed):
    rng = Xoshiro((seed * 31 + 5) & _M64)
    r = [np.ldexp(1.0, int(rng.below(span + 1)) - span // 2) for _ in range(n)]
    c = [np.ldexp(1.0, int(rng.below(span + 1)) - span // 2) for _ in range(n)]
    return np.array(r), np.array(c)


def unit_columns(n, nrhs, seed):
    rng = Xoshiro((seed * 17 + 3) & _M64)
    return [int(rng.below(n)) for _ in range(nrhs)]


def dyadic(n, seed, m, s, nrhs=1):
    """A = 2^-s C, C_ij = hash3(seed, i, j) mod (2m+1) - m; x*_j = +-1 from hash3(seed' ^ 0xA5A5A5A5, j, 7),
    seed' = 3 seed + 11 + c; b = 2^-s C x* (exact)."""
    C = np.array([[int(hash3(seed, i, j) % (2 * m + 1)) - m for j in range(n)] for i in range(n)], dtype=np.int64)
    X = np.empty((n, nrhs))
    for c in range(nrhs):
        rs = (seed * 3 + 11 + c) & _M64
        X[:, c] = [1.0 if hash3(rs ^ 0xA5A5A5A5, j, 7) & 1 else -1.0 for j in range(n)]
    B = (C @ X.astype(np.int64)).astype(np.float64) * 2.0 ** -s
    return np.asfortranarray(C.astype(np.float64) * 2.0 ** -s), np.asfortranarray(B), X
