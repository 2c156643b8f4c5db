"""The E1 scalar track: an operation-rounded arithmetic model of mixed-precision LU refinement on the CPU, with
a GPU leg that runs the matching method of this library through its C ABI.

The reference's E1 track (e1_scalar/run_e1.py:155-373 over e1_scalar/three_precision_lab/tpir) studies the
precision roles (u_f, u, u_r) of LU-based iterative refinement in software arithmetic: every elementary
operation of the factorization, the substitutions and the residual is rounded separately to its precision (no
fused multiply-add, no blocked or library kernel), on small stored FP32 systems.  That model is the subject of
the track, so it is restated here in code of its own:

  OpRounded          tpir/arithmetic.py:25-59      one rounding per add / sub / mul / div / power-of-two scale;
                                                   binary16 operands are exact in binary64, the binary64 result is
                                                   rounded to binary16 (2 * 11 + 2 <= 53, so this is the correctly
                                                   rounded binary16 operation, gradual underflow included)
  residual           tpir/arithmetic.py:62-82      s_i = fl(s_i + fl(a_ij x_j)) for j = 0 .. n-1, then fl(b - s)
  lu_factor          tpir/arithmetic.py:111-141    right-looking rank-one GEPP, first largest |a| as pivot, whole
                                                   rows swapped, every division / product / difference rounded
  lu_solve           tpir/arithmetic.py:93-108     column-oriented forward and back substitution
  scaled_solve       tpir/arithmetic.py:144-159    right-hand side scaled by 2^-e (e from its largest entry) before
                                                   the solve, the result scaled back in the working precision
  refine             tpir/solver.py:44-107         the classical loop and its terminal statuses: iteration_limit,
                                                   computed_residual_zero, unchanged_working_update,
                                                   working_precision_cycle, numerical_failure, direct
  make_problem       tpir/problems.py:8-37         spectral family: A = U diag(geomspace(1, 1/kappa, n)) V^T from
                                                   the Q factors of seeded Gaussian matrices, stored in FP32; random
                                                   or manufactured right-hand side
The method table (METHODS) is the track's role contract (tpir/solver.py:25-43).

The GPU leg (gpu_trial) solves the same stored FP32 system with this library's method of the same roles through
the C ABI (ctypes), so the E1 driver can set the model beside what the GPU computes.  The GPU's low-precision
factorization is not the scalar model: its panel is FP32 and its trailing update multiplies FP16 operands with
FP32 accumulation on MFMA, so the two are compared by their errors, not bit for bit.

Run:  python -m cpsc402_midterm_project_amd.e1 --ns 64 128 --kappas 1e2 1e4 --seeds 9200 --out OUT [--gpu]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

DT = {16: np.float16, 32: np.float32, 64: np.float64}
UNIT = {16: 2.0 ** -11, 32: 2.0 ** -24, 64: 2.0 ** -53}


class ModelFailure(RuntimeError):
    """A numerical outcome of the model (non-finite value, zero pivot); reported as a status, not an error."""


def _check(x, where):
    if not np.all(np.isfinite(x)):
        raise ModelFailure(f"nonfinite:{where}")


class OpRounded:
    """Elementary operations of one precision, each rounded once to nearest, ties to even."""

    def __init__(self, bits: int):
        if bits not in DT:
            raise ValueError("precision must be 16, 32 or 64 bits")
        self.bits, self.t = bits, DT[bits]

    def val(self, x):
        with np.errstate(over="ignore", under="ignore", invalid="ignore"):
            return np.asarray(x, dtype=self.t)

    def _op(self, ufunc, x, y):
        x, y = self.val(x), self.val(y)
        with np.errstate(over="ignore", under="ignore", invalid="ignore", divide="ignore"):
            if self.bits == 16:      # exact (add, sub, mul) or correctly rounded (div) in binary64, then one rounding
                return ufunc(x.astype(np.float64), y.astype(np.float64)).astype(np.float16)
            return ufunc(x, y, dtype=self.t)

    def add(self, x, y):
        return self._op(np.add, x, y)

    def sub(self, x, y):
        return self._op(np.subtract, x, y)

    def mul(self, x, y):
        return self._op(np.multiply, x, y)

    def div(self, x, y):
        return self._op(np.divide, x, y)

    def scale2(self, x, e: int):
        """x * 2^e (exact unless it leaves the range)."""
        with np.errstate(over="ignore", under="ignore", invalid="ignore"):
            if self.bits == 16:
                return np.ldexp(self.val(x).astype(np.float64), e).astype(np.float16)
            return np.ldexp(self.val(x), e).astype(self.t)


def residual(a, b, x, bits: int):
    """b - A x with the products and sums of each row accumulated in column order, every operation rounded."""
    R = OpRounded(bits)
    A, X = R.val(a), R.val(x)
    s = np.zeros(A.shape[0], dtype=R.t)
    for j in range(A.shape[1]):
        s = R.add(s, R.mul(A[:, j], X[j]))
    return R.sub(R.val(b), s)


@dataclass
class Factors:
    lu: np.ndarray            # L (unit lower, below the diagonal) and U, in the factor precision
    perm: np.ndarray          # row i of P A is row perm[i] of A (int64)
    bits: int
    growth: float             # largest |entry| seen during elimination over max |A|

    def packed_hash(self) -> str:
        return hashlib.sha256(self.lu.tobytes() + self.perm.tobytes()).hexdigest()


def lu_factor(a, bits: int) -> Factors:
    R = OpRounded(bits)
    w = R.val(a).copy()
    if w.ndim != 2 or w.shape[0] != w.shape[1]:
        raise ValueError("A must be square")
    _check(w, "factor_input")
    n = w.shape[0]
    perm = np.arange(n, dtype=np.int64)
    amax0 = float(np.abs(w.astype(np.float64)).max()) if n else 0.0
    seen = amax0
    for k in range(n):
        p = k + int(np.argmax(np.abs(w[k:, k])))          # first largest magnitude
        if w[p, k] == 0:
            raise ModelFailure(f"zero_pivot_factor:{k}")
        if p != k:
            w[[k, p], :] = w[[p, k], :]
            perm[[k, p]] = perm[[p, k]]
        if k + 1 < n:
            w[k + 1:, k] = R.div(w[k + 1:, k], w[k, k])
            w[k + 1:, k + 1:] = R.sub(w[k + 1:, k + 1:], R.mul(w[k + 1:, k, None], w[k, None, k + 1:]))
            _check(w[k + 1:, k:], f"factor_update:{k}")
            seen = max(seen, float(np.abs(w[k + 1:, k + 1:].astype(np.float64)).max()))
    return Factors(w, perm, bits, seen / amax0 if amax0 else float("inf"))


def lu_solve(F: Factors, rhs, bits: int | None = None):
    """(L U)^-1 P rhs by column-oriented substitution in `bits` (default: the factor precision)."""
    R = OpRounded(F.bits if bits is None else bits)
    M = R.val(F.lu)
    z = R.val(rhs)[F.perm].copy()
    n = M.shape[0]
    for j in range(n - 1):
        z[j + 1:] = R.sub(z[j + 1:], R.mul(M[j + 1:, j], z[j]))
    _check(z, "forward_substitution")
    for j in range(n - 1, -1, -1):
        if M[j, j] == 0:
            raise ModelFailure(f"zero_pivot_solve:{j}")
        z[j] = R.div(z[j], M[j, j])
        z[:j] = R.sub(z[:j], R.mul(M[:j, j], z[j]))
    _check(z, "back_substitution")
    return z


def scaled_solve(F: Factors, rhs, work_bits: int, rhs_bits: int, normalize: bool = True, solve_bits: int | None = None):
    """d = 2^e (L U)^-1 P (2^-e r): r in the right-hand side's precision scaled so its largest entry lies in
    [1/2, 1), solved, scaled back and rounded to the working precision.  Returns (d, e)."""
    Rr, Rw = OpRounded(rhs_bits), OpRounded(work_bits)
    r = Rr.val(rhs)
    _check(r, "rhs")
    peak = float(np.abs(r).max()) if r.size else 0.0
    e = int(np.frexp(peak)[1]) if (normalize and peak) else 0
    z = lu_solve(F, Rr.scale2(r, -e), solve_bits)
    d = Rw.scale2(Rw.val(z), e)
    _check(d, "rescaled_solution")
    return d, e


@dataclass(frozen=True)
class Roles:
    name: str
    factor: int
    work: int
    resid: int
    direct: bool = False
    input16: bool = False                # factor the FP16 rounding of A in the factor precision
    solve_bits: int | None = None        # substitutions in another precision than the factors'
    normalize: bool = True

    def __post_init__(self):
        if not self.factor <= self.work <= self.resid:
            raise ValueError("the roles need factor bits <= working bits <= residual bits")


METHODS = {r.name: r for r in [
    Roles("hss", 16, 32, 32), Roles("hsd", 16, 32, 64), Roles("sss", 32, 32, 32), Roles("ssd", 32, 32, 64),
    Roles("hdd", 16, 64, 64), Roles("sdd", 32, 64, 64), Roles("ddd", 64, 64, 64),
    Roles("direct16", 16, 32, 32, direct=True), Roles("direct32", 32, 32, 32, direct=True),
    Roles("direct64", 64, 64, 64, direct=True), Roles("input16_ssd", 32, 32, 64, input16=True),
    Roles("half_factors_solve32", 16, 32, 64, solve_bits=32), Roles("hsd_no_norm", 16, 32, 64, normalize=False),
]}


def refine(a_stored, b_stored, method, max_iterations: int = 30):
    """The model's refinement run (no reference solution is used to stop).  Returns (record, iterates): the
    iterates in the working precision, the initial solve first and the last one returned."""
    m = METHODS[method] if isinstance(method, str) else method
    if max_iterations < 0:
        raise ValueError("max_iterations must be non-negative")
    W = OpRounded(m.work)
    A, b = W.val(a_stored), W.val(b_stored)
    if A.shape != (b.size, b.size) or b.ndim != 1:
        raise ValueError("expected a square A and a vector b")
    if not (np.all(np.isfinite(A)) and np.all(np.isfinite(b))):
        raise ValueError("the stored system must be finite")
    rec = dict(method=m.name, roles=asdict(m), iterations=0, factorizations=0, initialization="low_precision_solve",
               exponents=[], status="not_started", failure=None, factor_hash=None, growth=None)
    it = []
    t0 = time.perf_counter()
    try:
        Af = OpRounded(16).val(A).astype(np.float32) if m.input16 else A
        F = lu_factor(Af, m.factor)
        rec.update(factorizations=1, factor_hash=F.packed_hash(), growth=F.growth)
        try:
            x, e = scaled_solve(F, b, m.work, m.work, m.normalize, m.solve_bits)
            rec["exponents"].append(e)
        except ModelFailure as err:
            if m.direct:
                raise
            x = np.zeros(b.size, dtype=W.t)          # an explicit zero start, never a silent refactorization
            rec["initialization"] = "zero_after_failed_initial_solve"
            rec["initial_solve_failure"] = str(err)
        it.append(x.copy())
        if m.direct:
            rec["status"] = "direct"
        else:
            rec["status"] = "iteration_limit"
            for k in range(max_iterations):
                r = residual(A, b, x, m.resid)
                _check(r, "residual")
                if not np.any(r):
                    rec["status"] = "computed_residual_zero"
                    break
                d, e = scaled_solve(F, r, m.work, m.resid, m.normalize, m.solve_bits)
                rec["exponents"].append(e)
                new = W.add(x, d)
                _check(new, "working_update")
                it.append(new.copy())
                rec["iterations"] = k + 1
                if np.array_equal(new, x):
                    x = new
                    rec["status"] = "unchanged_working_update"
                    break
                if any(np.array_equal(new, old) for old in it[max(0, len(it) - 6):-2]):   # a cycle of period <= 4
                    x = new
                    rec["status"] = "working_precision_cycle"
                    break
                x = new
    except ModelFailure as err:
        rec["status"] = "numerical_failure"
        rec["failure"] = str(err)
    rec["seconds"] = time.perf_counter() - t0
    xs = np.stack(it) if it else np.empty((0, b.size), dtype=W.t)
    return rec, xs


def make_problem(n: int, kappa: float, seed: int, family: str = "spectral", rhs: str = "random"):
    """The stored FP32 system of the spectral family: numpy's default_rng(seed) draws the two Gaussian n x n
    matrices whose Q factors (numpy's QR, no sign normalization) are U and V, then the manufactured solution, then
    (random right-hand side) b.  A = U diag(geomspace(1, 1/kappa, n)) V^T is formed in FP64 and stored in FP32;
    b = A x (manufactured, from the stored A) or a Gaussian vector, stored in FP32."""
    if n < 2 or kappa < 1:
        raise ValueError("n >= 2 and kappa >= 1 required")
    if family != "spectral":
        raise ValueError("only the spectral family is restated")
    rng = np.random.default_rng(seed)
    sig = np.geomspace(1.0, 1.0 / kappa, n)
    U = np.linalg.qr(rng.standard_normal((n, n)))[0]
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A32 = ((U * sig) @ V.T).astype(np.float32)
    xm = rng.standard_normal(n)
    if rhs == "manufactured":
        b32 = (A32.astype(np.float64) @ xm).astype(np.float32)
    else:
        b32 = rng.standard_normal(n).astype(np.float32)
    return A32, b32


def reference_solution(a32, b32):
    """FP64 LU solution refined with residuals in extended precision (numpy longdouble) until the update
    vanishes (at most 6 steps): the diagnostic reference of the stored system."""
    import scipy.linalg as sla
    A = np.asarray(a32, dtype=np.float64); b = np.asarray(b32, dtype=np.float64)
    lu = sla.lu_factor(A)
    Al, bl = A.astype(np.longdouble), b.astype(np.longdouble)
    x = sla.lu_solve(lu, b).astype(np.longdouble)
    for _ in range(6):
        d = sla.lu_solve(lu, (bl - Al @ x).astype(np.float64)).astype(np.longdouble)
        nx = x + d
        if np.array_equal(nx, x):
            break
        x = nx
    return x.astype(np.float64)


def errors(a32, b32, x, xref):
    """eta, omega (FP64 residual of the stored system), and the forward error against the reference."""
    A = np.asarray(a32, dtype=np.float64); b = np.asarray(b32, dtype=np.float64); x = np.asarray(x, dtype=np.float64)
    r = b - A @ x
    den = np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()
    t = np.abs(A) @ np.abs(x) + np.abs(b)
    q = np.where(t > 0, np.abs(r) / np.where(t > 0, t, 1), np.where(r != 0, np.inf, 0.0))
    nx = np.abs(xref).max()
    return dict(eta=float(np.abs(r).max() / den) if den else 0.0, omega=float(q.max()),
                finf=float(np.abs(x - xref).max() / (nx if nx else 1.0)))


# the library's method of the same precision roles (factor / working / residual) for the GPU leg
GPU_METHOD = {"hsd": "hetero16_ir_w32_r64", "sss": "native_ir32_w32_r32", "ssd": "native_ir32_w32_r64",
              "hdd": "hetero16_ir_w64_r64", "sdd": "native_ir32_w64_r64", "ddd": "native_ir64_w64_r64",
              "direct16": "hetero16_lu_direct", "direct32": "native_lu32", "direct64": "native_lu64"}


def gpu_trial(ctx, a32, b32, method: str, xref, max_outer: int = 30) -> dict:
    """Solve the stored FP32 system with the library's method of the same roles (through the C ABI) on the GPU
    of `ctx`.  Like the model, which does not stop on a small backward error, the GPU leg refines until its
    iterate stops improving: the backward-error target is 0, so the loop ends by its stagnation rule (eta above
    0.9 x the previous one three times) or at max_outer."""
    mid = GPU_METHOD[method]
    target = "fp64" if METHODS[method].work == 64 else "fp32"
    P = ctx.problem(np.asfortranarray(a32, dtype=np.float32), np.asarray(b32, dtype=np.float64).reshape(-1, 1),
                    x_ref=np.asarray(xref, dtype=np.float64).reshape(-1, 1))
    try:
        X, rec = P.solve(mid, target=target, tau=0.0, max_outer=max_outer, nb=64)
    finally:
        P.close()
    x = X[:, 0].astype(np.float32 if target == "fp32" else np.float64)
    e = errors(a32, b32, x, xref)
    return dict(method=method, gpu_method=mid, terminal=rec["terminal_reason"], ok=rec["ok"], outer=rec["outer_iters"],
                device_s=rec["device_ready_wall_s"], **e)


def run_case(n: int, kappa: float, seed: int, methods=None, ctx=None, max_iterations: int = 30) -> tuple[dict, list]:
    """One case of the campaign: the stored system, its reference, the model's run of every method and, with a
    library context, the GPU leg of every method that has one."""
    A32, b32 = make_problem(n, kappa, seed)
    xref = reference_solution(A32, b32)
    case_id = hashlib.sha256(A32.tobytes() + b32.tobytes() + f"|n={n}|seed={seed}".encode()).hexdigest()[:16]
    case = dict(case_id=case_id, n=n, kappa_nominal=kappa, seed=seed, family="spectral", rhs="random",
                a_sha256=hashlib.sha256(A32.tobytes()).hexdigest(), b_sha256=hashlib.sha256(b32.tobytes()).hexdigest(),
                kappa2=float(np.linalg.cond(A32.astype(np.float64), 2)))
    trials = []
    for mname in methods or ["hss", "hsd", "sss", "ssd", "input16_ssd", "half_factors_solve32"]:
        rec, xs = refine(A32, b32, mname, max_iterations)
        t = dict(case_id=case_id, leg="model", method=mname, status=rec["status"], failure=rec["failure"],
                 iterations=rec["iterations"], factor_hash=rec["factor_hash"], growth=rec["growth"], seconds=rec["seconds"])
        if len(xs):
            t.update(errors(A32, b32, xs[-1], xref))
            u = UNIT[METHODS[mname].work]
            t["finf_le_10u"] = bool(t["finf"] <= 10 * u)
        trials.append(t)
        if ctx is not None and mname in GPU_METHOD:
            g = gpu_trial(ctx, A32, b32, mname, xref)
            g.update(case_id=case_id, leg="gpu")
            trials.append(g)
    return case, trials


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--ns", type=int, nargs="+", default=[64, 128])
    ap.add_argument("--kappas", type=float, nargs="+", default=[1e2, 1e4])
    ap.add_argument("--seeds", type=int, nargs="+", default=[9200])
    ap.add_argument("--methods", nargs="+", default=None)
    ap.add_argument("--max-corrections", type=int, default=30)
    ap.add_argument("--gpu", action="store_true", help="add the GPU leg through the library's C ABI")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    out = Path(a.out); out.mkdir(parents=True, exist_ok=True)
    ctx = None
    if a.gpu:
        from . import Context
        ctx = Context(0)
    try:
        with open(out / "cases.jsonl", "w") as fc, open(out / "trials.jsonl", "w") as ft:
            for n in a.ns:
                for kappa in a.kappas:
                    for seed in a.seeds:
                        case, trials = run_case(n, kappa, seed, a.methods, ctx, a.max_corrections)
                        fc.write(json.dumps(case) + "\n")
                        for t in trials:
                            ft.write(json.dumps(t) + "\n")
                        print(f"[e1] n={n} kappa={kappa:g} seed={seed}: " +
                              ", ".join(f"{t['leg']}:{t['method']}={t.get('status', t.get('terminal'))}" for t in trials), flush=True)
    finally:
        if ctx is not None:
            ctx.close()


if __name__ == "__main__":
    main()
