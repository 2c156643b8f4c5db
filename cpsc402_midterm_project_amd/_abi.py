"""ctypes binding of libmpsolve (include/mpsolve.h).

The library is the product path: there is no CPU fallback.  Importing this module raises ImportError when
libmpsolve.so has not been built, and every call raises MPSolveError on a negative status.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmpsolve.so"

MP_OK = 0
STATUS = {-1: "MP_ERR_INVALID", -2: "MP_ERR_HIP", -3: "MP_ERR_ALLOC", -4: "MP_ERR_UNSUPPORTED",
          -5: "MP_ERR_INTERNAL", -6: "MP_ERR_NO_DEVICE"}
FMT = {"fp16": 0, "bf16": 1, "fp8": 2, "fp8e4m3": 2, "fp32": 3, "fp16x3": 4}

# every function declared in include/mpsolve.h
EXPORTS = [
    "mp_create", "mp_destroy", "mp_last_error", "mp_caps", "mp_synchronize", "mp_profile", "mp_profile_json", "mp_malloc", "mp_free",
    "mp_memcpy_h2d", "mp_memcpy_d2h", "mp_memcpy_d2d", "mp_residual_fp64", "mp_residual_dd", "mp_residual_noalloc", "mp_getrf",
    "mp_getrs", "mp_pack_operand", "mp_problem_factor_bytes", "mp_problem_from_host", "mp_problem_generate", "mp_problem_reference", "mp_problem_free", "mp_problem_copy_out", "mp_problem_copy_ref", "mp_problem_case_json", "mp_solve",
    "mp_last_record_json", "mp_last_hist_json", "mp_methods_json", "mp_eval_metrics_json", "mp_audit_json", "mp_sha256_hex", "mp_batch_solve",
    "mp_batch_solve_ptrs", "mp_generate_into", "mp_problem_from_device", "mp_problem_set_rhs", "mp_server_clock",
    "mp_server_create", "mp_server_submit", "mp_server_poll", "mp_server_stats_json", "mp_server_destroy", "mp_getrf64",
    "mp_factor_cache_put", "mp_factor_cache_get", "mp_factor_cache_evict",
]
SERVER_TERMINAL = ["converged", "nonfinite", "max_outer", "stagnation", "singular_factor", "error", "other"]
BATCH_TERMINAL = ["converged", "nonfinite", "max_outer", "stagnation", "singular_factor"]


class MPSolveError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class GetrfParams(C.Structure):
    _fields_ = [("nb", C.c_int32), ("fmt", C.c_int32), ("tile_scale", C.c_int32)]


class GetrfInfo(C.Structure):
    _fields_ = [("info", C.c_int32), ("panels", C.c_int32), ("conv_overflow", C.c_uint64), ("conv_subnormal", C.c_uint64)]


class TrialOut(C.Structure):
    _fields_ = [("ok", C.c_int32), ("outer_iters", C.c_int32), ("inner_iters_total", C.c_int32),
                ("refactorizations", C.c_int32), ("device_ready_wall_s", C.c_double), ("factor_total_s", C.c_double),
                ("eta_max", C.c_double), ("omega_max", C.c_double), ("terminal_reason", C.c_char * 48),
                ("final_stage", C.c_char * 48)]


class BatchResult(C.Structure):
    _fields_ = [("terminal", C.c_int32), ("outer_iters", C.c_int32), ("info", C.c_int32), ("reserved", C.c_int32),
                ("eta_max", C.c_double), ("omega_max", C.c_double)]


class ServerCompletion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("terminal", C.c_int32), ("path", C.c_int32), ("cache_hit", C.c_int32),
                ("outer_iters", C.c_int32), ("eta_max", C.c_double), ("omega_max", C.c_double), ("queue_s", C.c_double),
                ("service_s", C.c_double), ("latency_s", C.c_double)]


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m cpsc402_midterm_project_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    vp, i64, i32, dp, sz = C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_size_t
    sig = {
        "mp_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "mp_destroy": ([vp], None),
        "mp_last_error": ([], C.c_char_p),
        "mp_caps": ([vp, C.c_char_p, sz], C.c_int),
        "mp_synchronize": ([vp], C.c_int),
        "mp_profile": ([vp, C.c_int], C.c_int),
        "mp_profile_json": ([vp, C.c_char_p, sz], C.c_int),
        "mp_malloc": ([vp, C.POINTER(vp), sz], C.c_int),
        "mp_free": ([vp, vp], C.c_int),
        "mp_memcpy_h2d": ([vp, vp, vp, sz], C.c_int),
        "mp_memcpy_d2h": ([vp, vp, vp, sz], C.c_int),
        "mp_memcpy_d2d": ([vp, vp, vp, sz], C.c_int),
        "mp_residual_fp64": ([vp, i64, i32, i32, vp, i64, dp, i64, dp, i64, dp, i64, dp, i64, dp, vp], C.c_int),
        "mp_residual_dd": ([vp, i64, i32, i32, vp, i64, dp, i64, dp, i64, dp, dp, i64, dp, i64, dp, vp], C.c_int),
        "mp_residual_noalloc": ([vp, i64, i32, i32, vp, i64, dp, i64, dp, i64, dp, dp, i64, dp, i64, dp, vp], C.c_int),
        "mp_getrf": ([vp, i64, vp, i64, vp, vp, C.POINTER(GetrfParams), C.POINTER(GetrfInfo)], C.c_int),
        "mp_getrs": ([vp, i64, i32, vp, i64, vp, dp, i64, dp, i64, C.c_double], C.c_int),
        "mp_pack_operand": ([vp, i32, i32, vp, i64, i64, i32, i32, vp, vp, vp, vp], C.c_int),
        "mp_problem_factor_bytes": ([vp, C.POINTER(C.c_uint64)], C.c_int),
        "mp_problem_from_host": ([vp, i64, i32, i32, vp, i64, dp, i64, dp, dp, C.c_char_p, C.POINTER(vp)], C.c_int),
        "mp_problem_free": ([vp], None),
        "mp_problem_generate": ([vp, C.c_char_p, C.POINTER(vp)], C.c_int),
        "mp_problem_reference": ([vp, vp], C.c_int),
        "mp_problem_copy_out": ([vp, vp, vp, dp, dp], C.c_int),
        "mp_problem_copy_ref": ([vp, dp], C.c_int),
        "mp_problem_case_json": ([vp, C.c_char_p, sz], C.c_int),
        "mp_solve": ([vp, vp, C.c_char_p, C.c_char_p, C.POINTER(TrialOut), dp], C.c_int),
        "mp_last_record_json": ([vp, C.c_char_p, sz], C.c_int),
        "mp_last_hist_json": ([vp, C.c_char_p, sz], C.c_int),
        "mp_methods_json": ([C.c_char_p, sz], C.c_int),
        "mp_eval_metrics_json": ([vp, vp, dp, C.c_char_p, sz], C.c_int),
        "mp_audit_json": ([vp, vp, dp, dp, dp, C.c_char_p, sz], C.c_int),
        "mp_sha256_hex": ([vp, sz, C.c_char_p], C.c_int),
        "mp_batch_solve": ([vp, i32, vp, vp, dp, vp, dp, vp, dp, C.c_char_p, vp], C.c_int),
        "mp_batch_solve_ptrs": ([vp, i32, vp, vp, vp, vp, vp, C.c_char_p, vp], C.c_int),
        "mp_generate_into": ([vp, C.c_char_p, dp, dp, dp, vp], C.c_int),
        "mp_problem_from_device": ([vp, i64, i32, dp, dp, i32, C.c_char_p, C.POINTER(vp)], C.c_int),
        "mp_problem_set_rhs": ([vp, vp, i32, dp], C.c_int),
        "mp_server_clock": ([], C.c_double),
        "mp_server_create": ([C.c_int, C.c_char_p, C.POINTER(vp)], C.c_int),
        "mp_server_submit": ([vp, C.c_uint64, i64, i32, dp, dp, dp, C.c_double, C.POINTER(C.c_uint64)], C.c_int),
        "mp_server_poll": ([vp, vp, i32, C.c_double, C.POINTER(i32)], C.c_int),
        "mp_server_stats_json": ([vp, C.c_char_p, sz], C.c_int),
        "mp_server_destroy": ([vp], C.c_int),
        "mp_getrf64": ([vp, i64, dp, i64, vp, i32, C.POINTER(i32)], C.c_int),
        "mp_factor_cache_put": ([vp, C.c_uint64, vp], C.c_int),
        "mp_factor_cache_get": ([vp, C.c_uint64, C.POINTER(vp), C.POINTER(i32)], C.c_int),
        "mp_factor_cache_evict": ([vp, C.c_uint64], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


lib = _load()


def check(status: int) -> None:
    if status != MP_OK:
        raise MPSolveError(status, (lib.mp_last_error() or b"").decode())


def _ptr(a) -> C.c_void_p | None:
    """Device pointer of a torch tensor / DeviceBuffer, or host pointer of a numpy array."""
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if isinstance(a, DeviceBuffer):
        return C.c_void_p(a.ptr)
    raise TypeError(f"cannot take a pointer of {type(a)}")


def _json_call(fn, *args, cap: int = 1 << 16) -> str:
    while True:
        buf = C.create_string_buffer(cap)
        st = fn(*args, buf, cap)
        if st == MP_OK:
            return buf.value.decode()
        msg = (lib.mp_last_error() or b"").decode()
        if "too small" in msg and cap < (1 << 28):
            cap *= 8
            continue
        raise MPSolveError(st, msg)


def methods() -> list[dict]:
    return json.loads(_json_call(lib.mp_methods_json, cap=1 << 18))


class DeviceBuffer:
    """Device memory owned by a context (mp_malloc)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = C.c_void_p()
        check(lib.mp_malloc(ctx.h, C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx: "Context", a: np.ndarray) -> "DeviceBuffer":
        a = np.asfortranarray(a)
        b = cls(ctx, a.nbytes)
        check(lib.mp_memcpy_h2d(ctx.h, C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_array(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype, order="F")
        check(lib.mp_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            check(lib.mp_free(self.ctx.h, C.c_void_p(self.ptr)))
            self.ptr = None


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.mp_create(C.byref(h), device))
        self.h = h

    def audit(self, X: np.ndarray, x_ref: np.ndarray | None = None, x_exact: np.ndarray | None = None) -> dict:
        """FP64 and double-double evaluation of a solution with forward errors (mp_audit_json)."""
        f = lambda a: None if a is None else np.asfortranarray(a, dtype=np.float64).reshape(self.n, self.nrhs, order="F")
        X, R, E = f(X), f(x_ref), f(x_exact)
        return json.loads(_json_call(lib.mp_audit_json, self.ctx.h, self.h, _ptr(X), _ptr(R), _ptr(E)))

    def close(self) -> None:
        if self.h:
            lib.mp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def caps(self) -> dict:
        return json.loads(_json_call(lib.mp_caps, self.h))

    def sync(self) -> None:
        check(lib.mp_synchronize(self.h))

    def profile(self, enable: bool = True) -> None:
        check(lib.mp_profile(self.h, 1 if enable else 0))

    def profile_report(self) -> dict:
        return json.loads(_json_call(lib.mp_profile_json, self.h))

    def buffer(self, a: np.ndarray) -> DeviceBuffer:
        return DeviceBuffer.from_array(self, a)

    def empty(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    # ---- standalone kernels (device buffers)
    def residual(self, n, nrhs, A, lda, X, ldx, B, ldb, R, ldr, T=None, ldt=None, stats=None, worst=None,
                 a_bits=64, dd=False, Rlo=None):
        if dd:
            check(lib.mp_residual_dd(self.h, n, nrhs, a_bits, _ptr(A), lda, _ptr(X), ldx, _ptr(B), ldb if B is not None else n,
                                     _ptr(R), _ptr(Rlo), ldr, _ptr(T), ldt or n, _ptr(stats), _ptr(worst)))
        else:
            check(lib.mp_residual_fp64(self.h, n, nrhs, a_bits, _ptr(A), lda, _ptr(X), ldx, _ptr(B), ldb if B is not None else n,
                                       _ptr(R), ldr, _ptr(T), ldt or n, _ptr(stats), _ptr(worst)))

    def residual_noalloc(self, n, nrhs, A, lda, X, ldx, B, ldb, R, ldr, T=None, ldt=None, stats=None, worst=None,
                         a_bits=64, Rlo=None):
        """mp_residual_noalloc: Rlo given -> the double-double residual, else FP64."""
        check(lib.mp_residual_noalloc(self.h, n, nrhs, a_bits, _ptr(A), lda, _ptr(X), ldx, _ptr(B),
                                      ldb if B is not None else n, _ptr(R), _ptr(Rlo), ldr, _ptr(T), ldt or n,
                                      _ptr(stats), _ptr(worst)))

    def getrf(self, n, LU, lda, ipiv, perm, nb=256, fmt="fp16", tile_scale=True) -> GetrfInfo:
        p = GetrfParams(nb, FMT[fmt], 1 if tile_scale else 0)
        info = GetrfInfo()
        check(lib.mp_getrf(self.h, n, _ptr(LU), lda, _ptr(ipiv), _ptr(perm), C.byref(p), C.byref(info)))
        return info

    def factor_cache_put(self, matrix_id: int, problem: "Problem") -> None:
        check(lib.mp_factor_cache_put(self.h, matrix_id, problem.h))

    def factor_cache_has(self, matrix_id: int) -> bool:
        """True when the id is cached and its problem holds factors; MPSolveError when the id is unknown."""
        p = C.c_void_p(); hf = C.c_int32()
        check(lib.mp_factor_cache_get(self.h, matrix_id, C.byref(p), C.byref(hf)))
        return bool(hf.value)

    def factor_cache_evict(self, matrix_id: int) -> None:
        check(lib.mp_factor_cache_evict(self.h, matrix_id))

    def getrf64(self, n, A, lda, ipiv, nb=128) -> int:
        info = C.c_int32()
        check(lib.mp_getrf64(self.h, n, _ptr(A), lda, _ptr(ipiv), nb, C.byref(info)))
        return info.value

    def getrs(self, n, nrhs, LU, lda, perm, R, ldr, D, ldd, scale=1.0):
        check(lib.mp_getrs(self.h, n, nrhs, _ptr(LU), lda, _ptr(perm), _ptr(R), ldr, _ptr(D), ldd, float(scale)))

    def pack_operand(self, S, s_idx, s_k, nidx, K, out, exps, counts, fmt="fp16", tile_scale=True, out_lo=None):
        check(lib.mp_pack_operand(self.h, FMT[fmt], 1 if tile_scale else 0, _ptr(S), s_idx, s_k, nidx, K, _ptr(out),
                                  _ptr(out_lo), _ptr(exps), _ptr(counts)))

    def batch_solve_packed(self, n, nrhs, A, a_off, B, b_off, X, **params):
        """mp_batch_solve on packed device buffers; n, nrhs, a_off, b_off are host integer sequences."""
        count = len(n)
        n_ = np.ascontiguousarray(n, dtype=np.int32); r_ = np.ascontiguousarray(nrhs, dtype=np.int32)
        ao = np.ascontiguousarray(a_off, dtype=np.int64); bo = np.ascontiguousarray(b_off, dtype=np.int64)
        res = (BatchResult * max(count, 1))()
        check(lib.mp_batch_solve(self.h, count, _ptr(n_), _ptr(r_), _ptr(A), _ptr(ao), _ptr(B), _ptr(bo), _ptr(X),
                                 json.dumps(params).encode(), res))
        out = [dict(terminal=BATCH_TERMINAL[r.terminal], outer_iters=r.outer_iters, info=r.info, eta_max=r.eta_max,
                    omega_max=r.omega_max) for r in res[:count]]
        return out, json.loads(_json_call(lib.mp_last_record_json, self.h))

    def batch_solve(self, systems, **params):
        """Solve a list of (A, B) host systems (n <= 512, nrhs <= 16) in one launch.
        Returns ([X_i], [per-system result], batch record)."""
        mats = [np.asfortranarray(A, dtype=np.float64) for A, _ in systems]
        rhs = [np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(A.shape[0], -1, order="F")) for A, B in systems]
        n = [A.shape[0] for A in mats]; nrhs = [B.shape[1] for B in rhs]
        a_off = np.cumsum([0] + [k * k for k in n])[:-1]; b_off = np.cumsum([0] + [k * r for k, r in zip(n, nrhs)])[:-1]
        Ah = np.concatenate([A.ravel(order="F") for A in mats]) if mats else np.zeros(1)
        Bh = np.concatenate([B.ravel(order="F") for B in rhs]) if rhs else np.zeros(1)
        dA, dB, dX = self.buffer(Ah), self.buffer(Bh), self.empty(max(Bh.nbytes, 8))
        try:
            res, rec = self.batch_solve_packed(n, nrhs, dA, a_off, dB, b_off, dX, **params)
            Xh = dX.to_array((Bh.size,), np.float64)
        finally:
            for b in (dA, dB, dX):
                b.free()
        X = [Xh[o:o + k * r].reshape(k, r, order="F") for o, k, r in zip(b_off, n, nrhs)]
        return X, res, rec

    def generate_into(self, A, B, x_exact=None, **case) -> bool:
        """Seeded system generated straight into device buffers (mp_generate_into); returns has_exact."""
        he = C.c_int32(0)
        check(lib.mp_generate_into(self.h, json.dumps(case).encode(), _ptr(A), _ptr(B), _ptr(x_exact), C.byref(he)))
        return bool(he.value)

    def problem_from_device(self, n, nrhs, A, B, borrow_a=True, case=None) -> "Problem":
        self_p = Problem.__new__(Problem)
        h = C.c_void_p()
        check(lib.mp_problem_from_device(self.h, n, nrhs, _ptr(A), _ptr(B), 1 if borrow_a else 0,
                                         json.dumps(case or {}).encode(), C.byref(h)))
        self_p.ctx, self_p.h, self_p.n, self_p.nrhs = self, h, n, nrhs
        return self_p

    # ---- problems and solves
    def problem(self, A: np.ndarray, B: np.ndarray, x_exact=None, x_ref=None, case: dict | None = None) -> "Problem":
        return Problem(self, A, B, x_exact, x_ref, case)


class Problem:
    @classmethod
    def generate(cls, ctx: "Context", **case) -> "Problem":
        """Device-generated seeded system (mp_problem_generate)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(lib.mp_problem_generate(ctx.h, json.dumps(case).encode(), C.byref(h)))
        self.ctx, self.h = ctx, h
        rec = self.case_record()
        self.n, self.nrhs = rec["n"], rec["nrhs"]
        return self

    def host_copy(self, want_a: bool = True):
        """(A, B, x_exact or None) copied back from the device."""
        rec = self.case_record()
        A = np.empty((self.n, self.n), dtype=np.float32 if rec["dtype"] == "fp32" else np.float64, order="F") if want_a else None
        B = np.empty((self.n, self.nrhs), dtype=np.float64, order="F")
        X = np.empty((self.n, self.nrhs), dtype=np.float64, order="F") if rec["has_exact_solution"] else None
        check(lib.mp_problem_copy_out(self.ctx.h, self.h, _ptr(A), _ptr(B), _ptr(X)))
        return A, B, X

    def reference(self) -> dict:
        check(lib.mp_problem_reference(self.ctx.h, self.h))
        return self.case_record()

    def __init__(self, ctx: Context, A: np.ndarray, B: np.ndarray, x_exact=None, x_ref=None, case=None):
        A = np.asfortranarray(A)
        if A.dtype not in (np.float64, np.float32):
            raise TypeError("A must be float64 or float32")
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B.reshape(-1, 1, order="F")
        n, nrhs = A.shape[0], B.shape[1]
        xe = None if x_exact is None else np.asfortranarray(np.asarray(x_exact, dtype=np.float64).reshape(n, nrhs, order="F"))
        xr = None if x_ref is None else np.asfortranarray(np.asarray(x_ref, dtype=np.float64).reshape(n, nrhs, order="F"))
        self.ctx, self.n, self.nrhs = ctx, n, nrhs
        h = C.c_void_p()
        cj = json.dumps(case or {}).encode()
        check(lib.mp_problem_from_host(ctx.h, n, nrhs, 32 if A.dtype == np.float32 else 64, _ptr(A), n, _ptr(B), n,
                                       _ptr(xe), _ptr(xr), cj, C.byref(h)))
        self.h = h

    def case_record(self) -> dict:
        return json.loads(_json_call(lib.mp_problem_case_json, self.h))

    def set_rhs(self, B_dev, nrhs: int) -> None:
        check(lib.mp_problem_set_rhs(self.ctx.h, self.h, nrhs, _ptr(B_dev)))
        self.nrhs = nrhs

    def solve(self, method: str, **params) -> tuple[np.ndarray, dict]:
        X = np.zeros((self.n, self.nrhs), dtype=np.float64, order="F")
        out = TrialOut()
        check(lib.mp_solve(self.ctx.h, self.h, method.encode(), json.dumps(params).encode(), C.byref(out), _ptr(X)))
        rec = json.loads(_json_call(lib.mp_last_record_json, self.ctx.h, cap=1 << 20))
        return X, rec

    def history(self) -> list[dict]:
        txt = _json_call(lib.mp_last_hist_json, self.ctx.h, cap=1 << 20)
        return [json.loads(l) for l in txt.splitlines() if l.strip()]

    def eval_metrics(self, X: np.ndarray) -> dict:
        X = np.asfortranarray(X, dtype=np.float64)
        return json.loads(_json_call(lib.mp_eval_metrics_json, self.ctx.h, self.h, _ptr(X)))

    def close(self) -> None:
        if self.h:
            lib.mp_problem_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (mp_server_*).  A and B are device buffers that stay alive until their
    request completes; X is a float64 Fortran-ordered host array filled on completion."""

    def __init__(self, device: int = 0, **params):
        h = C.c_void_p()
        check(lib.mp_server_create(device, json.dumps(params).encode(), C.byref(h)))
        self.h = h
        self._pending = {}

    @staticmethod
    def clock() -> float:
        return lib.mp_server_clock()

    def submit(self, matrix_id: int, n: int, nrhs: int, A, B, X: np.ndarray, arrival: float = -1.0) -> int:
        assert X.dtype == np.float64 and X.flags.f_contiguous and X.size == n * nrhs
        rid = C.c_uint64()
        check(lib.mp_server_submit(self.h, matrix_id, n, nrhs, _ptr(A), _ptr(B), _ptr(X), float(arrival), C.byref(rid)))
        self._pending[rid.value] = X
        return rid.value

    def poll(self, max_count: int = 1024, wait: float = 0.05) -> list[dict]:
        buf = (ServerCompletion * max_count)()
        cnt = C.c_int32()
        check(lib.mp_server_poll(self.h, buf, max_count, float(wait), C.byref(cnt)))
        out = []
        for c in buf[:cnt.value]:
            self._pending.pop(c.request_id, None)
            out.append(dict(id=c.request_id, terminal=SERVER_TERMINAL[c.terminal], path=("batch", "large")[c.path],
                            cache_hit=bool(c.cache_hit), outer_iters=c.outer_iters, eta_max=c.eta_max,
                            omega_max=c.omega_max, queue_s=c.queue_s, service_s=c.service_s, latency_s=c.latency_s))
        return out

    def stats(self) -> dict:
        buf = C.create_string_buffer(4096)
        check(lib.mp_server_stats_json(self.h, buf, len(buf)))
        return json.loads(buf.value)

    def close(self) -> None:
        if self.h:
            check(lib.mp_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
