"""The trial record contract: every key the reference's collector reads from trials_*.jsonl, with the type
its use needs (collect.py:91-136 indexes these keys directly, compares several numerically, e.g.
factor_total_s > 0, rep >= 0, irs_niters != -9999, and passes the per-right-hand-side ones to a list cleaner).
The records of a converged solve, of a vendor method ID that is unsupported here, and of a trial skipped by
the memory admission are checked; a key that disappears or changes type fails the test.
"""
import json

import numpy as np
import pytest

from cpsc402_midterm_project_amd.problems import randsvd

pytestmark = pytest.mark.gpu

NUM = (int, float)
# key -> accepted types (None: the value may be null); the collector's direct indexing d[key]
REQUIRED = {
    "tag": (str,), "warmup": (bool,), "mode": (str,), "method_id": (str,), "rep": (int,), "tau": NUM, "target": (str,),
    "ok": (bool,), "skipped": (bool,), "eta": (list,), "omega": (list,), "terminal_reason": (str,),
    "host_e2e_s": NUM, "device_ready_wall_s": NUM, "method_internal_total_s": NUM, "factor_total_s": NUM,
    "initial_solve_s": NUM, "refine_s": NUM, "eval_s": NUM, "wall_including_eval_s": NUM, "x_dtype": (str,),
    "timestamp": (str,), "working": (int,), "final_stage": (str,), "used_fallback": (bool, int), "skip_reason": (str,),
    "outer_iters": (int,), "inner_iters_total": (int,), "matvecs": (int,), "precond_applies": (int,),
    "triangular_solves": (int,), "restarts": (int,), "true_residual_checks": (int,), "refactorizations": (int,),
    "stages": (list,), "stage_times_s": (list,), "irs_niters": (int,), "irs_outer_niters": (int,), "magma_iter": (int,),
    "magma_facto_time_s": NUM, "factor_info": (int,), "growth_proxy": NUM + (type(None),), "conv_overflow": (int,),
    "conv_subnormal": (int,), "scale_mode": (int,), "nb": (int,), "relres_inf": (list,), "relres_2": (list,),
    "rinf": (list,), "r2": (list,), "xinf": (list,), "worst_row": (list,), "solver_final_eta_max": NUM + (type(None),),
    "solver_final_omega_max": NUM + (type(None),), "peak_device_bytes_tracked": (int,), "device_free_after_bytes": (int,),
    "method_extra_bytes_tracked": (int,), "irs_workspace_bytes": (int,),
}
# read through d.get(...) or only when a flag is set
OPTIONAL = {"error": (str,), "x_path": (str,), "x_sha256": (str,), "panel": (str,), "b_integrity_ok": (bool,),
            "metrics_finite": (bool,), "has_exact": (bool,), "has_fwd": (bool,), "reuse": (bool,)}
FWD = {"has_fwd": ("finf", "f2", "fcomp_nonzero", "fzero_count"), "has_exact": ("einf", "e2", "ecomp_nonzero", "ezero_count")}


def _check(rec):
    for k, ty in REQUIRED.items():
        assert k in rec, f"missing key {k}"
        assert isinstance(rec[k], ty) and not (isinstance(rec[k], bool) and bool not in ty and int not in ty), (k, rec[k])
    for k, ty in OPTIONAL.items():
        if k in rec and rec[k] is not None:
            assert isinstance(rec[k], ty), (k, rec[k])
    for flag, keys in FWD.items():
        if rec.get(flag):
            for k in keys:
                assert k in rec, f"{flag} set but {k} missing"
    assert rec["rep"] >= 0 or rec["warmup"]
    for k in ("eta", "omega", "rinf", "xinf", "relres_inf", "relres_2", "r2"):
        assert all(v is None or isinstance(v, NUM) for v in rec[k]), k
    json.dumps(rec, allow_nan=False)                # non-finite numbers are written as null


def test_record_contract_of_a_converged_solve(ctx):
    A, b, xt, _ = randsvd(300, 1e3, 4, gpu=False)
    P = ctx.problem(A, b, x_exact=xt)
    P.reference()
    X, rec = P.solve("hetero16_gmresir_w64_r64", nb=64)
    assert rec["ok"] and rec["terminal_reason"] == "converged"
    assert rec["has_exact"] and rec["has_fwd"] and rec["factor_total_s"] > 0 and rec["eta"][0] <= rec["tau"]
    _check(rec)
    X, rec = P.solve("native_lu64")
    assert rec["terminal_reason"] == "direct_solve"
    _check(rec)
    P.close()


def test_record_contract_of_unsupported_and_skipped_trials(ctx):
    A, b, xt, _ = randsvd(128, 1e2, 5, gpu=False)
    P = ctx.problem(A, b)
    X, rec = P.solve("cusolver_irs_gmres_r16f")
    assert not rec["ok"] and rec["terminal_reason"] == "unsupported_on_platform"
    _check(rec)
    X, rec = P.solve("mx16_gmresir_w64_rdd", nb=64, hbm_budget_bytes=1 << 10)
    assert rec["skipped"] and rec["terminal_reason"] == "memory_skip"
    _check(rec)
    P.close()
