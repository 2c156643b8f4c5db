"""The `mpsolve` command-line boundary (src/mpsolve.cu:130-240): commands, flags, output files and records.

CPU tests check the argument handling that needs no device; GPU tests run small trial campaigns through the
binary and check the files a campaign driver reads: the case record, the binary vectors, one trial record per
(mode, rep, method) with the reference's rotation order, solution files whose SHA-256 matches the record, and
the host-mode end-to-end time.
"""
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "cpsc402_midterm_project_amd" / "mpsolve"


def _run(*args, timeout=120):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built (python -m cpsc402_midterm_project_amd.build)")
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_cli_usage_without_command():
    r = _run()
    assert r.returncode == 2
    assert "usage: mpsolve run|regen|caps|selftest|audit" in r.stderr


def test_cli_caps_reports_device_or_fatal_json():
    r = _run("caps")
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    if r.returncode == 0:
        assert rec["device"] and rec["total_mem_bytes"] > 0
    else:
        assert r.returncode == 1 and "fatal" in rec and r.stderr.startswith("FATAL:")


@pytest.mark.gpu
def test_cli_run_campaign(tmp_path, ctx):
    # `ctx` keeps a context of this process open on the device while the CLI runs: the runtime then refuses
    # cooperative launches, and the factorization must fall back to ordinary ones
    n, nrhs, methods = 256, 2, ["native_lu64", "mx16_ir_w64_r64"]
    print(f"[cli run] n={n} nrhs={nrhs} methods={methods}", flush=True)
    r = _run("run", "--n", n, "--nrhs", nrhs, "--family", "dyadic", "--rhs", "exact_sign", "--seed", 17,
             "--methods", ",".join(methods), "--modes", "resident,host", "--reps", 2, "--warmup", 1,
             "--case-id", "c1", "--tag", "t", "--out", tmp_path, "--nb", 64, "--embed-a")
    assert r.returncode == 0, r.stderr + r.stdout
    case = json.loads((tmp_path / "case_c1.json").read_text())
    assert case["n"] == n and case["nrhs"] == nrhs and case["family"] == "dyadic"
    assert case["kappa_nominal"] == 100 and case["target"] == "fp64" and case["tau"] == 100 * 2.0 ** -53
    assert case["has_exact_solution"] and case["mem_budget_bytes"] > 0 and len(case["a_sha256"]) == 64
    xe = np.fromfile(tmp_path / "Xexact_c1.f64")
    xr = np.fromfile(tmp_path / "Xref_c1.f64")
    b = np.fromfile(tmp_path / "B_c1.f64")
    assert xe.size == xr.size == b.size == n * nrhs
    A = np.fromfile(tmp_path / "A_c1.f64").reshape(n, n, order="F")
    B = b.reshape(n, nrhs, order="F")

    def backward_error(x):                                   # componentwise, in FP64 on the host
        X = x.reshape(n, nrhs, order="F")
        return (np.abs(B - A @ X) / (np.abs(A) @ np.abs(X) + np.abs(B))).max()

    assert backward_error(xe) <= 4 * n * 2.0 ** -53 and backward_error(xr) <= 4 * n * 2.0 ** -53
    assert hashlib.sha256(b.tobytes()).hexdigest() == case["b_sha256"]

    recs = [json.loads(l) for l in (tmp_path / "trials_c1.jsonl").read_text().splitlines()]
    assert len(recs) == 2 * 3 * len(methods)
    i = 0
    for mode in ("resident", "host"):
        for rep in (-1, 0, 1):
            rot = (1 + rep + 1) % len(methods)                 # order-seed 1, warmup 1
            for mi in range(len(methods)):
                rec = recs[i]; i += 1
                assert (rec["mode"], rec["rep"], rec["warmup"]) == (mode, rep, rep < 0)
                assert rec["method_id"] == methods[(mi + rot) % len(methods)]
                assert rec["ok"], (rec["method_id"], rec["mode"], rec["rep"], rec["terminal_reason"], rec["error"])
                assert rec["case_id"] == "c1" and rec["tag"] == "t" and rec["nb"] == 64
                if rec["method_id"] == "mx16_ir_w64_r64":
                    # this process's context shares the device: every panel ran, and a refused cooperative launch
                    # is recorded with the runtime's code
                    assert rec["panel_launches"] >= 1
                    assert (rec["panel_coop_refused"] == 0) == (rec["panel_coop_error"] == ""), rec["panel_coop_error"]
                    print(f"[cli run] panels {rec['panel_launches']}, refused {rec['panel_coop_refused']} "
                          f"({rec['panel_coop_error'] or 'none'})", flush=True)
                assert rec["eta_max"] <= 100 * 2.0 ** -53, rec["eta_max"]
                if mode == "host":
                    assert rec["host_e2e_s"] > rec["device_ready_wall_s"] > 0
                else:
                    assert rec["host_e2e_s"] == -1
                if rep < 0:
                    assert rec["x_path"] == ""
                    continue
                assert rec["x_path"] == f"X_c1_{rec['method_id']}_{mode}_r{rep}_t.f64"
                xb = (tmp_path / rec["x_path"]).read_bytes()
                assert hashlib.sha256(xb).hexdigest() == rec["x_sha256"]
                assert backward_error(np.frombuffer(xb, np.float64)) <= 4 * n * 2.0 ** -53
    hist = [json.loads(l) for l in (tmp_path / "hist_c1.jsonl").read_text().splitlines()]
    assert any(h["method_id"] == "mx16_ir_w64_r64" and h["kind"] == "outer" for h in hist)
    assert all(h["rep"] >= 0 for h in hist)

    # regen reproduces the same system from the same flags
    g = _run("regen", "--n", n, "--nrhs", nrhs, "--family", "dyadic", "--rhs", "exact_sign", "--seed", 17)
    assert g.returncode == 0, g.stderr
    rg = json.loads(g.stdout)
    assert rg["a_sha256"] == case["a_sha256"] and rg["b_sha256"] == case["b_sha256"]


@pytest.mark.gpu
def test_cli_selftest_and_unsupported_flags(tmp_path):
    print("[cli selftest] n=192", flush=True)
    r = _run("selftest", "--n", 192, "--kappa", 1)             # identity spectrum: an orthogonal matrix
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout)
    assert rec["selftest"] == "generator" and rec["n"] == 192
    assert rec["max_orthogonality_deviation"] < 1e-12
    r = _run("selftest", "--n", 64, "--dtype", "fp32")
    assert r.returncode == 1 and "fatal" in json.loads(r.stdout)


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(256, 9100), (384, 9101)])
def test_cli_selftest_reference_anchor(n, seed):
    """The reference's packaged generator check (finalize.py:80: `mpsolve selftest --n 256|384 --family
    prescribed --spectrum identity --seed 9100|9101`) reports an orthogonality deviation of about 1.5e-15 and no
    zero entries (finalize.py:85); the rotation-stage generator reproduces both."""
    r = _run("selftest", "--n", n, "--family", "prescribed", "--spectrum", "identity", "--seed", seed)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout)
    assert 5e-16 <= rec["max_orthogonality_deviation"] <= 4e-15, rec
    assert rec["zero_entries"] == 0 and rec["min_abs"] > 0


@pytest.mark.gpu
def test_cli_fp32_storage_and_rowcol_scaling(tmp_path):
    """--dtype fp32 stores A and B in FP32 (files .f32); --rowcol-span scales rows and columns by powers of two;
    both systems are solved to the FP32 target by the default method."""
    n = 300
    print(f"[cli fp32 rowcol] n={n}", flush=True)
    r = _run("run", "--n", n, "--kappa", 1e3, "--dtype", "fp32", "--rowcol-span", 6, "--methods", "mx16_ir_w64_r64",
             "--reps", 1, "--warmup", 0, "--case-id", "f", "--out", tmp_path, "--embed-a")
    assert r.returncode == 0, r.stderr + r.stdout
    case = json.loads((tmp_path / "case_f.json").read_text())
    assert case["dtype"] == "fp32" and case["rowcol_span"] == 6 and case["target"] == "fp32"
    A = np.fromfile(tmp_path / "A_f.f32", dtype=np.float32).reshape(n, n, order="F")
    b = np.fromfile(tmp_path / "B_f.f32", dtype=np.float32)
    assert A.size == n * n and b.size == n
    # row and column scales are powers of two spanning up to 2^6: max/min of row maxima is well above 1
    rm = np.abs(A).max(axis=1)
    assert rm.max() / rm.min() > 4
    rec = json.loads((tmp_path / "trials_f.jsonl").read_text().splitlines()[0])
    assert rec["ok"] and rec["terminal_reason"] == "converged" and rec["eta_max"] <= 100 * 2.0 ** -24


@pytest.mark.gpu
def test_mpaudit_regenerates_and_reevaluates(tmp_path, oracle):
    """mpaudit (src/mpaudit.cu:8-39) in a fresh process: the regenerated hashes equal the case record's, every
    listed solution gets its SHA-256 and FP64 metrics equal to the trial record's, the double-double eta agrees
    with the oracle's, forward errors come from --xexact, and unreadable entries are reported, not fatal."""
    n, nrhs = 200, 2
    case = ["--n", n, "--nrhs", nrhs, "--family", "prescribed", "--kappa", 1e4, "--rhs", "random", "--seed", 5]
    print(f"[mpaudit] n={n}", flush=True)
    r = _run("run", *case, "--methods", "mx16_ir_w64_r64,native_lu32", "--reps", 1, "--warmup", 0, "--case-id", "a1",
             "--out", tmp_path, "--nb", 64, "--embed-a", "--save-x", 1)
    assert r.returncode == 0, r.stderr + r.stdout
    crec = json.loads((tmp_path / "case_a1.json").read_text())
    trials = [json.loads(l) for l in (tmp_path / "trials_a1.jsonl").read_text().splitlines() if l.strip()]
    sols = [t for t in trials if t.get("x_path")]
    assert sols, "no saved solutions"
    lst = tmp_path / "xlist.txt"
    lst.write_text("".join(f"{tmp_path / t['x_path']} {t['x_dtype']}\n" for t in sols) + f"{tmp_path / 'missing.f64'} float64\n")
    xe = np.linspace(-1, 1, n * nrhs)
    (tmp_path / "xe.f64").write_bytes(xe.tobytes())
    aud = ROOT / "cpsc402_midterm_project_amd" / "mpaudit"
    p = subprocess.run([str(aud), "--case-id", "a1", "--x-list", str(lst), "--xexact", str(tmp_path / "xe.f64"),
                        "--dump-a", str(tmp_path / "dump.f64"), *map(str, case)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr + p.stdout
    lines = [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]
    head = lines[0]
    assert head["regenerated_a_sha256"] == crec["a_sha256"] and head["regenerated_b_sha256"] == crec["b_sha256"]
    assert head["regenerated_tile_sha256"] == crec["tile_sha256"] and head["audit_process_pid"]
    A = np.fromfile(tmp_path / "dump.f64").reshape(n, n, order="F")
    assert hashlib.sha256(A.tobytes(order="F")).hexdigest() == crec["a_sha256"]
    B = np.fromfile(tmp_path / "B_a1.f64").reshape(n, nrhs, order="F")
    by_path = {l["x_path"]: l for l in lines[1:]}
    assert by_path[str(tmp_path / "missing.f64")]["error"] == "cannot open"
    for t in sols:
        rec = by_path[str(tmp_path / t["x_path"])]
        raw = (tmp_path / t["x_path"]).read_bytes()
        assert rec["x_sha256"] == hashlib.sha256(raw).hexdigest() == t["x_sha256"]
        X = np.frombuffer(raw, dtype=np.float32 if t["x_dtype"] == "float32" else np.float64).astype(np.float64).reshape(n, nrhs, order="F")
        assert max(rec["eta_fp64"]) == t["eta_max"]
        m = oracle.metrics(A, B, X, dd=True)
        np.testing.assert_allclose(rec["eta_dd"], m["eta"], rtol=1e-6, atol=1e-30)
        assert rec["has_exact"] and np.allclose(rec["finf_exact"], np.abs(X - xe.reshape(n, nrhs, order="F")).max(axis=0)
                                                 / np.abs(xe.reshape(n, nrhs, order="F")).max(axis=0), rtol=1e-12)


@pytest.mark.gpu
def test_cli_reuse_keeps_factors_across_trials(tmp_path):
    """--reuse (src/mpsolve.cu:173): the factors of the first trial of a method serve its later trials."""
    r = _run("run", "--n", 300, "--kappa", 1e3, "--methods", "mx16_gmresir_w64_rdd", "--reps", 3, "--warmup", 0,
             "--case-id", "ru", "--out", tmp_path, "--nb", 64, "--reuse")
    assert r.returncode == 0, r.stderr + r.stdout
    recs = [json.loads(l) for l in (tmp_path / "trials_ru.jsonl").read_text().splitlines() if l.strip()]
    assert [x["refactorizations"] for x in recs] == [1, 0, 0], [x["refactorizations"] for x in recs]
    assert all(x["reuse"] and x["terminal_reason"] == "converged" for x in recs)
    assert len({x["x_sha256"] for x in recs}) == 1           # the same factors give the same solution
