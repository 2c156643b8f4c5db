"""Writes tests/golden/e1_vectors.json: input/output vectors of the reference's E1 scalar model (tpir, the
operation-rounded arithmetic of e1_scalar/three_precision_lab/tpir/arithmetic.py:25-159, its problem generator
tpir/problems.py:8-37 and the classical refinement loop tpir/solver.py:44-107) for small systems (n <= 16).

Run in the build container only, where the reference is present:  python tests/golden/make_e1_vectors.py
The reference's code is imported to produce data; nothing of it is kept: the fixture holds inputs, bit patterns
of outputs and terminal statuses.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, "/root/reference/e1_scalar/three_precision_lab")
from tpir.arithmetic import lu_factor, residual, scaled_solve        # noqa: E402
from tpir.problems import make_problem                                # noqa: E402
from tpir.solver import run_solver                                    # noqa: E402


def hexa(a):
    return np.ascontiguousarray(a).tobytes().hex()


out = {"problems": [], "factors": [], "solves": [], "residuals": [], "runs": []}
for n, kappa, seed in [(8, 1e2, 9200), (16, 1e4, 9201), (16, 1e2, 0)]:
    spec = dict(n=n, kappa=kappa, seed=seed, family="spectral", rhs="random")
    a, b, _, _ = make_problem(spec)
    out["problems"].append(dict(spec=spec, a32=hexa(a), b32=hexa(b),
                                a_sha256=hashlib.sha256(a.tobytes()).hexdigest(), b_sha256=hashlib.sha256(b.tobytes()).hexdigest()))
    for bits in (16, 32):
        f = lu_factor(a, bits)
        out["factors"].append(dict(n=n, seed=seed, kappa=kappa, bits=bits, lu=hexa(f.lu), perm=f.permutation.tolist(),
                                   growth=float(f.growth)))
        for work, rbits, norm in [(32, 32, True), (32, 64, True), (32, 64, False)]:
            d, ex = scaled_solve(f, b, work, rbits, norm)
            out["solves"].append(dict(n=n, seed=seed, kappa=kappa, bits=bits, work=work, rhs_bits=rbits, normalize=norm,
                                      d=hexa(d), exponent=int(ex)))
    x = np.linalg.solve(a.astype(np.float64), b.astype(np.float64)).astype(np.float32)
    for rbits in (32, 64):
        out["residuals"].append(dict(n=n, seed=seed, kappa=kappa, bits=rbits, x=hexa(x), r=hexa(residual(a, b, x, rbits))))
    for m in ["hss", "hsd", "sss", "ssd", "hdd", "sdd", "input16_ssd", "half_factors_solve32", "direct16", "hsd_no_norm"]:
        rec, xs = run_solver(a, b, m, 30)
        out["runs"].append(dict(n=n, seed=seed, kappa=kappa, method=m, status=rec["status"], iterations=rec["iterations"],
                                factor_hash=rec.get("factor_hash"), initialization=rec["initialization"],
                                exponents=[int(e) for e in rec["normalization_exponents"]],
                                x_last=hexa(xs[-1]) if len(xs) else None, x_dtype=str(xs.dtype), n_iterates=int(len(xs))))
# failure cases of the loop: singular input, zero right-hand side
for name, a, b in [("singular", np.ones((3, 3), np.float32), np.ones(3, np.float32)),
                   ("zero_rhs", np.eye(4, dtype=np.float32), np.zeros(4, np.float32))]:
    rec, xs = run_solver(a, b, "hsd", 30)
    out["runs"].append(dict(case=name, a32=hexa(a), b32=hexa(b), method="hsd", status=rec["status"], iterations=rec["iterations"],
                            failure=rec.get("failure"), n_iterates=int(len(xs))))
Path(__file__).with_name("e1_vectors.json").write_text(json.dumps(out, indent=0) + "\n")
print("wrote", Path(__file__).with_name("e1_vectors.json"), sum(len(v) for v in out.values()), "entries")
