"""The E1 track's GPU leg: the driver solves its stored FP32 systems with the library's methods of the same
precision roles through the C ABI.  The GPU factorizations are not the scalar model (FP32 panel, FP16 operands
with FP32 accumulation on MFMA), so the legs are compared by their errors: each GPU method meets its working
precision's backward-error target, and its forward error is within a small factor of the model's for the same
roles (or of 10 u sqrt(kappa) of the working precision when the model's is smaller).  Both legs refine until
their iterates stop improving (the model to its correction cap or an unchanged update, the GPU to its
stagnation rule), so their final forward errors are comparable."""
import pytest

from cpsc402_midterm_project_amd import e1

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,kappa", [(64, 1e2), (128, 1e3)])
def test_e1_gpu_leg(ctx, n, kappa):
    print(f"[e1] n={n} kappa={kappa:g}", flush=True)
    case, trials = e1.run_case(n, kappa, 9200, methods=["hsd", "ssd", "sdd", "direct64"], ctx=ctx)
    model = {t["method"]: t for t in trials if t["leg"] == "model"}
    gpu = {t["method"]: t for t in trials if t["leg"] == "gpu"}
    assert set(gpu) == {"hsd", "ssd", "sdd", "direct64"}
    for m, g in gpu.items():
        u = e1.UNIT[e1.METHODS[m].work]
        print(f"[e1] {m}: model {model[m]['status']} finf {model[m]['finf']:.2e}; gpu {g['gpu_method']} {g['terminal']} "
              f"eta {g['eta']:.2e} finf {g['finf']:.2e}", flush=True)
        assert g["ok"] and g["terminal"] in ("stagnation", "max_outer", "direct_solve"), g
        assert g["eta"] <= 100 * u
        assert g["finf"] <= 8 * max(model[m]["finf"], 10 * u * kappa ** 0.5), (g["finf"], model[m]["finf"])
