"""CPU tests of the boundary: libmpsolve.so loads, exports every symbol declared in include/mpsolve.h, reports
errors as statuses (no exception crosses the ABI), and the method registry mirrors the reference's IDs."""
import ctypes
import json
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    txt = (ROOT / "include" / "mpsolve.h").read_text()
    return sorted(set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import cpsc402_midterm_project_amd as pkg
    lib = ctypes.CDLL(str(pkg.LIB_PATH))
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(pkg.EXPORTS) == declared


def test_no_device_is_a_status_not_a_crash():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is visible")
    from cpsc402_midterm_project_amd import Context, MPSolveError
    with pytest.raises(MPSolveError) as e:
        Context(0)
    assert e.value.status == -6


REFERENCE_IDS = [  # src/mpsolve.cu:20-60
    "native_lu32", "native_lu64", "native_lu64_promoted32", "native_ir32_w32_r64", "native_ir32_w32_r32",
    "native_ir32_w64_r64", "native_ir64_w64_r64", "native_ir32_w64_r64_twosided", "native_ir32_w64_r64_global",
    "native_gmresir32_w64_r64", "native_gmresir32_w32_r64", "native_leftgmres32_app32_w64_r64",
    "native_leftgmres32_app64_w64_r64", "native_gmresir32_app64_w64_r64", "hetero16_lu_direct", "hetero16_ir_w64_r64",
    "hetero16_ir_w32_r64", "hetero16_ir_w64_r64_noscale", "hetero16_ir_w64_r64_twosided", "hetero16_gmresir_w64_r64",
    "hetero16_gmresir_w32_r64", "hetero16_gmresir_app64_w64_r64", "hetero16_leftgmres_app64_w64_r64",
    "custom32_lu_direct", "custom32_ir_w64_r64", "custom32_ir_w32_r64", "policy_hetero16_stationary_only",
    "policy_hetero16_gmres_only", "policy_hetero16_escalate_immediate", "policy_hetero16_cost_aware",
    "policy_native32_escalate_immediate"]


def test_method_registry_covers_reference_ids():
    from cpsc402_midterm_project_amd import methods
    cat = {m["method_id"]: m for m in methods()}
    for mid in REFERENCE_IDS:
        assert mid in cat, mid
    # contracts that the reference pins in make_method (src/mpsolve.cu:21-48)
    assert cat["native_lu64"]["kind"] == "direct" and cat["native_lu64"]["factorization"] == "cs64"
    assert cat["hetero16_gmresir_w64_r64"]["gmres"] and cat["hetero16_gmresir_w64_r64"]["scale_mode"] == 1
    assert cat["hetero16_ir_w64_r64_noscale"]["scale_mode"] == 0
    assert cat["hetero16_ir_w32_r64"]["working"] == 32
    assert cat["policy_hetero16_escalate_immediate"]["policy"] == "budget"
    assert cat["mx16_gmresir_w64_rdd"]["residual"].startswith("double-double")
    assert cat["policy_mx8_ladder"]["policy"] == "ladder"
    # options the build does not implement are declared, not silently changed
    assert cat["hetero16_ir_w64_r64_twosided"]["kind"] != "unsupported" and cat["hetero16_ir_w64_r64_twosided"]["scale_mode"] == 2
    for mid in ("native_leftgmres32_app32_w64_r64", "native_leftgmres32_app64_w64_r64",
                "native_gmresir32_app64_w64_r64", "hetero16_leftgmres_app64_w64_r64", "hetero16_gmresir_app64_w64_r64"):
        assert cat[mid]["kind"] != "unsupported" and cat[mid]["gmres"], mid
    assert cat["native_leftgmres32_app64_w64_r64"]["apply64"]
    json.dumps(cat)


def test_sha256_known_answers():
    """mp_sha256_hex (host-only) against the FIPS 180-2 examples and hashlib; the records' a_sha256 / x_sha256
    come from the same routine."""
    import hashlib
    import cpsc402_midterm_project_amd as pkg
    lib = ctypes.CDLL(str(pkg.LIB_PATH))
    lib.mp_sha256_hex.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p]

    def sha(b):
        out = ctypes.create_string_buffer(65)
        buf = ctypes.create_string_buffer(b, len(b))
        assert lib.mp_sha256_hex(buf, len(b), out) == 0
        return out.value.decode()

    assert sha(b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert sha(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    msg = b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq"
    assert sha(msg) == "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"
    for n in (55, 56, 63, 64, 65, 1000, 1 << 16):
        b = bytes((i * 7 + 3) & 0xFF for i in range(n))
        assert sha(b) == hashlib.sha256(b).hexdigest(), n
