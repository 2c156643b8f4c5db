"""GPU parity of the hot-path kernels against the CPU oracle (through the C ABI).

Bars: residual FP64 within the (n+2)u|A||X| rounding bound of the oracle's sequential-FMA result (the kernel
sums column chunks in a fixed order); double-double residual within 2^-100 |A||X| of the oracle's DD value;
getrf bitwise equal to the oracle (factors, ipiv, perm) on exact-arithmetic systems for every trailing format,
and on random systems a backward error at the format's level plus pivot agreement; getrs within an FP32
triangular-solve error bound of the oracle.
"""
import numpy as np
import pytest

from cpsc402_midterm_project_amd.problems import exact_lu_system, randsvd

pytestmark = pytest.mark.gpu
U64 = 2.0 ** -53


def _dev(ctx, a):
    return ctx.buffer(np.asfortranarray(a))


@pytest.mark.parametrize("n,nrhs,a_bits", [(1, 1, 64), (37, 3, 64), (1000, 1, 64), (777, 11, 64), (513, 2, 32), (2048, 8, 64)])
def test_residual_fp64_matches_oracle(ctx, oracle, n, nrhs, a_bits):
    print(f"[residual fp64] n={n} nrhs={nrhs} a_bits={a_bits}", flush=True)
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n)).astype(np.float32 if a_bits == 32 else np.float64, order="F")
    X = np.asfortranarray(rng.standard_normal((n, nrhs)))
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dA, dX, dB = _dev(ctx, A), _dev(ctx, X), _dev(ctx, B)
    dR, dT = ctx.empty(8 * n * nrhs), ctx.empty(8 * n * nrhs)
    dS, dW = ctx.empty(32 * nrhs), ctx.empty(4 * nrhs)
    ctx.residual(n, nrhs, dA, n, dX, n, dB, n, dR, n, dT, n, dS, dW, a_bits=a_bits)
    ctx.sync()
    R = dR.to_array((n, nrhs), np.float64); T = dT.to_array((n, nrhs), np.float64)
    S = dS.to_array((4 * nrhs,), np.float64).reshape(nrhs, 4); W = dW.to_array((nrhs,), np.int32)
    Ro, To = oracle.resid_fp64(A, X, B)
    bound = (n + 2) * U64 * To
    assert np.all(np.abs(R - Ro) <= bound), np.max(np.abs(R - Ro) / np.maximum(bound, 1e-300))
    assert np.allclose(T, To, rtol=(n + 2) * U64, atol=0)
    So, Wo = oracle.colstats(R, X, T)        # stats of the GPU's own R/T: exact reductions
    np.testing.assert_allclose(S[:, :3], So[:, :3], rtol=0, atol=0)
    np.testing.assert_allclose(S[:, 3], So[:, 3], rtol=1e-12)
    np.testing.assert_array_equal(W, Wo)


@pytest.mark.parametrize("n,nrhs", [(64, 1), (1000, 2), (1537, 4)])
def test_residual_dd_matches_oracle(ctx, oracle, n, nrhs):
    print(f"[residual dd] n={n} nrhs={nrhs}", flush=True)
    A, b, xt, _ = randsvd(n, 1e6, 7 + n, nrhs=nrhs, gpu=False)
    F = oracle.factor(A, "fp32", 64)
    X = oracle.getrs(F["LU"], F["perm"], b, F["g"])           # a good approximate solution: tiny residual
    dA, dX, dB = _dev(ctx, A), _dev(ctx, X), _dev(ctx, b)
    dH, dL, dT = ctx.empty(8 * n * nrhs), ctx.empty(8 * n * nrhs), ctx.empty(8 * n * nrhs)
    ctx.residual(n, nrhs, dA, n, dX, n, dB, n, dH, n, dT, n, dd=True, Rlo=dL)
    ctx.sync()
    H = dH.to_array((n, nrhs), np.float64); L = dL.to_array((n, nrhs), np.float64)
    Ho, Lo = oracle.resid_dd(A, X, b)
    _, To = oracle.resid_fp64(A, X, b)
    assert np.all(np.abs((H - Ho) + (L - Lo)) <= 2.0 ** -100 * To * n)
    # hi is the FP64 rounding of the pair
    assert np.all(np.abs(L) <= np.spacing(np.abs(H)) / 2 + 1e-320)


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "fp8", "fp32", "fp16x3"])
@pytest.mark.parametrize("n,nb", [(64, 32), (200, 64), (513, 128), (700, 256)])
def test_getrf_exact_bitwise(ctx, oracle, fmt, n, nb):
    print(f"[getrf exact] fmt={fmt} n={n} nb={nb}", flush=True)
    A, L, U, perm = exact_lu_system(n, 11 + n, fmt)
    dLU = _dev(ctx, A)
    dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
    info = ctx.getrf(n, dLU, n, dip, dpm, nb=nb, fmt=fmt, tile_scale=True)
    LU = dLU.to_array((n, n), np.float32); ipiv = dip.to_array((n,), np.int64); pm = dpm.to_array((n,), np.int32)
    LUo, ipo, pmo, info_o, _ = oracle.getrf(A, nb, fmt, True)
    assert info.info == info_o == 0
    np.testing.assert_array_equal(ipiv, ipo)
    np.testing.assert_array_equal(pm, pmo)
    np.testing.assert_array_equal(pm, perm)
    np.testing.assert_array_equal(LU, LUo)


@pytest.mark.parametrize("fmt", ["fp16", "bf16", "fp8", "fp32", "fp16x3"])
def test_getrf_random_backward_error(ctx, oracle, fmt):
    """Random Gaussian matrix: the GPU factors have the backward error of the same algorithm run by the oracle
    (within a factor 4 either way: too small would mean the operand rounding was skipped; the MFMA sums in
    FP32 where the oracle sums in FP64 then rounds), |L| <= 1, and the composed permutation equals the LAPACK
    swap sequence.  The first panel involves no trailing update, so its pivots must equal the oracle's; after
    that a near-tie resolved differently by the two summation orders changes every later pivot."""
    n, nb = 1024, 128
    print(f"[getrf random] fmt={fmt} n={n}", flush=True)
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((n, n)).astype(np.float32))
    dLU = _dev(ctx, A); dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
    info = ctx.getrf(n, dLU, n, dip, dpm, nb=nb, fmt=fmt)
    LU = dLU.to_array((n, n), np.float32); pm = dpm.to_array((n,), np.int32); ipiv = dip.to_array((n,), np.int64)
    L, U = oracle.unpack_lu(LU)
    err = np.abs(A[pm].astype(np.float64) - L @ U).max() / np.abs(A).max()
    LUo, ipo, pmo, _, _ = oracle.getrf(A, nb, fmt, True)
    Lo, Uo = oracle.unpack_lu(LUo)
    err_o = np.abs(A[pmo].astype(np.float64) - Lo @ Uo).max() / np.abs(A).max()
    assert info.info == 0
    assert 0.25 * err_o <= err <= 4.0 * err_o, (err, err_o)
    assert np.abs(np.tril(LU, -1)).max() <= 1.0
    np.testing.assert_array_equal(ipiv[:nb], ipo[:nb])
    p = np.arange(n)
    for i in range(n):
        j = ipiv[i] - 1
        p[i], p[j] = p[j], p[i]
    np.testing.assert_array_equal(p, pm)


@pytest.mark.parametrize("fmt,n,nb", [("fp16", 3000, 256), ("fp16x3", 2176, 128), ("fp8", 1409, 128)])
def test_getrf_lookahead_matches_sequential(ctx, fmt, n, nb, monkeypatch):
    """The look-ahead schedule (next panel's columns updated first, panel factorized on a second stream while
    the remaining columns are updated) runs the same kernels on the same column tiles: factors, ipiv and perm
    must be bit-identical to the one-stream schedule."""
    print(f"[getrf lookahead] fmt={fmt} n={n} nb={nb}", flush=True)
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n)).astype(np.float32))
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_OVERLAP", mode)
        dLU = _dev(ctx, A); dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
        info = ctx.getrf(n, dLU, n, dip, dpm, nb=nb, fmt=fmt)
        assert info.info == 0 and info.panels == -(-n // nb)
        out.append((dLU.to_array((n, n), np.float32), dip.to_array((n,), np.int64), dpm.to_array((n,), np.int32)))
        for b in (dLU, dip, dpm):
            b.free()
    for x, y in zip(out[0], out[1]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("method,fmt", [("mx16x3_gmresir_w64_rdd", "fp16x3"), ("mx16_gmresir_w64_rdd", "fp16")])
def test_solve_growth_and_split_copy(ctx, method, fmt, monkeypatch):
    """A solve's factor copy is converted in two parts (the first panel's columns, then the rest on the second
    stream beside that panel), and its growth factor max|LU| / max|copy| is measured on the second stream beside
    the refinement.  The growth must be the same bits as in the one-stream schedule and as max|LU| / max|g fl32(A)|
    formed on the host from getrf of the same scaled FP32 copy (getrf is deterministic, the power-of-two scale g
    and both maxima are exact)."""
    n, nb = 1000, 256
    print(f"[solve growth] method={method} n={n}", flush=True)
    A, b, _, _ = randsvd(n, 1e4, 11, gpu=False)
    P = ctx.problem(A, b)
    g = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_OVERLAP", mode)
        X, rec = P.solve(method, nb=nb)
        assert rec["ok"] and rec["terminal_reason"] == "converged", rec["terminal_reason"]
        g[mode] = (rec["growth_proxy"], rec["scale_mode"])
    P.close()
    assert g["0"] == g["1"], g
    growth, scale_mode = g["1"]
    assert scale_mode in (0, 1), scale_mode
    A32 = np.asarray(A, dtype=np.float64).astype(np.float32)
    amax = float(np.abs(A32).max())
    gs = 2.0 ** -np.frexp(amax)[1] if scale_mode == 1 else 1.0
    S = (A32 * np.float32(gs)).astype(np.float32)
    dLU = _dev(ctx, S); dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
    info = ctx.getrf(n, dLU, n, dip, dpm, nb=nb, fmt=fmt, tile_scale=True)
    assert info.info == 0
    LU = dLU.to_array((n, n), np.float32)
    for x in (dLU, dip, dpm):
        x.free()
    expect = float(np.abs(LU).max()) / (amax * gs)
    assert growth == expect, (growth, expect)


@pytest.mark.parametrize("fmt,n,nb,overlap", [("fp16x3", 4160, 256, "1"), ("fp16", 3000, 256, "0"), ("fp32", 1001, 40, "1"),
                                              ("fp16", 150, 7, "0")])
def test_getrf_deferred_left_interchanges(ctx, fmt, n, nb, overlap, monkeypatch):
    """The interchanges of the factored L columns applied all at once after the last panel (k_lperm, from the
    recorded inverse permutation of every step) move the same data as applying each step's interchanges as it
    comes: factors, ipiv and perm bit-identical, with and without look-ahead, for panel widths that are not a
    multiple of the kernel's 16 columns per workgroup and a narrower last block."""
    print(f"[getrf deferred L] fmt={fmt} n={n} nb={nb} overlap={overlap}", flush=True)
    rng = np.random.default_rng(n + nb)
    A = np.asfortranarray(rng.standard_normal((n, n)).astype(np.float32))
    monkeypatch.setenv("MPSOLVE_OVERLAP", overlap)
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("MPSOLVE_LPERM", mode)
        dLU = _dev(ctx, A); dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
        info = ctx.getrf(n, dLU, n, dip, dpm, nb=nb, fmt=fmt)
        assert info.info == 0 and info.panels == -(-n // nb)
        out.append((dLU.to_array((n, n), np.float32), dip.to_array((n,), np.int64), dpm.to_array((n,), np.int32)))
        for b in (dLU, dip, dpm):
            b.free()
    for x, y in zip(out[0], out[1]):
        np.testing.assert_array_equal(x, y)
    # and they factor the row-permuted matrix (a misplaced L row leaves an error of the order of |A| itself)
    LU, pm = out[1][0].astype(np.float64), out[1][2]
    L = np.tril(LU, -1) + np.eye(n); U = np.triu(LU)
    assert np.abs(A[pm].astype(np.float64) - L @ U).max() <= 0.25 * np.abs(A).max()


def test_getrf_singular_info(ctx, oracle):
    n = 96
    A = np.ones((n, n), dtype=np.float32, order="F")
    dLU = _dev(ctx, A); dip, dpm = ctx.empty(8 * n), ctx.empty(4 * n)
    info = ctx.getrf(n, dLU, n, dip, dpm, nb=32, fmt="fp16")
    _, _, _, info_o, _ = oracle.getrf(A, 32, "fp16", True)
    assert info.info == info_o == 2


@pytest.mark.parametrize("n,nrhs", [(64, 1), (1000, 1), (1500, 3), (4097, 1)])
def test_getrs_matches_oracle(ctx, oracle, n, nrhs):
    print(f"[getrs] n={n} nrhs={nrhs}", flush=True)
    A, b, xt, _ = randsvd(n, 1e3, 3 + n, nrhs=nrhs, gpu=False)
    F = oracle.factor(A, "fp32", 128)
    dLU = _dev(ctx, F["LU"]); dpm = _dev(ctx, F["perm"])
    dR = _dev(ctx, b); dD = ctx.empty(8 * n * nrhs)
    ctx.getrs(n, nrhs, dLU, n, dpm, dR, n, dD, n, scale=F["g"])
    ctx.sync()
    D = dD.to_array((n, nrhs), np.float64)
    Do = oracle.getrs(F["LU"], F["perm"], b, F["g"])
    # both are FP32 substitutions with the same factors and differ only in summation order: check each is
    # backward stable for (L U) x = P b (normwise, FP32 level)
    L, U = oracle.unpack_lu(F["LU"])
    Pb = b[F["perm"]]
    for sol in (D, Do):
        xs = sol / F["g"]
        res = np.abs(L @ (U @ xs) - Pb).max(axis=0)
        scale = (np.abs(L) @ (np.abs(U) @ np.abs(xs))).max(axis=0) + np.abs(Pb).max(axis=0)
        assert np.all(res / scale <= 4 * n * 2.0 ** -24), res / scale
    # forward errors against the FP64 solution x* of the factored system: the GPU's is within 8x of the
    # oracle's sequential substitution, or of kappa_S * u32 when that is larger (kappa_S = Skeel's condition
    # number || |(LU)^-1| |L| |U| |x*| || / ||x*||, the first-order growth of FP32 rounding in both)
    from scipy.linalg import solve_triangular
    xst = solve_triangular(U, solve_triangular(L, Pb, lower=True, unit_diagonal=True), lower=False)
    Linv = solve_triangular(L, np.eye(n), lower=True, unit_diagonal=True)
    Uinv = solve_triangular(U, np.eye(n), lower=False)
    kap = (np.abs(Uinv) @ (np.abs(Linv) @ (np.abs(L) @ (np.abs(U) @ np.abs(xst))))).max(axis=0) / np.abs(xst).max(axis=0)
    u32 = 2.0 ** -24
    err = lambda sol: (np.abs(sol / F["g"] - xst).max(axis=0) / np.abs(xst).max(axis=0))
    eg, eo = err(D), err(Do)
    print(f"[getrs] forward error gpu {eg.max():.2e} oracle {eo.max():.2e} kappa_S*u32 {(kap * u32).max():.2e}", flush=True)
    assert np.all(eg <= 8 * np.maximum(eo, kap * u32)), (eg, eo, kap * u32)
    assert np.all(eo <= n * kap * u32)


def test_getrs_exchange_timeout_is_an_error(ctx, oracle, monkeypatch):
    """A wait of the substitutions' block exchange that reaches its bound is reported (MP_ERR_INTERNAL), not
    returned as a result.  The host forces it (MPSOLVE_GS_FAIL: the waits look for a tag that is never
    published, with a one-poll bound), so every block with a predecessor gives up; every workgroup still exits
    normally.  Without it the same call succeeds again (the flag is cleared per call)."""
    from cpsc402_midterm_project_amd import MPSolveError
    n = 8192
    A, b, _, _ = randsvd(1024, 1e2, 3, gpu=False)
    F = oracle.factor(A, "fp32", 128)
    LU = np.zeros((n, n), np.float32, order="F"); LU[:, :] = np.eye(n, dtype=np.float32)
    LU[:1024, :1024] = F["LU"]
    LU[1024:, :1024] = np.float32(1e-3)                      # a dense lower part: every block row depends on the last
    pm = np.arange(n, dtype=np.int32); pm[:1024] = F["perm"]
    bb = np.ones(n)
    dLU = _dev(ctx, LU); dpm = _dev(ctx, pm); dR = _dev(ctx, bb); dD = ctx.empty(8 * n)
    monkeypatch.setenv("MPSOLVE_GS_FAIL", "all")
    with pytest.raises(MPSolveError) as ei:
        ctx.getrs(n, 1, dLU, n, dpm, dR, n, dD, n)
    assert ei.value.status == -5 and "timed out" in str(ei.value)
    monkeypatch.delenv("MPSOLVE_GS_FAIL")
    ctx.getrs(n, 1, dLU, n, dpm, dR, n, dD, n)


def test_substitution_timeout_survives_refactorization(ctx, monkeypatch):
    """A ladder refactorizes between substitutions.  Only the first substitution of the trial (the initial
    solve of the FP8 stage) is forced to time out (MPSOLVE_GS_FAIL=1); the ladder then escalates through own
    factorizations (FP8 fails at kappa=1e8), and the trial must still end as an error: a factorization clears
    only its own info words.  The next solve, with the normal bound, starts clean and converges."""
    n = 2048
    print(f"[timeout sticky] n={n}", flush=True)
    A, b, xt, _ = randsvd(n, 1e8, 17, gpu=False)
    P = ctx.problem(A, b)
    monkeypatch.setenv("MPSOLVE_GS_FAIL", "1")
    X, rec = P.solve("policy_mx8_ladder", nb=256, tau=1e-15, stop_on="eta")
    print(f"[timeout sticky] ok={rec['ok']} stages={rec['stages']} error={rec['error']!r}", flush=True)
    assert len(rec["stages"]) >= 2, rec["stages"]                # a refactorization happened after the timeout
    assert not rec["ok"] and "timed out" in rec["error"], (rec["ok"], rec["error"], rec["stages"])
    monkeypatch.delenv("MPSOLVE_GS_FAIL")
    X, rec = P.solve("policy_mx8_ladder", nb=256, tau=1e-15, stop_on="eta")
    assert rec["ok"] and rec["terminal_reason"] == "converged", (rec["error"], rec["terminal_reason"])
    P.close()


def _e4m3_table():
    v = np.zeros(256, dtype=np.float32)
    for b8 in range(256):
        s, e, m = b8 >> 7, (b8 >> 3) & 15, b8 & 7
        x = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        if e == 15 and m == 7:
            x = np.nan
        v[b8] = -x if s else x
    return v


def _unpack(raw, nidx, K, fmt):
    """The fragment layout of the pack (one 8-element fragment per lane and 16-index k-block) back to nidx x K."""
    nkb = (K + 15) // 16
    nT = (nidx + 31) // 32
    t, b, l, e = np.meshgrid(np.arange(nT), np.arange(nkb), np.arange(64), np.arange(8), indexing="ij")
    idx = t * 32 + l % 32
    kk = b * 16 + (2 * e + l // 32 if fmt == "fp32" else (l // 32) * 8 + e)
    if fmt in ("fp16", "fp16x3"):
        vals = raw.view(np.float16).astype(np.float32)
    elif fmt == "bf16":
        vals = (raw.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    elif fmt == "fp8":
        vals = _e4m3_table()[raw.view(np.uint8)]
    else:
        vals = raw.view(np.float32)
    out = np.zeros((nT * 32, nkb * 16), dtype=np.float32)
    out[idx.ravel(), kk.ravel()] = vals.reshape(-1)
    assert np.all(out[nidx:] == 0) and np.all(out[:, K:] == 0)   # padding positions hold zeros
    return out[:nidx, :K]


def _pack_input(fmt, tile_scale, nidx, K, seed):
    """Random values over a wide range, plus in every tile: a fixed maximum (sets the tile scale), exact
    rounding ties of the format after that scale, values in its subnormal range, FP32 subnormals, signed
    zeros and, without tile scaling, values above the format's maximum."""
    rng = np.random.default_rng(seed)
    S = (rng.standard_normal((nidx, K)) * 2.0 ** rng.integers(-12, 6, size=(nidx, K))).astype(np.float32)
    top = {"fp16": 14, "fp16x3": 14, "bf16": 14, "fp8": 7, "fp32": 0}[fmt]
    for t in range(0, nidx, 32):
        E = int(rng.integers(-8, 8))
        S[t, 0] = np.float32(1.5 * 2.0 ** E)                  # tile maximum: scale 2^(top - E)
        s = (top - E) if (tile_scale and fmt != "fp32") else 0
        rows = slice(t, min(t + 32, nidx))
        blk = S[rows].copy()
        nr = blk.shape[0]
        k = rng.integers(1, K, size=(nr, K // 4))
        ties = np.zeros((nr, K // 4), dtype=np.float32)
        if fmt in ("fp16", "fp16x3"):
            h = (rng.uniform(2.0 ** -14, 2.0 ** 13, size=ties.shape) * rng.choice([-1, 1], size=ties.shape)).astype(np.float16)
            h2 = np.nextafter(h, np.float16(np.inf) * np.sign(h).astype(np.float16), dtype=np.float16)
            ties = (h.astype(np.float32) + h2.astype(np.float32)) / np.float32(2)
        elif fmt == "bf16":
            hb = (rng.uniform(2.0 ** -20, 2.0 ** 13, size=ties.shape) * rng.choice([-1, 1], size=ties.shape)).astype(np.float32)
            ties = ((hb.view(np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x8000)).view(np.float32)
        elif fmt == "fp8":
            tab = _e4m3_table()[:0x7F]                        # positive finite grid, increasing
            i = rng.integers(0, len(tab) - 1, size=ties.shape)
            ties = ((tab[i] + tab[i + 1]) / np.float32(2)) * rng.choice([-1, 1], size=ties.shape).astype(np.float32)
        if fmt != "fp32":
            blk[np.arange(nr)[:, None], k] = ties * np.float32(2.0 ** -s)
        sub_rel = {"fp16": -20, "fp16x3": -20, "bf16": -10, "fp8": -8, "fp32": -10}[fmt]
        blk[:, K - 1] = np.float32(2.0 ** (sub_rel - s)) * np.float32(1.25)   # subnormal range of the format
        blk[0, K - 2] = np.float32(2.0 ** -140)                  # an FP32 subnormal input
        blk[min(1, nr - 1), K - 3] = np.float32(-0.0)
        if not tile_scale and fmt in ("fp16", "fp16x3", "fp8"):
            blk[min(2, nr - 1), 1] = np.float32(70000.0 if fmt != "fp8" else 500.0)
        S[rows] = blk
    return np.ascontiguousarray(S)


@pytest.mark.parametrize("fmt,tile_scale", [("fp16", True), ("fp16", False), ("bf16", True), ("fp8", True), ("fp8", False),
                                            ("fp32", True), ("fp16x3", True)])
@pytest.mark.parametrize("nidx,K,layout", [(100, 40, "idx_major"), (64, 256, "k_major"), (33, 17, "k_major")])
def test_pack_bitwise_vs_oracle(ctx, oracle, fmt, tile_scale, nidx, K, layout):
    """The trailing update's operand conversion (k_pack) against the oracle's restatement: every packed value
    bit for bit (round to nearest even at exact ties, subnormal grids, saturation of e4m3, FP16 overflow to
    inf), the tile exponents and both counters."""
    print(f"[pack] fmt={fmt} tile_scale={tile_scale} nidx={nidx} K={K} {layout}", flush=True)
    S = _pack_input(fmt, tile_scale, nidx, K, seed=nidx * K)
    if layout == "idx_major":          # element (i, q) at i * K + q: s_idx = K, s_k = 1
        src, s_idx, s_k = S, K, 1
    else:                              # column-major L21 style: element (i, q) at i + q * nidx
        src, s_idx, s_k = np.asfortranarray(S), 1, nidx
    es = {"fp16": 2, "fp16x3": 2, "bf16": 2, "fp8": 1, "fp32": 4}[fmt]
    nkb = (K + 15) // 16; nT = (nidx + 31) // 32
    nbytes = nT * nkb * 64 * 8 * es
    dS = ctx.buffer(np.ascontiguousarray(src.ravel(order="K")))
    dO, dOl = ctx.empty(nbytes), ctx.empty(nbytes)
    dE = ctx.empty(4 * nT); dC = ctx.buffer(np.zeros(2, dtype=np.uint64))
    ctx.pack_operand(dS, s_idx, s_k, nidx, K, dO, dE, dC, fmt=fmt, tile_scale=tile_scale,
                     out_lo=dOl if fmt == "fp16x3" else None)
    ctx.sync()
    raw = dO.to_array((nbytes,), np.uint8)
    q = _unpack(raw, nidx, K, fmt)
    qo, qlo, exo, cnto = oracle.pack(S, fmt, tile_scale)
    np.testing.assert_array_equal(dE.to_array((nT,), np.int32), exo)
    np.testing.assert_array_equal(q.view(np.uint32), qo.view(np.uint32))
    if fmt == "fp16x3":
        ql = _unpack(dOl.to_array((nbytes,), np.uint8), nidx, K, fmt)
        np.testing.assert_array_equal(ql.view(np.uint32), qlo.view(np.uint32))
    np.testing.assert_array_equal(dC.to_array((2,), np.uint64), cnto)
    assert cnto[1] > 0                                           # the subnormal range was exercised
