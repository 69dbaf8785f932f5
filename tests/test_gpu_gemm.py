"""GPU parity of the panel packs (FP16, scaled FP8) and of the MFMA Schur GEMM
against the oracle.  FP16/FP8 products are exact in FP32, so the GEMM is
checked against an FP64 product of the same rounded operands within the FP32
accumulation bound k * 2^-24 * |X||Y|^T."""
import numpy as np
import pytest

from tests.gpu_util import dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def pk16_index(R, K):
    Rp = -(-R // 256) * 256
    r = np.arange(Rp)[:, None]
    k = np.arange(K)[None, :]
    return (((r >> 8) * (K >> 6) + (k >> 6)) * 16384 + ((((r & 255) >> 5) << 2) + ((k & 63) >> 4)) * 512
            + (((r & 31) + 32 * ((k & 15) >> 3)) << 3) + (k & 7))


def pk8_index(R, K):
    Rp = -(-R // 256) * 256
    r = np.arange(Rp)[:, None]
    k = np.arange(K)[None, :]
    return (((r >> 8) * (K >> 7) + (k >> 7)) * 32768 + ((((r & 255) >> 5) << 1) + ((k & 127) >> 6)) * 2048
            + ((k & 31) >> 4) * 1024 + (((r & 31) + 32 * ((k & 63) >> 5)) << 4) + (k & 15))


def pack16(Xh):
    """numpy FP16 (R x K) -> packed uint16 array (reference packing)."""
    R, K = Xh.shape
    idx = pk16_index(R, K)
    out = np.zeros(idx.size, dtype=np.uint16)
    full = np.zeros((idx.shape[0], K), dtype=np.uint16)
    full[:R] = Xh.view(np.uint16)
    out[idx.ravel()] = full.ravel()
    return out


def test_pack_index_is_a_permutation():
    for R, K in [(256, 64), (300, 128), (512, 2560)]:
        idx = pk16_index(R, K).ravel()
        assert np.array_equal(np.sort(idx), np.arange(idx.size))
    for R, K in [(256, 128), (300, 512)]:
        idx = pk8_index(R, K).ravel()
        assert np.array_equal(np.sort(idx), np.arange(idx.size))


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("R,K", [(300, 128), (512, 256)])
def test_pack_f16_bitexact(R, K, side):
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(R + K)
    S = (rng.standard_normal((R, K)) * 10.0 ** rng.uniform(-6, 2, (R, K))).astype(np.float32)
    if side == "left":   # column-major R x K block with ld = R + 7
        ld = R + 7
        src = np.zeros((K, ld), dtype=np.float32)
        src[:, :R] = S.T
        rs, ls = 1, ld
    else:                # S(r, k) = M(k, r) of a column-major K x R block, ld = K + 4
        ld = K + 4
        src = np.zeros((R, ld), dtype=np.float32)
        src[:, :K] = S
        rs, ls = ld, 1
    sd = dev(src.ravel())
    nbytes = L.omx_panel_bytes(R, K, 0)
    P = empty(nbytes // 2, torch.int16)
    cm = empty((-(-R // 256)) * (K // 64), torch.float32)
    assert L.omx_pack_f16(R, K, ptr(sd), rs, ls, 0, ptr(P), ptr(cm), None) == 0
    got = host(P).view(np.uint16)
    want = pack16(S.astype(np.float16))
    assert np.array_equal(got, want)
    # chunk max-abs statistics (of the FP32 source values)
    cmx = host(cm).reshape(-(-R // 256), K // 64)
    Sp = np.zeros((cmx.shape[0] * 256, K), np.float32)
    Sp[:R] = np.abs(S)
    assert np.array_equal(cmx, Sp.reshape(cmx.shape[0], 256, K // 64, 64).max(axis=(1, 3)))


@pytest.mark.parametrize("mask", [1, 2, 3])
def test_pack_f16_masks(mask):
    torch = torch_cuda()
    L = _L()
    K = 256
    rng = np.random.default_rng(mask)
    W = rng.uniform(-1, 1, (K, K)).astype(np.float32)
    sd = dev(np.asfortranarray(W).ravel(order="F"))
    P = empty(L.omx_panel_bytes(K, K, 0) // 2, torch.int16)
    assert L.omx_pack_f16(K, K, ptr(sd), 1, K, mask, ptr(P), None, None) == 0
    r = np.arange(K)[:, None]
    k = np.arange(K)[None, :]
    if mask == 1:
        M = np.where(k <= r, W, 0)
    elif mask == 2:
        M = np.where(k > r, W, np.where(k == r, 1.0, 0.0))
    else:
        M = np.where(r > k, W, np.where(r == k, 1.0, 0.0))
    assert np.array_equal(host(P).view(np.uint16), pack16(M.astype(np.float16)))


@pytest.mark.parametrize("side", ["left", "right"])
def test_pack_f8_bitexact(orc, side):
    torch = torch_cuda()
    L = _L()
    R, K = 300, 512
    rng = np.random.default_rng(11)
    S = (rng.standard_normal((R, K)) * 1e-5).astype(np.float32)
    S[:, 256:] *= 3e4  # a second scale tile with a different exponent
    if side == "left":
        src, rs, ls = np.asfortranarray(S).ravel(order="F"), 1, R
    else:
        src, rs, ls = np.ascontiguousarray(S).ravel(), K, 1
    sd = dev(src)
    P = empty(L.omx_panel_bytes(R, K, 1), torch.uint8)
    ne = L.omx_panel_exps_count(R, K)
    ex = empty(ne, torch.int32)
    tmx = empty(ne, torch.float32)
    assert L.omx_pack_f8(R, K, ptr(sd), rs, ls, ptr(P), ptr(ex), ptr(tmx), None) == 0
    e = host(ex).reshape(2, 2)
    for i in range(2):
        for j in range(2):
            blk = S[i * 256:(i + 1) * 256, j * 256:(j + 1) * 256]
            assert e[i, j] == orc.pow2_scale_exp(float(np.abs(blk).max()))
    scaled = np.zeros((512, K), np.float32)
    for i in range(2):
        for j in range(2):
            blk = S[i * 256:(i + 1) * 256, j * 256:(j + 1) * 256]
            scaled[i * 256:i * 256 + blk.shape[0], j * 256:(j + 1) * 256] = np.ldexp(blk, e[i, j])
    q = orc.f32_to_e4m3_bits(scaled)
    idx = pk8_index(R, K)
    want = np.zeros(idx.size, np.uint8)
    want[idx.ravel()] = q.ravel()
    assert np.array_equal(host(P), want)


def _gemm_case(m, n, k, alpha, beta, ldc_pad=3, x_row0=0, y_row0=0, seed=0):
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(seed)
    Xh = rng.uniform(-1, 1, (x_row0 + m, k)).astype(np.float16)
    Yh = rng.uniform(-1, 1, (y_row0 + n, k)).astype(np.float16)
    ldc = m + ldc_pad
    C0 = rng.uniform(-1, 1, (n, ldc)).astype(np.float32)  # column-major m x n, ld = ldc
    Xd, Yd, Cd = dev(pack16(Xh).view(np.int16)), dev(pack16(Yh).view(np.int16)), dev(C0.ravel())
    rc = L.omx_gemm_f16_packed(m, n, k, alpha, ptr(Xd), x_row0, ptr(Yd), y_row0, beta, ptr(Cd), ldc, None)
    assert rc == 0
    got = host(Cd).reshape(n, ldc)
    X64 = Xh[x_row0:].astype(np.float64)
    Y64 = Yh[y_row0:].astype(np.float64)
    Cm = C0[:, :m].T.astype(np.float64)
    exact = (beta * Cm if beta != 0 else 0) + alpha * (X64 @ Y64.T)
    bound = (k + 2) * 2.0**-24 * np.abs(alpha) * (np.abs(X64) @ np.abs(Y64).T) + 2.0**-23 * np.abs(exact)
    assert np.all(np.abs(got[:, :m].T - exact) <= bound)
    assert np.array_equal(got[:, m:], C0[:, m:])  # padding rows of C untouched


@pytest.mark.parametrize("m,n,k", [(256, 256, 64), (512, 768, 256), (300, 200, 128), (2560, 2560, 2560)])
def test_gemm_f16_schur(m, n, k):
    print(f"gemm_f16 {m}x{n}x{k}", flush=True)
    _gemm_case(m, n, k, -1.0, 1.0, seed=m + n)


def test_gemm_f16_beta_zero_and_offsets():
    _gemm_case(512, 256, 256, 1.0, 0.0, x_row0=256, y_row0=512, seed=5)
    _gemm_case(1024, 1280, 512, -1.0, 1.0, x_row0=768, y_row0=256, ldc_pad=100, seed=6)


@pytest.mark.parametrize("tri,m,n,k,beta", [(1, 1000, 768, 768, 0.0), (1, 2560, 2560, 2560, 0.0),
                                             (1, 700, 1100, 512, 1.0), (2, 768, 1000, 768, 0.0),
                                             (2, 2560, 2816, 2560, 0.0), (2, 1100, 700, 512, 1.0)])
def test_gemm_f16_triangular_factor_equals_full_product(tri, m, n, k, beta):
    """omx_gemm_f16_packed_tri against omx_gemm_f16_packed on operands that
    hold the triangular factor's zeros (tri 1: Y(j, l) = 0 for l > j, the
    column panel times U^-1; tri 2: X(i, l) = 0 for l > i, L^-1 times the row
    panel): every product a tile skips is zero, so C is the same bit for
    bit, edge tiles and tiles past the factor's order included."""
    print(f"gemm_f16_tri tri={tri} {m}x{n}x{k} beta={beta}", flush=True)
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(100 * tri + m)
    Xh = rng.uniform(-1, 1, (m, k)).astype(np.float16)
    Yh = rng.uniform(-1, 1, (n, k)).astype(np.float16)
    if tri == 1:
        Yh[np.arange(n)[:, None] < np.arange(k)[None, :]] = 0
    else:
        Xh[np.arange(m)[:, None] < np.arange(k)[None, :]] = 0
    ldc = m + 5
    C0 = rng.uniform(-1, 1, (n, ldc)).astype(np.float32)
    Xd, Yd = dev(pack16(Xh).view(np.int16)), dev(pack16(Yh).view(np.int16))
    C_full, C_tri = dev(C0.ravel()), dev(C0.ravel())
    assert L.omx_gemm_f16_packed(m, n, k, 1.0, ptr(Xd), 0, ptr(Yd), 0, beta, ptr(C_full), ldc, None) == 0
    assert L.omx_gemm_f16_packed_tri(m, n, k, 1.0, ptr(Xd), ptr(Yd), tri, beta, ptr(C_tri), ldc, None, None) == 0
    full, got = host(C_full), host(C_tri)
    assert np.array_equal(got, full)  # (an exact zero may differ in sign)
    # and the product is the triangular one
    X64, Y64 = Xh.astype(np.float64), Yh.astype(np.float64)
    want = (beta * C0[:, :m].T.astype(np.float64) if beta else 0) + X64 @ Y64.T
    bound = (k + 2) * 2.0**-24 * (np.abs(X64) @ np.abs(Y64).T) + 2.0**-23 * np.abs(want)
    assert np.all(np.abs(got.reshape(n, ldc)[:, :m].T - want) <= bound)
    assert L.omx_gemm_f16_packed_tri(m, n, k, 1.0, ptr(Xd), ptr(Yd), 3, beta, ptr(C_tri), ldc, None, None) == -1


@pytest.mark.parametrize("tri,m,n", [(1, 1000, 768), (1, 2560, 2560), (2, 768, 1000), (2, 2560, 2816), (0, 700, 512)])
def test_gemm_f16_tri_tile_maxima_equal_the_panel_pass(tri, m, n):
    """The tile maxima the beta = 0 product leaves (tile_max) against
    omx_panel_tile_maxabs over the C it wrote, bit for bit: C's rows are the
    panel's rows for tri 0 and 1, its columns for tri 2 (the row panel);
    edge tiles take only the entries inside C."""
    print(f"gemm_f16_tri tile maxima tri={tri} {m}x{n}", flush=True)
    torch = torch_cuda()
    L = _L()
    k = 512
    rng = np.random.default_rng(7 * m + n + tri)
    Xh = (rng.standard_normal((m, k)) * rng.uniform(0.1, 4.0, (m, 1))).astype(np.float16)
    Yh = rng.uniform(-1, 1, (n, k)).astype(np.float16)
    if tri == 1:
        Yh[np.arange(n)[:, None] < np.arange(k)[None, :]] = 0
    elif tri == 2:
        Xh[np.arange(m)[:, None] < np.arange(k)[None, :]] = 0
    ldc = m + 3
    Xd, Yd = dev(pack16(Xh).view(np.int16)), dev(pack16(Yh).view(np.int16))
    Cd = dev(np.full((n, ldc), np.nan, np.float32).ravel())
    rows, kk, rs, ls = (m, n, 1, ldc) if tri != 2 else (n, m, ldc, 1)
    cnt = L.omx_panel_exps_count(rows, kk)
    got = empty(cnt, torch.float32).fill_(float("nan"))
    want = empty(cnt, torch.float32)
    assert L.omx_gemm_f16_packed_tri(m, n, k, -0.5, ptr(Xd), ptr(Yd), tri, 0.0, ptr(Cd), ldc, ptr(got), None) == 0
    assert L.omx_panel_tile_maxabs(rows, kk, ptr(Cd), rs, ls, ptr(want), None) == 0
    assert np.array_equal(host(got), host(want))
    assert np.all(host(got) > 0)
    # tile maxima need beta = 0
    assert L.omx_gemm_f16_packed_tri(m, n, k, 1.0, ptr(Xd), ptr(Yd), tri, 1.0, ptr(Cd), ldc, ptr(got), None) == -1


def test_gemm_f8_schur(orc):
    torch = torch_cuda()
    L = _L()
    m, n, k = 512, 768, 512
    rng = np.random.default_rng(3)
    Xs = (rng.standard_normal((m, k)) * 1e-5).astype(np.float32)
    Ys = rng.uniform(-0.5, 0.5, (n, k)).astype(np.float32)
    packs = []
    for S in (Xs, Ys):
        R = S.shape[0]
        sd = dev(np.asfortranarray(S).ravel(order="F"))
        P = empty(L.omx_panel_bytes(R, k, 1), torch.uint8)
        ne = L.omx_panel_exps_count(R, k)
        ex = empty(ne, torch.int32)
        tmx = empty(ne, torch.float32)
        assert L.omx_pack_f8(R, k, ptr(sd), 1, R, ptr(P), ptr(ex), ptr(tmx), None) == 0
        packs.append((P, ex, sd))
    C0 = rng.uniform(-1, 1, (n, m)).astype(np.float32)
    Cd = dev(C0.ravel())
    (XP, xe, _), (YP, ye, _) = packs
    assert L.omx_gemm_f8_packed(m, n, k, -1.0, ptr(XP), ptr(xe), 0, ptr(YP), ptr(ye), 0, 1.0, ptr(Cd), m, None) == 0
    got = host(Cd).reshape(n, m).T.astype(np.float64)
    # decode the packed operands through the oracle and form the exact product
    dec = []
    for (P, ex, _), S in zip(packs, (Xs, Ys)):
        R = S.shape[0]
        q = host(P)[pk8_index(R, k)][:R]
        e = host(ex).reshape(-(-R // 256), k // 256)
        scale = np.ldexp(1.0, -np.repeat(np.repeat(e, 256, 0), 256, 1)[:R])
        dec.append(orc.e4m3_bits_to_f32(q).astype(np.float64) * scale)
    exact = C0.T.astype(np.float64) - dec[0] @ dec[1].T
    bound = (k + 2) * 2.0**-24 * (np.abs(dec[0]) @ np.abs(dec[1]).T) + 2.0**-23 * np.abs(exact)
    assert np.all(np.abs(got - exact) <= bound)
    # and the FP8 result approximates the FP32 product to e4m3 accuracy
    full = C0.T.astype(np.float64) - Xs.astype(np.float64) @ Ys.astype(np.float64).T
    assert np.max(np.abs(got - full)) < 0.1 * np.max(np.abs(Xs.astype(np.float64) @ Ys.T.astype(np.float64)))


def _pack16_scaled_dev(S):
    """Column panel S (R x K, column-major on the device) packed with
    per-strip scales; returns (packed, exponents) device tensors."""
    torch = torch_cuda()
    L = _L()
    R, K = S.shape
    sd = dev(np.asfortranarray(S).ravel(order="F"))
    P = empty(L.omx_panel_bytes(R, K, 0) // 2, torch.int16)
    se = empty(-(-R // 256), torch.int32)
    tm = empty(L.omx_panel_exps_count(R, K), torch.float32)
    assert L.omx_pack_f16_scaled(R, K, ptr(sd), 1, R, 0, ptr(P), ptr(se), ptr(tm), None) == 0
    return P, se


def test_pack_f16_scaled_bitexact():
    """Strip exponent = largest e with max|strip| 2^e < 2^15; values
    fp16(S 2^e) in the packed layout."""
    rng = np.random.default_rng(12)
    R, K = 700, 512
    S = (rng.standard_normal((R, K)) * 10.0 ** rng.uniform(-9, 7, (R, 1))).astype(np.float32)
    P, se = _pack16_scaled_dev(S)
    e = host(se)
    scaled = np.zeros_like(S)
    for s in range(e.size):
        blk = S[s * 256:(s + 1) * 256]
        m = float(np.abs(blk).max())
        assert m * 2.0 ** e[s] < 2.0**15 <= m * 2.0 ** (e[s] + 1)
        scaled[s * 256:(s + 1) * 256] = np.ldexp(blk, e[s])
    assert np.array_equal(host(P).view(np.uint16), pack16(scaled.astype(np.float16)))


def test_gemm_f16_scaled_equals_unscaled_in_range():
    """Power-of-two scaling is exact: for panels whose plain FP16 cast
    neither overflows nor goes subnormal, the scaled GEMM is bit-identical
    to the plain one."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(13)
    m, n, k = 768, 512, 512
    X = (rng.uniform(-1, 1, (m, k)) * 4.0 ** rng.integers(-3, 4, (m, 1))).astype(np.float16).astype(np.float32)
    Y = rng.uniform(-1, 1, (n, k)).astype(np.float16).astype(np.float32)
    XP, xe = _pack16_scaled_dev(X)
    YP, ye = _pack16_scaled_dev(Y)
    Xd = dev(pack16(X.astype(np.float16)).view(np.int16))
    Yd = dev(pack16(Y.astype(np.float16)).view(np.int16))
    C0 = rng.uniform(-1, 1, (n, m)).astype(np.float32)
    C1, C2 = dev(C0.ravel()), dev(C0.ravel())
    assert L.omx_gemm_f16_packed(m, n, k, -1.0, ptr(Xd), 0, ptr(Yd), 0, 1.0, ptr(C1), m, None) == 0
    assert L.omx_gemm_f16_packed_scaled(m, n, k, -1.0, ptr(XP), ptr(xe), 0, ptr(YP), ptr(ye), 0, 1.0, ptr(C2), m,
                                        None) == 0
    assert np.array_equal(host(C1), host(C2))


def test_gemm_f16_scaled_out_of_fp16_range():
    """Panels of magnitude 1e7 and 1e-9: the plain FP16 cast overflows /
    flushes, the scaled path stays within the FP32 accumulation bound of the
    exact product of its (scaled-rounded) operands."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(14)
    m, n, k = 512, 256, 256
    X = (rng.uniform(-1, 1, (m, k)) * 1e7).astype(np.float32)
    Y = (rng.uniform(-1, 1, (n, k)) * 1e-9).astype(np.float32)
    XP, xe = _pack16_scaled_dev(X)
    YP, ye = _pack16_scaled_dev(Y)
    Cd = dev(np.zeros(n * m, np.float32))
    assert L.omx_gemm_f16_packed_scaled(m, n, k, 1.0, ptr(XP), ptr(xe), 0, ptr(YP), ptr(ye), 0, 0.0, ptr(Cd), m,
                                        None) == 0
    got = host(Cd).reshape(n, m).T.astype(np.float64)
    ex, ey = host(xe), host(ye)
    Xr = np.concatenate([np.ldexp(np.ldexp(X[s * 256:(s + 1) * 256], ex[s]).astype(np.float16).astype(np.float64),
                                  -ex[s]) for s in range(ex.size)])
    Yr = np.concatenate([np.ldexp(np.ldexp(Y[s * 256:(s + 1) * 256], ey[s]).astype(np.float16).astype(np.float64),
                                  -ey[s]) for s in range(ey.size)])
    exact = Xr @ Yr.T
    bound = (k + 2) * 2.0**-24 * (np.abs(Xr) @ np.abs(Yr).T)
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - exact) <= bound)
    assert np.max(np.abs(exact - X.astype(np.float64) @ Y.astype(np.float64).T)) <= 2.0**-10 * np.max(np.abs(exact))


@pytest.mark.parametrize("fp8", [False, True])
def test_window_pair_product_equals_two_products(orc, fp8):
    """Two k-column panels packed side by side in a 2k-column buffer (the
    second starting at packed row `shift`, as the LU packs a pair): the
    product over both windows equals C - X1 Y1^T - X2 Y2^T within the FP32
    accumulation bound, and the product over the second window alone is
    bit-identical to the same panel packed on its own."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(21 + fp8)
    k, shift = 512, 256
    m, n = 768, 512  # target rows/cols counted from packed row `shift`
    R = shift + m
    X1 = rng.uniform(-1, 1, (R, k)).astype(np.float32)
    Y1 = rng.uniform(-1, 1, (shift + n, k)).astype(np.float32)
    X2 = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    Y2 = rng.uniform(-1, 1, (n, k)).astype(np.float32)
    ldk = 2 * k
    fp = 1 if fp8 else 0
    PX = empty(L.omx_panel_bytes(R, ldk, fp), torch.uint8)
    PY = empty(L.omx_panel_bytes(shift + n, ldk, fp), torch.uint8)
    ex = empty(L.omx_panel_exps_count(R, ldk), torch.int32)
    ey = empty(L.omx_panel_exps_count(shift + n, ldk), torch.int32)
    tmax = empty(L.omx_panel_exps_count(R, k) + 64, torch.float32)
    strip_bytes = (ldk // (128 if fp8 else 64)) * 32768
    ne = ldk // 256

    def pack(S, P, E, row_shift, k0):
        sd = dev(np.asfortranarray(S).ravel(order="F"))
        base = ptr(P) + (row_shift // 256) * strip_bytes
        if fp8:
            eb = ptr(E) + 4 * (row_shift // 256) * ne
            assert L.omx_pack_f8_window(S.shape[0], k, ptr(sd), 1, S.shape[0], 0, 0, base, eb, ldk, k0,
                                        ptr(tmax), None) == 0
        else:
            assert L.omx_pack_f16_window(S.shape[0], k, ptr(sd), 1, S.shape[0], 0, base, None, None, ldk, k0,
                                         None) == 0
        return sd

    keep = [pack(X1, PX, ex, 0, 0), pack(X2, PX, ex, shift, k), pack(Y1, PY, ey, 0, 0), pack(Y2, PY, ey, shift, k)]
    C0 = rng.uniform(-1, 1, (n, m)).astype(np.float32)
    Cd = dev(C0.ravel())
    if fp8:
        rc = L.omx_gemm_f8_packed_window(m, n, ldk, -1.0, ptr(PX), ptr(ex), shift, ldk, 0, ptr(PY), ptr(ey), shift,
                                         ldk, 0, 1.0, ptr(Cd), m, None)
    else:
        rc = L.omx_gemm_f16_packed_window(m, n, ldk, -1.0, ptr(PX), None, shift, ldk, 0, ptr(PY), None, shift,
                                          ldk, 0, 1.0, ptr(Cd), m, None)
    assert rc == 0
    got = host(Cd).reshape(n, m).T.astype(np.float64)
    if fp8:
        def dq(S):
            out = np.zeros_like(S, dtype=np.float64)
            for s0 in range(0, S.shape[0], 256):
                for t in range(0, k, 256):
                    blk = S[s0:s0 + 256, t:t + 256]
                    e = orc.pow2_scale_exp(float(np.abs(blk).max()))
                    out[s0:s0 + 256, t:t + 256] = orc.e4m3_bits_to_f32(orc.f32_to_e4m3_bits(np.ldexp(blk, e))) * 2.0 ** -e
            return out
        Xa, Ya, Xb, Yb = dq(X1)[shift:], dq(Y1)[shift:], dq(X2), dq(Y2)
    else:
        h = lambda S: S.astype(np.float16).astype(np.float64)
        Xa, Ya, Xb, Yb = h(X1)[shift:], h(Y1)[shift:], h(X2), h(Y2)
    exact = C0.T.astype(np.float64) - Xa @ Ya.T - Xb @ Yb.T
    bound = (ldk + 2) * 2.0**-24 * (np.abs(Xa) @ np.abs(Ya).T + np.abs(Xb) @ np.abs(Yb).T) + 2.0**-23 * np.abs(exact)
    assert np.all(np.abs(got - exact) <= bound)
    # second window alone == the panel packed by itself
    C1, C2 = dev(C0.ravel()), dev(C0.ravel())
    if fp8:
        P2 = empty(L.omx_panel_bytes(m, k, 1), torch.uint8)
        Q2 = empty(L.omx_panel_bytes(n, k, 1), torch.uint8)
        e2 = empty(L.omx_panel_exps_count(m, k), torch.int32)
        f2 = empty(L.omx_panel_exps_count(n, k), torch.int32)
        t2 = empty(L.omx_panel_exps_count(max(m, n), k), torch.float32)
        assert L.omx_pack_f8(m, k, ptr(keep[1]), 1, m, ptr(P2), ptr(e2), ptr(t2), None) == 0
        assert L.omx_pack_f8(n, k, ptr(keep[3]), 1, n, ptr(Q2), ptr(f2), ptr(t2), None) == 0
        assert L.omx_gemm_f8_packed(m, n, k, -1.0, ptr(P2), ptr(e2), 0, ptr(Q2), ptr(f2), 0, 1.0, ptr(C1), m,
                                    None) == 0
        assert L.omx_gemm_f8_packed_window(m, n, k, -1.0, ptr(PX), ptr(ex), shift, ldk, k, ptr(PY), ptr(ey), shift,
                                           ldk, k, 1.0, ptr(C2), m, None) == 0
    else:
        P2 = dev(pack16(X2.astype(np.float16)).view(np.int16))
        Q2 = dev(pack16(Y2.astype(np.float16)).view(np.int16))
        assert L.omx_gemm_f16_packed(m, n, k, -1.0, ptr(P2), 0, ptr(Q2), 0, 1.0, ptr(C1), m, None) == 0
        assert L.omx_gemm_f16_packed_window(m, n, k, -1.0, ptr(PX), None, shift, ldk, k, ptr(PY), None, shift,
                                            ldk, k, 1.0, ptr(C2), m, None) == 0
    assert np.array_equal(host(C1), host(C2))
