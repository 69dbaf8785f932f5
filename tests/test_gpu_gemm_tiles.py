"""The two trailing-update GEMM kernel families against each other: 256 x 256
tiles one workgroup per CU (the default) and 256 x 128 tiles two per CU
(OMX_GEMM_TILE=half).  Both accumulate each C element over k in the same
order with the same MFMA instructions, so their results are bit-identical;
edge tiles (m, n not multiples of the tile), short products (the FP16
kernel's checked path at k = 64), scaled FP16, beta = 0 and the FP8 scale
tiles are all covered.  The default kernel's accuracy against the exact
product of the rounded operands is checked by test_gpu_gemm.py."""
import numpy as np
import pytest

from tests.gpu_util import dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def _run_both(monkeypatch, launch, C0):
    outs = []
    for tile in ("256", "half"):
        if tile == "256":
            monkeypatch.delenv("OMX_GEMM_TILE", raising=False)
        else:
            monkeypatch.setenv("OMX_GEMM_TILE", tile)
        Cd = dev(C0)
        assert launch(Cd) == 0
        outs.append(host(Cd))
    return outs


def _panel(L, R, K, fp8, rng, col_scale):
    torch = torch_cuda()
    S = (rng.standard_normal((R, K)) * col_scale).astype(np.float32)
    sd = dev(np.asfortranarray(S).ravel(order="F"))
    fp = 1 if fp8 else 0
    P = empty(L.omx_panel_bytes(R, K, fp), torch.uint8)
    ne = L.omx_panel_exps_count(R, K)
    ex = empty(max(ne, 1), torch.int32)
    tm = empty(max(ne, 1), torch.float32)
    if fp8:
        assert L.omx_pack_f8(R, K, ptr(sd), 1, R, ptr(P), ptr(ex), ptr(tm), None) == 0
    else:
        assert L.omx_pack_f16(R, K, ptr(sd), 1, R, 0, ptr(P), None, None) == 0
    return P, ex, sd


@pytest.mark.parametrize("m,n,k,beta", [(256, 128, 64, 1.0), (700, 300, 128, 1.0), (1024, 1280, 512, 0.0),
                                        (3072, 2944, 5120, 1.0)])
def test_f16_tiles_bitidentical(monkeypatch, m, n, k, beta):
    torch_cuda()
    L = _L()
    print(f"f16 tiles m={m} n={n} k={k} beta={beta}", flush=True)
    rng = np.random.default_rng(m + k)
    X, _, xs = _panel(L, m, k, False, rng, 0.3)
    Y, _, ys = _panel(L, n, k, False, rng, 0.3)
    ldc = m + 5
    C0 = rng.uniform(-1, 1, n * ldc).astype(np.float32)
    a, b = _run_both(monkeypatch, lambda Cd: L.omx_gemm_f16_packed(m, n, k, -1.0, ptr(X), 0, ptr(Y), 0, beta,
                                                                    ptr(Cd), ldc, None), C0)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, C0)


def test_f16_scaled_tiles_bitidentical(monkeypatch):
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(7)
    m, n, k = 768, 640, 256
    packs = []
    for R in (m, n):
        S = (rng.standard_normal((R, k)) * 1e3).astype(np.float32)
        sd = dev(np.asfortranarray(S).ravel(order="F"))
        P = empty(L.omx_panel_bytes(R, k, 0), torch.uint8)
        se = empty(-(-R // 256), torch.int32)
        tm = empty(L.omx_panel_exps_count(R, k), torch.float32)
        assert L.omx_pack_f16_scaled(R, k, ptr(sd), 1, R, 0, ptr(P), ptr(se), ptr(tm), None) == 0
        packs.append((P, se, sd))
    (X, xe, _), (Y, ye, _) = packs
    C0 = rng.uniform(-1, 1, n * m).astype(np.float32)
    a, b = _run_both(monkeypatch, lambda Cd: L.omx_gemm_f16_packed_scaled(m, n, k, -1.0, ptr(X), ptr(xe), 0, ptr(Y),
                                                                           ptr(ye), 0, 1.0, ptr(Cd), m, None), C0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("m,n,k,beta", [(512, 384, 256, 1.0), (900, 200, 512, 0.0), (2560, 2816, 10240, 1.0)])
def test_f8_tiles_bitidentical(monkeypatch, m, n, k, beta):
    torch_cuda()
    L = _L()
    print(f"f8 tiles m={m} n={n} k={k} beta={beta}", flush=True)
    rng = np.random.default_rng(m + n)
    X, xe, xs = _panel(L, m, k, True, rng, 1e-5)
    Y, ye, ys = _panel(L, n, k, True, rng, 0.3)
    C0 = rng.uniform(-1, 1, n * m).astype(np.float32)
    a, b = _run_both(monkeypatch, lambda Cd: L.omx_gemm_f8_packed(m, n, k, -1.0, ptr(X), ptr(xe), 0, ptr(Y), ptr(ye),
                                                                   0, beta, ptr(Cd), m, None), C0)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, C0)
