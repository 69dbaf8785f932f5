"""Pins the oracle's FP16 and FP8 e4m3 conversions: against numpy's float16
(IEEE round-to-nearest-even), the reference's fp16sim.hpp compiled from its
sources, torch.float8_e4m3fn, and exhaustive enumeration of code points."""
import numpy as np
import pytest


def _edge_floats():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e30, -1e30, np.inf, -np.inf,
                  2.0**-14, 2.0**-15, 2.0**-24, 2.0**-25, 1.5 * 2.0**-25, 2.0**-26, 6.1e-5,
                  1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 1e-5, -1e-5], dtype=np.float32),
    ]).astype(np.float32)
    return x


def test_f16_encode_matches_numpy(orc):
    x = _edge_floats()
    h = orc.f32_to_f16_bits(x)
    assert np.array_equal(h, x.astype(np.float16).view(np.uint16))


def test_f16_ties_to_even_all_midpoints(orc):
    """Every midpoint between adjacent finite float16 values must round to the
    even neighbour."""
    allh = np.arange(0, 0x7c00, dtype=np.uint16)
    lo = allh[:-1].view(np.float16).astype(np.float64)
    hi = allh[1:].view(np.float16).astype(np.float64)
    mid = ((lo + hi) / 2).astype(np.float32)
    exact = (mid.astype(np.float64) == (lo + hi) / 2)
    h = orc.f32_to_f16_bits(mid[exact])
    even = np.where((allh[:-1][exact] & 1) == 0, allh[:-1][exact], allh[1:][exact])
    assert np.array_equal(h, even)


def test_f16_decode_exhaustive(orc):
    allh = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    f = orc.f16_bits_to_f32(allh)
    npf = allh.view(np.float16).astype(np.float32)
    nan = np.isnan(npf)
    assert np.array_equal(np.isnan(f), nan)
    assert np.array_equal(f[~nan], npf[~nan])


def test_f16_matches_reference_fp16sim(orc, ref):
    x = _edge_floats()[:4000]
    h = orc.f32_to_f16_bits(x)
    hr = np.array([ref.ref_f32_to_f16(float(v)) for v in x], dtype=np.uint16)
    finite = np.isfinite(x)
    assert np.array_equal(h[finite], hr[finite])


def test_e4m3_decode_exhaustive_matches_torch(orc):
    torch = pytest.importorskip("torch")
    q = np.arange(256, dtype=np.uint8)
    f = orc.e4m3_bits_to_f32(q)
    t = torch.from_numpy(q).view(torch.float8_e4m3fn).float().numpy()
    nan = np.isnan(t)
    assert np.array_equal(np.isnan(f), nan)
    assert np.array_equal(f[~nan], t[~nan])
    assert np.nanmax(f) == 448.0
    assert f[1] == 2.0**-9  # smallest subnormal


def test_e4m3_encode_matches_torch_in_range(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(50000) * 10.0 ** rng.uniform(-4, 2.6, 50000)).astype(np.float32)
    x = x[np.abs(x) <= 448.0]
    # all midpoints between neighbouring codes
    codes = np.arange(0, 0x7e, dtype=np.uint8)
    v = orc.e4m3_bits_to_f32(codes).astype(np.float64)
    mids = ((v[:-1] + v[1:]) / 2).astype(np.float32)
    x = np.concatenate([x, mids, -mids, np.float32([0.0, -0.0, 448.0, -448.0])])
    q = orc.f32_to_e4m3_bits(x)
    qt = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(q, qt)


def test_e4m3_saturates(orc):
    q = orc.f32_to_e4m3_bits(np.float32([449.0, 1e6, np.inf, -500.0]))
    assert list(q) == [0x7e, 0x7e, 0x7e, 0xfe]


@pytest.mark.parametrize("maxabs", [1e-5, 3.7e-3, 0.49, 1.0, 448.0, 500.0, 51200.0])
def test_pow2_scale_exponent(orc, maxabs):
    e = orc.pow2_scale_exp(maxabs)
    assert maxabs * 2.0**e <= 448.0
    assert maxabs * 2.0 ** (e + 1) > 448.0
