"""GPU parity of the refinement kernels (residual GEMV, norms, vector ops,
blocked triangular solves) against the oracle."""
import sys

import numpy as np
import pytest

from tests.gpu_util import dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


@pytest.mark.parametrize("n", [1000, 2048, 10240])
def test_otf_gemv_matches_oracle(orc, n):
    torch = torch_cuda()
    print(f"otf_gemv n={n}", flush=True)
    L = _L()
    d = orc.diag_gpu_order(n, 42)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    yo = orc.otf_gemv(n, 42, d, -1.0, x, y0)
    dd, xd, yd = dev(d), dev(x), dev(y0)
    wb = L.omx_otf_gemv_workspace_bytes(n)
    w = empty(max(wb // 8, 1), torch.float64)
    assert L.omx_otf_gemv_f64(n, 42, ptr(dd), -1.0, ptr(xd), ptr(yd), ptr(w), wb, None) == 0
    yg = host(yd)
    # |A_ij| < 0.5 off the diagonal: row bound of sum |a_ij x_j|
    bound = 4 * n * 2.0**-53 * (0.5 * np.abs(x).sum() + d * np.abs(x))
    assert np.all(np.abs(yg - yo) <= bound + 1e-300)
    # deterministic
    yd2 = dev(y0)
    assert L.omx_otf_gemv_f64(n, 42, ptr(dd), -1.0, ptr(xd), ptr(yd2), ptr(w), wb, None) == 0
    assert np.array_equal(host(yd2), yg)


def test_otf_gemv_workspace_check():
    L = _L()
    n = 10240
    wb = L.omx_otf_gemv_workspace_bytes(n)
    assert wb > 0
    assert L.omx_otf_gemv_f64(n, 42, 1, -1.0, 1, 1, None, 0, None) == -3


def test_norm_and_vector_ops_bitexact():
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(1)
    n = 100003
    a = rng.standard_normal(n)
    b = rng.uniform(0.5, 2, n)
    ad, bd = dev(a), dev(b)
    out = empty(1, torch.float64)
    assert L.omx_norm_inf_f64(n, ptr(ad), ptr(out), None) == 0
    assert host(out)[0] == np.max(np.abs(a))
    c = empty(n, torch.float64)
    assert L.omx_copy_f64(n, ptr(ad), ptr(c), None) == 0
    assert np.array_equal(host(c), a)
    assert L.omx_div_f64(n, ptr(ad), ptr(bd), ptr(c), None) == 0
    assert np.array_equal(host(c), a / b)
    assert L.omx_axpy_f64(n, 0.75, ptr(ad), ptr(bd), None) == 0
    got = host(bd)
    exact = b + 0.75 * a
    # the device contracts y + alpha*x into one fma: one rounding instead of
    # numpy's two, so they agree to within an ulp of the summands
    assert np.all(np.abs(got - exact) <= 2.0**-52 * (np.abs(b) + 0.75 * np.abs(a)))


@pytest.mark.parametrize("n,b", [(1024, 256), (2048, 512)])
def test_trsv_matches_oracle(orc, n, b):
    torch = torch_cuda()
    print(f"trsv n={n} b={b}", flush=True)
    L = _L()
    A, d = orc.matgen_full(n, 42)
    LU, _ = orc.lu_factor(A, b, alt=1)
    rng = np.random.default_rng(2)
    r = rng.uniform(-1, 1, n)
    Ad = dev(np.asfortranarray(LU).ravel(order="F"))
    xd = dev(r)
    wb = L.omx_trsv_workspace_bytes(n)
    w = empty(wb // 8, torch.float64)
    assert L.omx_trsv_lower_unit_f32f64(n, b, ptr(Ad), n, ptr(xd), ptr(w), wb, None) == 0
    yo = orc.trsv_lower_unit(LU, b, r)
    yg = host(xd)
    assert np.max(np.abs(yg - yo)) <= 1e-13 * np.max(np.abs(yo))
    assert L.omx_trsv_upper_f32f64(n, b, ptr(Ad), n, ptr(xd), ptr(w), wb, None) == 0
    zo = orc.trsv_upper(LU, b, yg)
    zg = host(xd)
    assert np.max(np.abs(zg - zo)) <= 1e-12 * np.max(np.abs(zo))
