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


@pytest.mark.parametrize("n,lda", [(5, 7), (256, 256), (300, 303), (777, 780), (1000, 1000), (2304, 2305)])
def test_trsv_ragged_and_strided_match_lapack(n, lda):
    """Orders that are not multiples of the 256-row solve block, leading
    dimensions with and without 16-byte column alignment; against LAPACK's
    FP64 triangular solve of the same FP32 factors."""
    import scipy.linalg as sla
    torch = torch_cuda()
    print(f"trsv ragged n={n} lda={lda}", flush=True)
    L = _L()
    rng = np.random.default_rng(n)
    F = (np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) + np.triu(rng.uniform(-1, 1, (n, n)) / n, 1)
         + np.diag(rng.uniform(1, 2, n))).astype(np.float32)
    store = np.zeros((n, lda), np.float32)  # column j at store[j, :lda]
    store[:, :n] = F.T
    Ad = dev(store.ravel())
    r = rng.uniform(-1, 1, n)
    xd = dev(r)
    assert L.omx_trsv_lower_unit_f32f64(n, n, ptr(Ad), lda, ptr(xd), None, 0, None) == 0
    F64 = F.astype(np.float64)
    y = sla.solve_triangular(np.tril(F64, -1) + np.eye(n), r, lower=True, unit_diagonal=True)
    yg = host(xd)
    assert np.max(np.abs(yg - y)) <= 1e-13 * np.max(np.abs(y))
    assert L.omx_trsv_upper_f32f64(n, n, ptr(Ad), lda, ptr(xd), None, 0, None) == 0
    z = sla.solve_triangular(np.triu(F64), yg, lower=False)
    zg = host(xd)
    assert np.max(np.abs(zg - z)) <= 1e-13 * np.max(np.abs(z))


def test_trsv_argument_errors():
    torch = torch_cuda()
    L = _L()
    a = empty(64, torch.float32)
    x = empty(8, torch.float64)
    assert L.omx_trsv_lower_unit_f32f64(8, 3, ptr(a), 8, ptr(x), None, 0, None) < 0   # n % b
    assert L.omx_trsv_upper_f32f64(8, 8, ptr(a), 4, ptr(x), None, 0, None) < 0        # lda < n
    assert L.omx_trsv_upper_f32f64(0, 8, ptr(a), 8, ptr(x), None, 0, None) < 0
