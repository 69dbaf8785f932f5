"""GPU parity of the on-device generators against the oracle (bit-exact), and
the reference-held known answer KAT-0 (doc/crusher_example_32x32.out:186)."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests.gpu_util import colmajor_to_numpy, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,seed", [(512, 42), (1536, 7), (2560, 42)])
def test_matgen_full_bitexact(orc, n, seed):
    torch = torch_cuda()
    _say(f"matgen_full n={n} seed={seed}")
    lda = n + 64
    A = empty(lda * n, torch.float32)
    d = empty(n, torch.float64)
    assert _L().omx_matgen_f32(n, seed, ptr(A), lda, ptr(d), None) == 0
    Ag = colmajor_to_numpy(A, n, n, lda)
    Ao, do = orc.matgen_full(n, seed)
    assert np.array_equal(host(d), do)
    assert np.array_equal(Ag, Ao)


@pytest.mark.parametrize("i0,j0,m,k", [(0, 0, 300, 5), (204544, 100000, 256, 9), (123, 204790, 77, 10)])
def test_matgen_block_at_full_size(orc, i0, j0, m, k):
    torch = torch_cuda()
    n = 204800
    _say(f"matgen_block n={n} i0={i0} j0={j0} m={m} k={k}")
    A = empty(m * k, torch.float32)
    assert _L().omx_matgen_block_f32(n, 42, i0, j0, m, k, ptr(A), m, None) == 0
    assert np.array_equal(colmajor_to_numpy(A, m, k), orc.matgen_block(n, 42, i0, j0, m, k))


def test_diag_rows_at_full_size(orc):
    torch = torch_cuda()
    n = 204800
    _say("diag rows n=204800")
    for i0, m in [(0, 24), (204800 - 24, 24), (102400, 8)]:
        d = empty(m, torch.float64)
        assert _L().omx_diag_f64(n, 42, i0, m, ptr(d), None) == 0
        assert np.array_equal(host(d), orc.diag_gpu_order(n, 42, i0, m))


def test_rhs_at_full_size(orc):
    torch = torch_cuda()
    n = 204800
    b = empty(n, torch.float64)
    assert _L().omx_rhs_f64(n, 42, ptr(b), None) == 0
    assert np.array_equal(host(b), orc.rhs(n, 42))


def test_generator_argument_errors():
    L = _L()
    assert L.omx_matgen_f32(0, 42, None, 0, None, None) < 0
    assert L.omx_diag_f64(100, 42, 90, 20, None, None) < 0


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_kat0_step0_residual_and_harness():
    """N=4014080, seed 42, x0 = b./d: the reference's run printed
    residual=8.778080e-04 and hpl-harness=1311160.547096 for IR step 0.  This
    involves no LU: it pins generator, diagonal, rhs, on-the-fly GEMV, norms
    and the harness formula.  The reference summed in another order (32x32
    grid), so agreement is to the printed digits."""
    torch = torch_cuda()
    n = 4014080
    L = _L()
    _say("KAT-0: diag of N=4014080")
    d = empty(n, torch.float64)
    b = empty(n, torch.float64)
    x = empty(n, torch.float64)
    r = empty(n, torch.float64)
    nrm = empty(3, torch.float64)
    assert L.omx_diag_f64(n, 42, 0, n, ptr(d), None) == 0
    assert L.omx_rhs_f64(n, 42, ptr(b), None) == 0
    assert L.omx_div_f64(n, ptr(b), ptr(d), ptr(x), None) == 0
    assert L.omx_copy_f64(n, ptr(b), ptr(r), None) == 0
    torch.cuda.synchronize()
    _say("KAT-0: residual GEMV")
    wb = L.omx_otf_gemv_workspace_bytes(n)
    w = empty(max(wb // 8, 1), torch.float64)
    assert L.omx_otf_gemv_f64(n, 42, ptr(d), -1.0, ptr(x), ptr(r), ptr(w), wb, None) == 0
    assert L.omx_norm_inf_f64(n, ptr(r), ptr(nrm), None) == 0
    assert L.omx_norm_inf_f64(n, ptr(x), ptr(nrm) + 8, None) == 0
    assert L.omx_norm_inf_f64(n, ptr(b), ptr(nrm) + 16, None) == 0
    normr, normx, normb = host(nrm)
    norma = 2.0 * float(host(d).max())
    h = L.omx_harness(normr, norma, normx, normb, n)
    _say(f"KAT-0: residual={normr:e} harness={h:f}")
    assert f"{normr:e}" == "8.778080e-04"
    assert abs(h - 1311160.547096) < 2e-6 * 1311160.547096
    assert f"{h:.3f}" == "1311160.547"
