"""The allocation-free residual GEMV: omx_otf_gemv_f64_noalloc (partial sums
in the library's static device scratch) against omx_otf_gemv_f64 (partial
sums in a caller workspace) on the same inputs.  The two run the same
kernels over the same column segments, so they agree bit for bit.  Two calls
in a row, the second with a smaller order, so the second reads none of the
first call's partial sums; the second call must leave the device's free
memory unchanged (nothing allocated)."""
import numpy as np
import pytest

from tests.gpu_util import dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def test_noalloc_gemv_matches_workspace_entry():
    torch = torch_cuda()
    from openmxp_amd import _lib
    L = _lib.lib()
    seed = 42
    for call, n in enumerate((204800, 10240)):
        print(f"noalloc gemv call {call} n={n}", flush=True)
        rng = np.random.default_rng(n)
        d = empty(n, torch.float64)
        assert L.omx_diag_f64(n, seed, 0, n, ptr(d), None) == 0
        x = dev(rng.uniform(-1.0, 1.0, n))
        y0 = rng.uniform(-1.0, 1.0, n)
        y_ws = dev(y0)
        y_na = dev(y0)
        wb = L.omx_otf_gemv_workspace_bytes(n)
        assert wb > 0
        w = empty(wb, torch.uint8)
        assert L.omx_otf_gemv_f64(n, seed, ptr(d), -1.0, ptr(x), ptr(y_ws), ptr(w), wb, None) == 0
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        assert L.omx_otf_gemv_f64_noalloc(n, seed, ptr(d), -1.0, ptr(x), ptr(y_na), None) == 0
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        a, b = host(y_ws), host(y_na)
        assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
        assert not np.array_equal(a, y0)
        if call > 0:
            assert free1 == free0


def test_noalloc_gemv_rejects_bad_arguments():
    torch_cuda()
    from openmxp_amd import _lib
    L = _lib.lib()
    assert L.omx_otf_gemv_f64_noalloc(0, 42, 1, -1.0, 1, 1, None) == -1
    assert L.omx_otf_gemv_f64_noalloc(16, 42, None, -1.0, 1, 1, None) == -1
