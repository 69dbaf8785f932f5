"""Device-memory plumbing for the GPU tests: torch tensors hold the buffers,
the computation goes through the C ABI of libomxp.so."""
import numpy as np


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        import pytest
        pytest.skip("no GPU")
    return torch


def dev(a):
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def empty(n, dtype):
    torch = torch_cuda()
    return torch.empty(n, dtype=dtype, device="cuda")


def ptr(t):
    return t.data_ptr()


def host(t):
    torch = torch_cuda()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def colmajor_to_numpy(t, m, n, ld=None):
    """A device buffer holding an m x n column-major matrix -> numpy (m, n)."""
    ld = m if ld is None else ld
    h = host(t).reshape(n, ld)[:, :m]
    return h.T
