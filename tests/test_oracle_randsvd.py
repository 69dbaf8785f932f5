"""The randsvd test matrices: the oracle's dense construction against what
the mathematics fixes (singular values, condition number, orthogonality),
and the library's low-rank factors against the oracle (CPU only: the factor
computation is host code)."""
import ctypes as C

import numpy as np
import pytest


@pytest.mark.parametrize("n,kappa", [(64, 1e2), (200, 1e4), (301, 1e6)])
def test_dense_has_the_prescribed_singular_values(orc, n, kappa):
    A = orc.randsvd_dense(n, 42, kappa, 2)
    s = np.linalg.svd(A, compute_uv=False)
    sigma = kappa ** (-np.arange(n) / (n - 1))
    assert np.max(np.abs(s - sigma)) <= 1e-13
    assert abs((s[0] / s[-1]) / kappa - 1) <= 1e-6
    # not a diagonal matrix: the reflectors mix every row and column
    off = A - np.diag(np.diag(A))
    assert np.count_nonzero(np.abs(off) > 1e-12) > n * (n - 1) // 2


def test_reflectors_are_orthogonal_and_seeded(orc):
    n = 96
    A1 = orc.randsvd_dense(n, 7, 1e3, 2)
    A2 = orc.randsvd_dense(n, 8, 1e3, 2)
    assert not np.allclose(A1, A2)
    # no reflectors: the diagonal of singular values itself
    D = orc.randsvd_dense(n, 7, 1e3, 0)
    want = np.diag(1e3 ** (-np.arange(n) / (n - 1)))
    assert np.allclose(D, want, rtol=4e-16, atol=0)  # pow() may differ by an ulp
    # with reflectors A^T A has the same spectrum as diag(sigma)^2 and the
    # same Frobenius norm
    assert abs(np.linalg.norm(A1) - np.linalg.norm(D)) <= 1e-13 * np.linalg.norm(D)


def test_transpose_is_a_different_matrix(orc):
    """A and A^T share the singular values, so the SVD check cannot tell
    the left and right reflector sets apart; they do differ, and the
    library comparison below pins which is which."""
    n = 80
    A = orc.randsvd_dense(n, 3, 1e2, 2)
    assert not np.allclose(A, A.T)


@pytest.mark.parametrize("n,nrefl", [(1, 2), (5, 1), (128, 2), (333, 3)])
def test_library_lowrank_factors_match_oracle(orc, n, nrefl):
    from openmxp_amd import _lib
    L = _lib.lib()
    R = 2 * nrefl
    sigma = np.zeros(n)
    P = np.zeros((R, n))
    Q = np.zeros((R, n))
    rc = L.omx_randsvd_factors_host(n, 42, 1e4, nrefl, sigma.ctypes.data, P.ctypes.data, Q.ctypes.data)
    assert rc == 0
    A = np.diag(sigma) + P.T @ Q  # P.T is n x R (column-major n x R storage)
    Ao = orc.randsvd_dense(n, 42, 1e4, nrefl)
    assert np.max(np.abs(A - Ao)) <= 1e-14 * max(1.0, np.max(np.abs(Ao)))


def test_library_rejects_bad_arguments():
    from openmxp_amd import _lib
    L = _lib.lib()
    s = np.zeros(8)
    P = np.zeros(8 * 34)
    assert L.omx_randsvd_factors_host(8, 1, 0.5, 1, s.ctypes.data, P.ctypes.data, P.ctypes.data) < 0
    assert L.omx_randsvd_factors_host(8, 1, 10.0, 9, s.ctypes.data, P.ctypes.data, P.ctypes.data) < 0
    assert L.omx_randsvd_factors_host(0, 1, 10.0, 1, s.ctypes.data, P.ctypes.data, P.ctypes.data) < 0
