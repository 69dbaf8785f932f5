"""The Fasi-Higham test matrices: the oracle against what the mathematics
fixes (A = L U exactly, cond_inf(A) = kappa for the paper's form), against
the reference's own routines compiled from higham_mat_impl.cpp, and the
library's host routines (beta search, cond estimate, infinity norm)
against the oracle.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import higham as H

RHO = 0.5


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


@pytest.mark.parametrize("n,kappa", [(50, 1e2), (200, 1e3), (301, 1e4), (128, 1e6)])
def test_matrix_is_the_product_of_its_factors(n, kappa):
    b = H.beta(n, kappa, RHO)
    a = RHO * b
    for sa, sb in ((a, b), (-a, -b)):  # the paper's form and the reference's panel form
        A = H.matrix(n, sa, sb)
        L, U = H.factors(n, sa, sb)
        assert np.max(np.abs(L @ U - A)) <= 4 * n * 2.0**-53 * np.max(np.abs(A))
    # the reference's panel form (matgen.hpp:157-208) is L = I + alpha N,
    # U = I + beta N: above the diagonal beta + alpha beta i
    P = H.matrix(n, -a, -b)
    i, j = 3, n - 1
    assert P[i, j] == b + (a * b) * i and P[j, i] == a + (a * b) * i and P[i, i] == 1 + (a * b) * i


@pytest.mark.parametrize("n,kappa", [(100, 1e2), (300, 1e3), (500, 1e4), (200, 1e6)])
def test_beta_gives_the_requested_condition_number(n, kappa):
    """For the paper's form the estimate is the exact cond_inf (LAPACK's
    inverse agrees to ~1e-12 here): beta solves cond_inf(A) = kappa."""
    b = H.beta(n, kappa, RHO)
    assert 0 < RHO * b <= 1
    assert abs(H.cond_est(n, RHO * b, b) / kappa - 1) <= 1e-12
    A = H.matrix(n, RHO * b, b)
    assert abs(np.linalg.cond(A, np.inf) / kappa - 1) <= 1e-9


def test_cond_estimate_closed_form_pieces():
    """n = 2: A = [[1, -b], [-a, 1 + ab]], A^-1 = [[1 + ab, b], [a, 1]];
    row sums 1 + b, a + 1 + ab and 1 + ab + b, a + 1."""
    a, b = 0.25, 0.5
    A = H.matrix(2, a, b)
    assert np.array_equal(A, [[1, -b], [-a, 1 + a * b]])
    want = max(1 + b, a + 1 + a * b) * max(1 + a * b + b, a + 1)
    assert abs(H.cond_est(2, a, b) - want) <= 1e-15 * want


@pytest.mark.parametrize("n", [2, 5, 100, 777])
def test_norm_inf_matches_row_sums(n):
    L = _L()
    b = H.beta(max(n, 3), 1e3, RHO) if n > 2 else 0.3
    for sa, sb in ((RHO * b, b), (-RHO * b, -b)):
        want = np.abs(H.matrix(n, sa, sb)).sum(axis=1).max()
        assert abs(H.norm_inf(n, sa, sb) - want) <= 1e-15 * want
        assert abs(L.omx_higham_norm_inf(n, sa, sb) - want) <= 4 * n * 2.0**-53 * want


@pytest.mark.parametrize("n,kappa", [(100, 1e2), (2048, 1e3), (61440, 1e4), (1000, 1e8), (204800, 1e3),
                                     (12288, 1e6)])
def test_library_beta_and_estimate_equal_the_oracle(n, kappa):
    L = _L()
    bo = H.beta(n, kappa, RHO)
    assert L.omx_higham_beta(n, kappa, RHO) == bo
    for x in (bo, 0.5 * bo, 1e-7, 0.3):
        assert L.omx_higham_cond_est(n, RHO * x, x) == H.cond_est(n, RHO * x, x)


def test_library_rejects_bad_arguments():
    L = _L()
    assert np.isnan(L.omx_higham_beta(1, 1e3, RHO))
    assert np.isnan(L.omx_higham_beta(100, 0.5, RHO))
    assert np.isnan(L.omx_higham_beta(100, 1e3, 0.0))
    assert np.isnan(L.omx_higham_cond_est(1, 0.1, 0.2))
    assert np.isnan(L.omx_higham_norm_inf(0, 0.1, 0.2))
    # generation checks its arguments before any device call
    assert L.omx_higham_matgen_f32(0, 0.1, 0.2, None, 1, None, None) != 0
    assert L.omx_higham_matgen_f32(8, 0.1, 0.2, None, 4, None, None) != 0
    assert L.omx_higham_gemv_f64(0, 0.1, 0.2, 1.0, None, None, None) != 0


# --- against the reference's routines, compiled from its own source ---------

@pytest.mark.parametrize("n,kappa", [(100, 1e2), (2048, 1e3), (30000, 1e4), (1000, 1e8), (46000, 1e3)])
def test_oracle_beta_equals_the_reference(ref, n, kappa):
    """higham_mat_comp_beta and fhpl (higham_mat_impl.cpp:36-128) at orders
    where their 32-bit integer products do not overflow (n <= 46340)."""
    bo = H.beta(n, kappa, RHO)
    assert ref.higham_mat_comp_beta(n, kappa, RHO) == bo
    for x in (bo, 0.5 * bo, 1e-7, 0.3):
        assert ref.ref_higham_fhpl(n, RHO * x, x) == H.cond_est(n, RHO * x, x)


@pytest.mark.parametrize("n,lda", [(64, 64), (300, 303)])
def test_oracle_matrix_equals_the_reference(ref, n, lda):
    """hplai_matrix_impl (higham_mat_impl.cpp:130-141), column-major."""
    b = H.beta(n, 1e3, RHO)
    a = RHO * b
    out = np.zeros(lda * n)
    ref.hplai_matrix_impl(n, out.ctypes.data, lda, a, b)
    got = out.reshape(n, lda)[:, :n].T
    assert np.array_equal(got, H.matrix(n, a, b))
