"""Pins the oracle's generator against the reference's own hpl_rand.hpp
(compiled from its sources into oracle/_ref), an independent closed form with
Python integers, and exact summation."""
import math

import numpy as np
import pytest

CASES = [(10240, 0, 0), (10240, 5, 7), (10240, 10239, 10239), (61440, 123, 45678),
         (204800, 204799, 204799), (204800, 17, 204800), (4014080, 4014079, 4014080)]


@pytest.mark.parametrize("n,i,j", CASES)
def test_lcg_value_matches_reference_and_closed_form(orc, ref, n, i, j):
    s = orc.lcg_state(42, n, i, j)
    assert s == orc.py_lcg_state(42, n, i, j)
    v = orc.lcg_value(s)
    assert v == ref.ref_lcg_value(42, n, i, j)
    assert v == orc.py_lcg_value(s)
    assert -0.5 <= v < 0.5


def test_lcg_jump_matches_both_reference_powering_routines(orc, ref):
    import ctypes as C
    for steps in [0, 1, 2, 3, 511, 512 * 204800, 2**40 + 12345, 2**63 + 7]:
        a, c = orc.lcg_jump(steps)
        ra, rc = C.c_uint64(), C.c_uint64()
        ref.ref_lcg_jump(steps, C.byref(ra), C.byref(rc))
        assert (a, c) == (ra.value, rc.value)
        if steps < 2**32:
            ref.ref_lcg_jump_fugaku(steps, C.byref(ra), C.byref(rc))
            assert (a, c) == (ra.value, rc.value)


def test_seed_changes_stream(orc):
    assert orc.lcg_value(orc.lcg_state(42, 1000, 3, 4)) != orc.lcg_value(orc.lcg_state(43, 1000, 3, 4))


@pytest.mark.parametrize("n,i0,j0,m,k", [(1000, 3, 5, 64, 33), (10240, 10000, 9000, 240, 17),
                                         (204800, 204544, 100000, 256, 9)])
def test_matgen_block_matches_reference(orc, ref, n, i0, j0, m, k):
    A = orc.matgen_block(n, 42, i0, j0, m, k)
    Ar = np.zeros((k, m), dtype=np.float32)
    ref.ref_fill_block(n, i0, j0, m, k, Ar.ctypes.data, m, 42)
    assert np.array_equal(A, Ar.T)


def test_diag_sequential_order_matches_reference_calc_diag(orc, ref):
    n = 1536
    d = orc.diag_seq_order(n, 42)
    dr = np.array([ref.ref_calc_diag(i, n, 42) for i in range(n)])
    assert np.array_equal(d, dr)
    # the diagonal block of panel_fill_one_with_rand stores float(calc_diag)
    b = 256
    Ab = np.zeros((b, b), dtype=np.float32)
    ref.ref_fill_block_diag(n, 512, b, Ab.ctypes.data, b, 42)
    assert np.array_equal(np.diag(Ab.T), d[512:512 + b].astype(np.float32))


def test_diag_gpu_order_is_the_exact_row_sum_to_rounding(orc):
    """The 512-partial order differs from calc_diag only in the last bits; both
    are within n*eps of the exactly rounded sum (math.fsum)."""
    n = 3000
    dg = orc.diag_gpu_order(n, 7)
    ds = orc.diag_seq_order(n, 7)
    for i in [0, 1, 511, 512, 1499, 2999]:
        row = orc.matgen_block(n, 7, i, 0, 1, n)[0].astype(np.float64)
        # values are FP64 before the FP32 cast: regenerate them exactly
        vals = [abs(orc.lcg_value(orc.lcg_state(7, n, i, j))) for j in range(n) if j != i]
        exact = math.fsum(vals)
        assert abs(dg[i] - exact) <= n * 2.0**-53 * exact
        assert abs(ds[i] - exact) <= n * 2.0**-53 * exact
        assert row.shape == (n,)
    assert np.max(np.abs(dg - ds) / ds) < 1e-13
    # a dropped or doubled term is far outside that band
    assert np.min(dg) > 0.2 * n


def test_diag_gpu_order_tree_on_a_tiny_case(orc):
    """n < 512: every partial holds one term, so the tree is a fixed pairing
    that can be written out by hand."""
    n = 8
    vals = [[abs(orc.lcg_value(orc.lcg_state(3, n, i, j))) if i != j else 0.0 for j in range(n)]
            for i in range(n)]
    d = orc.diag_gpu_order(n, 3)
    for i in range(n):
        sh = vals[i] + [0.0] * (512 - n)
        w = 256
        while w >= 4:
            sh = [sh[t] + sh[t + w] if t < w else sh[t] for t in range(len(sh))]
            w //= 2
        assert d[i] == ((sh[0] + sh[1]) + sh[2]) + sh[3]


def test_rhs_matches_reference(orc, ref):
    n = 4096
    b = orc.rhs(n, 42)
    br = np.zeros(n)
    ref.ref_rhs(n, 0, n, br.ctypes.data, 42)
    assert np.array_equal(b, br)
    part = orc.rhs(n, 42, 1000, 96)
    assert np.array_equal(part, b[1000:1096])


def test_matgen_full_is_diagonally_dominant(orc):
    n = 512
    A, d = orc.matgen_full(n, 42)
    off = np.abs(A.astype(np.float64)).sum(axis=1) - np.abs(np.diag(A))
    assert np.array_equal(np.diag(A), d.astype(np.float32))
    assert np.allclose(off, d, rtol=1e-6)
    assert np.all(np.abs(np.diag(A)) >= off * (1 - 1e-6))
