"""Pins the oracle's refinement: harness closed form, classical IR and GMRES-IR
against an FP64 LAPACK solve (scipy dgesv) of the same generated system."""
import numpy as np
import pytest
import scipy.linalg as sla


def _system(orc, n, seed):
    A, d = orc.matgen_full(n, seed)
    A64 = np.array([[orc.lcg_value(orc.lcg_state(seed, n, i, j)) for j in range(n)] for i in range(n)])
    A64[np.diag_indices(n)] = d
    b = orc.rhs(n, seed)
    return A, d, A64, b


def _fwd_bound(A64, harness):
    """Forward error implied by the stopping test: ||dx||/||x|| <= kappa_inf *
    (harness * n * u) * 2, plus LAPACK's own kappa * n * u."""
    n = A64.shape[0]
    kappa = np.linalg.cond(A64, np.inf)
    u = 2.0**-53
    return kappa * n * u * (2.0 * max(harness, 1.0) + 1.0)


def test_harness_formula(orc):
    # iterative_refinement.hpp:156-157 with u = DBL_EPSILON / 2
    h = orc.harness(1e-10, 2.0, 3.0, 0.5, 1000)
    assert h == 1e-10 / (2.0 * 3.0 + 0.5) / (1000 * np.finfo(np.float64).eps / 2)


@pytest.fixture(scope="module")
def sys512(orc):
    return _system(orc, 512, 42)


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_ir_classic_converges_to_lapack_solution(orc, sys512, alt):
    A, d, A64, b = sys512
    n, blk = 512, 256
    LU, _ = orc.lu_factor(A, blk, alt=alt)
    res = orc.ir_classic(LU, blk, 42, d, b)
    assert 0 <= res["iters"] <= 5
    assert res["harness"] < 16.0
    xs = sla.solve(A64, b)
    fwd = np.max(np.abs(res["x"] - xs)) / np.max(np.abs(xs))
    assert fwd <= _fwd_bound(A64, res["harness"])
    # the harness of step 0 is the x0 = b./d residual: recompute it directly
    x0 = b / d
    r0 = b - A64 @ x0
    h0 = np.max(np.abs(r0)) / (2 * d.max() * np.max(np.abs(x0)) + np.max(np.abs(b))) / (n * 2.0**-53)
    assert abs(res["trace"][0, 1] - h0) < 1e-9 * h0


def test_gmres_ir_converges(orc, sys512):
    A, d, A64, b = sys512
    LU, _ = orc.lu_factor(A, 256, alt=1)
    res = orc.gmres_ir(LU, 256, 42, d, b)
    assert 0 <= res["iters"] <= 3
    assert res["harness"] < 16.0
    xs = sla.solve(A64, b)
    assert np.max(np.abs(res["x"] - xs)) / np.max(np.abs(xs)) <= _fwd_bound(A64, res["harness"])


def test_ir_detects_bad_factors(orc, sys512):
    """With the LU replaced by its diagonal only (a planted defect) classical IR
    still contracts for this dominant matrix but needs many more steps; with
    garbage factors it fails to converge within maxit."""
    A, d, A64, b = sys512
    bad = np.asfortranarray(np.diag(np.diag(A)).astype(np.float32))
    res = orc.ir_classic(bad, 256, 42, d, b, maxit=50)
    good = orc.ir_classic(orc.lu_factor(A, 256)[0], 256, 42, d, b)
    assert res["iters"] == -1 or res["iters"] > good["iters"] + 3
    junk = np.asfortranarray(np.eye(512, dtype=np.float32) * 1e-3)
    res2 = orc.ir_classic(junk, 256, 42, d, b, maxit=5)
    assert res2["iters"] == -1 and res2["harness"] == -1.0
