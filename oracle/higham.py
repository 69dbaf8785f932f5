"""CPU restatement of the Fasi-Higham HPL-AI test matrix (the reference's
--highammat family) -- TEST INFRASTRUCTURE ONLY: tests use it as the checker
of the library's generator, residual and beta search; nothing in the
library imports it.

Reference: src/higham_mat_impl.cpp (after github.com/higham/hpl-ai-matrix)
and src/highammgen.hpp.  The matrix is A = L U with L unit lower, -alpha
below the diagonal, and U unit upper, -beta above it; alpha = rho beta with
beta chosen so that the paper's estimate of cond_inf(A) equals kappa.
"""
from __future__ import annotations

import math
import sys

import numpy as np

EPS = sys.float_info.epsilon
DBL_MAX = sys.float_info.max


def _pow(x: float, y: float) -> float:
    """C's pow: +inf on overflow where Python raises."""
    try:
        return math.pow(x, y)
    except OverflowError:
        return math.inf


def cond_est(n: int, a: float, b: float) -> float:
    """The paper's cond_inf estimate of A(n, a, b): the largest of three row
    sums of A (the first row, row i' = min(floor(1/a), n), the last row)
    times the larger of the first and last row sums of A^-1 (fhpl,
    higham_mat_impl.cpp:37-66; the integer products exactly, where the
    reference's 32-bit ints overflow above n = 46340)."""
    if math.isnan(a):
        a = b / 2
    l1 = 1 + (n - 1) * b
    idash = min(int(math.floor(1.0 / a)), n)
    k = min(int(math.floor((1 + b) / b)), n - 1)
    lid = 1 + float(2 * k - idash + 1) * a + float(n - idash) * b \
        + float(-k * k + k + 3 * idash * (idash - 1) // 2 - n * idash + n) * a * b
    ln = 1 + float(2 * k - n + 1) * a + float(-k * k + k + n * (n - 1) // 2) * a * b
    norm_a = max(max(l1, lid), ln)
    r = (1 + a) * (1 + b)
    d1 = (1 + a) * (1.0 / (1 + a) + (0.0 if r == 0.0 else b * (1 - _pow(r, n - 1)) / (1 - r)))
    dn = _pow(1 + a, n) * (1.0 / (1 + a))
    v = norm_a * max(d1, dn)
    return DBL_MAX if math.isinf(v) else v


def _bisect(f, lo: float, hi: float) -> float:
    """Bisection (higham_mat_impl.cpp:68-93): the upper end once the bracket
    is below (2|hi| + 1/2) eps or its midpoint stops moving."""
    if f(lo) > 0.0 or f(hi) < 0.0:
        return math.nan
    while True:
        if hi - lo < (2.0 * abs(hi) + 0.5) * EPS:
            break
        mid = (lo + hi) / 2
        if mid == lo or mid == hi:
            break
        fm = f(mid)
        if fm == 0.0:
            return mid
        if fm < 0.0:
            lo = mid
        else:
            hi = mid
    return hi


def beta(n: int, kappa: float, rho: float) -> float:
    """higham_mat_comp_beta (higham_mat_impl.cpp:95-128)."""
    def f(x):
        return cond_est(n, rho * x, x) - kappa
    lo = EPS
    hi = 1.0 / rho
    for _ in range(1, 100):
        v = f(hi)
        if math.isfinite(v) and v > 0.0:
            break
        hi *= 0.5
    bt = _bisect(f, lo, hi)
    while rho * bt > 1.0:
        hi *= 0.25
        bt = _bisect(f, lo, hi)
    return bt


def matrix(n: int, alpha: float, beta_: float) -> np.ndarray:
    """The FP64 matrix (hplai_matrix_impl, higham_mat_impl.cpp:130-141):
    a_ij = -beta + (alpha beta) i above, 1 + (alpha beta) i on, -alpha +
    (alpha beta) j below the diagonal, each a product then a sum."""
    ab = alpha * beta_
    i = np.arange(n, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    up = -beta_ + ab * i
    dg = 1.0 + ab * i
    lo = -alpha + ab * j
    A = np.where(i < j, up, np.where(i == j, dg, lo))
    return np.ascontiguousarray(A)


def factors(n: int, alpha: float, beta_: float):
    """The exact no-pivot LU factors: L unit lower with -alpha below the
    diagonal, U unit upper with -beta above it."""
    L = np.eye(n) + np.tril(np.full((n, n), -alpha), -1)
    U = np.eye(n) + np.triu(np.full((n, n), -beta_), 1)
    return L, U


def norm_inf(n: int, alpha: float, beta_: float) -> float:
    return float(np.abs(matrix(n, alpha, beta_)).sum(axis=1).max())
