"""Pins the oracle's dense kernels and LU driver: the reference's own
sgetrf_nopiv_tuned and shgemm (compiled from its sources), LAPACK through
scipy, and closed forms (exact FP16 products, L*U = A)."""
import numpy as np
import pytest
import scipy.linalg as sla


def _dd_matrix(n, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-0.5, 0.5, (n, n)).astype(np.float32)
    A[np.diag_indices(n)] = np.abs(A).sum(axis=1) + 1.0
    return np.asfortranarray(A)


def test_getrf_matches_reference_tuned_lu(orc, ref):
    n = 384
    A = _dd_matrix(n, 1)
    LU, info = orc.getrf_nopiv(A)
    assert info == 0
    R = np.array(A, order="F", copy=True)
    ref.ref_sgetrf_nopiv_tuned(n, R.ctypes.data, n)
    # same algorithm, different blocking: rounding-level differences only
    assert np.max(np.abs(LU - R) / (np.abs(R) + 1e-3)) < 1e-4


@pytest.mark.parametrize("n", [64, 300])
def test_getrf_matches_reference_naive_lu_bit_exact(orc, ref, n):
    """The oracle follows sgetrf_nopiv_naive's operation order (scale the
    column by the reciprocal pivot, then column-by-column updates, no fused
    multiply-add): identical bits."""
    A = _dd_matrix(n, 3)
    LU, info = orc.getrf_nopiv(A)
    assert info == 0
    R = np.array(A, order="F", copy=True)
    ref.ref_sgetrf_nopiv_naive(n, n, R.ctypes.data, n)
    assert np.array_equal(LU, R)


def test_getrf_matches_lapack_on_dominant_matrix(orc):
    """Partial pivoting never swaps rows of a column-dominant matrix, so LAPACK's
    getrf is the no-pivot LU; compare in FP64."""
    n = 256
    A = _dd_matrix(n, 2)
    A64 = np.asfortranarray(A.T.astype(np.float64))  # make it column-dominant
    lu, piv = sla.lu_factor(A64)
    assert np.all(piv == np.arange(n))
    LU, info = orc.getrf_nopiv(A64.astype(np.float32))
    assert np.max(np.abs(LU - lu)) < 2e-5 * np.max(np.abs(lu))


def test_getrf_reports_zero_pivot(orc):
    A = np.asfortranarray(np.eye(4, dtype=np.float32))
    A[2, 2] = 0.0
    _, info = orc.getrf_nopiv(A)
    assert info == 3


def test_trsm_variants_match_scipy(orc):
    rng = np.random.default_rng(3)
    b, m = 128, 300
    U = np.triu(rng.uniform(-1, 1, (b, b))) + np.eye(b) * 8
    L = np.tril(rng.uniform(-1, 1, (b, b)), -1) / 4 + np.eye(b)
    B1 = rng.uniform(-1, 1, (m, b))
    X = orc.trsm_runn(U.astype(np.float32), B1.astype(np.float32))  # X U = B
    Xs = sla.solve_triangular(U.T, B1.T, lower=True).T
    assert np.max(np.abs(X - Xs)) < 1e-5
    B2 = rng.uniform(-1, 1, (b, m))
    Y = orc.trsm_llnu(L.astype(np.float32), B2.astype(np.float32))  # L Y = B
    Ys = sla.solve_triangular(L, B2, lower=True, unit_diagonal=True)
    assert np.max(np.abs(Y - Ys)) < 1e-5
    Z = orc.trsm_rlnu(L.astype(np.float32), B1.astype(np.float32))  # Z L = B
    Zs = sla.solve_triangular(L.T, B1.T, lower=False, unit_diagonal=True).T
    assert np.max(np.abs(Z - Zs)) < 1e-5


def test_trtri_matches_inverse(orc):
    rng = np.random.default_rng(4)
    b = 96
    M = rng.uniform(-1, 1, (b, b)).astype(np.float32)
    M[np.diag_indices(b)] += 6
    Ui, info = orc.trtri_upper_nonunit(M)
    assert info == 0
    U = np.triu(M).astype(np.float64)
    assert np.max(np.abs(np.triu(Ui) - np.linalg.inv(U))) < 1e-6
    assert np.array_equal(np.tril(Ui, -1), np.tril(M, -1))  # lower part untouched
    Li = orc.trtri_lower_unit(M)
    L = np.tril(M, -1).astype(np.float64) / 1.0 + np.eye(b)
    Lr = np.tril(M.astype(np.float64), -1) + np.eye(b)
    inv = np.linalg.inv(Lr)
    assert np.max(np.abs(np.tril(Li, -1) - np.tril(inv, -1))) < 1e-3 * np.max(np.abs(inv))
    assert np.array_equal(np.triu(Li), np.triu(M))  # upper part untouched
    assert L.shape == Lr.shape


@pytest.mark.parametrize("beta", [0.0, 0.5, -2.0])
def test_gemm_f16_matches_reference_shgemm_bitwise(orc, ref, beta):
    rng = np.random.default_rng(5)
    m, n, k = 70, 45, 96
    X = rng.uniform(-2, 2, (m, k)).astype(np.float16)
    Y = rng.uniform(-2, 2, (n, k)).astype(np.float16)
    Cm = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    got = orc.gemm_f16(-1.0, X.view(np.uint16), Y.view(np.uint16), beta, Cm)
    # shgemm(m, n, k, alpha, a (m x k, lda m), b (k x n, ldb k), beta, c)
    a = np.asfortranarray(X).view(np.uint16)
    bt = np.asfortranarray(Y.T).view(np.uint16)
    c = np.array(Cm, order="F", copy=True)
    ref.ref_shgemm(m, n, k, -1.0, a.ctypes.data, m, bt.ctypes.data, k, beta, c.ctypes.data, m)
    assert np.array_equal(got, c)


def test_gemm_f16_within_fp32_accumulation_bound(orc):
    rng = np.random.default_rng(6)
    m, n, k = 128, 64, 2560
    X = (rng.standard_normal((m, k)) * 1e-3).astype(np.float16)
    Y = rng.uniform(-0.5, 0.5, (n, k)).astype(np.float16)
    Cm = rng.uniform(-0.5, 0.5, (m, n)).astype(np.float32)
    got = orc.gemm_f16(-1.0, X.view(np.uint16), Y.view(np.uint16), 1.0, Cm).astype(np.float64)
    exact = Cm.astype(np.float64) - X.astype(np.float64) @ Y.astype(np.float64).T
    bound = k * 2.0**-24 * (np.abs(X.astype(np.float64)) @ np.abs(Y.astype(np.float64)).T) \
        + 2.0**-24 * np.abs(exact)
    assert np.all(np.abs(got - exact) <= bound * 1.01)


def test_gemm_f8_scaled_is_exact_products(orc):
    rng = np.random.default_rng(7)
    m, n, k, tile = 300, 260, 512, 256
    xf = (rng.standard_normal((m, k)) * 1e-5).astype(np.float32)
    yf = rng.uniform(-0.5, 0.5, (n, k)).astype(np.float32)

    def quant(a):
        r, c = a.shape
        rt, ct = -(-r // tile), -(-c // tile)
        e = np.zeros((rt, ct), dtype=np.int32)
        q = np.zeros_like(a, dtype=np.uint8)
        for i in range(rt):
            for j in range(ct):
                blk = a[i * tile:(i + 1) * tile, j * tile:(j + 1) * tile]
                e[i, j] = orc.pow2_scale_exp(float(np.abs(blk).max()))
                q[i * tile:(i + 1) * tile, j * tile:(j + 1) * tile] = orc.f32_to_e4m3_bits(
                    np.ldexp(blk, e[i, j]).astype(np.float32))
        return q, e

    qx, ex = quant(xf)
    qy, ey = quant(yf)
    Cm = np.zeros((m, n), dtype=np.float32)
    got = orc.gemm_f8(-1.0, qx, ex, qy, ey, tile, 1.0, Cm).astype(np.float64)
    dx = orc.e4m3_bits_to_f32(qx).astype(np.float64) * np.ldexp(1.0, -np.repeat(np.repeat(ex, tile, 0), tile, 1)[:m, :k])
    dy = orc.e4m3_bits_to_f32(qy).astype(np.float64) * np.ldexp(1.0, -np.repeat(np.repeat(ey, tile, 0), tile, 1)[:n, :k])
    exact = -(dx @ dy.T)
    bound = k * 2.0**-24 * (np.abs(dx) @ np.abs(dy).T)
    assert np.all(np.abs(got - exact) <= bound)
    # quantisation error itself is ~2^-4 relative per element, far above FP32
    rel = np.abs(dx - xf) / np.abs(xf).max()
    assert rel.max() < 2.0**-4


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_lu_factor_reconstructs_matrix(orc, alt):
    n, b = 1024, 256
    A, d = orc.matgen_full(n, 42)
    LU, info = orc.lu_factor(A, b, alt=alt)
    assert info == 0
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    R = L @ U - A.astype(np.float64)
    # FP16 Schur inputs: relative error ~ u16 * |L||U| per element
    bound = 2.0**-9 * (np.abs(L) @ np.abs(U))
    assert np.max(np.abs(R) / bound) < 1.0
    # a dropped Schur term would leave errors ~|l||r| >> the bound on the off-diagonal blocks
    LUs = orc.lu_factor(A, b, alt=alt)[0]
    assert np.array_equal(LU, LUs)  # deterministic


def test_lu_factor_alt_variants_agree(orc):
    n, b = 768, 256
    A, _ = orc.matgen_full(n, 5)
    f0 = orc.lu_factor(A, b, alt=0)[0].astype(np.float64)
    f1 = orc.lu_factor(A, b, alt=1)[0].astype(np.float64)
    f2 = orc.lu_factor(A, b, alt=2)[0].astype(np.float64)
    low = np.tril(np.ones((n, n)), -1).astype(bool)
    lmax = np.abs(f0[low]).max()
    # alt 1/2 form L and U panels with FP16 inputs: a few FP16 ulps apart from
    # the FP32 TRSM of alt 0; alt 1 and 2 differ only by the inverse's rounding
    for fa, fb, tol in [(f0, f1, 2.0**-8), (f1, f2, 2.0**-12)]:
        d = np.abs(fa - fb)
        assert np.max(d[low]) <= tol * lmax
        assert np.max(d[~low] / (np.abs(fa[~low]) + 0.05)) <= tol


def test_lu_factor_fp8_schedule(orc):
    n, b = 1024, 256
    A, _ = orc.matgen_full(n, 9)
    f16 = orc.lu_factor(A, b, alt=1)[0].astype(np.float64)
    f8 = orc.lu_factor(A, b, alt=1, sched=[1, 1, 1, 1])[0].astype(np.float64)
    L = np.tril(f8, -1) + np.eye(n)
    U = np.triu(f8)
    R = L @ U - A.astype(np.float64)
    assert np.max(np.abs(R)) < 0.05  # FP8 Schur updates: larger but bounded error
    assert not np.array_equal(f16, f8)


def test_otf_gemv_matches_dense_product(orc):
    n = 1024
    A, d = orc.matgen_full(n, 42)
    # exact FP64 matrix: off-diagonal LCG values in FP64, diagonal = d
    A64 = np.array([[orc.lcg_value(orc.lcg_state(42, n, i, j)) for j in range(n)] for i in range(0, n, 97)])
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    y = orc.otf_gemv(n, 42, d, -1.0, x, y0)
    for r, i in enumerate(range(0, n, 97)):
        row = A64[r].copy()
        row[i] = d[i]
        exact = y0[i] - np.dot(row, x)
        assert abs(y[i] - exact) < 1e-13 * (np.abs(row) @ np.abs(x))


def test_trsv_matches_scipy(orc):
    n, b = 1024, 256
    A, d = orc.matgen_full(n, 42)
    LU, _ = orc.lu_factor(A, b, alt=1)
    rng = np.random.default_rng(9)
    r = rng.uniform(-1, 1, n)
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    y = orc.trsv_lower_unit(LU, b, r)
    ys = sla.solve_triangular(L, r, lower=True, unit_diagonal=True)
    assert np.max(np.abs(y - ys)) < 1e-13 * np.max(np.abs(ys))
    z = orc.trsv_upper(LU, b, y)
    zs = sla.solve_triangular(U, y, lower=False)
    assert np.max(np.abs(z - zs)) < 1e-12 * np.max(np.abs(zs))


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_lu_factor_recovers_known_factors_and_catches_a_dropped_block(orc, alt):
    """A = L0 U0 with well-conditioned, coherently signed factors: the LU must
    return L0, U0 to rounding (< 1e-3 on U), and a planted defect -- one
    trailing block missing the step-0 update -- moves U by > 0.05."""
    n, b = 1280, 256
    rng = np.random.default_rng(alt + 10)
    L0 = np.tril(rng.uniform(0, 1, (n, n)) / n, -1) + np.eye(n)
    U0 = np.triu(rng.uniform(0, 1, (n, n)), 1)
    U0 += np.diag(1.0 + U0.sum(axis=1))
    A64 = L0 @ U0
    LU, info = orc.lu_factor(A64.astype(np.float32), b, alt=alt)
    assert info == 0
    Lg = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    Ug = np.triu(LU.astype(np.float64))
    assert np.abs(Lg - L0).max() < 2e-5
    assert np.abs(Ug - U0).max() < 5e-3
    Ab = A64.copy()
    Ab[512:768, 512:768] += L0[512:768, 0:256] @ U0[0:256, 512:768]
    LUb, _ = orc.lu_factor(Ab.astype(np.float32), b, alt=alt)
    assert np.abs(np.triu(LUb.astype(np.float64)) - U0).max() > 0.05


def test_lu_factor_fp32_class_recovers_factors_to_fp32(orc):
    """Schedule value -1 (the FP32 Schur class): panels by FP32 TRSM, Schur
    products of the unrounded FP32 panels.  On A = L0 U0 the unit lower
    factor comes back to FP32 accuracy, more than 50 times closer than with
    the FP16 class (which rounds every Schur operand to 11 bits), and a mixed
    schedule sits in between; U's error is the rounding of A to FP32 either
    way."""
    n, b = 1280, 256
    rng = np.random.default_rng(77)
    L0 = np.tril(rng.uniform(0, 1, (n, n)) / n, -1) + np.eye(n)
    U0 = np.triu(rng.uniform(0, 1, (n, n)), 1)
    U0 += np.diag(1.0 + U0.sum(axis=1))
    A = (L0 @ U0).astype(np.float32)
    errs = {}
    for name, sched in (("fp32", [-1] * 5), ("fp16", None), ("mixed", [-1, 0, -1, 0, 0])):
        for alt in (0, 2):
            LU, info = orc.lu_factor(A, b, alt=alt, sched=sched)
            assert info == 0
            Ug = np.triu(LU.astype(np.float64))
            Lg = np.tril(LU.astype(np.float64), -1) + np.eye(n)
            errs[name, alt] = (np.abs(Ug - U0).max() / np.abs(U0).max(), np.abs(Lg - L0).max())
    for alt in (0, 2):
        eu32, el32 = errs["fp32", alt]
        eu16, el16 = errs["fp16", alt]
        _, elm = errs["mixed", alt]
        assert eu32 < 4e-6 and eu16 < 4e-6 and el32 < 1e-7, errs
        assert el16 > 50 * el32, errs
        assert el32 < elm < el16 * 1.01, errs


def test_lu_factor_fp32_class_resolves_what_fp16_cannot(orc):
    """randsvd with kappa_2 = 1e4 (8 reflectors per factor): the FP16 Schur
    class perturbs the LU by ~u16 |L||U| ~ 5e-4, more than the smallest
    singular value 1e-4, so its factors are no preconditioner; the FP32 class
    keeps the backward error two orders below sigma_min."""
    n, b = 1024, 256
    A64 = orc.randsvd_dense(n, 42, 1e4, 8)
    A = A64.astype(np.float32)
    res = {}
    for name, sched in (("fp32", [-1] * 4), ("fp16", [0] * 4)):
        LU, info = orc.lu_factor(A, b, alt=1, sched=sched)
        F = LU.astype(np.float64)
        L = np.tril(F, -1) + np.eye(n)
        U = np.triu(F)
        res[name] = np.abs(L @ U - A64).max() if info == 0 else np.inf
    smin = np.linalg.svd(A64, compute_uv=False)[-1]
    assert res["fp32"] < 1e-2 * smin, (res, smin)
    assert res["fp16"] > smin, (res, smin)
