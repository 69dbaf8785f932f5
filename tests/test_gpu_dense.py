"""GPU parity of the FP32 panel kernels (getrf, the three TRSM variants,
SGEMM, identity, block copy) against the oracle and numpy."""
import numpy as np
import pytest

from tests.gpu_util import colmajor_to_numpy, dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def _fdev(M, ld=None):
    m, n = M.shape
    ld = m if ld is None else ld
    buf = np.zeros((n, ld), dtype=np.float32)
    buf[:, :m] = M.T
    return dev(buf.ravel()), ld


def _dd(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-0.5, 0.5, (n, n)).astype(np.float32)
    A[np.diag_indices(n)] = np.abs(A).sum(axis=1)
    return A


def test_sgemm_matches_numpy():
    L = _L()
    rng = np.random.default_rng(0)
    m, n, k = 200, 130, 64
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    Ad, lda = _fdev(A, m + 3)
    Bd, ldb = _fdev(B, k + 5)
    Cd, ldc = _fdev(C, m + 1)
    assert L.omx_sgemm_f32(m, n, k, -1.0, ptr(Ad), lda, ptr(Bd), ldb, 1.0, ptr(Cd), ldc, None) == 0
    got = colmajor_to_numpy(Cd, m, n, ldc)
    exact = C.astype(np.float64) - A.astype(np.float64) @ B.astype(np.float64)
    bound = (k + 2) * 2.0**-24 * (np.abs(C) + np.abs(A).astype(np.float64) @ np.abs(B))
    assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize("m,n,k,beta", [(2100, 2050, 80, 1.0), (2176, 1930, 208, 0.0), (4100, 1030, 16, 0.5),
                                        (600, 520, 64, 1.0), (4200, 4100, 64, 1.0), (4300, 4000, 96, 0.0)])
def test_sgemm_large_matches_numpy(m, n, k, beta):
    """Products of at least 512 x 512 entries with 16-byte aligned A and B
    run on the matrix cores, in 64-tiles below 1024 tiles of 128 and in
    128-tiles from there (the last two cases): ragged edges in both
    dimensions, k in stages of 64 with a shorter last stage (or a single
    stage), and beta = 0 with C full of NaN (C must not be read)."""
    torch_cuda()
    print(f"sgemm large m={m} n={n} k={k} beta={beta}", flush=True)
    L = _L()
    rng = np.random.default_rng(m + k)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32) if beta else np.full((m, n), np.nan, np.float32)
    Ad, lda = _fdev(A, m + 4)
    Bd, ldb = _fdev(B, k + 4)
    Cd, ldc = _fdev(C, m + 1)
    assert L.omx_sgemm_f32(m, n, k, -1.0, ptr(Ad), lda, ptr(Bd), ldb, beta, ptr(Cd), ldc, None) == 0
    got = colmajor_to_numpy(Cd, m, n, ldc)
    ab = A.astype(np.float64) @ B.astype(np.float64)
    c0 = C.astype(np.float64) if beta else np.zeros((m, n))
    exact = beta * c0 - ab
    bound = (k + 2) * 2.0**-24 * (np.abs(beta * c0) + np.abs(A).astype(np.float64) @ np.abs(B))
    assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize("n", [256, 512, 2560])
def test_getrf_matches_oracle(orc, n):
    torch = torch_cuda()
    print(f"getrf n={n}", flush=True)
    L = _L()
    A = _dd(n, n)
    Ad, lda = _fdev(A, n + 64)
    info = empty(4, torch.int32)
    assert L.omx_getrf_nopiv_f32(n, ptr(Ad), lda, ptr(info), None) == 0
    got = colmajor_to_numpy(Ad, n, n, lda).astype(np.float64)
    assert host(info)[0] == 2**31 - 1
    want, _ = orc.getrf_nopiv(A)
    scale = np.abs(want.astype(np.float64)) + 1e-3 * np.abs(want).max() / n
    assert np.max(np.abs(got - want) / scale) < 1e-3
    Lf = np.tril(got, -1) + np.eye(n)
    Uf = np.triu(got)
    res = np.abs(Lf @ Uf - A.astype(np.float64))
    assert np.max(res / (np.abs(Lf) @ np.abs(Uf))) < n * 2.0**-23


def test_getrf_zero_pivot_info():
    torch = torch_cuda()
    L = _L()
    n = 128
    A = np.eye(n, dtype=np.float32)
    A[70, 70] = 0.0
    Ad, lda = _fdev(A)
    info = empty(4, torch.int32)
    assert L.omx_getrf_nopiv_f32(n, ptr(Ad), lda, ptr(info), None) == 0
    assert host(info)[0] == 71


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_trsm_matches_oracle(orc, variant):
    L = _L()
    rng = np.random.default_rng(variant)
    b, m = 256, 700
    if variant == 0:
        T = (np.triu(rng.uniform(-1, 1, (b, b))) + np.eye(b) * 16).astype(np.float32)
        B = rng.uniform(-1, 1, (m, b)).astype(np.float32)
        want = orc.trsm_runn(T, B)
    elif variant == 1:
        T = (np.tril(rng.uniform(-1, 1, (b, b)), -1) / 8 + np.eye(b)).astype(np.float32)
        B = rng.uniform(-1, 1, (b, m)).astype(np.float32)
        want = orc.trsm_llnu(T, B)
    else:
        T = (np.tril(rng.uniform(-1, 1, (b, b)), -1) / 8 + np.eye(b)).astype(np.float32)
        B = rng.uniform(-1, 1, (m, b)).astype(np.float32)
        want = orc.trsm_rlnu(T, B)
    Td, ldt = _fdev(T)
    Bd, ldb = _fdev(B, B.shape[0] + 2)
    assert L.omx_trsm_f32(variant, B.shape[0], B.shape[1], ptr(Td), ldt, ptr(Bd), ldb, None) == 0
    got = colmajor_to_numpy(Bd, B.shape[0], B.shape[1], ldb)
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.abs(want).max())


def test_identity_and_copy():
    torch = torch_cuda()
    L = _L()
    n = 300
    I = empty(n * (n + 4), torch.float32)
    assert L.omx_identity_f32(n, ptr(I), n + 4, None) == 0
    assert np.array_equal(colmajor_to_numpy(I, n, n, n + 4), np.eye(n, dtype=np.float32))
    rng = np.random.default_rng(1)
    M = rng.uniform(-1, 1, (50, 40)).astype(np.float32)
    Md, ldm = _fdev(M, 60)
    O = empty(40 * 50, torch.float32)
    assert L.omx_copy_block_f32(50, 40, ptr(Md), ldm, ptr(O), 50, None) == 0
    assert np.array_equal(colmajor_to_numpy(O, 50, 40), M)


@pytest.mark.parametrize("variant,n", [(0, 256), (0, 640), (2, 256), (2, 640)])
def test_trsm_identity_matches_identity_then_trsm(variant, n):
    """The triangular inverse that solves only the rows that can be
    non-zero equals the identity followed by the full TRSM, bit for bit
    (numpy compares -0 and +0 as equal: the skipped rows hold the identity's
    zeros where the full solve may have formed -0); the product with the
    triangle is the identity to FP32 accuracy."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(n + variant)
    if variant == 0:
        T = (np.triu(rng.uniform(-1, 1, (n, n))) + np.eye(n) * 8).astype(np.float32)
    else:
        T = (np.tril(rng.uniform(-1, 1, (n, n)), -1) / 8 + np.eye(n)).astype(np.float32)
    Td, ldt = _fdev(T, n + 8)
    Xa = empty(n * n, torch.float32)
    Xb = empty(n * n, torch.float32)
    assert L.omx_trsm_identity_f32(variant, n, ptr(Td), ldt, ptr(Xa), n, None) == 0
    assert L.omx_identity_f32(n, ptr(Xb), n, None) == 0
    assert L.omx_trsm_f32(variant, n, n, ptr(Td), ldt, ptr(Xb), n, None) == 0
    a = colmajor_to_numpy(Xa, n, n, n)
    assert np.array_equal(a, colmajor_to_numpy(Xb, n, n, n))
    assert np.array_equal(a, np.triu(a) if variant == 0 else np.tril(a))
    Tt = T.astype(np.float64) if variant == 0 else np.tril(T.astype(np.float64), -1) + np.eye(n)
    assert np.max(np.abs(a.astype(np.float64) @ Tt - np.eye(n))) < 1e-4


def test_trsm_identity_argument_errors():
    torch = torch_cuda()
    L = _L()
    t = empty(128 * 128, torch.float32)
    x = empty(128 * 128, torch.float32)
    assert L.omx_trsm_identity_f32(1, 128, ptr(t), 128, ptr(x), 128, None) < 0   # LLNU not offered
    assert L.omx_trsm_identity_f32(0, 100, ptr(t), 128, ptr(x), 128, None) < 0   # n % 64
    assert L.omx_trsm_identity_f32(0, 128, ptr(t), 64, ptr(x), 128, None) < 0    # ldt < n


@pytest.mark.parametrize("n", [64, 320, 640, 2560])
def test_trtri_lu_inverts_both_triangles(n):
    """U^-1 and L^-1 of a factored block by diagonal-block substitution and
    level-by-level merges (n / 64 = 1, 5, 10, 40 blocks: odd counts leave a
    range for a later level): against FP64 inverses of the same FP32
    triangles, the other triangle exactly zero (outputs and workspace start
    as NaN), L^-1's diagonal exactly 1."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(n)
    Lt = np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) + np.eye(n)
    Ut = np.triu(rng.uniform(-1, 1, (n, n)) / 4, 1)
    Ut += np.diag(1.0 + np.abs(Ut).sum(axis=1))
    D = (np.tril(Lt, -1) + Ut).astype(np.float32)
    Dd, ldd = _fdev(D, n + 16)
    # every entry of the outputs must be written (zeros included): start from NaN
    Xu = empty(n * n, torch.float32).fill_(float("nan"))
    Xl = empty(n * n, torch.float32).fill_(float("nan"))
    W = empty(n * n, torch.float32).fill_(float("nan"))
    assert L.omx_trtri_lu_f32(n, ptr(Dd), ldd, ptr(Xu), ptr(Xl), n, ptr(W), None) == 0
    xu = colmajor_to_numpy(Xu, n, n).astype(np.float64)
    xl = colmajor_to_numpy(Xl, n, n).astype(np.float64)
    U32 = np.triu(D).astype(np.float64)
    L32 = np.tril(D, -1).astype(np.float64) + np.eye(n)
    iu = np.linalg.inv(U32)
    il = np.linalg.inv(L32)
    print(f"  n={n}: |Xu - U^-1| {np.abs(xu - iu).max():.2e}, |Xl - L^-1| {np.abs(xl - il).max():.2e}", flush=True)
    assert np.abs(xu - iu).max() <= 1e-5 * np.abs(iu).max()
    assert np.abs(xl - il).max() <= 1e-5 * np.abs(il).max()
    assert not np.any(np.tril(xu, -1)) and not np.any(np.triu(xl, 1))
    assert np.all(np.diag(xl) == 1.0)


def test_trtri_lu_argument_errors():
    torch = torch_cuda()
    L = _L()
    a = empty(128 * 128, torch.float32)
    assert L.omx_trtri_lu_f32(100, ptr(a), 128, ptr(a), ptr(a), 128, ptr(a), None) < 0   # n % 64
    assert L.omx_trtri_lu_f32(128, ptr(a), 64, ptr(a), ptr(a), 128, ptr(a), None) < 0    # ldd < n
    assert L.omx_trtri_lu_f32(128, ptr(a), 128, ptr(a), ptr(a), 128, None, None) < 0     # no work
