"""ctypes bindings of the CPU oracle (``liboracle.so``) and of the reference's
own host routines compiled into ``_ref/libref.so``.

TEST INFRASTRUCTURE ONLY (see the package docstring).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")
REF_SRC = "/root/reference/src"

_lib = None
_ref = None

i64 = C.c_int64
u64 = C.c_uint64
dbl = C.c_double
flt = C.c_float
vp = C.c_void_p


def build(with_ref: bool = True) -> None:
    """Compile liboracle.so, and _ref/libref.so where the reference tree exists."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if with_ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_SRC={REF_SRC}"], check=True)


def _p(a):
    return a.ctypes.data_as(vp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
            "orc_lcg_jump": (None, [u64, C.POINTER(u64), C.POINTER(u64)]),
            "orc_lcg_init": (u64, [u64]),
            "orc_lcg_state": (u64, [u64, u64, u64, u64]),
            "orc_lcg_value": (dbl, [u64]),
            "orc_lcg_value_f32": (flt, [u64]),
            "orc_matgen_block": (None, [i64, u64, i64, i64, i64, i64, vp, i64]),
            "orc_diag_gpu_order": (None, [i64, u64, i64, i64, vp]),
            "orc_diag_seq_order": (None, [i64, u64, i64, i64, vp]),
            "orc_rhs": (None, [i64, u64, i64, i64, vp]),
            "orc_matgen_full": (None, [i64, u64, vp, i64, vp]),
            "orc_f32_to_f16_array": (None, [i64, vp, vp]),
            "orc_f16_to_f32_array": (None, [i64, vp, vp]),
            "orc_f32_to_e4m3_array": (None, [i64, vp, vp]),
            "orc_e4m3_to_f32_array": (None, [i64, vp, vp]),
            "orc_cast_f16": (None, [i64, i64, vp, i64, vp, i64]),
            "orc_pow2_scale_exp": (C.c_int, [flt, flt]),
            "orc_getrf_nopiv": (C.c_int, [i64, vp, i64]),
            "orc_trsm_runn": (None, [i64, i64, vp, i64, vp, i64]),
            "orc_trsm_llnu": (None, [i64, i64, vp, i64, vp, i64]),
            "orc_trsm_rlnu": (None, [i64, i64, vp, i64, vp, i64]),
            "orc_trtri_upper_nonunit": (C.c_int, [i64, vp, i64]),
            "orc_trtri_lower_unit": (None, [i64, vp, i64]),
            "orc_gemm_f16": (None, [i64, i64, i64, flt, vp, i64, vp, i64, flt, vp, i64]),
            "orc_gemm_f8": (None, [i64, i64, i64, flt, vp, i64, vp, vp, i64, vp, i64, flt, vp, i64]),
            "orc_lu_factor": (C.c_int, [i64, i64, C.c_int, vp, i64, vp, i64, vp]),
            "orc_otf_gemv": (None, [i64, u64, vp, dbl, vp, vp]),
            "orc_norm_inf": (dbl, [i64, vp]),
            "orc_harness": (dbl, [dbl, dbl, dbl, dbl, i64]),
            "orc_trsv_lower_unit": (None, [i64, i64, vp, i64, vp]),
            "orc_trsv_upper": (None, [i64, i64, vp, i64, vp]),
            "orc_ir_classic": (C.c_int, [i64, i64, u64, vp, i64, vp, vp, C.c_int, vp,
                                         C.POINTER(dbl), C.POINTER(dbl), vp]),
            "orc_gmres_ir": (C.c_int, [i64, i64, u64, vp, i64, vp, vp, C.c_int, C.c_int, dbl, vp,
                                       C.POINTER(dbl), C.POINTER(dbl), vp, C.POINTER(C.c_int)]),
            "orc_randsvd_dense": (None, [i64, u64, dbl, C.c_int, vp]),
            "orc_mlp_forward": (None, [i64, vp, vp, vp, vp, vp]),
            "orc_mlp_backward": (flt, [i64, vp, vp, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's routines compiled from its sources, or None."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        sig = {
            "ref_lcg_value": (dbl, [u64, C.c_int, u64, u64]),
            "ref_lcg_jump": (None, [u64, C.POINTER(u64), C.POINTER(u64)]),
            "ref_lcg_jump_fugaku": (None, [u64, C.POINTER(u64), C.POINTER(u64)]),
            "ref_fill_block": (None, [C.c_int] * 5 + [vp, C.c_size_t, u64]),
            "ref_fill_block_diag": (None, [C.c_int] * 3 + [vp, C.c_size_t, u64]),
            "ref_calc_diag": (dbl, [C.c_int, C.c_int, u64]),
            "ref_rhs": (None, [C.c_int, C.c_int, C.c_int, vp, u64]),
            "ref_f32_to_f16": (C.c_uint16, [flt]),
            "ref_f16_to_f32": (flt, [C.c_uint16]),
            "ref_shgemm": (None, [C.c_int, C.c_int, C.c_int, flt, vp, C.c_int, vp, C.c_int, flt, vp, C.c_int]),
            "ref_sgetrf_nopiv_tuned": (None, [C.c_int, vp, C.c_size_t]),
            "ref_sgetrf_nopiv_naive": (None, [C.c_int, C.c_int, vp, C.c_size_t]),
            "ref_higham_fhpl": (dbl, [C.c_int, dbl, dbl]),
            "higham_mat_comp_beta": (dbl, [C.c_int, dbl, dbl]),
            "hplai_matrix_impl": (None, [C.c_int, vp, C.c_int, dbl, dbl]),
        }
        for name, (res, args) in sig.items():
            f = getattr(R, name)
            f.restype = res
            f.argtypes = args
        _ref = R
    return _ref


# ---------------------------------------------------------------------------
# numpy-level wrappers
# ---------------------------------------------------------------------------
def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def lcg_jump(steps: int):
    a, c = u64(), u64()
    lib().orc_lcg_jump(steps, C.byref(a), C.byref(c))
    return a.value, c.value


def lcg_state(seed, n, i, j) -> int:
    return int(lib().orc_lcg_state(seed, n, i, j))


def lcg_value(state) -> float:
    return float(lib().orc_lcg_value(state))


def matgen_block(n, seed, i0, j0, m, k) -> np.ndarray:
    A = np.empty((k, m), dtype=np.float32)  # column-major m x k
    lib().orc_matgen_block(n, seed, i0, j0, m, k, _p(A), m)
    return A.T


def diag_gpu_order(n, seed, i0=0, m=None) -> np.ndarray:
    m = n - i0 if m is None else m
    d = np.empty(m, dtype=np.float64)
    lib().orc_diag_gpu_order(n, seed, i0, m, _p(d))
    return d


def diag_seq_order(n, seed, i0=0, m=None) -> np.ndarray:
    m = n - i0 if m is None else m
    d = np.empty(m, dtype=np.float64)
    lib().orc_diag_seq_order(n, seed, i0, m, _p(d))
    return d


def rhs(n, seed, i0=0, m=None) -> np.ndarray:
    m = n - i0 if m is None else m
    b = np.empty(m, dtype=np.float64)
    lib().orc_rhs(n, seed, i0, m, _p(b))
    return b


def matgen_full(n, seed):
    """(A column-major as an F-ordered float32 array, FP64 diagonal)."""
    A = np.empty((n, n), dtype=np.float32, order="F")
    d = np.empty(n, dtype=np.float64)
    lib().orc_matgen_full(n, seed, _p(A), n, _p(d))
    return A, d


def f32_to_f16_bits(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_f16_array(x.size, _p(x), _p(y))
    return y


def f16_bits_to_f32(h) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    y = np.empty(h.shape, dtype=np.float32)
    lib().orc_f16_to_f32_array(h.size, _p(h), _p(y))
    return y


def f32_to_e4m3_bits(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(x.shape, dtype=np.uint8)
    lib().orc_f32_to_e4m3_array(x.size, _p(x), _p(y))
    return y


def e4m3_bits_to_f32(q) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.uint8)
    y = np.empty(q.shape, dtype=np.float32)
    lib().orc_e4m3_to_f32_array(q.size, _p(q), _p(y))
    return y


def pow2_scale_exp(maxabs: float, target: float = 448.0) -> int:
    return int(lib().orc_pow2_scale_exp(maxabs, target))


def _fcopy32(A):
    return np.array(A, dtype=np.float32, order="F", copy=True)


def getrf_nopiv(A):
    A = _fcopy32(A)
    info = lib().orc_getrf_nopiv(A.shape[0], _p(A), A.shape[0])
    return A, int(info)


def trsm_runn(U, B):
    B = _fcopy32(B)
    U = _fcopy32(U)
    lib().orc_trsm_runn(B.shape[0], B.shape[1], _p(U), U.shape[0], _p(B), B.shape[0])
    return B


def trsm_llnu(L, B):
    B = _fcopy32(B)
    L = _fcopy32(L)
    lib().orc_trsm_llnu(B.shape[0], B.shape[1], _p(L), L.shape[0], _p(B), B.shape[0])
    return B


def trsm_rlnu(L, B):
    B = _fcopy32(B)
    L = _fcopy32(L)
    lib().orc_trsm_rlnu(B.shape[0], B.shape[1], _p(L), L.shape[0], _p(B), B.shape[0])
    return B


def trtri_upper_nonunit(A):
    A = _fcopy32(A)
    info = lib().orc_trtri_upper_nonunit(A.shape[0], _p(A), A.shape[0])
    return A, int(info)


def trtri_lower_unit(A):
    A = _fcopy32(A)
    lib().orc_trtri_lower_unit(A.shape[0], _p(A), A.shape[0])
    return A


def gemm_f16(alpha, X16, Y16, beta, Cm):
    """C = beta*C + alpha * X * Y^T; X (m x k), Y (n x k) as float16 bit arrays."""
    X16 = np.asfortranarray(X16, dtype=np.uint16)
    Y16 = np.asfortranarray(Y16, dtype=np.uint16)
    Cm = _fcopy32(Cm)
    m, k = X16.shape
    n = Y16.shape[0]
    lib().orc_gemm_f16(m, n, k, alpha, _p(X16), m, _p(Y16), n, beta, _p(Cm), m)
    return Cm


def gemm_f8(alpha, X8, xexp, Y8, yexp, tile, beta, Cm):
    X8 = np.asfortranarray(X8, dtype=np.uint8)
    Y8 = np.asfortranarray(Y8, dtype=np.uint8)
    xexp = np.asfortranarray(xexp, dtype=np.int32)
    yexp = np.asfortranarray(yexp, dtype=np.int32)
    Cm = _fcopy32(Cm)
    m, k = X8.shape
    n = Y8.shape[0]
    lib().orc_gemm_f8(m, n, k, alpha, _p(X8), m, _p(xexp), _p(Y8), n, _p(yexp), tile, beta, _p(Cm), m)
    return Cm


def lu_factor(A, b, alt=1, sched=None, fp8_tile=256, want_stats=False):
    A = _fcopy32(A)
    n = A.shape[0]
    nb = n // b
    s = None
    sp = None
    if sched is not None:
        s = np.ascontiguousarray(sched, dtype=np.int32)
        sp = _p(s)
    st = np.zeros(4 * max(nb, 1), dtype=np.float64)
    info = lib().orc_lu_factor(n, b, alt, _p(A), n, sp, fp8_tile, _p(st))
    if want_stats:
        return A, int(info), st.reshape(-1, 4)
    return A, int(info)


def otf_gemv(n, seed, diag, alpha, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64, copy=True)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    lib().orc_otf_gemv(n, seed, _p(diag), alpha, _p(x), _p(y))
    return y


def harness(normr, norma, normx, normb, n) -> float:
    return float(lib().orc_harness(normr, norma, normx, normb, n))


def trsv_lower_unit(LU, b, x):
    LU = np.asfortranarray(LU, dtype=np.float32)
    x = np.array(x, dtype=np.float64, copy=True)
    lib().orc_trsv_lower_unit(LU.shape[0], b, _p(LU), LU.shape[0], _p(x))
    return x


def trsv_upper(LU, b, x):
    LU = np.asfortranarray(LU, dtype=np.float32)
    x = np.array(x, dtype=np.float64, copy=True)
    lib().orc_trsv_upper(LU.shape[0], b, _p(LU), LU.shape[0], _p(x))
    return x


def ir_classic(LU, b, seed, diag, rhs_, maxit=50):
    LU = np.asfortranarray(LU, dtype=np.float32)
    n = LU.shape[0]
    x = np.empty(n, dtype=np.float64)
    r, h = dbl(), dbl()
    trace = np.zeros(2 * maxit, dtype=np.float64)
    it = lib().orc_ir_classic(n, b, seed, _p(LU), n, _p(np.ascontiguousarray(diag)),
                              _p(np.ascontiguousarray(rhs_)), maxit, _p(x), C.byref(r), C.byref(h), _p(trace))
    return dict(iters=int(it), x=x, resid=r.value, harness=h.value,
                trace=trace.reshape(-1, 2)[: (it + 1 if it >= 0 else maxit)])


def gmres_ir(LU, b, seed, diag, rhs_, maxit=20, restart=50, inner_tol=1e-10):
    LU = np.asfortranarray(LU, dtype=np.float32)
    n = LU.shape[0]
    x = np.empty(n, dtype=np.float64)
    r, h = dbl(), dbl()
    inner = C.c_int()
    trace = np.zeros(2 * maxit, dtype=np.float64)
    it = lib().orc_gmres_ir(n, b, seed, _p(LU), n, _p(np.ascontiguousarray(diag)),
                            _p(np.ascontiguousarray(rhs_)), maxit, restart, inner_tol, _p(x),
                            C.byref(r), C.byref(h), _p(trace), C.byref(inner))
    return dict(iters=int(it), x=x, resid=r.value, harness=h.value, inner=inner.value,
                trace=trace.reshape(-1, 2)[: (it + 1 if it >= 0 else maxit)])


def randsvd_dense(n, seed, kappa, nrefl=2):
    """FP64 dense randsvd matrix (n x n numpy, C order rows = matrix rows)."""
    A = np.empty((n, n), dtype=np.float64, order="F")
    lib().orc_randsvd_dense(n, seed, float(kappa), nrefl, _p(A))
    return A


MLP_NPARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3


def mlp_forward(params, X):
    """(H1, H2, Z) of the 16-64-64-3 MLP for a batch X (B x 16)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    B = X.shape[0]
    H1 = np.empty((B, 64), np.float32)
    H2 = np.empty((B, 64), np.float32)
    Z = np.empty((B, 3), np.float32)
    lib().orc_mlp_forward(B, _p(params), _p(X), _p(H1), _p(H2), _p(Z))
    return H1, H2, Z


def mlp_backward(params, X, y, t, m):
    """(mean loss, gradient) of the MLP loss."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    m = np.ascontiguousarray(m, dtype=np.float32)
    g = np.empty(MLP_NPARAMS, np.float32)
    loss = lib().orc_mlp_backward(X.shape[0], _p(params), _p(X), _p(y), _p(t), _p(m), _p(g))
    return float(loss), g


# ---------------------------------------------------------------------------
# independent pure-Python restatements (tiny cases only)
# ---------------------------------------------------------------------------
LCG_A = 6364136223846793005
LCG_C = 1
MASK64 = (1 << 64) - 1


def py_lcg_state(seed: int, n: int, i: int, j: int) -> int:
    """Closed form x_k = a^k x_0 + c (a^k - 1)/(a - 1) mod 2^64 with
    x_0 = a*seed + c and k = i + n*j, computed with Python integers."""
    k = i + n * j
    x0 = (LCG_A * seed + LCG_C) & MASK64
    mod = (LCG_A - 1) << 64
    ak = pow(LCG_A, k, mod)
    geo = ((ak - 1) % mod) // (LCG_A - 1)
    return (ak * x0 + LCG_C * geo) & MASK64


def py_lcg_value(state: int) -> float:
    s = state - (1 << 64) if state >= (1 << 63) else state
    return float(s) * float.fromhex("0x1.fffffffffffffp-65")
