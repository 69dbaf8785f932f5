"""The headline block size b = 2560 on the GPU.

- LU parity against the oracle's reference-order LU (panel_trf_gdirect.hpp:
  593-1192) at b = 2560: n = 12800 (5 panels, pairs: the trailing update is a
  product of depth 2b = 5120 over the pair buffer's two K windows) and
  n = 23040 (9 panels, groups of four: depth 4b = 10240, 40 FP8 scale tiles
  per 256-row strip), FP16 and FP8 schedules, -alt 0/1/2, and the FP32
  Schur class (all steps, and mixed with FP8 and FP16 steps).
- The windowed Schur GEMMs at the bench's depths (FP16 K = 5120, FP8
  K = 10240) against the exact product of their rounded operands.
- Whole solves at N = 61440 and N = 204800 (b = 2560): harness < 16 and the
  forward error against a further FP64 refinement to the floor.
"""
import ctypes as C

import numpy as np
import pytest

from tests.gpu_util import dev, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu

B = 2560


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def _solver(n, b, alt=1, fp8=False, group4=False):
    h = C.c_void_p()
    flags = (1 if fp8 else 0) | (4 if group4 else 0)
    assert _L().omx_solver_create(n, b, alt, flags, C.byref(h)) == 0
    return h


def _factor_vs_oracle(orc, n, alt, group4, sched):
    torch = torch_cuda()
    L = _L()
    seed = 42
    print(f"lu n={n} b={B} alt={alt} group={'4' if group4 else '2'} sched={sched}", flush=True)
    A = empty(n * n, torch.float32)
    d = empty(n, torch.float64)
    assert L.omx_matgen_f32(n, seed, ptr(A), n, ptr(d), None) == 0
    h = _solver(n, B, alt, fp8=sched is not None and any(sched), group4=group4)
    tim = np.zeros(5)
    sp = None
    if sched is not None:
        s = np.ascontiguousarray(sched, dtype=np.int32)
        sp = s.ctypes.data
    info = L.omx_lu_factor(h, ptr(A), n, sp, tim.ctypes.data)
    used = np.zeros(n // B, np.int32)
    assert L.omx_solver_last_schedule(h, used.ctypes.data, None) == 0
    L.omx_solver_destroy(h)
    assert info == 0
    if sched is not None:
        assert np.array_equal(used[:-1], np.asarray(sched)[:-1])
    got = host(A).reshape(n, n).T  # FP32 (row, col)
    del A
    Ao, _ = orc.matgen_full(n, seed)
    want, oinfo = orc.lu_factor(Ao, B, alt=alt, sched=sched)
    assert oinfo == 0
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    diff = np.abs(got - want)
    lmax = float(np.abs(want[low]).max())
    dl = float(diff[low].max())
    du = float(np.max(diff[~low] / (np.abs(want[~low]) + np.float32(0.05))))
    print(f"  max |dL|/max|L| = {dl / lmax:.3e}, max |dU|/(|U|+.05) = {du:.3e}, "
          f"trailing launches {int(tim[2])}", flush=True)
    # FP16 Schur inputs: a few FP16 ulps of the panel entries; FP8 steps: the
    # same against the oracle's identical FP8 rounding (the two sides differ
    # only in FP32 accumulation order, then in the roundings that follow)
    tol = 2.0**-8 if sched is None or not any(sched) else 2.0**-6
    assert dl <= tol * lmax
    assert du <= tol
    # backward error |A - L U| / (|L||U|) on 64 sampled rows: the GPU
    # factors against the oracle's (the reference's method has this error
    # too: the -alt products multiply FP16-rounded panels by FP16-rounded
    # inverses of the diagonal blocks, so it grows with their condition)
    rng = np.random.default_rng(n)
    rows = np.sort(rng.choice(n, 64, replace=False))

    def backward(F):
        F64 = F.astype(np.float64)
        # row r of the unit lower factor: columns j < r, 1 at j = r
        Lr = np.where(np.arange(n)[None, :] < rows[:, None], F64[rows], 0.0)
        Lr[np.arange(64), rows] = 1.0
        Uf = np.triu(F64)
        R = Lr @ Uf - Ao[rows].astype(np.float64)
        S = np.abs(Lr) @ np.abs(Uf)
        return float(np.max(np.abs(R) / S))

    bg, bo = backward(got), backward(want)
    print(f"  sampled backward error / (|L||U|): GPU {bg:.3e}, oracle {bo:.3e}", flush=True)
    assert bg <= 1.5 * bo + 2.0**-20
    return tim


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_lu_b2560_pairs_fp16(orc, alt):
    """n = 12800: pair 0's trailing update is A[4b:, 4b:] -= [L0 L1][R0 R1]^T
    (depth 5120); the look-ahead products cover blocks 2, 3."""
    tim = _factor_vs_oracle(orc, 12800, alt, False, None)
    assert int(tim[2]) == 1


def test_lu_b2560_pairs_mixed_fp8(orc):
    """A pair with one FP8 and one FP16 panel (two products over one buffer's
    K windows) and an all-FP8 pair."""
    _factor_vs_oracle(orc, 12800, 1, False, [1, 0, 1, 1, 0])


def test_lu_b2560_groups4_fp16(orc):
    """n = 23040, groups of four: group 0's trailing update has depth
    4b = 10240 over A[8b:, 8b:]."""
    tim = _factor_vs_oracle(orc, 23040, 1, True, None)
    assert int(tim[2]) == 1


def test_lu_b2560_groups4_fp8(orc):
    """The learned solver's configuration (FP8 in groups of four, -alt 2
    masks) with one FP16 panel inside group 1."""
    _factor_vs_oracle(orc, 23040, 2, True, [1, 1, 1, 1, 1, 0, 1, 1, 0])


def test_lu_b2560_groups4_fp32_class(orc):
    """Every step in the FP32 class (schedule -1) at b = 2560 in groups of
    four: each group's trailing update is one FP32 product of depth 4b =
    10240 straight from the FP32 panels in A; the factors match the
    oracle's FP32-class LU."""
    _factor_vs_oracle(orc, 23040, 1, True, [-1] * 9)


def test_lu_b2560_pairs_three_classes(orc):
    """FP8, FP32 and FP16 steps in one factorisation (pairs at b = 2560):
    a pair of one FP32 and one FP16 panel is two products, one per class."""
    _factor_vs_oracle(orc, 12800, 1, False, [1, -1, 0, -1, 0])


def _pack_window(L, S, fp8, P, E, ldk, k0, tmax):
    """Pack the FP32 column panel S (R x b) into K window [k0, k0 + b) of a
    ldk-column packed buffer."""
    R, b = S.shape
    sd = dev(np.asfortranarray(S).ravel(order="F"))
    if fp8:
        rc = L.omx_pack_f8_window(R, b, ptr(sd), 1, R, 0, 0, ptr(P), ptr(E), ldk, k0, ptr(tmax), None)
    else:
        rc = L.omx_pack_f16_window(R, b, ptr(sd), 1, R, 0, ptr(P), None, None, ldk, k0, None)
    assert rc == 0
    return sd


def _decode(orc, S, fp8, e_dev, k0, ldk):
    """The GEMM's operand values for panel S in window k0: FP16 rounding, or
    e4m3 with the pack's per-tile power-of-two exponents."""
    if not fp8:
        return S.astype(np.float16).astype(np.float64)
    R, b = S.shape
    nt = ldk // 256
    e = host(e_dev).reshape(-(-R // 256), nt)[:, k0 // 256:(k0 + b) // 256]
    E = np.repeat(np.repeat(e, 256, 0), 256, 1)[:R]
    # exponents: the largest e with max|tile| 2^e <= 448 (pow2 scale to e4m3 range)
    for s0 in range(0, R, 256):
        for t in range(0, b, 256):
            blk = S[s0:s0 + 256, t:t + 256]
            assert E[s0, t] == orc.pow2_scale_exp(float(np.abs(blk).max()))
    q = orc.f32_to_e4m3_bits(np.ldexp(S, E).astype(np.float32))
    return orc.e4m3_bits_to_f32(q).astype(np.float64) * np.ldexp(1.0, -E)


@pytest.mark.parametrize("fp8", [False, True])
def test_gemm_b2560_group_window(orc, fp8):
    """The trailing update as the solvers launch it at b = 2560: FP16 pairs
    (ldk = 2b = 5120) and FP8 groups of four (ldk = 4b = 10240, 40 scale
    tiles per strip), m = n = 3072 rows counted from packed row 512 (the
    group's next blocks), against C - sum_i X_i Y_i^T of the rounded
    operands within the FP32 accumulation bound."""
    torch = torch_cuda()
    L = _L()
    G = 4 if fp8 else 2
    ldk = G * B
    shift, m, n = 512, 3072, 3072
    R = shift + max(m, n)
    print(f"gemm window fp8={fp8} m={m} n={n} ldk={ldk}", flush=True)
    rng = np.random.default_rng(40 + G)
    fp = 1 if fp8 else 0
    PX = empty(L.omx_panel_bytes(R, ldk, fp), torch.uint8)
    PY = empty(L.omx_panel_bytes(R, ldk, fp), torch.uint8)
    ex = empty(L.omx_panel_exps_count(R, ldk), torch.int32)
    ey = empty(L.omx_panel_exps_count(R, ldk), torch.int32)
    tmax = empty(L.omx_panel_exps_count(R, B) + 64, torch.float32)
    Xs, Ys = [], []
    keep = []
    for g in range(G):
        # column panels at the LU's magnitudes: multipliers ~1e-5 (FP16
        # subnormal), row panel entries in [-0.5, 0.5)
        X = (rng.standard_normal((R, B)) * 10.0 ** rng.uniform(-6, -4)).astype(np.float32)
        Y = rng.uniform(-0.5, 0.5, (R, B)).astype(np.float32)
        keep.append(_pack_window(L, X, fp8, PX, ex, ldk, g * B, tmax))
        keep.append(_pack_window(L, Y, fp8, PY, ey, ldk, g * B, tmax))
        Xs.append(_decode(orc, X, fp8, ex, g * B, ldk)[shift:shift + m])
        Ys.append(_decode(orc, Y, fp8, ey, g * B, ldk)[shift:shift + n])
    C0 = rng.uniform(-1, 1, (n, m)).astype(np.float32)  # C^T: column j of C is row j here
    Cd = dev(C0.ravel())
    if fp8:
        rc = L.omx_gemm_f8_packed_window(m, n, ldk, -1.0, ptr(PX), ptr(ex), shift, ldk, 0, ptr(PY), ptr(ey), shift,
                                         ldk, 0, 1.0, ptr(Cd), m, None)
    else:
        rc = L.omx_gemm_f16_packed_window(m, n, ldk, -1.0, ptr(PX), None, shift, ldk, 0, ptr(PY), None, shift,
                                          ldk, 0, 1.0, ptr(Cd), m, None)
    assert rc == 0
    got = host(Cd).reshape(n, m).T.astype(np.float64)
    Xa = np.concatenate(Xs, axis=1)
    Ya = np.concatenate(Ys, axis=1)
    exact = C0.T.astype(np.float64) - Xa @ Ya.T
    bound = (ldk + 2) * 2.0**-24 * (np.abs(Xa) @ np.abs(Ya).T) + 2.0**-23 * np.abs(exact)
    err = np.abs(got - exact)
    print(f"  max err / bound = {np.max(err / bound):.3f}", flush=True)
    assert np.all(err <= bound)


def _full_solve(n, label, fp8, group4, learned, method):
    """Generate, factor and refine the order-n LCG system; returns the
    harness, the corrections, the forward error against 3 more FP64
    corrections and the schedule's FP8 step count."""
    torch = torch_cuda()
    L = _L()
    from openmxp_amd import scheduler
    seed = 42
    print(f"solve n={n} b={B} {label}", flush=True)
    A = empty(n * n, torch.float32)
    d = empty(n, torch.float64)
    rhs = empty(n, torch.float64)
    x = empty(n, torch.float64)
    assert L.omx_matgen_f32(n, seed, ptr(A), n, ptr(d), None) == 0
    assert L.omx_rhs_f64(n, seed, ptr(rhs), None) == 0
    h = _solver(n, B, 1, fp8=fp8, group4=group4)
    if learned:
        p = np.ascontiguousarray(scheduler.load_params(), np.float32)
        assert L.omx_solver_set_policy(h, p.ctypes.data) == 0
    sched = None
    if fp8 and not learned:
        sched = np.ones(n // B, np.int32)
    tim = np.zeros(5)
    info = L.omx_lu_factor(h, ptr(A), n, None if sched is None else sched.ctypes.data, tim.ctypes.data)
    used = np.zeros(n // B, np.int32)
    assert L.omx_solver_last_schedule(h, used.ctypes.data, None) == 0
    out = np.zeros(4)
    trace = np.zeros(100)
    rc = L.omx_ir_solve(h, ptr(A), n, seed, ptr(d), ptr(rhs), ptr(x), method, 50, 50, 1e-10, 16.0, 0,
                        out.ctypes.data, trace.ctypes.data)
    torch.cuda.synchronize()
    xs = x.clone()
    out2 = np.zeros(4)
    rc2 = L.omx_ir_solve(h, ptr(A), n, seed, ptr(d), ptr(rhs), ptr(x), 0, 4, 50, 1e-10, 0.0, 1,
                         out2.ctypes.data, np.zeros(8).ctypes.data)
    torch.cuda.synchronize()
    fwd = float((xs - x).abs().max() / x.abs().max())
    L.omx_solver_destroy(h)
    del A
    torch.cuda.empty_cache()
    fp8_steps = int((used[:-1] > 0).sum())
    print(f"  info {info} rc {rc} harness {out[1]:.4g} corrections {int(out[2])} inner {int(out[3])} "
          f"fp8 steps {fp8_steps}/{n // B - 1} LU {tim[0]:.0f} ms forward error {fwd:.3e}", flush=True)
    assert info == 0 and rc == 0 and rc2 == 0
    return out, fwd, fp8_steps


@pytest.mark.parametrize("label,fp8,group4,learned,method", [
    ("plain FP16 pairs, classical IR", False, False, False, 0),
    ("FP8 groups of four, GMRES-IR", True, True, False, 1),
    ("learned schedule, GMRES-IR", True, True, True, 1)])
def test_solve_n61440(label, fp8, group4, learned, method):
    """The plain method on one GPU at N = 61440, B = 2560, and the FP8 and
    learned variants: harness < 16, forward error against FP64 refinement."""
    out, fwd, _ = _full_solve(61440, label, fp8, group4, learned, method)
    assert out[1] < 16.0
    assert fwd < 1e-12


@pytest.mark.parametrize("label,fp8,group4,learned,method", [
    ("plain FP16 pairs, classical IR", False, False, False, 0),
    ("learned schedule, GMRES-IR", True, True, True, 1)])
def test_solve_n204800(label, fp8, group4, learned, method):
    """The memory-filling size N = 204800, B = 2560 (168 GB FP32 matrix):
    the plain method and the learned schedule refine to harness < 16."""
    out, fwd, fp8_steps = _full_solve(204800, label, fp8, group4, learned, method)
    assert out[1] < 16.0
    assert fwd < 1e-12
    if learned:
        assert fp8_steps > 0
