"""End-to-end GPU parity: the look-ahead LU against the oracle's LU in the
reference step order, and the GPU refinement (classical and GMRES-IR)
against the oracle and an FP64 LAPACK solve."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

from tests.gpu_util import colmajor_to_numpy, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


class Solver:
    def __init__(self, n, b, alt=1, fp8=False, group4=False):
        self.L = _L()
        self.h = C.c_void_p()
        flags = (1 if fp8 else 0) | (4 if group4 else 0)
        rc = self.L.omx_solver_create(n, b, alt, flags, C.byref(self.h))
        assert rc == 0, rc
        self.n, self.b = n, b

    def lu(self, A, lda, sched=None):
        tim = np.zeros(5)
        sp = None
        if sched is not None:
            s = np.ascontiguousarray(sched, dtype=np.int32)
            sp = s.ctypes.data
        info = self.L.omx_lu_factor(self.h, ptr(A), lda, sp, tim.ctypes.data)
        return info, tim

    def ir(self, LU, lda, seed, d, b, x, method=0, maxit=50, restart=50, tol=1e-10):
        out = np.zeros(4)
        trace = np.zeros(2 * maxit)
        rc = self.L.omx_ir_solve(self.h, ptr(LU), lda, seed, ptr(d), ptr(b), ptr(x), method, maxit, restart,
                                 tol, 16.0, 0, out.ctypes.data, trace.ctypes.data)
        return rc, out, trace

    def close(self):
        self.L.omx_solver_destroy(self.h)


def _generate(n, seed, lda):
    torch = torch_cuda()
    L = _L()
    A = empty(lda * n, torch.float32)
    d = empty(n, torch.float64)
    b = empty(n, torch.float64)
    assert L.omx_matgen_f32(n, seed, ptr(A), lda, ptr(d), None) == 0
    assert L.omx_rhs_f64(n, seed, ptr(b), None) == 0
    return A, d, b


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_lu_matches_oracle(orc, alt):
    torch = torch_cuda()
    n, b, seed = 1536, 256, 42
    print(f"lu n={n} b={b} alt={alt}", flush=True)
    lda = n + 64
    A, d, rhs = _generate(n, seed, lda)
    S = Solver(n, b, alt)
    info, tim = S.lu(A, lda)
    S.close()
    assert info == 0
    got = colmajor_to_numpy(A, n, n, lda).astype(np.float64)
    Ao, _ = orc.matgen_full(n, seed)
    want, _ = orc.lu_factor(Ao, b, alt=alt)
    want = want.astype(np.float64)
    low = np.tril(np.ones((n, n)), -1).astype(bool)
    dl = np.abs(got - want)[low]
    du = np.abs(got - want)[~low]
    lmax = np.abs(want[low]).max()
    print(f"  max |dL|/max|L| = {dl.max() / lmax:.3e}, max |dU|/(|U|+.05) = "
          f"{np.max(du / (np.abs(want[~low]) + 0.05)):.3e}", flush=True)
    # FP16 Schur inputs: factors agree to a few FP16 ulps of the panel entries
    assert dl.max() <= 2.0**-8 * lmax
    assert np.max(du / (np.abs(want[~low]) + 0.05)) <= 2.0**-8
    # backward error of the GPU factors within the FP16 bound
    Lf = np.tril(got, -1) + np.eye(n)
    Uf = np.triu(got)
    R = Lf @ Uf - Ao.astype(np.float64)
    # (FP16 panel inputs and, for -alt 1/2, an FP16 inverse: a few units of 2^-11)
    assert np.max(np.abs(R) / (2.0**-9 * (np.abs(Lf) @ np.abs(Uf)))) < 1.0
    # the trailing matrix beyond the next pair's blocks takes one product
    # per pair of panels (2p, 2p+1) while blocks 2p+4.. exist
    assert tim[2] == max(0, (n // b - 3) // 2) and tim[1] > 0


def known_factors(n, seed):
    """Well-conditioned L0 (unit lower, entries in [0, 1/n)) and diagonally
    dominant U0 with coherent signs: each Schur block update is large enough
    that dropping or misplacing one changes U by ~0.06, against rounding
    differences below 1e-3."""
    rng = np.random.default_rng(seed)
    L0 = np.tril(rng.uniform(0, 1, (n, n)) / n, -1) + np.eye(n)
    U0 = np.triu(rng.uniform(0, 1, (n, n)), 1)
    U0 += np.diag(1.0 + U0.sum(axis=1))
    return L0, U0


@pytest.mark.parametrize("alt", [0, 1, 2])
def test_lu_recovers_known_factors(alt):
    torch = torch_cuda()
    from tests.gpu_util import dev
    n, b = 1280, 256
    L0, U0 = known_factors(n, alt + 10)
    A64 = L0 @ U0
    Ad = dev(np.asfortranarray(A64.astype(np.float32)).ravel(order="F"))
    S = Solver(n, b, alt)
    info, _ = S.lu(Ad, n)
    S.close()
    assert info == 0
    got = colmajor_to_numpy(Ad, n, n).astype(np.float64)
    Lg = np.tril(got, -1) + np.eye(n)
    Ug = np.triu(got)
    print(f"  alt={alt}: max|L-L0| = {np.abs(Lg - L0).max():.2e}, max|U-U0| = {np.abs(Ug - U0).max():.2e}",
          flush=True)
    assert np.abs(Lg - L0).max() < 2e-5
    assert np.abs(Ug - U0).max() < 5e-3


@pytest.mark.parametrize("method", [0, 1])
def test_refinement_matches_oracle_and_lapack(orc, method):
    torch = torch_cuda()
    n, b, seed = 1024, 256, 7
    lda = n
    A, d, rhs = _generate(n, seed, lda)
    S = Solver(n, b, 1)
    info, _ = S.lu(A, lda)
    assert info == 0
    x = empty(n, torch.float64)
    rc, out, trace = S.ir(A, lda, seed, d, rhs, x, method=method)
    S.close()
    assert rc == 0, (out, trace)
    assert out[1] < 16.0
    LU = colmajor_to_numpy(A, n, n, lda)
    dd = host(d)
    bb = host(rhs)
    res = (orc.ir_classic if method == 0 else orc.gmres_ir)(LU, b, seed, dd, bb)
    assert res["iters"] >= 0
    assert abs(int(out[2]) - res["iters"]) <= 1
    # FP64 LAPACK solution of the same generated system
    A64 = np.array([[orc.lcg_value(orc.lcg_state(seed, n, i, j)) for j in range(n)] for i in range(n)])
    A64[np.diag_indices(n)] = dd
    xs = sla.solve(A64, bb)
    xg = host(x)
    kappa = np.linalg.cond(A64, np.inf)
    fwd = np.max(np.abs(xg - xs)) / np.max(np.abs(xs))
    print(f"  method={method} steps={int(out[2])} harness={out[1]:.4f} fwd={fwd:.3e}", flush=True)
    assert fwd <= kappa * n * 2.0**-53 * (2 * max(out[1], 1.0) + 1)


def test_configs0_solve_converges():
    """N=10240, B=256 (the reference's CPU-runnable case), classical IR."""
    torch = torch_cuda()
    n, b, seed = 10240, 256, 42
    print("solve n=10240 b=256", flush=True)
    A, d, rhs = _generate(n, seed, n)
    S = Solver(n, b, 1)
    info, tim = S.lu(A, n)
    x = empty(n, torch.float64)
    rc, out, trace = S.ir(A, n, seed, d, rhs, x)
    S.close()
    print(f"  lu {tim[0]:.1f} ms, steps {int(out[2])}, harness {out[1]:.4f}", flush=True)
    assert info == 0 and rc == 0 and out[1] < 16.0 and out[2] <= 5


def test_fp8_schedule_converges():
    torch = torch_cuda()
    n, b, seed = 2048, 256, 42
    A, d, rhs = _generate(n, seed, n)
    S = Solver(n, b, 1, fp8=True)
    info, _ = S.lu(A, n, sched=np.ones(n // b, dtype=np.int32))
    x = empty(n, torch.float64)
    rc, out, trace = S.ir(A, n, seed, d, rhs, x, method=1)
    S.close()
    print(f"  fp8: steps {int(out[2])} harness {out[1]:.4f} inner {int(out[3])}", flush=True)
    assert info == 0 and rc == 0 and out[1] < 16.0


def test_scaled_fp16_lu_matches_plain():
    """OMX_SOLVER_SCALE16 at n = 1536: the scaled factors differ from the
    plain ones only where a plain FP16 panel entry is subnormal (|L| below
    2^-14 here); they agree to the FP16 tolerance and their backward error is
    no worse."""
    torch = torch_cuda()
    n, b, seed = 1536, 256, 3
    outs = []
    for flags in (0, 2):
        A, d, rhs = _generate(n, seed, n)
        L = _L()
        h = C.c_void_p()
        assert L.omx_solver_create(n, b, 1, flags, C.byref(h)) == 0
        tim = np.zeros(5)
        assert L.omx_lu_factor(h, ptr(A), n, None, tim.ctypes.data) == 0
        L.omx_solver_destroy(h)
        outs.append(colmajor_to_numpy(A, n, n).astype(np.float64))
    from oracle import matgen_full
    A0, _ = matgen_full(n, seed)
    A0 = A0.astype(np.float64)
    berr = []
    for F in outs:
        Lf = np.tril(F, -1) + np.eye(n)
        berr.append(np.abs(Lf @ np.triu(F) - A0).max())
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    assert np.abs(outs[0] - outs[1])[low].max() <= 2.0**-8 * np.abs(outs[0][low]).max()
    print(f"  backward error plain {berr[0]:.3e} scaled {berr[1]:.3e}", flush=True)
    assert berr[1] <= 1.05 * berr[0]


@pytest.mark.parametrize("n,alt", [(768, 1), (1280, 2), (2560, 1), (3584, 0), (3584, 1)])
def test_lu_group4_matches_oracle(orc, n, alt):
    """OMX_SOLVER_GROUP4: the trailing matrix takes one product per four
    panels (depth 4b); the factors agree with the oracle's reference-order
    LU to the FP16-update tolerance, also with fewer panels than one group
    (n / b = 3), a last group of one or two panels (5, 10, 14)."""
    torch = torch_cuda()
    b, seed = 256, 42
    print(f"lu group4 n={n} b={b} alt={alt}", flush=True)
    lda = n + 64
    A, d, rhs = _generate(n, seed, lda)
    S = Solver(n, b, alt, group4=True)
    info, tim = S.lu(A, lda)
    S.close()
    assert info == 0
    got = colmajor_to_numpy(A, n, n, lda).astype(np.float64)
    Ao, _ = orc.matgen_full(n, seed)
    want, _ = orc.lu_factor(Ao, b, alt=alt)
    want = want.astype(np.float64)
    low = np.tril(np.ones((n, n)), -1).astype(bool)
    dl = np.abs(got - want)[low]
    du = np.abs(got - want)[~low]
    assert dl.max() <= 2.0**-8 * np.abs(want[low]).max()
    assert np.max(du / (np.abs(want[~low]) + 0.05)) <= 2.0**-8
    # one bulk product per group g while blocks 4 (g + 2) .. exist
    nb = n // b
    assert tim[2] == sum(1 for g in range(nb) if 4 * (g + 1) < nb and 4 * (g + 2) < nb)


def test_lu_group4_mixed_precision_schedule():
    """Groups whose panels mix FP8 and FP16 are applied as one product per
    run of equal precision: the factors match the pairwise solver's under
    the same schedule to FP8 accuracy, and the GMRES refinement converges."""
    torch = torch_cuda()
    n, b, seed = 3584, 256, 5
    sched = np.array([1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 0, 1], dtype=np.int32)
    facs = []
    for g4 in (False, True):
        A, d, rhs = _generate(n, seed, n)
        S = Solver(n, b, 1, fp8=True, group4=g4)
        info, _ = S.lu(A, n, sched=sched)
        assert info == 0
        x = empty(n, torch.float64)
        rc, out, _ = S.ir(A, n, seed, d, rhs, x, method=1)
        S.close()
        print(f"  group4={g4}: rc {rc} corrections {int(out[2])} inner {int(out[3])} harness {out[1]:.4f}",
              flush=True)
        assert rc == 0 and out[1] < 16.0
        facs.append(colmajor_to_numpy(A, n, n).astype(np.float64))
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    d2 = np.abs(facs[0] - facs[1])
    assert d2[low].max() <= 2.0**-4 * np.abs(facs[0][low]).max()
    assert np.max(d2[~low] / (np.abs(facs[0][~low]) + 0.05)) <= 2.0**-4


@pytest.mark.parametrize("alt,group4,mixed", [(0, False, False), (1, False, False), (2, False, False),
                                               (1, True, False), (2, True, True), (1, False, True)])
def test_lu_fp32_class_matches_oracle(orc, alt, group4, mixed):
    """Schedule value -1: the step's trailing products run in FP32 on its
    panels as they stand in A, and its row panel comes from the FP32 TRSM
    (no FP16 inverse).  All steps in FP32: the factors agree with the
    oracle's (same schedule, reference step order) to FP32 accumulation
    differences and their backward error is at FP32 level, far below the
    FP16 schedule's; a schedule mixing -1 and 0 (runs of one class per
    group product) agrees to the FP16 tolerance."""
    torch = torch_cuda()
    n, b, seed = 2048, 256, 7
    nb = n // b
    sched = np.array([-1, 0, 0, -1, -1, 0, -1, 0] if mixed else [-1] * nb, dtype=np.int32)
    print(f"lu fp32 class n={n} b={b} alt={alt} group4={group4} sched={sched.tolist()}", flush=True)
    lda = n + 64
    A, d, rhs = _generate(n, seed, lda)
    S = Solver(n, b, alt, group4=group4)
    info, _ = S.lu(A, lda, sched=sched)
    S.close()
    assert info == 0
    got = colmajor_to_numpy(A, n, n, lda).astype(np.float64)
    Ao, _ = orc.matgen_full(n, seed)
    want, oinfo = orc.lu_factor(Ao, b, alt=alt, sched=sched)
    assert oinfo == 0
    want = want.astype(np.float64)
    low = np.tril(np.ones((n, n)), -1).astype(bool)
    dl = np.abs(got - want)[low].max() / np.abs(want[low]).max()
    du = np.max(np.abs(got - want)[~low] / (np.abs(want[~low]) + 0.05))
    Lf = np.tril(got, -1) + np.eye(n)
    Uf = np.triu(got)
    berr = np.max(np.abs(Lf @ Uf - Ao.astype(np.float64)) / (np.abs(Lf) @ np.abs(Uf)))
    Lo = np.tril(want, -1) + np.eye(n)
    Uo = np.triu(want)
    bo = np.max(np.abs(Lo @ Uo - Ao.astype(np.float64)) / (np.abs(Lo) @ np.abs(Uo)))
    print(f"  dL {dl:.3e} dU {du:.3e} backward {berr:.3e} (oracle {bo:.3e})", flush=True)
    tol = 2.0**-8 if mixed else 2.0**-16
    assert dl <= tol and du <= tol
    # componentwise backward error: the oracle's under the same schedule
    # (FP16 steps: ~2.5e-3 at this seed, also with every step in FP16), and
    # FP32 level when every step is FP32
    assert berr <= 1.5 * bo + 2.0**-20
    if not mixed:
        assert berr < 2.0**-18


@pytest.mark.parametrize("alt,method", [(1, 0), (1, 1), (0, 0)])
def test_phase_timers(alt, method):
    """With OMX_SOLVER_TIMERS the solver reports GPU time per phase under the
    reference's Timer items; the operation counts are exact: the Schur
    products sum to sum_k 2 b (n - (k+1) b)^2 whatever the grouping, the
    diagonal LUs (and -alt inverses) to 2/3 b^3 each, each residual to 2 n^2,
    each pair of triangular solves to 2 n^2."""
    from openmxp_amd import _lib
    torch = torch_cuda()
    n, b, seed = 2048, 256, 42
    print(f"phase timers alt={alt} method={method}", flush=True)
    L = _L()
    h = C.c_void_p()
    assert L.omx_solver_create(n, b, alt, _lib.SOLVER_TIMERS, C.byref(h)) == 0
    A, d, rhs = _generate(n, seed, n)
    x = empty(n, torch.float64)
    tim = np.zeros(5)
    assert L.omx_lu_factor(h, ptr(A), n, None, tim.ctypes.data) == 0
    out = np.zeros(4)
    trace = np.zeros(100)
    assert L.omx_ir_solve(h, ptr(A), n, seed, ptr(d), ptr(rhs), ptr(x), method, 50, 50, 1e-10, 16.0, 0,
                          out.ctypes.data, trace.ctypes.data) == 0
    sec = np.zeros(_lib.TIMER_ITEMS)
    ops = np.zeros(_lib.TIMER_ITEMS)
    assert L.omx_solver_timers(h, sec.ctypes.data, ops.ctypes.data) == 0
    L.omx_solver_destroy(h)
    name = {L.omx_timer_name(i).decode(): i for i in range(_lib.TIMER_ITEMS)}
    for item in ("DIAG_LU", "TRSM_L", "TRSM_R", "CONV_L", "CONV_R", "GEMM_UPDATE", "IR_GEMV", "IR_TRSV", "TOTAL"):
        assert sec[name[item]] > 0.0, item
    for item in ("DIAG_BCAST", "LCOL_BCAST", "RROW_BCAST", "IR_GEMV_COMM", "IR_TRSV_COMM", "INIT"):
        assert sec[name[item]] == 0.0 and ops[name[item]] == 0.0, item
    nb = n // b
    schur = sum(2.0 * b * (n - (k + 1) * b) ** 2 for k in range(nb))
    assert ops[name["GEMM_UPDATE"]] == pytest.approx(schur, rel=1e-12)
    inverses = (nb - 2) if alt else 0  # every panel but the first and the last
    assert ops[name["DIAG_LU"]] == pytest.approx(2.0 / 3.0 * b ** 3 * (nb + inverses), rel=1e-12)
    # out[2] corrections, out[3] inner GMRES steps (0 for classical IR): a
    # residual per check and per Arnoldi step, a pair of solves per
    # correction (GMRES: its first basis vector) and per Arnoldi step
    gemvs = out[2] + 1 + out[3]
    pairs = out[2] + out[3]
    assert ops[name["IR_GEMV"]] == pytest.approx(2.0 * n * n * gemvs, rel=1e-12)
    assert ops[name["IR_TRSV"]] == pytest.approx(2.0 * n * n * pairs, rel=1e-12)
    # MISC closes the sum to TOTAL
    assert sec[name["MISC"]] == pytest.approx(sec[name["TOTAL"]] - sec[:name["TOTAL"]].sum(), abs=1e-12)


def test_phase_timers_off_by_default():
    L = _L()
    h = C.c_void_p()
    assert L.omx_solver_create(1024, 256, 1, 0, C.byref(h)) == 0
    buf = np.zeros(32)
    assert L.omx_solver_timers(h, buf.ctypes.data, buf.ctypes.data) < 0
    L.omx_solver_destroy(h)



def test_ir_rejects_unknown_method_and_unreserved_restart():
    """omx_ir_solve checks its arguments before queueing any work: a method
    other than 0 (classical) or 1 (GMRES) is OMX_ERR_ARG; a GMRES restart
    above the basis omx_solver_create reserved (50) is OMX_ERR_WORKSPACE
    until omx_solver_reserve_gmres enlarges it, after which it refines."""
    torch = torch_cuda()
    n, b, seed = 1024, 256, 11
    A, d, rhs = _generate(n, seed, n)
    S = Solver(n, b, 1)
    info, _ = S.lu(A, n)
    assert info == 0
    x = empty(n, torch.float64)
    for method in (2, -1, 7):
        rc, out, _ = S.ir(A, n, seed, d, rhs, x, method=method)
        assert rc == -1, (method, rc)
    rc, out, _ = S.ir(A, n, seed, d, rhs, x, method=1, restart=60)
    assert rc == -3
    assert S.L.omx_solver_reserve_gmres(S.h, 60) == 0
    assert S.L.omx_solver_reserve_gmres(S.h, 513) == -1
    rc, out, _ = S.ir(A, n, seed, d, rhs, x, method=1, restart=60)
    S.close()
    assert rc == 0 and out[1] < 16.0
