"""GPU randsvd generator, its low-rank residual product and LU + GMRES-IR
on randsvd systems, against the oracle's dense FP64 construction and
LAPACK."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

from tests.gpu_util import colmajor_to_numpy, empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


class RandSVD:
    """Device factors, FP32 matrix, FP64 diagonal and row sums."""

    def __init__(self, n, seed, kappa, nrefl=2, lda=None):
        torch = torch_cuda()
        L = _L()
        self.n, self.R = n, 2 * nrefl
        self.lda = lda or n
        self.sigma = empty(n, torch.float64)
        self.P = empty(max(1, n * self.R), torch.float64)
        self.Q = empty(max(1, n * self.R), torch.float64)
        assert L.omx_randsvd_factors(n, seed, kappa, nrefl, ptr(self.sigma), ptr(self.P), ptr(self.Q), None) == 0
        self.A = empty(self.lda * n, torch.float32)
        self.diag = empty(n, torch.float64)
        self.rowabs = empty(n, torch.float64)
        wb = L.omx_randsvd_workspace_bytes(n)
        w = empty(wb // 8 + 1, torch.float64)
        assert L.omx_randsvd_f32(n, self.R, ptr(self.sigma), ptr(self.P), ptr(self.Q), ptr(self.A), self.lda,
                                 ptr(self.diag), ptr(self.rowabs), ptr(w), wb, None) == 0
        torch.cuda.synchronize()
        self.norm_inf = float(host(self.rowabs).max())

    def system(self):
        from openmxp_amd import _lib
        return _lib.System(kind=_lib.SYSTEM_LOWRANK, seed=0, diag=ptr(self.diag), norm_inf=self.norm_inf,
                           rank=self.R, sigma=ptr(self.sigma), P=ptr(self.P), Q=ptr(self.Q))


@pytest.mark.parametrize("n,lda", [(1000, 1000), (2048, 2050)])
def test_fill_matches_oracle(orc, n, lda):
    print(f"randsvd fill n={n}", flush=True)
    M = RandSVD(n, 11, 1e4, lda=lda)
    Ao = orc.randsvd_dense(n, 11, 1e4, 2)
    got = colmajor_to_numpy(M.A, n, n, lda).astype(np.float64)
    want32 = Ao.astype(np.float32).astype(np.float64)
    # the FP64 entries agree to ~1e-16 relative, so their FP32 roundings agree
    # except next to a rounding boundary: at most one FP32 ulp
    assert np.all(np.abs(got - want32) <= 2.0**-23 * np.abs(Ao) + 1e-300)
    assert np.mean(got == want32) > 0.999
    assert np.max(np.abs(host(M.diag) - np.diag(Ao))) <= 1e-14
    rs = np.abs(Ao).sum(axis=1)
    assert np.max(np.abs(host(M.rowabs) - rs) / rs) <= 1e-13


def test_lowrank_gemv_matches_dense(orc):
    torch = torch_cuda()
    L = _L()
    n = 1500
    M = RandSVD(n, 5, 1e3)
    Ao = orc.randsvd_dense(n, 5, 1e3, 2)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y0.copy()).cuda()
    wb = L.omx_lowrank_gemv_workspace_bytes(M.R)
    w = empty(wb // 8 + 1, torch.float64)
    assert L.omx_lowrank_gemv_f64(n, M.R, ptr(M.sigma), ptr(M.P), ptr(M.Q), -0.5, ptr(xd), ptr(yd), ptr(w), wb,
                                  None) == 0
    want = y0 - 0.5 * (Ao @ x)
    assert np.max(np.abs(host(yd) - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("kappa,mode", [(1e2, "fp16"), (1e4, "fp16"), (1e2, "fp8"), (1e4, "fp8"), (1e4, "fp16s"),
                                        (1e6, "fp16"), (1e6, "fp8"), (1e6, "fp16s"), (1e4, "fp32"),
                                        (1e6, "fp32")])
def test_lu_gmres_ir_on_randsvd(orc, kappa, mode):
    """No-pivot LU (FP16, FP16 with per-strip scales, or FP8 Schur) +
    GMRES-IR on a randsvd system.
    FP16, kappa <= 1e4: harness < 16 and the forward error against LAPACK's
    FP64 solve within the conditioning bound.  kappa = 1e6 is beyond a
    low-precision no-pivot factorisation (kappa * u16 ~ 500: the late pivots,
    ~1e-6, drown in the FP16 rounding of the Schur update and the factors
    overflow), and FP8 (3 mantissa bits) is beyond it on these
    non-dominant matrices at any kappa (the oracle's FP8 factors of the
    kappa = 100 case have a backward error ~35 |A|): the solve must then
    report non-convergence, never a false success.  The FP32 Schur class
    (schedule -1 on every step; kappa * u32 <= 0.06) must converge at both
    kappa = 1e4 and 1e6."""
    torch = torch_cuda()
    L = _L()
    n, b = 2048, 256
    print(f"randsvd solve kappa={kappa:g} {mode}", flush=True)
    fp8 = mode == "fp8"
    M = RandSVD(n, 21, kappa)
    rhs = empty(n, torch.float64)
    assert L.omx_rhs_f64(n, 21, ptr(rhs), None) == 0
    h = C.c_void_p()
    flags = 1 if fp8 else (2 if mode == "fp16s" else 0)  # OMX_SOLVER_FP8 / OMX_SOLVER_SCALE16
    assert L.omx_solver_create(n, b, 1, flags, C.byref(h)) == 0
    sched = np.full(n // b, 1 if fp8 else (-1 if mode == "fp32" else 0), dtype=np.int32)
    tim = np.zeros(5)
    info = L.omx_lu_factor(h, ptr(M.A), n, sched.ctypes.data, tim.ctypes.data)
    x = empty(n, torch.float64)
    out = np.zeros(4)
    trace = np.zeros(100)
    sysd = M.system()
    rc = L.omx_ir_solve_system(h, ptr(M.A), n, C.byref(sysd), ptr(rhs), ptr(x), 1, 50, 50, 1e-12, 16.0, 0,
                               out.ctypes.data, trace.ctypes.data)
    L.omx_solver_destroy(h)
    print(f"  info {info} rc {rc} steps {int(out[2])} inner {int(out[3])} harness {out[1]:.4g}", flush=True)
    if (kappa <= 1e4 and not fp8) or mode == "fp32":
        assert info == 0 and rc == 0 and out[1] < 16.0
    if rc != 0:
        assert rc == -4 and out[2] == -1  # OMX_ERR_NOCONV, no corrections claimed
        return
    assert out[1] < 16.0
    Ao = orc.randsvd_dense(n, 21, kappa, 2)
    bb = host(rhs)
    xs = sla.solve(Ao, bb)
    fwd = np.max(np.abs(host(x) - xs)) / np.max(np.abs(xs))
    kinf = np.linalg.cond(Ao, np.inf)
    print(f"  fwd {fwd:.3e} kappa_inf {kinf:.3e}", flush=True)
    assert fwd <= kinf * n * 2.0**-53 * (2 * max(out[1], 1.0) + 1)
