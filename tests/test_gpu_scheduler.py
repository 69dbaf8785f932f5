"""The learned schedule's device kernels against the CPU oracle (MLP
forward bit-exact, gradients, Adam), the feature kernel against a numpy
restatement, and the policy inside omx_lu_factor against fixed schedules."""
import ctypes as C

import numpy as np
import pytest

from tests.gpu_util import empty, host, ptr, torch_cuda

pytestmark = pytest.mark.gpu

NP = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3


def _L():
    from openmxp_amd import _lib
    return _lib.lib()


def problem(seed, B):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(NP) * 0.2).astype(np.float32)
    X = rng.standard_normal((B, 16)).astype(np.float32)
    y = rng.integers(0, 2, B).astype(np.int32)
    t = rng.uniform(0, 3, B).astype(np.float32)
    m = (rng.uniform(0, 1, B) < 0.5).astype(np.float32)
    return params, X, y, t, m


@pytest.mark.parametrize("B", [1, 300])
def test_mlp_forward_bitexact_and_backward_vs_oracle(orc, B):
    torch = torch_cuda()
    L = _L()
    params, X, y, t, m = problem(B, B)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    P, Xd, Yd, Td, Md = d(params), d(X), d(y), d(t), d(m)
    H1 = empty(B * 64, torch.float32)
    H2 = empty(B * 64, torch.float32)
    Z = empty(B * 3, torch.float32)
    assert L.omx_mlp_forward_f32(B, ptr(P), ptr(Xd), ptr(H1), ptr(H2), ptr(Z), None) == 0
    h1, h2, z = orc.mlp_forward(params, X)
    assert np.array_equal(host(H1).reshape(B, 64), h1)
    assert np.array_equal(host(H2).reshape(B, 64), h2)
    assert np.array_equal(host(Z).reshape(B, 3), z)
    G = empty(NP, torch.float32)
    loss = empty(1, torch.float32)
    wb = L.omx_mlp_backward_workspace_bytes(B)
    W = empty(wb // 4 + 1, torch.float32)
    assert L.omx_mlp_backward_f32(B, ptr(P), ptr(Xd), ptr(H1), ptr(H2), ptr(Z), ptr(Yd), ptr(Td), ptr(Md),
                                  ptr(G), ptr(loss), ptr(W), wb, None) == 0
    lo, go = orc.mlp_backward(params, X, y, t, m)
    # identical sums; only expf/logf of the loss may differ by an ulp
    assert abs(host(loss)[0] - lo) <= 1e-6 * max(1, abs(lo))
    assert np.max(np.abs(host(G) - go)) <= 1e-6 * max(1e-3, np.abs(go).max())


def test_adam_step_matches_numpy():
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(4)
    n = 1000
    p = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    mo = rng.standard_normal(n).astype(np.float32) * 0.1
    ve = rng.uniform(0, 0.1, n).astype(np.float32)
    d = lambda a: torch.from_numpy(a.copy()).cuda()
    P, G, M, V = d(p), d(g), d(mo), d(ve)
    assert L.omx_mlp_adam_f32(n, ptr(P), ptr(G), ptr(M), ptr(V), 1e-2, 0.9, 0.999, 1e-8, 3, None) == 0
    # FP64 restatement of the update with the FP32 constants the kernel uses
    f = np.float32
    b1, b2 = float(f(0.9)), float(f(0.999))
    m1 = b1 * mo.astype(np.float64) + float(f(1) - f(0.9)) * g
    v1 = b2 * ve.astype(np.float64) + float(f(1) - f(0.999)) * g.astype(np.float64) ** 2
    want = p - 1e-2 * (m1 / (1 - b1 ** 3)) / (np.sqrt(v1 / (1 - b2 ** 3)) + 1e-8)
    assert np.allclose(host(P), want, rtol=1e-5, atol=1e-6)
    assert np.allclose(host(M), m1, rtol=1e-5, atol=1e-7) and np.allclose(host(V), v1, rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("mu", [0.0, 0.9])
def test_sgd_step_matches_numpy(mu):
    """Two SGD steps, FP32 as written (v = mu v + g; p -= lr v): bit-exact
    against numpy float32."""
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(7)
    n = 5443
    p = rng.standard_normal(n).astype(np.float32)
    g1 = rng.standard_normal(n).astype(np.float32)
    g2 = rng.standard_normal(n).astype(np.float32)
    d = lambda a: torch.from_numpy(a.copy()).cuda()
    P, V = d(p), d(np.zeros(n, np.float32))
    f = np.float32
    lr = f(0.05)
    want, v = p.copy(), np.zeros(n, np.float32)
    for g in (g1, g2):
        G = d(g)
        assert L.omx_mlp_sgd_f32(n, ptr(P), ptr(G), ptr(V) if mu else None, 0.05, mu, None) == 0
        if mu:
            v = f(mu) * v + g
            want = want - lr * v
        else:
            want = want - lr * g
    assert np.array_equal(host(P), want)
    assert L.omx_mlp_sgd_f32(n, ptr(P), ptr(G), None, 0.05, 0.9, None) < 0   # momentum without a buffer


def numpy_features(lmax, rmax, udiag, step, nb, n, b, prev):
    def stats(v):
        v = v[(v > 0) & np.isfinite(v)]
        if v.size == 0:
            return -64.0, -64.0, -64.0, 0.0
        l = np.log2(v.astype(np.float64))
        return l.max(), l.min(), l.mean(), l.std()
    Lmx, Lmn, Lme, Lsd = stats(lmax)
    Rmx, Rmn, Rme, Rsd = stats(rmax)
    u = np.abs(udiag.astype(np.float64))
    lu = np.where(u > 0, np.log2(np.where(u > 0, u, 1)), -64.0)
    umx, umn = lu.max(), lu.min()
    rest = n - (step + 1) * b
    f = [Lmx / 16, Lmn / 16, Lme / 16, Lsd / 8, Rmx / 16, Rmn / 16, Rme / 16, Rsd / 8, umx / 16, umn / 16,
         (umx - umn) / 16, step / nb, np.log2(max(rest, 1) / n) / 8, (Lmx - umx) / 16, (Rmx - umx) / 16, prev]
    return np.clip(np.array(f), -8, 8)


def test_features_match_numpy():
    torch = torch_cuda()
    L = _L()
    rng = np.random.default_rng(8)
    rows, b = 3000, 512
    col = (rng.standard_normal((b, rows)) * 2.0 ** rng.integers(-20, 5, (1, rows))).astype(np.float32)  # [k][r]
    colT = np.ascontiguousarray(col)              # column panel, rs = 1, ls = rows
    row = np.ascontiguousarray(col.T * 3)         # row panel read transposed: rs = b, ls = 1
    Cd = torch.from_numpy(colT.ravel()).cuda()
    Rd = torch.from_numpy(row.ravel()).cuda()
    nt = L.omx_panel_exps_count(rows, b)
    lm = empty(nt, torch.float32)
    rm = empty(nt, torch.float32)
    assert L.omx_panel_tile_maxabs(rows, b, ptr(Cd), 1, rows, ptr(lm), None) == 0
    assert L.omx_panel_tile_maxabs(rows, b, ptr(Rd), b, 1, ptr(rm), None) == 0
    # tile maxima: strips of 256 rows x 256 columns
    want_l = np.zeros(nt, np.float32)
    kt = b // 256
    for s in range(-(-rows // 256)):
        for q in range(kt):
            want_l[s * kt + q] = np.abs(col[q * 256:(q + 1) * 256, s * 256:(s + 1) * 256]).max()
    assert np.array_equal(host(lm), want_l)
    assert np.array_equal(host(rm), want_l * 3)
    Akk = (rng.standard_normal((b, b)) * 10).astype(np.float32)
    Ad = torch.from_numpy(np.asfortranarray(Akk).ravel(order="F")).cuda()
    f = empty(16, torch.float32)
    assert L.omx_sched_features_f32(nt, ptr(lm), nt, ptr(rm), ptr(Ad), b, b, 3, 20, 10240, 1, ptr(f), None) == 0
    want = numpy_features(want_l, want_l * 3, np.diag(Akk), 3, 20, 10240, b, 1)
    # FP32 log2 and one-pass variance on the GPU against FP64 numpy
    assert np.allclose(host(f), want, rtol=1e-4, atol=2e-3)


def _const_policy(fp8, headroom=0):
    p = np.zeros(NP, np.float32)
    p[-3:] = [0.0, 1.0, float(headroom)] if fp8 else [1.0, 0.0, 0.0]
    return p


@pytest.mark.parametrize("fp8,headroom", [(False, 0), (True, 0), (True, 1)])
def test_constant_policy_equals_fixed_schedule(fp8, headroom):
    """A policy that always answers the same must reproduce the matching
    fixed schedule bit for bit (the policy only chooses; it computes
    nothing else)."""
    torch = torch_cuda()
    L = _L()
    n, b, seed = 2048, 256, 42
    res = []
    for mode in ("fixed", "policy"):
        A = empty(n * n, torch.float32)
        d = empty(n, torch.float64)
        assert L.omx_matgen_f32(n, seed, ptr(A), n, ptr(d), None) == 0
        h = C.c_void_p()
        assert L.omx_solver_create(n, b, 1, 1, C.byref(h)) == 0
        tim = np.zeros(5)
        if mode == "fixed":
            sched = np.full(n // b, (1 + headroom) if fp8 else 0, np.int32)
            info = L.omx_lu_factor(h, ptr(A), n, sched.ctypes.data, tim.ctypes.data)
        else:
            p = _const_policy(fp8, headroom)
            assert L.omx_solver_set_policy(h, p.ctypes.data) == 0
            info = L.omx_lu_factor(h, ptr(A), n, None, tim.ctypes.data)
        used = np.zeros(n // b, np.int32)
        assert L.omx_solver_last_schedule(h, used.ctypes.data, None) == 0
        L.omx_solver_destroy(h)
        assert info == 0
        res.append((host(A), used))
    assert np.array_equal(res[0][0], res[1][0])
    expect = (1 + headroom) if fp8 else 0
    assert np.all(res[1][1][:-1] == expect) and res[1][1][-1] == 0


def test_trained_policy_against_fixed_schedules():
    """What the trained network gains over the methods without it, on
    seeded systems outside its training set (n = 16384, b = 512, GMRES-IR):
    the diagonally dominant LCG matrix and randsvd matrices (8 reflectors per
    factor) with kappa = 30, 300, 3000, each factored all-FP16, all-FP8 and
    with the learned schedule:
      - safer than all-FP8: the learned schedule refines wherever all-FP16
        does, while all-FP8 fails (or exceeds the FP16 cost + 3) on at least
        one randsvd system;
      - more FP8 than all-FP16: FP8 on (nearly) every step of the LCG system
        and on more than half of the four systems' steps, with the four
        refinements' total cost within the labelling rule's FP16 cost + 3
        per system.
    (At this order the LU is bound by its diagonal-block chain, not by the
    Schur products, so its time hardly depends on their precision; the time
    gain shows at n = 204800, in the bench's learned / all-FP8 / FP16 legs.)"""
    torch_cuda()
    from openmxp_amd import scheduler
    try:
        params = scheduler.load_params()
    except (OSError, ValueError):
        pytest.skip("no trained scheduler committed")
    n, b = 16384, 512
    nb = n // b
    S = scheduler.Solver(n, b)
    rows = []
    try:
        for kind, seed, kappa in (("lcg", 4242, 1.0), ("randsvd", 4343, 30.0), ("randsvd", 4344, 300.0),
                                  ("randsvd", 4345, 3000.0)):
            sysm = scheduler.System(kind, n, seed, kappa=kappa)
            S.set_policy(None)
            S.solve(sysm, sched=np.zeros(nb, np.int32))  # warm-up
            f16 = S.solve(sysm, sched=np.zeros(nb, np.int32))
            all8 = S.solve(sysm, sched=np.ones(nb, np.int32))
            S.set_policy(params)
            learned = S.solve(sysm)
            S.set_policy(None)
            n8 = int((learned["sched"][:-1] > 0).sum())
            print(f"  {kind} kappa={kappa:g}: fp16 cost {f16['cost']} LU {f16['lu_s'] * 1e3:.1f} ms | all-fp8 "
                  f"{'cost ' + str(all8['cost']) if all8['ok'] else 'fails'} LU {all8['lu_s'] * 1e3:.1f} ms | learned "
                  f"{'cost ' + str(learned['cost']) if learned['ok'] else 'fails'} LU {learned['lu_s'] * 1e3:.1f} ms, "
                  f"{n8}/{nb - 1} FP8 steps", flush=True)
            rows.append((kind, kappa, f16, all8, learned, n8))
            del sysm
    finally:
        S.close()
    for kind, kappa, f16, all8, learned, n8 in rows:
        if f16["ok"]:
            assert learned["ok"], (kind, kappa)
    assert any(not (all8["ok"] and all8["cost"] <= f16["cost"] + 3)
               for kind, kappa, f16, all8, learned, n8 in rows if kind == "randsvd")
    kind, kappa, f16, all8, learned, n8 = rows[0]
    assert n8 >= 0.9 * (nb - 1) and learned["cost"] <= f16["cost"] + 3
    assert sum(r[5] for r in rows) > 0.5 * len(rows) * (nb - 1)
    assert sum(r[4]["cost"] for r in rows) <= sum(r[2]["cost"] for r in rows) + 3 * len(rows)
