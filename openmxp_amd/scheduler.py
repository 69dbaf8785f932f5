"""Learned per-step FP8/FP16 precision schedule for the LU.

The network is the library's MLP 16-64-64-3 (include/omxp.h, "Learned
precision schedule"): its inputs are the 16 statistics omx_lu_factor
computes for every panel step (tile maxima of the step's column and row
panels, the pivots of the diagonal block, the step's position), its outputs
the FP16/FP8 logits and the FP8 tile-scale headroom.

Labels come from the solver's own runs on seeded systems (the reference's
diagonally dominant LCG matrices and randsvd matrices with condition numbers
log-uniform in [1e1, 1e6]): a step is labelled FP8 when the refinement of
the factorisation that uses FP8 for it costs at most `slack` more
iterations than the all-FP16 factorisation.  Training runs on the GPU with
the library's forward, backward and Adam kernels.  At solve time the policy
is evaluated by omx_lu_factor itself, one forward pass per step.

The trained parameters live in data/sched_mlp.f32 (raw little-endian FP32,
OMX_MLP_NPARAMS values), written by `python -m openmxp_amd.scheduler train`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib

PKG = os.path.dirname(os.path.abspath(__file__))
PARAMS_FILE = os.path.join(PKG, "data", "sched_mlp.f32")
NPARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3
NFEAT = 16


def load_params(path: str = PARAMS_FILE) -> np.ndarray:
    p = np.fromfile(path, dtype="<f4")
    if p.size != NPARAMS:
        raise ValueError(f"{path}: {p.size} values, expected {NPARAMS}")
    return p.astype(np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 3)):
        parts.append((rng.standard_normal(fan_in * fan_out) * math.sqrt(2.0 / fan_in)).astype(np.float32))
        parts.append(np.zeros(fan_out, np.float32))
    p = np.concatenate(parts)
    assert p.size == NPARAMS
    return p


# ---------------------------------------------------------------------------
# systems and solves (device buffers through torch; compute through the C ABI)
# ---------------------------------------------------------------------------
class System:
    """One generated system: A (FP32, device), its FP64 diagonal, rhs, and
    for randsvd the low-rank description."""

    def __init__(self, kind: str, n: int, seed: int, kappa: float = 1.0, nrefl: int = 8):
        import torch
        self.torch = torch
        L = _lib.lib()
        self.kind, self.n, self.seed, self.kappa = kind, n, seed, kappa
        dev = "cuda"
        self.A0 = None
        self.A = torch.empty(n * n, dtype=torch.float32, device=dev)
        self.diag = torch.empty(n, dtype=torch.float64, device=dev)
        self.rhs = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.check(L.omx_rhs_f64(n, seed, self.rhs.data_ptr(), None), "rhs")
        if kind == "lcg":
            self.R = 0
        else:
            self.R = 2 * nrefl
            self.sigma = torch.empty(n, dtype=torch.float64, device=dev)
            self.P = torch.empty(n * self.R, dtype=torch.float64, device=dev)
            self.Q = torch.empty(n * self.R, dtype=torch.float64, device=dev)
            self.rowabs = torch.empty(n, dtype=torch.float64, device=dev)
            _lib.check(L.omx_randsvd_factors(n, seed, float(kappa), nrefl, self.sigma.data_ptr(),
                                             self.P.data_ptr(), self.Q.data_ptr(), None), "randsvd factors")
            wb = L.omx_randsvd_workspace_bytes(n)
            self.work = torch.empty(wb // 8 + 1, dtype=torch.float64, device=dev)
        self.generate()
        if kind != "lcg":
            self.norm_inf = float(self.rowabs.max().item())

    def generate(self):
        L = _lib.lib()
        n = self.n
        if self.kind == "lcg":
            _lib.check(L.omx_matgen_f32(n, self.seed, self.A.data_ptr(), n, self.diag.data_ptr(), None), "matgen")
        else:
            _lib.check(L.omx_randsvd_f32(n, self.R, self.sigma.data_ptr(), self.P.data_ptr(), self.Q.data_ptr(),
                                         self.A.data_ptr(), n, self.diag.data_ptr(), self.rowabs.data_ptr(),
                                         self.work.data_ptr(), self.work.numel() * 8, None), "randsvd fill")
        self.torch.cuda.synchronize()

    def desc(self):
        if self.kind == "lcg":
            return _lib.System(kind=_lib.SYSTEM_LCG, seed=self.seed, diag=self.diag.data_ptr())
        return _lib.System(kind=_lib.SYSTEM_LOWRANK, seed=0, diag=self.diag.data_ptr(), norm_inf=self.norm_inf,
                           rank=self.R, sigma=self.sigma.data_ptr(), P=self.P.data_ptr(), Q=self.Q.data_ptr())


class Solver:
    def __init__(self, n: int, b: int, alt: int = 1):
        L = _lib.lib()
        self.h = C.c_void_p()
        _lib.check(L.omx_solver_create(n, b, alt, 1, C.byref(self.h)), "solver create")
        self.n, self.b, self.nb = n, b, n // b
        self.solves = 0  # factorisations + refinements run by solve()

    def set_policy(self, params):
        L = _lib.lib()
        if params is None:
            _lib.check(L.omx_solver_set_policy(self.h, None), "clear policy")
        else:
            p = np.ascontiguousarray(params, dtype=np.float32)
            _lib.check(L.omx_solver_set_policy(self.h, p.ctypes.data), "set policy")

    def solve(self, system: System, sched=None, method: int = 1, maxit: int = 50, restart: int = 50,
              inner_tol: float = 1e-12, regenerate: bool = True):
        """LU (sched: None = policy/FP16, else per-step ints) + refinement.
        Returns a dict with the cost (outer + inner iterations), the LU and
        IR times, the schedule used and the step features."""
        import torch
        L = _lib.lib()
        if regenerate:
            system.generate()
        self.solves += 1
        n = self.n
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        sp = None
        if sched is not None:
            s = np.ascontiguousarray(sched, dtype=np.int32)
            sp = s.ctypes.data
        tim = np.zeros(5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = L.omx_lu_factor(self.h, system.A.data_ptr(), n, sp, tim.ctypes.data)
        t1 = time.perf_counter()
        out = np.zeros(4)
        trace = np.zeros(2 * maxit)
        d = system.desc()
        rc = L.omx_ir_solve_system(self.h, system.A.data_ptr(), n, C.byref(d), system.rhs.data_ptr(), x.data_ptr(),
                                   method, maxit, restart, inner_tol, 16.0, 0, out.ctypes.data, trace.ctypes.data)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        used = np.zeros(self.nb, dtype=np.int32)
        feats = np.zeros((self.nb, NFEAT), dtype=np.float32)
        _lib.check(L.omx_solver_last_schedule(self.h, used.ctypes.data, feats.ctypes.data), "last schedule")
        ok = (info == 0 and rc == 0 and out[1] < 16.0)
        cost = int(out[2]) + int(out[3]) if ok else 10 ** 6
        return dict(ok=bool(ok), info=int(info), rc=int(rc), cost=cost, outer=int(out[2]), inner=int(out[3]),
                    harness=float(out[1]), lu_s=t1 - t0, ir_s=t2 - t1, sched=used, features=feats, x=x)

    def close(self):
        _lib.lib().omx_solver_destroy(self.h)


# ---------------------------------------------------------------------------
# labels
# ---------------------------------------------------------------------------
def label_system(solver: Solver, system: System, slack: int = 3, log=print):
    """Per-step labels for one system: (features, y, t, m) arrays of nb-1
    rows (the last step has no panels), plus a summary."""
    nb = solver.nb
    steps = nb - 1
    base = solver.solve(system, sched=np.zeros(nb, np.int32))
    if not base["ok"]:
        log(f"  {system.kind} n={system.n} seed={system.seed}: all-FP16 does not converge; skipped")
        return None
    budget = base["cost"] + slack
    all8 = solver.solve(system, sched=np.ones(nb, np.int32))
    if all8["ok"] and all8["cost"] <= budget:
        y = np.ones(steps, np.int32)
    else:
        # one step at a time in FP8, the rest FP16
        single = np.zeros(steps)
        for k in range(steps):
            s = np.zeros(nb, np.int32)
            s[k] = 1
            r = solver.solve(system, sched=s)
            single[k] = r["cost"] if r["ok"] else 10 ** 6
        y = (single <= base["cost"] + 1).astype(np.int32)
        # the combination must stay within the budget: give back the most
        # expensive FP8 steps until it does
        order = np.argsort(-single)
        while y.any():
            s = np.zeros(nb, np.int32)
            s[:steps] = y
            r = solver.solve(system, sched=s)
            if r["ok"] and r["cost"] <= budget:
                break
            for k in order:
                if y[k]:
                    y[k] = 0
                    break
    # headroom: one extra bit on every FP8 step, kept if it lowers the cost
    t = np.zeros(steps, np.float32)
    if y.any():
        s0 = np.zeros(nb, np.int32)
        s0[:steps] = y
        r0 = solver.solve(system, sched=s0)
        s1 = s0.copy()
        s1[s1 == 1] = 2
        r1 = solver.solve(system, sched=s1)
        if r1["ok"] and r0["ok"] and r1["cost"] < r0["cost"]:
            t[y == 1] = 1.0
    F = base["features"][:steps].copy()
    # the previous-step input as the labelled schedule would have set it
    F[1:, 15] = y[:-1]
    F[0, 15] = 0
    m = y.astype(np.float32)
    log(f"  {system.kind} n={system.n} seed={system.seed} kappa={system.kappa:.3g}: fp16 cost {base['cost']}, "
        f"all-fp8 cost {all8['cost'] if all8['ok'] else 'fail'}, fp8 steps {int(y.sum())}/{steps}")
    return F, y, t, m, dict(kind=system.kind, n=system.n, seed=system.seed, kappa=system.kappa,
                            fp16_cost=base["cost"], fp8_cost=all8["cost"] if all8["ok"] else None,
                            fp8_steps=int(y.sum()), steps=steps)


# ---------------------------------------------------------------------------
# training on the GPU
# ---------------------------------------------------------------------------
def train(F, y, t, m, iters: int = 3000, lr: float = 3e-3, seed: int = 0, log=print):
    """Full-batch Adam with the library's kernels.  Returns (params, losses)."""
    import torch
    L = _lib.lib()
    B = F.shape[0]
    dev = "cuda"
    X = torch.from_numpy(np.ascontiguousarray(F, np.float32)).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(y, np.int32)).to(dev)
    T = torch.from_numpy(np.ascontiguousarray(t, np.float32)).to(dev)
    M = torch.from_numpy(np.ascontiguousarray(m, np.float32)).to(dev)
    P = torch.from_numpy(init_params(seed)).to(dev)
    G = torch.zeros(NPARAMS, dtype=torch.float32, device=dev)
    Mo = torch.zeros_like(G)
    V = torch.zeros_like(G)
    H1 = torch.empty(B * 64, dtype=torch.float32, device=dev)
    H2 = torch.empty(B * 64, dtype=torch.float32, device=dev)
    Z = torch.empty(B * 3, dtype=torch.float32, device=dev)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    wb = L.omx_mlp_backward_workspace_bytes(B)
    W = torch.empty(wb // 4 + 1, dtype=torch.float32, device=dev)
    losses = []
    for it in range(1, iters + 1):
        _lib.check(L.omx_mlp_forward_f32(B, P.data_ptr(), X.data_ptr(), H1.data_ptr(), H2.data_ptr(), Z.data_ptr(),
                                         None), "mlp forward")
        _lib.check(L.omx_mlp_backward_f32(B, P.data_ptr(), X.data_ptr(), H1.data_ptr(), H2.data_ptr(), Z.data_ptr(),
                                          Y.data_ptr(), T.data_ptr(), M.data_ptr(), G.data_ptr(), loss.data_ptr(),
                                          W.data_ptr(), wb, None), "mlp backward")
        _lib.check(L.omx_mlp_adam_f32(NPARAMS, P.data_ptr(), G.data_ptr(), Mo.data_ptr(), V.data_ptr(), lr, 0.9,
                                      0.999, 1e-8, it, None), "adam")
        if it % 250 == 0 or it == 1:
            losses.append((it, float(loss.item())))
            log(f"  train iter {it}: loss {losses[-1][1]:.5f}")
    _lib.check(L.omx_mlp_forward_f32(B, P.data_ptr(), X.data_ptr(), H1.data_ptr(), H2.data_ptr(), Z.data_ptr(),
                                     None), "mlp forward")
    z = Z.view(B, 3).cpu().numpy()
    acc = float(np.mean((z[:, 1] > z[:, 0]).astype(np.int32) == y))
    return P.cpu().numpy(), losses, acc


TRAIN_SIZES = (8192, 16384, 32768)


def training_systems(count: int, seed: int = 1000, sizes=TRAIN_SIZES):
    """Seeded training systems: every third the diagonally dominant LCG
    matrix, the others randsvd with kappa log-uniform in [1e1, 1e6]; n drawn
    from `sizes`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(list(sizes)))
        if i % 3 == 0:
            out.append(("lcg", n, seed + i, 1.0))
        else:
            out.append(("randsvd", n, seed + i, float(10 ** rng.uniform(1, 6))))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned FP8/FP16 schedule on the GPU")
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--systems", type=int, default=400, help="most systems to label")
    ap.add_argument("--solves", type=int, default=2000,
                    help="stop labelling new systems once this many solves have run")
    ap.add_argument("--sizes", default=",".join(map(str, TRAIN_SIZES)))
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--slack", type=int, default=3)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--out", default=PARAMS_FILE)
    ap.add_argument("--summary", default=None, help="JSON summary of the data and the training")
    ap.add_argument("--data", default=None,
                    help="train on a labelled set saved by an earlier run (its *_data.npz) instead of labelling; "
                         "the summary's system list is kept from the JSON beside it")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    if args.data:
        z = np.load(args.data)
        F, y, t, m = z["F"], z["y"], z["t"], z["m"]
        prev = json.load(open(args.data.replace("_data.npz", ".json")))
        log(f"dataset {args.data}: {F.shape[0]} steps, {int(y.sum())} labelled FP8 (labelled earlier: "
            f"{prev.get('solves')} solves)")
        params, losses, acc = train(F, y, t, m, iters=args.iters, log=log)
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        params.astype("<f4").tofile(args.out)
        log(f"trained: accuracy {acc:.3f} on the training steps; parameters -> {args.out}")
        if args.summary:
            prev.update(losses=losses, train_accuracy=acc, iters=args.iters, train_seconds=time.time() - t0)
            with open(args.summary, "w") as f:
                json.dump(prev, f, indent=1)
        return
    rows = []
    summ = []
    solvers = {}
    nsolves = 0
    sizes = tuple(int(v) for v in args.sizes.split(","))
    skipped = 0
    for kind, n, seed, kappa in training_systems(args.systems, sizes=sizes):
        if sum(s.solves for s in solvers.values()) >= args.solves:
            break
        if n not in solvers:
            solvers[n] = Solver(n, args.b)
        sysm = System(kind, n, seed, kappa)
        lab = label_system(solvers[n], sysm, args.slack, log)
        del sysm
        if lab is None:
            skipped += 1
            continue
        F, y, t, m, info = lab
        rows.append((F, y, t, m))
        summ.append(info)
    nsolves = sum(s.solves for s in solvers.values())
    for s in solvers.values():
        s.close()
    F = np.concatenate([r[0] for r in rows])
    y = np.concatenate([r[1] for r in rows])
    t = np.concatenate([r[2] for r in rows])
    m = np.concatenate([r[3] for r in rows])
    log(f"dataset: {F.shape[0]} steps from {len(rows)} systems ({skipped} skipped: all-FP16 did not converge), "
        f"{nsolves} solves, {int(y.sum())} labelled FP8 ({time.time() - t0:.0f} s)")
    params, losses, acc = train(F, y, t, m, iters=args.iters, log=log)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    params.astype("<f4").tofile(args.out)
    log(f"trained: accuracy {acc:.3f} on the training steps; parameters -> {args.out}")
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(dict(systems=summ, samples=int(F.shape[0]), fp8_labels=int(y.sum()), losses=losses,
                           train_accuracy=acc, b=args.b, slack=args.slack, iters=args.iters, sizes=list(sizes),
                           solves=nsolves, skipped_systems=skipped, nrefl=8,
                           seconds=time.time() - t0), f, indent=1)
        np.savez(args.summary.replace(".json", "_data.npz"), F=F, y=y, t=t, m=m)


if __name__ == "__main__":
    main()
