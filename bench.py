#!/usr/bin/env python3
"""LU + iterative-refinement throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 204800] [--b 2560]

One step is one full solve of the generated order-n system: the matrix and
right-hand side are generated on the device first (not timed, as in the
reference, main.cpp:529-554/:621), then the no-pivot LU with look-ahead and
the FP64 refinement run inside a timed region bracketed by device
synchronisations.  Throughput is (2/3 n^3 + 3/2 n^2) / time
(main.cpp:663).  Each step is timed separately because the factorisation
overwrites the matrix; a rate is steps * flops / sum of step times.

Two methods are measured on the same system:
  plain    FP16 Schur update, classical IR (the reference's method; the
           trailing matrix updated per pair of panels);
  learned  the MLP policy (openmxp_amd/data/sched_mlp.f32) picks FP8 or FP16
           per LU step, GMRES-IR, trailing update per group of four panels;
           if the refinement fails the system is generated again and
           factored in FP16, inside the timed region.
`value` is the learned method's rate; the line carries the plain rate and
the ratio.  Two attribution legs separate the ratio's parts (--attr-steps
timed solves each): plain FP16 with groups of four (the grouping gain) and
FP8 on every step with groups of four and GMRES-IR (the FP8 gain; the
learned leg over it is the policy's own gain).  `roofline` is that of the headline run's dominant kernel, the
FP8 trailing-update GEMM of the learned run (`roofline_plain_fp16`: the FP16
one of the plain run), each timed with HIP events on the stream its launches
run on, with the HBM bytes per launch from the committed rocprofv3 counter
summary.  The line also carries the forward errors against an FP64
refinement to the floor and the CPU oracle timed on this host as the
baseline.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP16_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
PEAK_FP8_TFLOPS = 5000.0   # MI355X dense FP8 MFMA
PMC_SUMMARY = os.path.join(ROOT, "profiles", "gemm_traffic_n204800.txt")


def hpl_flops(n: int) -> float:
    return 2.0 / 3.0 * n ** 3 + 1.5 * n ** 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=204800)
    ap.add_argument("--b", type=int, default=2560)
    ap.add_argument("--alt", type=int, default=1)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-learned", action="store_true", help="measure the plain method only")
    ap.add_argument("--group", type=int, choices=(2, 4), default=4,
                    help="panels per trailing update for the learned (FP8) solver")
    ap.add_argument("--group-plain", type=int, choices=(2, 4), default=2,
                    help="panels per trailing update for the plain FP16 solver")
    ap.add_argument("--attr-steps", type=int, default=2,
                    help="timed solves of each attribution leg (plain FP16 in groups of four; FP8 on every step)")
    ap.add_argument("--no-attr", action="store_true", help="skip the attribution legs")
    ap.add_argument("--cpu-n", type=int, default=10240, help="order of the CPU baseline sample")
    ap.add_argument("--cpu-b", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ir-kernels", action="store_true", help="skip the triangular-solve / residual timings")
    ap.add_argument("--randsvd-kappas", default="1e2,1e4,1e6",
                    help="condition numbers of the randsvd legs (comma separated; empty: none)")
    ap.add_argument("--no-randsvd-fp32", action="store_true",
                    help="skip the FP32-class solves of the randsvd legs beyond FP16 (kappa * 2^-11 >= 1)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (the reference kernel for PMC byte counters)")
    ap.add_argument("--gemm-sweep", action="store_true",
                    help="run only the trailing-update GEMMs of one factorisation, FP16 in pairs then FP8 in groups of four "
                         "on the default stream, for counter collection")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("TRACE_CSV", "LEGS"),
                    help="print per-kernel totals and per-method trailing-update rates from a rocprofv3 "
                         "--kernel-trace CSV of a bench run (LEGS: the solves in order, leg:group,...), and exit")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the Schur GEMM traffic from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                         "counter_collection files of a --copy-calibration run, and exit")
    ap.add_argument("--ir-sweep", action="store_true",
                    help="run only the refinement's kernels (lower and upper triangular solves, residual GEMV) on a "
                         "generated n x n matrix, three times each, print their timings, and exit (for counter "
                         "collection; the matrix stands in for the factors, so the values are not a solve)")
    ap.add_argument("--summarize-ir", nargs=2, metavar=("FETCH_CSV", "SQ_CSV"),
                    help="print the triangular solves' traffic and the residual GEMV's VALU figures from rocprofv3 "
                         "--pmc counter_collection files of an --ir-sweep run (FETCH_SIZE; IR_SQ_COUNTERS), and exit")
    ap.add_argument("--summarize-sq", metavar="SQ_CSV",
                    help="print clock, MFMA busy, waves per SIMD and wait fractions of the trailing-update GEMMs "
                         "from a rocprofv3 --pmc counter_collection file of a --gemm-sweep run, and exit")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class GpuSolve:
    def __init__(self, n, b, alt, fp8, policy=None, group=2):
        import numpy as np
        import torch
        from openmxp_amd import _lib
        self.torch = torch
        self.L = _lib.lib()
        self.n, self.b = n, b
        self.lda = n
        self.A = torch.empty(self.lda * n, dtype=torch.float32, device="cuda")
        self.d = torch.empty(n, dtype=torch.float64, device="cuda")
        self.rhs = torch.empty(n, dtype=torch.float64, device="cuda")
        self.x = torch.empty(n, dtype=torch.float64, device="cuda")
        self.h = C.c_void_p()
        flags = (_lib.SOLVER_FP8 if fp8 else 0) | (_lib.SOLVER_GROUP4 if group == 4 else 0)
        rc = self.L.omx_solver_create(n, b, alt, flags, C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"omx_solver_create failed: {rc}")
        if policy is not None:
            self.params = np.ascontiguousarray(policy, dtype=np.float32)
            assert self.L.omx_solver_set_policy(self.h, self.params.ctypes.data) == 0

    def generate(self, seed):
        L = self.L
        assert L.omx_matgen_f32(self.n, seed, self.A.data_ptr(), self.lda, self.d.data_ptr(), None) == 0
        assert L.omx_rhs_f64(self.n, seed, self.rhs.data_ptr(), None) == 0
        self.torch.cuda.synchronize()

    def _lu_ir(self, seed, sched, method):
        import numpy as np
        tim = np.zeros(5)
        out = np.zeros(4)
        trace = np.zeros(100)
        sp = None
        if sched is not None:
            s = np.ascontiguousarray(sched, dtype=np.int32)
            sp = s.ctypes.data
        t0 = time.perf_counter()
        info = self.L.omx_lu_factor(self.h, self.A.data_ptr(), self.lda, sp, tim.ctypes.data)
        t1 = time.perf_counter()
        rc = self.L.omx_ir_solve(self.h, self.A.data_ptr(), self.lda, seed, self.d.data_ptr(), self.rhs.data_ptr(),
                                 self.x.data_ptr(), method, 50, 50, 1e-10, 16.0, 0, out.ctypes.data,
                                 trace.ctypes.data)
        self.torch.cuda.synchronize()
        t2 = time.perf_counter()
        return info, rc, tim, out, t1 - t0, t2 - t1

    def solve(self, seed, sched=None, method=0, fallback=False):
        """One timed LU + IR (+ the FP16 fallback when asked and needed)."""
        import numpy as np
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        info, rc, tim, out, lu_s, ir_s = self._lu_ir(seed, sched, method)
        used = np.zeros(self.n // self.b, np.int32)
        self.L.omx_solver_last_schedule(self.h, used.ctypes.data, None)
        fell_back = False
        if fallback and (rc != 0 or info != 0 or not out[1] < 16.0):
            fell_back = True
            self.generate(seed)
            info, rc, tim, out, lu_s, ir_s = self._lu_ir(seed, np.zeros(self.n // self.b, np.int32), method)
        t = time.perf_counter() - t0
        return dict(seconds=t, lu_s=lu_s, ir_s=ir_s, info=int(info), ir_rc=int(rc), gemm_ms=tim[1],
                    gemm_launches=int(tim[2]), gemm_flops=tim[3], resid=out[0], harness=out[1],
                    ir_steps=int(out[2]), inner=int(out[3]), fp8_steps=int((used[:-1] > 0).sum()),
                    steps_total=int(used.size - 1), fell_back=fell_back)

    def forward_error(self, seed):
        """FP64 refinement of the current solution to the floor (3 more
        corrections, no stopping test): the large-n FP64 oracle."""
        import numpy as np
        xs = self.x.clone()
        out = np.zeros(4)
        trace = np.zeros(8)
        rc = self.L.omx_ir_solve(self.h, self.A.data_ptr(), self.lda, seed, self.d.data_ptr(), self.rhs.data_ptr(),
                                 self.x.data_ptr(), 0, 4, 50, 1e-10, 0.0, 1, out.ctypes.data, trace.ctypes.data)
        self.torch.cuda.synchronize()
        if rc != 0:
            return None
        fwd = float((xs - self.x).abs().max() / self.x.abs().max())
        self.x.copy_(xs)
        return fwd

    def close(self):
        self.L.omx_solver_destroy(self.h)
        del self.A


def measure(gs, args, label, sched=None, method=0, fallback=False):
    res = []
    for w in range(args.warmup):
        gs.generate(args.seed)
        r = gs.solve(args.seed, sched, method, fallback)
        log(f"[{label}] warmup {w}: {r['seconds']:.3f} s harness {r['harness']:.4f}")
    for k in range(args.steps):
        gs.generate(args.seed)
        r = gs.solve(args.seed, sched, method, fallback)
        log(f"[{label}] step {k}: {r['seconds']:.3f} s (LU {r['lu_s']:.3f}, IR {r['ir_s']:.3f}, "
            f"{r['ir_steps']} corrections, {r['inner']} inner) harness {r['harness']:.4f}, FP8 steps "
            f"{r['fp8_steps']}/{r['steps_total']}, trailing GEMM "
            f"{r['gemm_flops'] / (r['gemm_ms'] * 1e-3) * 1e-12 if r['gemm_ms'] else 0:.1f} TFLOP/s"
            + (" [fell back to FP16]" if r["fell_back"] else ""))
        res.append(r)
    fwd = gs.forward_error(args.seed)
    return res, fwd


def agg(res, flops):
    t = sum(r["seconds"] for r in res)
    gm = sum(r["gemm_ms"] for r in res)
    gf = sum(r["gemm_flops"] for r in res)
    gl = sum(r["gemm_launches"] for r in res)
    ok = all(r["ir_rc"] == 0 and r["harness"] < 16.0 and r["info"] == 0 for r in res)
    last = res[-1]
    return dict(tflops=len(res) * flops / t * 1e-12, ms=t / len(res) * 1e3, gemm_ms=gm, gemm_flops=gf,
                gemm_launches=gl, ok=ok, harness=last["harness"], resid=last["resid"], ir_steps=last["ir_steps"],
                inner=last["inner"], lu_s=last["lu_s"], ir_s=last["ir_s"], fp8_steps=last["fp8_steps"],
                steps_total=last["steps_total"], fell_back=any(r["fell_back"] for r in res))


def randsvd_table(n, b, alt, kappas, policy, seed, group=4, nrefl=8, fp32=True):
    """The conditioned systems: randsvd A = U diag(sigma) V^T (sigma
    geometric from 1 to 1/kappa, U and V products of nrefl seeded Householder
    reflectors) for each kappa, solved with the Schur update in FP16 on every
    step, FP8 on every step, and the learned schedule (when its refinement
    fails: FP16, unless the learned schedule was all FP16, then the FP32
    class), all with groups of `group` panels and GMRES-IR; where
    kappa * 2^-11 >= 1 (beyond FP16) also with every step in the FP32 class
    (schedule -1: FP32 trailing products on the panels as they stand).  The
    forward error is against the same solution refined on with three more
    GMRES-IR corrections to the FP64 floor (residuals from the FP64 low-rank
    description)."""
    import numpy as np
    import torch
    from openmxp_amd import _lib, scheduler
    L = _lib.lib()
    flops = hpl_flops(n)
    rows = []
    for kappa in kappas:
        sysm = scheduler.System("randsvd", n, seed, kappa=kappa, nrefl=nrefl)
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        for meth in ("fp16", "fp8", "learned", "fp32"):
            if meth == "learned" and policy is None:
                continue
            if meth == "fp32" and (not fp32 or kappa * 2.0**-11 < 1.0):
                continue
            h = C.c_void_p()
            flags = _lib.SOLVER_FP8 | (_lib.SOLVER_GROUP4 if group == 4 else 0)
            _lib.check(L.omx_solver_create(n, b, alt, flags, C.byref(h)), "solver create")
            if meth == "learned":
                pp = np.ascontiguousarray(policy, dtype=np.float32)
                _lib.check(L.omx_solver_set_policy(h, pp.ctypes.data), "set policy")
            nb = n // b
            d = sysm.desc()

            def lu_ir(sched):
                sp = None
                if sched is not None:
                    sa = np.ascontiguousarray(sched, dtype=np.int32)
                    sp = sa.ctypes.data
                tim = np.zeros(5)
                info = L.omx_lu_factor(h, sysm.A.data_ptr(), n, sp, tim.ctypes.data)
                out = np.zeros(4)
                rc = L.omx_ir_solve_system(h, sysm.A.data_ptr(), n, C.byref(d), sysm.rhs.data_ptr(), x.data_ptr(),
                                           1, 50, 50, 1e-10, 16.0, 0, out.ctypes.data, None)
                return info, rc, out, (sa if sched is not None else None)

            sched = {"fp16": [0] * nb, "fp8": [1] * nb, "learned": None, "fp32": [-1] * nb}[meth]
            sysm.generate()
            t0 = time.perf_counter()
            info, rc, out, _ = lu_ir(sched)
            torch.cuda.synchronize()
            used = np.zeros(nb, np.int32)
            L.omx_solver_last_schedule(h, used.ctypes.data, None)
            fp8_steps = int((used[:-1] > 0).sum())
            ok = info == 0 and rc == 0 and out[1] < 16.0
            fell_back = None
            if meth == "learned" and not ok:
                # the fallback ladder: FP16 (skipped when the learned
                # schedule had no FP8 step, i.e. was the FP16 schedule), then
                # the FP32 class
                for rung, sv in (("fp16", 0), ("fp32", -1)):
                    if ok or (rung == "fp16" and fp8_steps == 0):
                        continue
                    fell_back = rung
                    sysm.generate()
                    info, rc, out, _ = lu_ir([sv] * nb)
                    torch.cuda.synchronize()
                    ok = info == 0 and rc == 0 and out[1] < 16.0
            t = time.perf_counter() - t0
            fwd = None
            if info == 0:
                xs = x.clone()
                o2 = np.zeros(4)
                rc2 = L.omx_ir_solve_system(h, sysm.A.data_ptr(), n, C.byref(d), sysm.rhs.data_ptr(), x.data_ptr(),
                                            1, 3, 50, 1e-14, 0.0, 1, o2.ctypes.data, None)
                torch.cuda.synchronize()
                if rc2 == 0:
                    fwd = float((xs - x).abs().max() / x.abs().max())
            L.omx_solver_destroy(h)
            row = {"kappa": kappa, "method": meth, "converged": bool(ok), "seconds": round(t, 3),
                   "tflops": round(flops / t * 1e-12, 2), "harness": float(out[1]), "corrections": int(out[2]),
                   "inner_iters": int(out[3]), "forward_error": fwd, "fp8_steps": fp8_steps, "steps": nb - 1}
            if meth == "learned":
                row["fell_back_to"] = fell_back
            log(f"[randsvd kappa={kappa:g} {meth}] {t:.3f} s converged {ok} harness {out[1]:.4g} corrections "
                f"{int(out[2])} inner {int(out[3])} forward error {fwd} FP8 steps {fp8_steps}/{nb - 1}"
                + (f" [fell back to {fell_back}]" if fell_back else ""))
            rows.append(row)
        del sysm, x
        torch.cuda.empty_cache()
    return rows


PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
PEAK_FP64_VALU_TFLOPS = 78.6  # MI355X FP64 vector


def ir_kernels(gs, seed, reps=3):
    """The refinement's two HBM-bound triangular solves with the FP32 factors
    (2 n^2 bytes of factor read each: half the n x n FP32 matrix) and its
    VALU-bound residual GEMV (n^2 LCG values regenerated, n^2 FP64 FMAs),
    each timed with HIP events on the null stream the C ABI launches on,
    on the factored matrix in gs.A."""
    import torch
    L = gs.L
    n, b = gs.n, gs.b
    x = gs.rhs.clone()
    y = torch.empty_like(x)
    wb = L.omx_otf_gemv_workspace_bytes(n)
    w = torch.empty(max(wb, 8) // 8, dtype=torch.float64, device="cuda")

    def timed(fn):
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            assert fn() == 0
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return min(ts)

    tl = timed(lambda: L.omx_trsv_lower_unit_f32f64(n, b, gs.A.data_ptr(), gs.lda, x.data_ptr(), None, 0, None))
    tu = timed(lambda: L.omx_trsv_upper_f32f64(n, b, gs.A.data_ptr(), gs.lda, x.data_ptr(), None, 0, None))
    tg = timed(lambda: L.omx_otf_gemv_f64(n, seed, gs.d.data_ptr(), -1.0, x.data_ptr(), y.data_ptr(), w.data_ptr(),
                                          wb, None))
    tri = 2.0 * n * n
    out = {}
    for name, t in (("trsv_lower", tl), ("trsv_upper", tu)):
        gbs = tri / t * 1e-9
        out[name] = {"ms": round(t * 1e3, 3), "bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                     "bytes": tri, "launches": n // 256}
    tf = 2.0 * n * n / tg * 1e-12
    out["otf_gemv"] = {"ms": round(tg * 1e3, 3), "bound": "valu", "elements_per_s": round(n * n / tg, 1),
                       "fp64_tflops": round(tf, 2), "peak_fp64_valu_tflops": PEAK_FP64_VALU_TFLOPS,
                       "fp64_frac": round(tf / PEAK_FP64_VALU_TFLOPS, 4),
                       "note": "each element is one LCG step (64-bit integer multiply-add), an int64->FP64 "
                               "conversion and scaling, and one FP64 FMA: the FP64 fraction understates the "
                               "VALU issue rate (9 VALU instructions per element, SIMDs issuing VALU ~96 % of "
                               "their cycles: profiles/ir_kernels_n204800.txt)"}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, b, alt, seed):
    import oracle
    threads = oracle.get_threads()
    A, d = oracle.matgen_full(n, seed)
    rhs = oracle.rhs(n, seed)
    t0 = time.perf_counter()
    LU, info = oracle.lu_factor(A, b, alt=alt)
    res = oracle.ir_classic(LU, b, seed, d, rhs)
    t = time.perf_counter() - t0
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return dict(value=hpl_flops(n) / t * 1e-12, seconds=t, threads=threads, harness=res["harness"],
                iters=res["iters"], info=info, model=cpu_model(), nproc=affinity,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def gpu_same_sample(n, b, alt, seed):
    """The plain GPU method on the CPU baseline's sample (n, b): generate,
    then LU + classical IR timed as the bench's steps are."""
    gs = GpuSolve(n, b, alt, fp8=False, group=2)
    rates = []
    for _ in range(3):
        gs.generate(seed)
        r = gs.solve(seed)
        rates.append(hpl_flops(n) / r["seconds"] * 1e-12)
    gs.close()
    return dict(tflops=max(rates), harness=r["harness"], ir_steps=r["ir_steps"])


def pmc_traffic(tag, n, b, group):
    """HBM bytes per launch of the FP16 ("f16") or FP8 ("f8") trailing-update
    GEMM from the committed PMC summary (rocprofv3 FETCH_SIZE / WRITE_SIZE,
    copy-calibrated), if any, and only for the workload it was measured on
    (its n and b; FP16 in pairs, FP8 in groups of four)."""
    try:
        txt = open(PMC_SUMMARY).read()
    except OSError:
        return None
    if not re.search(rf"\bn={n} b={b}\b", txt) or group != {"f16": 2, "f8": 4}[tag]:
        return None
    m = re.search(tag + r"\.traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


SQ_COUNTERS = ("GRBM_GUI_ACTIVE", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_WAIT_ANY",
               "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS", "SQ_LDS_BANK_CONFLICT")


IR_SQ_COUNTERS = ("GRBM_GUI_ACTIVE", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU",
                  "SQ_INSTS_SALU", "SQ_WAIT_ANY", "SQ_ACTIVE_INST_ANY")


def summarize_ir(fetch_csv, sq_csv, n):
    """The refinement's kernels from the counter passes of an --ir-sweep run:
    the triangular solves' bytes fetched past L2 (FETCH_SIZE, times the copy
    calibration factor 2 of the trailing-update summary) against their 2 n^2
    bytes of factor reads per triangle; the residual GEMV's VALU figures:
    VALU instructions per matrix element, and the fraction of the SIMDs'
    cycles with a VALU instruction in flight (SQ_ACTIVE_INST_VALU counts
    per-wave cycles in quads: 4 x it / (GRBM_GUI_ACTIVE / 8 XCDs x 1024
    SIMDs))."""
    import collections
    import csv

    def per_kernel(path):
        out = collections.defaultdict(lambda: collections.defaultdict(float))
        durs = collections.defaultdict(dict)
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"]
            key = ("trsv_lower" if "k_trsv_lower" in name else "trsv_upper" if "k_trsv_upper" in name
                   else "otf_gemv" if "k_otf_gemv" in name else None)
            if key is None:
                continue
            out[key][r["Counter_Name"]] += float(r["Counter_Value"])
            durs[key][r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        return out, durs
    fc, fd = per_kernel(fetch_csv)
    tri = 2.0 * n * n
    for key in ("trsv_lower", "trsv_upper"):
        if key not in fc:
            continue
        nd = len(fd[key])
        solves = nd / (n // 256) if n >= 256 else nd
        fetched = fc[key]["FETCH_SIZE"] * 1024 * 2.0
        sec = sum(fd[key].values()) * 1e-9
        print(f"{key}: {nd} launches ({solves:.0f} solves), {sec * 1e3:.2f} ms in all, "
              f"{tri * solves / sec * 1e-9:.1f} GB/s of factor reads (8000 peak); "
              f"fetched past L2 {fetched / solves:.4e} bytes per solve = {fetched / (tri * solves):.3f} x the 2 n^2 "
              "factor bytes")
    sc, sd = per_kernel(sq_csv)
    if "otf_gemv" in sc:
        t = sc["otf_gemv"]
        nd = len(sd["otf_gemv"])
        sec = sum(sd["otf_gemv"].values()) * 1e-9
        g = t["GRBM_GUI_ACTIVE"]
        elems = float(n) * n * nd
        print(f"otf_gemv: {nd} launches, {sec * 1e3:.2f} ms in all, {elems / sec:.4e} elements/s; "
              f"clock {g / 8 / sec * 1e-9:.3f} GHz")
        for k in IR_SQ_COUNTERS:
            print(f"  {k:22s} {t[k]:.4e}")
        print(f"  VALU instructions per element {t['SQ_INSTS_VALU'] * 64 / elems:.2f} (per lane-element: "
              f"wave instructions x 64 lanes / elements); VALU busy {4 * t['SQ_ACTIVE_INST_VALU'] / (g / 8 * 1024):.3f} "
              f"of SIMD cycles; waves waiting {t['SQ_WAIT_ANY'] / t['SQ_WAVE_CYCLES']:.3f} of wave cycles")


def summarize_sq(sq_csv):
    """Shader-sequencer counters of the sweep's beta = 1 trailing-update GEMMs
    (one rocprofv3 --pmc pass with SQ_COUNTERS): per kernel family the sums
    and clock = GRBM_GUI_ACTIVE / 8 XCDs / kernel time, MFMA busy =
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 x 1024 SIMDs), waves per
    SIMD = 4 SQ_WAVE_CYCLES / (GRBM_GUI_ACTIVE / 8 x 1024), waiting =
    SQ_WAIT_ANY / SQ_WAVE_CYCLES (LDS: SQ_WAIT_INST_LDS / SQ_WAVE_CYCLES)."""
    import collections
    import csv
    rows = list(csv.DictReader(open(sq_csv)))
    for pre in ("k_gemm_f16<false", "k_gemm_f8<false"):
        disp = {}
        for r in rows:
            if pre not in r["Kernel_Name"]:
                continue
            d = disp.setdefault(r["Dispatch_Id"], {"dur": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        if not disp:
            continue
        tot = collections.Counter()
        dur = 0
        for d in disp.values():
            dur += d.pop("dur")
            tot.update(d)
        g = tot["GRBM_GUI_ACTIVE"]
        print(f"{pre}: {len(disp)} launches, {dur / 1e6:.1f} ms")
        for k in SQ_COUNTERS:
            print(f"  {k:26s} {tot[k]:.4e}")
        print(f"  clock {g / 8 / (dur * 1e-9) / 1e9:.3f} GHz; MFMA busy {tot['SQ_VALU_MFMA_BUSY_CYCLES'] / (g / 8 * 1024):.3f}; "
              f"waves/SIMD {4 * tot['SQ_WAVE_CYCLES'] / (g / 8 * 1024):.2f}; "
              f"waiting {tot['SQ_WAIT_ANY'] / tot['SQ_WAVE_CYCLES']:.3f} "
              f"(LDS {tot['SQ_WAIT_INST_LDS'] / tot['SQ_WAVE_CYCLES']:.3f}); "
              f"LDS issuing {tot['SQ_ACTIVE_INST_LDS'] / tot['SQ_WAVE_CYCLES']:.3f}")


def summarize_pmc(fetch_csv, write_csv, n, b):
    """Schur GEMM HBM traffic per launch (FP16 and FP8 trailing updates) from
    two rocprofv3 PMC runs of the GEMM sweep, each counter scaled by what it
    reads for the 1 GiB copy of the same run."""
    import csv

    def rows(path, counter):
        out = []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                out.append((r["Kernel_Name"], float(r["Counter_Value"]) * 1024.0))
        return out

    copy_bytes = float(1 << 30)
    cal = {}
    data = {}
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rs = rows(path, counter)
        c = [v for k, v in rs if "copyBuffer" in k]
        if not c:
            raise SystemExit(f"{path}: no copy kernel to calibrate {counter}")
        # the copy reads and writes 1 GiB each
        cal[counter] = copy_bytes / c[0]
        data[counter] = rs
    print(f"workload: the trailing updates of one factorisation, n={n} b={b}: FP16 in pairs (m x m x 2b, "
          f"m = n - (2p+4) b), FP8 in groups of four (m x m x 4b, m = n - (4g+8) b)")
    print(f"fetch_calibration_factor: {cal['FETCH_SIZE']:.4f}  (bytes per counted byte, 1 GiB copy)")
    print(f"write_calibration_factor: {cal['WRITE_SIZE']:.4f}")
    # the beta = 1 launches of either GEMM family (256 x 256 or 256 x 128 tiles)
    families = {"f16": ("k_gemm_f16<false", "k_gemm_half<false, false"),
                "f8": ("k_gemm_f8<false", "k_gemm_half<true, false")}
    for tag, eb in (("f16", 2.0), ("f8", 1.0)):
        names = families[tag]
        fetch = [v * cal["FETCH_SIZE"] for k, v in data["FETCH_SIZE"] if any(f in k for f in names)]
        write = [v * cal["WRITE_SIZE"] for k, v in data["WRITE_SIZE"] if any(f in k for f in names)]
        kname = " / ".join(names)
        if not fetch or not write:
            continue
        nl = len(fetch)
        G = SWEEP_GROUPS["fp16" if tag == "f16" else "fp8"]
        ms = bulk_shapes(n, b, G)
        # algorithmic bytes: FP32 C read + write, the two packed panels of K = G b
        alg = sum(8.0 * m * m + 2 * eb * m * G * b for m in ms)
        print(f"{tag}.kernel: {kname} (trailing update, beta = 1)")
        print(f"{tag}.launches: {nl} (algorithmic model: {len(ms)})")
        print(f"{tag}.fetch_bytes_total: {sum(fetch):.4e}")
        print(f"{tag}.write_bytes_total: {sum(write):.4e}")
        print(f"{tag}.algorithmic_bytes_total: {alg:.4e}  (8 m^2 for C + {2 * eb:.0f} m {G}b for the two panels)")
        print(f"{tag}.traffic_over_algorithmic: {(sum(fetch) + sum(write)) / alg:.3f}")
        print(f"{tag}.traffic_bytes_per_launch: {(sum(fetch) + sum(write)) / max(nl, 1):.6e}")


def summarize_trace(trace_csv, n, b, legs):
    """Per-kernel totals and the trailing-update GEMM rates per method from a
    rocprofv3 --kernel-trace CSV of a bench run.  The trace is split into the
    bench's solves at each matrix generation (one k_matgen_block dispatch
    per solve); `legs` names them in order with their group size, e.g.
    "plain:2,plain:2,g4:4,g4:4,learned:4,learned:4,fp8:4,fp8:4".  A
    dispatch of the beta = 1 GEMM is a trailing update when its grid is the
    square tile grid of m = n - G (g + 2) b for the solve's G (look-ahead
    products are m x b)."""
    import csv
    from collections import defaultdict
    rows = list(csv.DictReader(open(trace_csv)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    plan = [(l.split(":")[0], int(l.split(":")[1])) for l in legs.split(",")]
    tot = defaultdict(lambda: [0, 0.0])
    rate = defaultdict(lambda: [0, 0.0, 0.0])  # launches, flops, seconds
    seg = -1
    for r in rows:
        name = r["Kernel_Name"]
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
        bare = name.replace("void ", "").replace("(anonymous namespace)::", "")
        key = bare.split("(")[0]
        if key.startswith("_Z"):  # unmangled names only
            key = bare[:60]
        tot[key][0] += 1
        tot[key][1] += dur
        if "k_matgen_block" in name:
            seg += 1
            continue
        if seg < 0 or seg >= len(plan):
            continue
        leg, G = plan[seg]
        beta1 = ("k_gemm_f16<false" in name or "k_gemm_f8<false" in name or "k_gemm_half<false, false" in name
                 or "k_gemm_half<true, false" in name)
        if not beta1:
            continue
        wg = int(r["Grid_Size_X"]) // int(r["Workgroup_Size_X"])
        half = "k_gemm_half" in name
        for g in range(n // (G * b)):
            m = n - G * (g + 2) * b
            if m <= 0:
                break
            tiles = (m // 256) * (m // (128 if half else 256))
            if wg == tiles:
                prec = "fp8" if ("k_gemm_f8" in name or "k_gemm_half<true" in name) else "fp16"
                k = rate[(leg, prec)]
                k[0] += 1
                k[1] += 2.0 * m * m * G * b
                k[2] += dur
                break
    grand = sum(v[1] for v in tot.values())
    print(f"workload: bench solves at n={n} b={b}, in order: {legs}")
    print("trailing-update GEMM per method (per-dispatch durations from the trace):")
    for (leg, prec), (nl, fl, sec) in sorted(rate.items()):
        print(f"  {leg:8s} {prec}: {nl} launches, {fl:.4e} flops, {sec * 1e3:.1f} ms, {fl / sec * 1e-12:.1f} TFLOP/s")
    print("| kernel | calls | total ms | avg us | share |")
    print("|---|---:|---:|---:|---:|")
    for key, (c, sec) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:24]:
        print(f"| `{key}` | {c} | {sec * 1e3:.1f} | {sec / c * 1e6:.1f} | {100.0 * sec / grand:.1f}% |")


def bulk_shapes(n, b, G=2):
    """Row counts m of the trailing updates of one factorisation with panels
    in groups of G: group g updates A[f:, f:] with f = G (g + 2) b, depth
    K = G b."""
    nb = n // b
    out = []
    g = 0
    while G * (g + 1) < nb:
        if G * (g + 2) < nb:
            out.append(n - G * (g + 2) * b)
        g += 1
    return out


# the solvers the bench measures: plain FP16 in pairs, the learned FP8
# schedule in groups of four
SWEEP_GROUPS = {"fp16": 2, "fp8": 4}


def gemm_sweep(n, b, seed, precs=("fp16", "fp8")):
    """The trailing updates of one factorisation as the bench's solvers make
    them (FP16: pairs, m x m x 2b; FP8: groups of four, m x m x 4b), one
    launch each, on the null stream, with packed panels of the right size
    (random contents) and the generated matrix as C; prints the event-timed
    rates."""
    import torch
    from openmxp_amd import _lib
    L = _lib.lib()
    A = torch.empty(n * n, dtype=torch.float32, device="cuda")
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    assert L.omx_matgen_f32(n, seed, A.data_ptr(), n, d.data_ptr(), None) == 0
    for prec in precs:
        fp8 = prec == "fp8"
        G = SWEEP_GROUPS[prec]
        ms = bulk_shapes(n, b, G)
        k = G * b
        pb = L.omx_panel_bytes(n - b, k, 1 if fp8 else 0)
        if fp8:
            # e4m3 codes below 0x78 (no NaN), unit scales
            X = torch.randint(0, 0x78, (pb,), dtype=torch.uint8, device="cuda")
            Y = torch.randint(0, 0x78, (pb,), dtype=torch.uint8, device="cuda")
            ex = torch.zeros(L.omx_panel_exps_count(n - b, k), dtype=torch.int32, device="cuda")
        else:
            X = (torch.rand(pb // 2, device="cuda") - 0.5).half()
            Y = (torch.rand(pb // 2, device="cuda") - 0.5).half()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        flops = 0.0
        e0.record()
        for m in ms:
            f = n - m
            Cp = A.data_ptr() + 4 * f * (n + 1)
            if fp8:
                rc = L.omx_gemm_f8_packed(m, m, k, -1.0, X.data_ptr(), ex.data_ptr(), 0, Y.data_ptr(),
                                          ex.data_ptr(), 0, 1.0, Cp, n, None)
            else:
                rc = L.omx_gemm_f16_packed(m, m, k, -1.0, X.data_ptr(), 0, Y.data_ptr(), 0, 1.0, Cp, n, None)
            assert rc == 0
            flops += 2.0 * m * m * k
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3
        log(f"gemm sweep {prec} n={n} b={b} depth {k}: {len(ms)} launches, {t * 1e3:.1f} ms, "
            f"{flops / t * 1e-12:.1f} TFLOP/s")
        del X, Y


def main():
    args = parse()
    if args.gemm_sweep:
        if args.copy_calibration:
            import torch
            src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
            log("copied 1 GiB device to device (PMC calibration)")
        gemm_sweep(args.n, args.b, args.seed)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace[0], args.n, args.b, args.summarize_trace[1])
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], args.n, args.b)
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq)
        return
    if args.summarize_ir:
        summarize_ir(args.summarize_ir[0], args.summarize_ir[1], args.n)
        return
    if args.ir_sweep:
        gs = GpuSolve(args.n, args.b, args.alt, fp8=False, group=2)
        gs.generate(args.seed)
        print(json.dumps({"ir_sweep": ir_kernels(gs, args.seed), "n": args.n, "b": args.b}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        log("bench.py measures one MI355X: multi-GPU runs are not part of this build")
        sys.exit(2)
    n, b = args.n, args.b
    flops = hpl_flops(n)
    policy = None
    if not args.no_learned:
        from openmxp_amd import scheduler
        try:
            policy = scheduler.load_params()
        except (OSError, ValueError) as e:
            log(f"no trained policy ({e}); measuring the plain method only")
    if args.copy_calibration:
        import torch
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copied 1 GiB device to device (PMC calibration)")
    log(f"n={n} b={b} alt={args.alt}: generating and solving on cuda:0")
    gs = GpuSolve(n, b, args.alt, fp8=False, group=args.group_plain)
    res16, fwd16 = measure(gs, args, "plain fp16 + IR")
    irk = None
    if not args.no_ir_kernels:
        irk = ir_kernels(gs, args.seed)
        log("IR kernels: " + ", ".join(f"{k} {v['ms']} ms" for k, v in irk.items()))
    gs.close()
    del gs
    a16 = agg(res16, flops)
    aL = a8 = None
    fwdL = fwd8 = None
    attr = argparse.Namespace(**vars(args))
    attr.steps, attr.warmup = max(1, args.attr_steps), min(args.warmup, 1)
    a16g = fwd16g = None
    if not args.no_attr and args.group_plain != 4:
        gs = GpuSolve(n, b, args.alt, fp8=False, group=4)
        res16g, fwd16g = measure(gs, attr, "plain fp16 groups of 4 + IR", None, 0)
        gs.close()
        del gs
        a16g = agg(res16g, flops)
    if policy is not None or not args.no_attr:
        gs = GpuSolve(n, b, args.alt, fp8=True, policy=policy, group=args.group)
        if policy is not None:
            resL, fwdL = measure(gs, args, "learned + GMRES-IR", None, 1, True)
            aL = agg(resL, flops)
        if not args.no_attr:
            res8, fwd8 = measure(gs, attr, "fp8 every step + GMRES-IR", [1] * (n // b), 1)
            a8 = agg(res8, flops)
        gs.close()
        del gs
    rsvd = None
    kappas = [float(v) for v in args.randsvd_kappas.split(",") if v.strip()]
    if kappas:
        rsvd = randsvd_table(n, b, args.alt, kappas, policy, args.seed, group=args.group,
                             fp32=not args.no_randsvd_fp32)
    achieved = a16["gemm_flops"] / (a16["gemm_ms"] * 1e-3) * 1e-12
    cpu = gpu_small = None
    if not args.no_cpu:
        log(f"CPU baseline: oracle solve at n={args.cpu_n} b={args.cpu_b}")
        cpu = cpu_baseline(args.cpu_n, args.cpu_b, args.alt, args.seed)
        gpu_small = gpu_same_sample(args.cpu_n, args.cpu_b, args.alt, args.seed)
    head = aL if aL else a16
    roof16 = {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
              "frac": round(achieved / PEAK_FP16_TFLOPS, 4), "traffic": pmc_traffic("f16", n, b, args.group_plain),
              "kernel": f"FP16 Schur GEMM (omx_gemm_f16_packed_window) trailing update per {args.group_plain} panels (depth {args.group_plain}b, "
                        "alpha=-1, beta=1), plain run",
              "launches": a16["gemm_launches"]}
    line = {
        "metric": "LU+IR TFLOP/s @N=204800 1 GPU, learned FP8/FP16 vs plain FP16; resid<16",
        "value": round(head["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp8/fp16 Schur (fp32 acc) per learned step / fp32 panels / fp64 GMRES-IR" if aL else
                  "fp16 Schur (fp32 acc) / fp32 panels / fp64 IR"),
        "data": "synthetic: reference LCG generator, seed %d, diagonally dominant" % args.seed,
        "config": {"workload": f"HPL-MxP LU+IR n={n} b={b} 1x1 grid alt={args.alt}", "n": n, "b": b,
                   "method": ("learned per-step FP8/FP16 schedule + GMRES-IR(50), FP16 fallback" if aL else
                              "plain FP16 + classical IR")},
        "valid": head["ok"],
        "harness": head["harness"],
        "harness_eps": head["harness"] / 2.0,
        "resid": head["resid"],
        "ir_steps": head["ir_steps"],
        "forward_error": fwdL if aL else fwd16,
        "lu_s": round(head["lu_s"], 4),
        "ir_s": round(head["ir_s"], 4),
        "roofline": roof16,
        "roofline_plain_fp16": roof16,
        "plain_fp16": {"tflops": round(a16["tflops"], 3), "ms_per_step": round(a16["ms"], 2), "valid": a16["ok"],
                       "harness": a16["harness"], "ir_steps": a16["ir_steps"], "forward_error": fwd16,
                       "lu_s": round(a16["lu_s"], 4), "ir_s": round(a16["ir_s"], 4)},
    }
    if aL:
        achL = aL["gemm_flops"] / (aL["gemm_ms"] * 1e-3) * 1e-12
        # the headline run's dominant kernel: its trailing updates (FP8 where
        # the schedule chose FP8; all of them when fp8_steps == steps)
        line["roofline"] = {"bound": "mfma", "achieved": round(achL, 1), "peak": PEAK_FP8_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(achL / PEAK_FP8_TFLOPS, 4), "traffic": pmc_traffic("f8", n, b, args.group),
                            "kernel": f"FP8 Schur GEMM (omx_gemm_f8_packed_window) trailing update per {args.group} panels (depth {args.group}b, "
                                      "alpha=-1, beta=1), learned run",
                            "launches": aL["gemm_launches"],
                            "fp8_steps": aL["fp8_steps"], "steps": aL["steps_total"]}
        line["learned"] = {"speedup_vs_plain_fp16": round(aL["tflops"] / a16["tflops"], 3),
                           "fp8_steps": aL["fp8_steps"], "steps": aL["steps_total"], "inner_iters": aL["inner"],
                           "policy_equals_fp8_all": aL["fp8_steps"] == aL["steps_total"],
                           "fell_back_to_fp16": aL["fell_back"], "trailing_gemm_tflops": round(achL, 1),
                           "trailing_gemm_frac_fp8_peak": round(achL / PEAK_FP8_TFLOPS, 4)}
        if a16g:
            line["learned"]["speedup_vs_plain_fp16_g4"] = round(aL["tflops"] / a16g["tflops"], 3)
        if a8:
            line["learned"]["speedup_vs_fp8_all"] = round(aL["tflops"] / a8["tflops"], 3)
    if a16g:
        achg = a16g["gemm_flops"] / (a16g["gemm_ms"] * 1e-3) * 1e-12
        line["plain_fp16_g4"] = {"tflops": round(a16g["tflops"], 3), "ms_per_step": round(a16g["ms"], 2),
                                 "steps": attr.steps, "valid": a16g["ok"], "harness": a16g["harness"],
                                 "ir_steps": a16g["ir_steps"], "forward_error": fwd16g,
                                 "lu_s": round(a16g["lu_s"], 4), "ir_s": round(a16g["ir_s"], 4),
                                 "trailing_gemm_tflops": round(achg, 1),
                                 "trailing_gemm_frac_fp16_peak": round(achg / PEAK_FP16_TFLOPS, 4)}
    if a8:
        ach8 = a8["gemm_flops"] / (a8["gemm_ms"] * 1e-3) * 1e-12
        line["fp8_all"] = {"tflops": round(a8["tflops"], 3), "ms_per_step": round(a8["ms"], 2), "valid": a8["ok"],
                           "steps": attr.steps, "harness": a8["harness"], "ir_steps": a8["ir_steps"],
                           "inner_iters": a8["inner"], "forward_error": fwd8,
                           "speedup_vs_plain_fp16": round(a8["tflops"] / a16["tflops"], 3),
                           "gemm_tflops": round(ach8, 1), "gemm_frac_fp8_peak": round(ach8 / PEAK_FP8_TFLOPS, 4)}
    if aL and a16g and a8:
        # value / plain = (G4 / G2 plain) x (all-FP8 / plain G4) x (learned / all-FP8)
        line["attribution"] = {"grouping_gain": round(a16g["tflops"] / a16["tflops"], 3),
                               "fp8_gain": round(a8["tflops"] / a16g["tflops"], 3),
                               "policy_gain": round(aL["tflops"] / a8["tflops"], 3),
                               "note": "value / plain_fp16 = grouping_gain x fp8_gain x policy_gain; the two middle "
                                       "legs are timed over attr steps"}
    if irk:
        line["ir_kernels"] = irk
    if rsvd:
        line["randsvd"] = {"cases": rsvd, "n": n, "b": b, "group": args.group, "nrefl": 8, "ir": "GMRES-IR(50)",
                           "forward_error_vs": "the same solution refined on with 3 more FP64 GMRES-IR corrections"}
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["value"], 5), "unit": "TFLOP/s", "cores": cpu["threads"],
                                "kind": "port",
                                "sample": f"oracle (C/OpenMP) full LU+IR at n={args.cpu_n} b={args.cpu_b}, "
                                          f"{cpu['seconds']:.1f} s, harness {cpu['harness']:.3f}",
                                "cpu_model": cpu["model"], "nproc": cpu["nproc"],
                                "omp_num_threads": cpu["omp_num_threads"],
                                "gpu_same_sample_tflops": round(gpu_small["tflops"], 3),
                                "projected_cpu_seconds_at_n": {
                                    "n": n, "seconds": round(hpl_flops(n) / (cpu["value"] * 1e12), 1),
                                    "note": "projected: the measured CPU rate applied to the headline n, not run"}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
