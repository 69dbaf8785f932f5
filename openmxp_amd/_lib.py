"""ctypes binding of libomxp.so (the C ABI declared in include/omxp.h).

This module only loads the native library and declares argument types; it
has no fallback.  If libomxp.so is missing the import of a function raises,
so a GPU run can never silently fall back to host code.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libomxp.so")

i64 = C.c_int64
u64 = C.c_uint64
i32 = C.c_int
dbl = C.c_double
flt = C.c_float
vp = C.c_void_p
sz = C.c_size_t

# name -> (restype, argtypes); kept in the order of include/omxp.h
SIGNATURES = {
    "omx_version": (C.c_char_p, []),
    "omx_matgen_f32": (i32, [i64, u64, vp, i64, vp, vp]),
    "omx_matgen_block_f32": (i32, [i64, u64, i64, i64, i64, i64, vp, i64, vp]),
    "omx_diag_f64": (i32, [i64, u64, i64, i64, vp, vp]),
    "omx_rhs_f64": (i32, [i64, u64, vp, vp]),
    "omx_otf_gemv_workspace_bytes": (sz, [i64]),
    "omx_otf_gemv_f64": (i32, [i64, u64, vp, dbl, vp, vp, vp, sz, vp]),
    "omx_otf_gemv_f64_noalloc": (i32, [i64, u64, vp, dbl, vp, vp, vp]),
    "omx_norm_inf_f64": (i32, [i64, vp, vp, vp]),
    "omx_copy_f64": (i32, [i64, vp, vp, vp]),
    "omx_div_f64": (i32, [i64, vp, vp, vp, vp]),
    "omx_axpy_f64": (i32, [i64, dbl, vp, vp, vp]),
    "omx_randsvd_factors_host": (i32, [i64, u64, dbl, i32, vp, vp, vp]),
    "omx_randsvd_factors": (i32, [i64, u64, dbl, i32, vp, vp, vp, vp]),
    "omx_randsvd_workspace_bytes": (sz, [i64]),
    "omx_randsvd_f32": (i32, [i64, i64, vp, vp, vp, vp, i64, vp, vp, vp, sz, vp]),
    "omx_higham_beta": (dbl, [i64, dbl, dbl]),
    "omx_higham_cond_est": (dbl, [i64, dbl, dbl]),
    "omx_higham_norm_inf": (dbl, [i64, dbl, dbl]),
    "omx_higham_matgen_f32": (i32, [i64, dbl, dbl, vp, i64, vp, vp]),
    "omx_higham_gemv_f64": (i32, [i64, dbl, dbl, dbl, vp, vp, vp]),
    "omx_lowrank_gemv_workspace_bytes": (sz, [i64]),
    "omx_lowrank_gemv_f64": (i32, [i64, i64, vp, vp, vp, dbl, vp, vp, vp, sz, vp]),
    "omx_trsv_workspace_bytes": (sz, [i64]),
    "omx_trsv_lower_unit_f32f64": (i32, [i64, i64, vp, i64, vp, vp, sz, vp]),
    "omx_trsv_upper_f32f64": (i32, [i64, i64, vp, i64, vp, vp, sz, vp]),
    "omx_panel_bytes": (sz, [i64, i64, i32]),
    "omx_panel_exps_count": (sz, [i64, i64]),
    "omx_pack_f16": (i32, [i64, i64, vp, i64, i64, i32, vp, vp, vp]),
    "omx_pack_f8": (i32, [i64, i64, vp, i64, i64, vp, vp, vp, vp]),
    "omx_pack_f16_scaled": (i32, [i64, i64, vp, i64, i64, i32, vp, vp, vp, vp]),
    "omx_pack_f8_ex": (i32, [i64, i64, vp, i64, i64, i32, i32, vp, vp, vp, vp]),
    "omx_pack_f16_window": (i32, [i64, i64, vp, i64, i64, i32, vp, vp, vp, i64, i64, vp]),
    "omx_pack_f8_window": (i32, [i64, i64, vp, i64, i64, i32, i32, vp, vp, i64, i64, vp, vp]),
    "omx_panel_tile_maxabs": (i32, [i64, i64, vp, i64, i64, vp, vp]),
    "omx_gemm_f16_packed": (i32, [i64, i64, i64, flt, vp, i64, vp, i64, flt, vp, i64, vp]),
    "omx_gemm_f16_packed_scaled": (i32, [i64, i64, i64, flt, vp, vp, i64, vp, vp, i64, flt, vp, i64, vp]),
    "omx_gemm_f16_packed_tri": (i32, [i64, i64, i64, flt, vp, vp, i32, flt, vp, i64, vp, vp]),
    "omx_gemm_f16_packed_window": (i32, [i64, i64, i64, flt, vp, vp, i64, i64, i64, vp, vp, i64, i64, i64, flt, vp,
                                         i64, vp]),
    "omx_gemm_f8_packed_window": (i32, [i64, i64, i64, flt, vp, vp, i64, i64, i64, vp, vp, i64, i64, i64, flt, vp,
                                        i64, vp]),
    "omx_gemm_f8_packed": (i32, [i64, i64, i64, flt, vp, vp, i64, vp, vp, i64, flt, vp, i64, vp]),
    "omx_sgemm_f32": (i32, [i64, i64, i64, flt, vp, i64, vp, i64, flt, vp, i64, vp]),
    "omx_getrf_nopiv_f32": (i32, [i64, vp, i64, vp, vp]),
    "omx_trsm_f32": (i32, [i32, i64, i64, vp, i64, vp, i64, vp]),
    "omx_trsm_identity_f32": (i32, [i32, i64, vp, i64, vp, i64, vp]),
    "omx_trtri_lu_f32": (i32, [i64, vp, i64, vp, vp, i64, vp, vp]),
    "omx_identity_f32": (i32, [i64, vp, i64, vp]),
    "omx_copy_block_f32": (i32, [i64, i64, vp, i64, vp, i64, vp]),
    "omx_mlp_forward_f32": (i32, [i64, vp, vp, vp, vp, vp, vp]),
    "omx_mlp_backward_workspace_bytes": (sz, [i64]),
    "omx_mlp_backward_f32": (i32, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
    "omx_mlp_sgd_f32": (i32, [i64, vp, vp, vp, flt, flt, vp]),
    "omx_mlp_adam_f32": (i32, [i64, vp, vp, vp, vp, flt, flt, flt, flt, i32, vp]),
    "omx_sched_features_f32": (i32, [i64, vp, i64, vp, vp, i64, i64, i64, i64, i64, i32, vp, vp]),
    "omx_solver_create": (i32, [i64, i64, i32, i32, C.POINTER(vp)]),
    "omx_solver_destroy": (i32, [vp]),
    "omx_solver_reserve_gmres": (i32, [vp, i32]),
    "omx_solver_set_skip": (i32, [vp, i32]),
    "omx_solver_step_times": (i32, [vp, vp]),
    "omx_lu_factor": (i32, [vp, vp, i64, vp, vp]),
    "omx_solver_set_policy": (i32, [vp, vp]),
    "omx_solver_last_schedule": (i32, [vp, vp, vp]),
    "omx_solver_timers": (i32, [vp, vp, vp]),
    "omx_timer_name": (C.c_char_p, [i32]),
    "omx_ir_solve": (i32, [vp, vp, i64, u64, vp, vp, vp, i32, i32, i32, dbl, dbl, i32, vp, vp]),
    "omx_ir_solve_system": (i32, [vp, vp, i64, vp, vp, vp, i32, i32, i32, dbl, dbl, i32, vp, vp]),
    "omx_harness": (dbl, [dbl, dbl, dbl, dbl, i64]),
}



class System(C.Structure):
    """struct omx_system (include/omxp.h)."""
    _fields_ = [("kind", C.c_int), ("seed", C.c_uint64), ("diag", vp), ("norm_inf", C.c_double),
                ("rank", C.c_int64), ("sigma", vp), ("P", vp), ("Q", vp), ("alpha", C.c_double),
                ("beta", C.c_double)]


SOLVER_FP8 = 1      # omx_solver_create flags
SOLVER_SCALE16 = 2
SOLVER_GROUP4 = 4
SOLVER_TIMERS = 8
TIMER_ITEMS = 23

SYSTEM_LCG = 0
SYSTEM_LOWRANK = 1
SYSTEM_HIGHAM = 2

_lib = None


class OmxError(RuntimeError):
    pass


def lib():
    """Load libomxp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OmxError(
                f"{LIB_PATH} is missing: build it with `python -m openmxp_amd.build` "
                "(the HIP path has no fallback)")
        # torch (device memory and streams in the tests and the benchmark)
        # ships its own HIP runtime; importing it first makes this library
        # bind to the same runtime instead of loading a second one
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        missing = [name for name in SIGNATURES if not hasattr(L, name)]
        if missing:
            raise OmxError(f"{LIB_PATH} does not export {', '.join(missing)}: it is older than "
                           "include/omxp.h; rebuild it with `python -m openmxp_amd.build`")
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(SIGNATURES)


def check(rc: int, what: str = "omx call") -> None:
    if rc != 0:
        raise OmxError(f"{what} returned {rc}")
