"""The C ABI library loads and exports every symbol include/omxp.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "omxp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(omx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "omx_matgen_f32" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    from openmxp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libomxp.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from openmxp_amd import _lib
    assert sorted(_lib.declared_symbols()) == _declared()


def test_version_and_host_harness():
    from openmxp_amd import _lib
    L = _lib.lib()
    assert L.omx_version().startswith(b"omxp")
    assert L.omx_harness(1e-10, 2.0, 3.0, 0.5, 1000) == 1e-10 / 6.5 / (1000 * 2.0**-53)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from openmxp_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.OmxError):
        _lib.lib()


# the reference's Timer item names in enumeration order (timer.hpp:57-81,
# :120-145)
REFERENCE_TIMER_ITEMS = [
    "DIAG_BCAST", "LCOL_BCAST", "RROW_BCAST", "TEST", "WAIT", "DIAG_LU", "TRSM_L", "TRSM_R", "CONV_L", "CONV_R",
    "GEMM_UPDATE", "GEMM_PROGRESS", "LAZY_INIT", "WRITE_BACK", "IR_GEMV", "IR_GEMV_COMM", "IR_TRSV", "IR_TRSV_MV",
    "IR_TRSV_COMM", "INIT", "MEMCPY", "TOTAL", "MISC"]


def test_timer_items_are_the_references():
    from openmxp_amd import _lib
    L = _lib.lib()
    assert _lib.TIMER_ITEMS == len(REFERENCE_TIMER_ITEMS)
    got = [L.omx_timer_name(i).decode() for i in range(_lib.TIMER_ITEMS)]
    assert got == REFERENCE_TIMER_ITEMS
    assert L.omx_timer_name(-1) is None and L.omx_timer_name(_lib.TIMER_ITEMS) is None
    # no solver, no timers
    buf = (ctypes.c_double * _lib.TIMER_ITEMS)()
    assert L.omx_solver_timers(None, buf, buf) < 0



def test_triangular_product_rejects_bad_arguments_before_any_launch():
    """omx_gemm_f16_packed_tri refuses an unknown tri and tile maxima with
    beta != 0 before it touches the GPU (so this runs without one)."""
    from openmxp_amd import _lib
    L = _lib.lib()
    p = 16  # never dereferenced: the checks return first
    assert L.omx_gemm_f16_packed_tri(256, 256, 256, 1.0, p, p, 3, 0.0, p, 256, None, None) == -1
    assert L.omx_gemm_f16_packed_tri(256, 256, 256, 1.0, p, p, -1, 0.0, p, 256, None, None) == -1
    assert L.omx_gemm_f16_packed_tri(256, 256, 256, 1.0, p, p, 1, 1.0, p, 256, p, None) == -1
