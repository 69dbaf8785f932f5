"""omxp: MI355X (gfx950) mixed-precision dense solver (HPL-MxP style).

The product is the C ABI in include/omxp.h, implemented by hand-written HIP
kernels and a C++ host driver in openmxp_amd/csrc (libomxp.so, bin/omxp).
This Python package only loads that library (``openmxp_amd._lib``) and wraps
it for tests and the benchmark (``openmxp_amd.api``).
"""
from ._lib import OmxError, check, lib  # noqa: F401

__all__ = ["lib", "check", "OmxError"]
