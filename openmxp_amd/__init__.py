"""omxp: MI355X (gfx950) mixed-precision dense solver (HPL-MxP style).

The product is the C ABI in include/omxp.h, implemented by hand-written HIP
kernels and a C++ host driver in openmxp_amd/csrc (libomxp.so, bin/omxp).
This Python package loads that library through ctypes (``openmxp_amd._lib``,
used by the tests and the benchmark), builds it (``openmxp_amd.build``) and
trains the learned precision schedule (``openmxp_amd.scheduler``).
"""
from ._lib import OmxError, check, lib  # noqa: F401

__all__ = ["lib", "check", "OmxError"]
