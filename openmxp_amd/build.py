"""Build the omxp HIP library (libomxp.so) and the `omxp` driver for gfx950.

    python -m openmxp_amd.build          # compile everything in-tree

Objects go to openmxp_amd/_build/, the library to openmxp_amd/libomxp.so and
the command-line driver to openmxp_amd/bin/omxp.  Every HIP source is
compiled with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libomxp.so")
BIN_DIR = os.path.join(PKG, "bin")
DRIVER = os.path.join(BIN_DIR, "omxp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

LIB_SOURCES = [
    "omx_matgen.hip",
    "omx_ir.hip",
    "omx_pack.hip",
    "omx_gemm.hip",
    "omx_dense.hip",
    "omx_randsvd.hip",
    "omx_mlp.hip",
    "omx_higham.hip",
    "omx_solver.cpp",
]
DRIVER_SOURCES = ["omxp_main.cpp"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *lang, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile every source (always) and link libomxp.so and bin/omxp."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    drv = [s for s in DRIVER_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if drv:
        dobjs = [_compile(s) for s in drv]
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", DRIVER, *dobjs, "-L", PKG, "-lomxp",
               f"-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver link failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}" + (f" and {DRIVER}" if drv else ""))
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
