"""CPU oracle for the mixed-precision no-pivot LU + iterative refinement path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or as the timed CPU baseline.  The product (``openmxp_amd``) never
imports it.

``liboracle.so`` (built from ``omx_oracle.c`` by ``oracle/Makefile``) holds the
C restatement; ``_ref/libref.so`` holds routines of the reference compiled from
its own sources (built only where the reference tree is present).
"""
from .oracle import *  # noqa: F401,F403
