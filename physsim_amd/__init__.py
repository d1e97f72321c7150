"""gfx950 (MI355X) backend for the per-substep physics core of Geon (a9lim/physsim).

The product is libgeon_hip.so (C ABI in include/geon.h).  This package holds its
HIP sources (csrc/), the in-tree build (build.py), the ctypes binding and the
`HipPhysics` host class (geon.py), and seeded workload generators (workloads.py).
"""
from .geon import HipPhysics, HipBatch, GeonError, default_params, load_library  # noqa: F401
from . import _abi as abi  # noqa: F401
