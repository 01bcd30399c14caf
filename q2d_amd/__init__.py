"""MI355X-native explicit finite-volume step for Q2D (drop-in for Explicit_Step).

The compute path is the HIP library q2d_amd/lib/libq2d_gpu.so behind the C-ABI
of include/q2d_gpu.h; this package is its Python host (ctypes binding, setup-time
geometry, seeded workloads, batching and serving loops).
"""
from . import capi, cases, geometry  # noqa: F401
from .capi import GpuSolver, Q2DError, load_library  # noqa: F401

__version__ = "0.1.0"
