"""camille_amd: MI355X-native linearized-Euler time advance of camille.

The compute path is the HIP library liblee_amd.so (C ABI in include/lee.h);
this package holds its sources (csrc/), the build script and the ctypes
binding that mirrors the reference's solver call sequence (lee.py).
"""
from .lee import FP32, FP64, GRID_RESIDENT, PATH_FUSED, PATH_PLAIN, Batch, LeeError, Solver, library, parse_bc_line, read_bc_file  # noqa: F401
