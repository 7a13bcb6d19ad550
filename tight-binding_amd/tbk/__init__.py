"""tbk -- MI355X-native NRL tight-binding k-point engine (host side).

The compute path lives in lib/libtbk.so (HIP kernels for gfx950 behind the
C ABI of include/tbk.h); this package marshals arrays and provides the
structure / parameter model and the seeded synthetic workloads.
"""
from . import structures
from .structures import Structure, NRLParams

__all__ = ["structures", "Structure", "NRLParams"]
