"""MI355X-native D3Q19 lattice-Boltzmann step and STL voxelizer.

The compute path is libfluidsim.so (hand-written HIP for gfx950, C ABI in
include/fluidsim.h); this package is its Python front end: ctypes binding
(``capi``), solver handles (``solver``), workload builders (``workloads``).
"""
from .capi import BGK, HOME, HYBRID, FLAG_BAND, FLAG_INLET, FLAG_OUTLET, FLAG_SOLID  # noqa: F401
from .solver import Adjoint, Batch, DeviceBuffer, Solver, band, load_stl, voxelize  # noqa: F401

__version__ = "0.1.0"
