"""wake_amd -- MI355X-native lattice-Boltzmann time step of wake.

Hand-written HIP kernels for gfx950 behind the reference's C ABI (the ISPC
kernel exports, `wake_amd.kernels`) and its grid API (`LatticeGrid2D`,
`LatticeGrid3D`).  The compute runs only through libwake_hip.so; importing
the package fails if the library has not been built.
"""
from . import _lib
from .batch import Batch
from .grid import (CS2, DEFAULT_SMAGORINSKY_CONSTANT, RHO_0, CellType, CollisionType, LatticeGrid2D,
                   LatticeGrid3D, Storage)
from .thermal import ThermalSolver2D

_lib.lib()  # fail loudly at import when the native library is missing

kernels = _lib

__all__ = [
    "Batch", "CellType", "CollisionType", "LatticeGrid2D", "LatticeGrid3D", "Storage", "RHO_0", "CS2",
    "DEFAULT_SMAGORINSKY_CONSTANT", "ThermalSolver2D", "kernels",
]
