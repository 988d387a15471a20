"""qpad_amd -- MI355X (gfx950) implementation of QPAD's per-xi-slice plasma
response and per-3D-step beam push, behind QPAD's simulation / species2d /
field / beam3d interfaces.

Compute runs only in libqpad_hip.so (HIP, FP64); see include/qpad_hip.h for
the C ABI.  Simulation() raises QpadError if the library is not built.
"""
from . import configs  # noqa: F401
from ._lib import LIB_PATH, QpadError, load  # noqa: F401
from .simulation import Batch, Beam3D, Field, SeedNet, Simulation, Species2D  # noqa: F401

__all__ = ["Simulation", "SeedNet", "Batch", "Species2D", "Beam3D", "Field", "QpadError", "load", "configs", "LIB_PATH"]
