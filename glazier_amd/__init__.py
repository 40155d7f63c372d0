"""glazier-hip: a Cellular Potts engine for one AMD Instinct MI355X (gfx950).

The device path of glazier (the checkerboard Metropolis sweep, the field
stencil, per-label moments, division/death) as hand-written HIP kernels behind
a C ABI (include/glazier_hip.h). This package is the Python side: the model
and Blueprint reader, and the ctypes binding of the engine.
"""

from .model import CellType, Model, Species, default_cell_types, tiled_labels
from .engine import GhipError, HipBatch, HipSimulation
from .director import mid_plane

__all__ = [
    "CellType",
    "Model",
    "Species",
    "GhipError",
    "HipBatch",
    "HipSimulation",
    "default_cell_types",
    "mid_plane",
    "tiled_labels",
]
