"""MI355X-native adaptive GIMP substep (drop-in for g1n0st/adaptive-gimp's simulator API).

The compute path is libagimp.so (HIP kernels for gfx950, C-ABI in include/agimp.h); this
package is the Python host side mirroring ``adaptive_gimp.py``'s ``AdaptiveGIMP`` class.
"""
from . import render, scenes
from ._lib import AgimpError
from .sim import ACTIVATED, CULLING, GHOST, T_JUNCTION, UNACTIVATED, AdaptiveGIMP, Batch

__all__ = ["AdaptiveGIMP", "Batch", "AgimpError", "scenes", "render", "UNACTIVATED", "ACTIVATED", "GHOST", "T_JUNCTION", "CULLING"]
