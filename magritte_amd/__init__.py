"""MI355X (gfx950) radiation-field and level-population library.

A drop-in for the hot path of UCL/Magritte's Simulation API:
compute_radiation_field, compute_Jeff, compute_level_populations_from_stateq
and compute_level_populations, running hand-written HIP kernels through the
C-ABI of libmagritte_hip.so (include/magritte_hip.h).
"""
from .model import Model  # noqa: F401

__all__ = ["Model"]
