"""Streaming patterns and ghost-layer periodicity (lbmpy advanced_streaming/__init__.py)."""
from .utility import (Timestep, streaming_patterns, is_inplace, get_accessor, get_timesteps, numeric_index,
                      numeric_offsets, inverse_dir_index, AccessPdfValues)
from .communication import get_communication_slices, LBMPeriodicityHandling, GpuArrays

__all__ = ["Timestep", "streaming_patterns", "is_inplace", "get_accessor", "get_timesteps", "numeric_index",
           "numeric_offsets", "inverse_dir_index", "AccessPdfValues", "get_communication_slices",
           "LBMPeriodicityHandling", "GpuArrays"]
