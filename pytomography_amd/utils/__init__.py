"""Geometry helpers of the SPECT path (utils/spatial.py:36-119, utils/misc.py:21-38)
and the energy-window scatter estimate (utils/scatter.py:10-99)."""
from .spatial import (compute_pad_size, compute_pad_size_padded, pad_object, unpad_object,
                      pad_proj, unpad_proj)
from .misc import get_distance, print_collimator_parameters
from .scatter import compute_EW_scatter, get_smoothed_scatter, get_1d_gaussian_taps

__all__ = ["compute_pad_size", "compute_pad_size_padded", "pad_object", "unpad_object",
           "pad_proj", "unpad_proj", "get_distance", "print_collimator_parameters", "compute_EW_scatter",
           "get_smoothed_scatter", "get_1d_gaussian_taps"]
