"""Geometry helpers of the SPECT path (utils/spatial.py:36-119, utils/misc.py:21-38)."""
from .spatial import (compute_pad_size, compute_pad_size_padded, pad_object, unpad_object,
                      pad_proj, unpad_proj)
from .misc import get_distance

__all__ = ["compute_pad_size", "compute_pad_size_padded", "pad_object", "unpad_object",
           "pad_proj", "unpad_proj", "get_distance"]
