"""Padding helpers with the reference's semantics (utils/spatial.py:36-119).

The kernels never materialise padded volumes (padding is folded into their
gather bounds); these functions exist for callers that work with padded
tensors directly, as the reference's users do.
"""
from __future__ import annotations

import math

import torch
from torch.nn.functional import pad


def compute_pad_size(width: int) -> int:
    """Pixels of padding per side so that a rotation keeps the whole object:
    ceil((sqrt(2) - 1) * width / 2)  (utils/spatial.py:36-45)."""
    return int(math.ceil((math.sqrt(2) * width - width) / 2))


def compute_pad_size_padded(width: int) -> int:
    """Inverse of compute_pad_size from a padded width (utils/spatial.py:47-62)."""
    a = (math.sqrt(2) - 1) / 2
    if width % 2 == 0:
        width_old = int(2 * math.floor((width / 2) / (1 + 2 * a)))
    else:
        width_old = int(2 * math.floor(((width - 1) / 2) / (1 + 2 * a)))
    return int((width - width_old) / 2)


def pad_object(object: torch.Tensor, mode: str = "constant") -> torch.Tensor:
    """Zero-pad x and y (utils/spatial.py:64-81); 'back_project' replicates along x."""
    p = compute_pad_size(object.shape[-2])
    if mode == "back_project":
        object = pad(object.unsqueeze(0), [0, 0, 0, 0, p, p], mode="replicate").squeeze(0)
        return pad(object, [0, 0, p, p], mode="constant")
    return pad(object, [0, 0, p, p, p, p], mode=mode)


def unpad_object(object: torch.Tensor) -> torch.Tensor:
    """utils/spatial.py:83-93."""
    p = compute_pad_size_padded(object.shape[-2])
    return object[p:-p, p:-p, :]


def pad_proj(proj: torch.Tensor, mode: str = "constant", value: float = 0) -> torch.Tensor:
    """utils/spatial.py:95-107."""
    p = compute_pad_size(proj.shape[-2])
    return pad(proj, [0, 0, p, p], mode=mode, value=value)


def unpad_proj(proj: torch.Tensor) -> torch.Tensor:
    """utils/spatial.py:109-119."""
    p = compute_pad_size_padded(proj.shape[-2])
    return proj[:, p:-p, :]
