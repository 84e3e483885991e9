"""pytomography_amd: MI355X-native drop-in for PyTomography's SPECT system-matrix path.

Module globals mirror ``pytomography/__init__.py:15-40`` (``device``,
``dtype``, ``delta``, ``verbose`` and their setters).  The compute path is the
HIP library ``libspect_hip.so`` (see ``pytomography_amd._lib``); there is no
CPU fallback.
"""
from __future__ import annotations

import torch

__version__ = "0.1.0"

device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
dtype = torch.float32
delta = 1e-11
verbose = False


def set_dtype(dt: torch.dtype) -> None:
    """pytomography/__init__.py:24-32 (also sets torch's default dtype)."""
    global dtype, delta
    dtype = dt
    torch.set_default_dtype(dt)
    if dt == torch.float16:
        delta = 1e-5
    elif dt == torch.float32:
        delta = 1e-11


def set_device(d) -> None:
    global device
    device = torch.device(d)


def set_verbose(b: bool) -> None:
    global verbose
    verbose = b
