"""utils/misc.py:21-66 of the reference."""
from __future__ import annotations

import numpy as np


def get_distance(Lx: int, r: float, dx: float) -> np.ndarray:
    """Distance (cm) from each plane parallel to the detector to a detector at
    +x at radius r, clipped at 0 (utils/misc.py:21-38)."""
    if Lx % 2 == 0:
        d = r + (Lx / 2 - np.arange(Lx) - 1 / 2) * dx
    else:
        d = r + ((Lx - 1) / 2 - np.arange(Lx)) * dx
    d[d < 0] = 0
    return d


def get_object_nearest_neighbour(object, shifts):
    """The object shifted by ``shifts`` voxels with replicate-clamped edges:
    out(v) = object(clamp(v + shifts)) (utils/misc.py:40-57).  Host/torch
    helper for transform set-up; the per-iteration stencils are HIP kernels."""
    import torch
    import torch.nn.functional as F
    shift_max = int(max(abs(int(s)) for s in shifts))
    if shift_max == 0:
        return object
    nb = F.pad(object.unsqueeze(0), 6 * [shift_max], mode="replicate").squeeze(0)
    nb = torch.roll(nb, shifts=[-int(s) for s in shifts], dims=(0, 1, 2))
    return nb[shift_max:-shift_max, shift_max:-shift_max, shift_max:-shift_max]


def print_collimator_parameters():
    """Prints every line of the shipped collimator database, whose records
    ``io.SPECT.dicom.get_psfmeta_from_scanner_params`` looks up
    (utils/misc.py:59-66)."""
    from pytomography_amd.io.SPECT.dicom import COLLIMATOR_FILE
    with open(COLLIMATOR_FILE) as f:
        for line in f.readlines():
            print(line)
