"""utils/misc.py:21-38 of the reference."""
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
