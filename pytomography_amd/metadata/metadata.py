"""metadata/metadata.py:3-27 of the reference."""
from __future__ import annotations


class ObjectMeta:
    """Object-space metadata: voxel sizes ``dr`` (cm) and ``shape``."""

    def __init__(self, dr, shape) -> None:
        self.dr = dr
        self.dx, self.dy, self.dz = dr[0], dr[1], dr[2]
        self.shape = shape

    def __repr__(self):
        return "".join(f"{k} = {v}\n" for k, v in sorted(vars(self).items()))


class ProjMeta:
    """Projection-space metadata."""

    def __init__(self, angles) -> None:
        self.angles = angles
        self.num_projections = len(angles)

    def __repr__(self):
        return "".join(f"{k} = {v}\n" for k, v in sorted(vars(self).items()))
