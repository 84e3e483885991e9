"""Gated motion correction (projectors/shared/motion_correction_system_matrix.py:7-31):
forward H_n M_n per gate, backward sum_n M_n^T H_n^T."""
from __future__ import annotations

from ..system_matrix import ExtendedSystemMatrix


class MotionSystemMatrix(ExtendedSystemMatrix):
    def __init__(self, system_matrices, motion_transforms) -> None:
        super().__init__(system_matrices=system_matrices, obj2obj_transforms=motion_transforms)
