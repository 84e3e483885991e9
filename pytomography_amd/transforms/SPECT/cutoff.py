"""Projection mask transform (transforms/SPECT/cutoff.py:5-20)."""
from __future__ import annotations

import torch

from pytomography_amd.utils import pad_proj
from ..transform import Transform


class CutOffTransform(Transform):
    """proj2proj transform that multiplies the padded projections by a mask.

    Args:
        mask: [Ltheta, Lr, Lz] (one mask per projection) or [Lr, Lz] (every
            projection), unpadded; padded with zeros along Lr as the reference
            does (cutoff.py:7).
    """

    def __init__(self, mask: torch.Tensor) -> None:
        super().__init__()
        self.mask = pad_proj(mask)

    def _mask(self, proj: torch.Tensor, ang_idx) -> torch.Tensor:
        mask = self.mask.to(device=proj.device, dtype=proj.dtype)
        if mask.dim() == 3 and ang_idx is not None and mask.shape[0] != proj.shape[0]:
            # the projections of a subset: the mask rows of its angles
            mask = mask[torch.as_tensor(ang_idx, device=mask.device).reshape(-1)]
        return mask

    @torch.no_grad()
    def forward(self, proj: torch.Tensor, ang_idx=None) -> torch.Tensor:
        """cutoff.py:9-14: proj * mask."""
        return proj * self._mask(proj, ang_idx)

    @torch.no_grad()
    def backward(self, proj: torch.Tensor, ang_idx=None) -> torch.Tensor:
        """cutoff.py:15-20: the mask is diagonal, so the same product."""
        return proj * self._mask(proj, ang_idx)
