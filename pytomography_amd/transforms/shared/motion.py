"""Deformation-vector-field motion transform (transforms/shared/motion.py:7-110):
the forward and backward warps are one HIP trilinear-sampling kernel each
(``spect_dvf_warp``); the volume ratio is 1 as in the reference."""
from __future__ import annotations

import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.transforms import Transform


class DVFMotionTransform(Transform):
    def __init__(self, dvf_forward: torch.Tensor | None = None,
                 dvf_backward: torch.Tensor | None = None) -> None:
        def place(d):
            if d is None:
                return None
            return d.to(pytomography_amd.device).to(pytomography_amd.dtype).contiguous()
        self.dvf_forward = place(dvf_forward)
        self.dvf_backward = place(dvf_backward)
        self.dvf_forward_vol_ratio = 1
        self.dvf_backward_vol_ratio = 1
        super().__init__()

    def _apply_dvf(self, DVF: torch.Tensor, vol_ratio, object_i: torch.Tensor) -> torch.Tensor:
        if object_i.dim() != 3 or tuple(DVF.shape) != (*object_i.shape, 3):
            raise ValueError(f"DVF shape {tuple(DVF.shape)} does not match object {tuple(object_i.shape)}")
        dt = DVF.dtype
        f = _lib.require_device_tensor(object_i, "object", dt)
        out = torch.empty_like(f)
        nx, ny, nz = f.shape
        _lib.check(_lib.lib().spect_dvf_warp(_lib.dtype_code(dt), f.data_ptr(), DVF.data_ptr(),
                                             out.data_ptr(), nx, ny, nz, _lib.stream_ptr(f.device)))
        return out if vol_ratio == 1 else out * vol_ratio

    def forward(self, object_i: torch.Tensor) -> torch.Tensor:
        if self.dvf_forward is None:
            return object_i
        return self._apply_dvf(self.dvf_forward, self.dvf_forward_vol_ratio, object_i)

    def backward(self, object_i: torch.Tensor) -> torch.Tensor:
        if self.dvf_backward is None:
            return object_i
        return self._apply_dvf(self.dvf_backward, self.dvf_backward_vol_ratio, object_i)
