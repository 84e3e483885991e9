"""Depth-dependent Gaussian collimator PSF (transforms/SPECT/psf.py:101-231).

The per-plane taps are built on the host exactly as the reference builds its
Conv1d weights (psf.py:62-83, 123-142, 170-197); the blur itself runs in the
fused HIP kernels with the taps staged per plane.
"""
from __future__ import annotations

import numpy as np
import torch

import pytomography_amd
from pytomography_amd.utils import compute_pad_size, get_distance
from ..transform import Transform


def gaussian_taps(sigmas: np.ndarray, kernel_size: int, dtype: torch.dtype) -> torch.Tensor:
    """Normalised Gaussian taps per plane, [len(sigmas), kernel_size], computed
    like get_1D_PSF_layer_gaussian (psf.py:62-83): exp(-x^2/(2 sigma^2 + delta))
    in ``dtype`` from float64 sigmas, divided by the per-plane sum."""
    n = len(sigmas)
    x = torch.arange(-int(kernel_size // 2), int(kernel_size // 2) + 1).unsqueeze(0).unsqueeze(0)
    x = x.repeat((n, 1, 1))
    s = torch.tensor(np.asarray(sigmas)).to(dtype).reshape((n, 1, 1))
    kern = torch.exp(-x ** 2 / (2 * s ** 2 + pytomography_amd.delta))
    kern = kern / kern.sum(axis=-1).unsqueeze(-1)
    return kern.to(dtype)[:, 0, :]


class SPECTPSFTransform(Transform):
    """obj2obj PSF transform configured from ``SPECTPSFMeta`` (Gaussian shape)."""

    def __init__(self, psf_meta=None, psf_operator=None, assume_padded: bool = True) -> None:
        super().__init__()
        if psf_operator is not None:
            raise NotImplementedError("psf_operator networks are not supported by the fused kernels")
        if psf_meta is None:
            raise ValueError("psf_meta is required")
        if psf_meta.shape != "gaussian":
            raise NotImplementedError("only the 'gaussian' PSF shape is supported")
        if not assume_padded:
            raise NotImplementedError("assume_padded=False is not supported by the fused kernels")
        self.psf_meta = psf_meta
        self.psf_operator = None
        self.assume_padded = assume_padded

    def configure(self, object_meta, proj_meta) -> None:
        super().configure(object_meta, proj_meta)
        if proj_meta.radii is None:
            raise ValueError("PSF modelling needs proj_meta.radii")
        self._tables = {}

    def _get_sigma(self, radius: float) -> np.ndarray:
        """psf.py:180-197 (assume_padded): sigma per padded x plane, float64."""
        dim = self.object_meta.shape[0] + 2 * compute_pad_size(self.object_meta.shape[0])
        distances = get_distance(dim, radius, self.object_meta.dx)
        return self.psf_meta.sigma_fit(distances, *self.psf_meta.sigma_fit_params)

    def _compute_kernel_size(self, radius: float, axis: int) -> int:
        """psf.py:170-178."""
        sigma_max = max(self._get_sigma(radius)) / self.object_meta.dr[axis]
        return int((np.ceil(sigma_max * self.psf_meta.min_sigmas) * 2 + 1).astype(int))

    def radius_taps(self, radius: float, dtype: torch.dtype):
        """(taps_r [P, Kr], taps_z [P, Kz] or None) for one detector radius."""
        key = (float(radius), dtype)
        if key not in self._tables:
            kr = self._compute_kernel_size(radius, axis=0)
            kz = self._compute_kernel_size(radius, axis=2)
            sigma = self._get_sigma(radius)
            taps_r = gaussian_taps(sigma / self.object_meta.dx, kr, dtype)
            taps_z = (gaussian_taps(sigma / self.object_meta.dz, kz, dtype)
                      if self.psf_meta.kernel_dimensions == "2D" else None)
            self._tables[key] = (taps_r, taps_z)
        return self._tables[key]

    def forward(self, object_i, ang_idx):
        raise NotImplementedError("SPECTPSFTransform is applied inside SPECTSystemMatrix's fused kernels")

    backward = forward
