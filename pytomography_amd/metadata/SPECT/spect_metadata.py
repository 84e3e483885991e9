"""SPECT metadata with the reference's fields (metadata/SPECT/spect_metadata.py:8-94)."""
from __future__ import annotations

from typing import Callable, Sequence

import torch

import pytomography_amd
from pytomography_amd.utils import compute_pad_size
from ..metadata import ObjectMeta, ProjMeta


class SPECTObjectMeta(ObjectMeta):
    """Object metadata: ``dr`` voxel size in cm, ``shape`` [Lx, Ly, Lz]
    (spect_metadata.py:8-32).  ``padded_shape`` is the rotation-safe frame."""

    def __init__(self, dr: list[float], shape: list[int]) -> None:
        super().__init__(dr, shape)
        self.compute_padded_shape()
        self.recon_method = None
        self.units = "counts"

    def compute_padded_shape(self) -> None:
        self.pad_size = compute_pad_size(self.shape[0])
        self.padded_shape = (int(self.shape[0] + 2 * self.pad_size),
                             int(self.shape[1] + 2 * self.pad_size), int(self.shape[2]))


class SPECTProjMeta(ProjMeta):
    """Projection metadata (spect_metadata.py:34-64): ``angles`` (degrees,
    stored as a tensor of the global dtype on the global device), ``radii``
    (cm, per angle; needed for PSF modelling)."""

    def __init__(self, projection_shape: Sequence, dr: list[float], angles: Sequence,
                 radii=None) -> None:
        self.angles = torch.tensor(angles).to(pytomography_amd.dtype).to(pytomography_amd.device)
        self.dr = dr
        self.radii = radii
        self.num_projections = len(angles)
        self.shape = (self.num_projections, projection_shape[0], projection_shape[1])
        self.compute_padded_shape()

    def compute_padded_shape(self) -> None:
        self.pad_size = compute_pad_size(self.shape[1])
        self.padded_shape = (int(self.shape[0]), int(self.shape[1] + 2 * self.pad_size),
                             int(self.shape[2]))


def _linear_sigma(r, a, b):
    return a * r + b


class SPECTPSFMeta:
    """Gaussian collimator PSF metadata (spect_metadata.py:66-94):
    sigma(d) = sigma_fit(d, *sigma_fit_params) (default a*d + b), truncated at
    ``min_sigmas`` sigma; '2D' blurs along r and z, '1D' along r only."""

    def __init__(self, sigma_fit_params: Sequence[float], sigma_fit: Callable = _linear_sigma,
                 kernel_dimensions: str = "2D", min_sigmas: float | None = 3,
                 shape: str = "gaussian") -> None:
        self.sigma_fit_params = sigma_fit_params
        self.sigma_fit = sigma_fit
        self.kernel_dimensions = kernel_dimensions
        self.min_sigmas = 1 if shape == "square" else min_sigmas
        self.shape = shape

    def __repr__(self):
        return "".join(f"{k} = {v}\n" for k, v in sorted(vars(self).items()))
