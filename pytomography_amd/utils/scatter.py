"""Energy-window scatter estimates: the additive term of the likelihood
(utils/scatter.py:10-99).  The window combination and each axis of the
Gaussian smoothing are HIP kernels (spect_ew_scatter, spect_smooth_axis)."""
from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

import pytomography_amd
from pytomography_amd import _lib

_MODES = {"circular": _lib.PAD_CIRCULAR, "replicate": _lib.PAD_REPLICATE, "zeros": _lib.PAD_ZEROS}


def get_1d_gaussian_taps(sigma: float, kernel_size: int, dtype: torch.dtype | None = None) -> torch.Tensor:
    """Normalised taps of get_1d_gaussian_kernel (utils/misc.py:78-94): integer
    offsets, exp(-x^2/(2 sigma^2)) in the global dtype, divided by their sum."""
    dt = pytomography_amd.dtype if dtype is None else dtype
    x = torch.arange(-int(kernel_size // 2), int(kernel_size // 2) + 1)
    k = torch.exp((-x ** 2 / (2 * sigma ** 2)).to(dt))
    return k / k.sum()


def _smooth_axis(x: torch.Tensor, axis: int, taps: torch.Tensor, mode: str) -> torch.Tensor:
    out = torch.empty_like(x)
    t = taps.to(x.device, x.dtype).contiguous()
    n0, n1, n2 = x.shape
    _lib.check(_lib.lib().spect_smooth_axis(
        _lib.dtype_code(x.dtype), x.data_ptr(), out.data_ptr(), n0, n1, n2, axis, t.data_ptr(),
        int(t.numel()), _MODES[mode], _lib.stream_ptr(x.device)))
    return out


@torch.no_grad()
def get_smoothed_scatter(scatter: torch.Tensor, proj_meta, sigma_theta: float = 0, sigma_r: float = 0,
                         sigma_z: float = 0, N_sigmas: int = 3) -> torch.Tensor:
    """Gaussian smoothing of [A, Lr, Lz] projections along theta (circular, sigma
    in degrees), r and z (replicate, sigma in cm) (utils/scatter.py:10-50)."""
    dr, dz = proj_meta.dr
    dtheta = torch.diff(proj_meta.angles)[0].item()
    dS = [dtheta, dr, dz]
    ksize = [2 * int(np.ceil(N_sigmas * sigma_theta / dtheta)) + 1,
             2 * int(np.ceil(N_sigmas * sigma_r / dr)) + 1,
             2 * int(np.ceil(N_sigmas * sigma_z / dz)) + 1]
    sigmas = [sigma_theta, sigma_r, sigma_z]
    modes = ["circular", "replicate", "replicate"]
    if scatter.dim() != 3:
        raise ValueError(f"scatter must be [angles, r, z]; got shape {tuple(scatter.shape)}")
    x = _lib.require_device_tensor(scatter, "scatter")
    for i in range(3):
        if sigmas[i] > pytomography_amd.delta:
            if modes[i] == "circular" and ksize[i] // 2 > x.shape[i]:
                raise ValueError("circular smoothing kernel wider than the angle axis")
            x = _smooth_axis(x, i, get_1d_gaussian_taps(sigmas[i] / dS[i], ksize[i], x.dtype), modes[i])
    return x


def compute_EW_scatter(projection_lower: torch.Tensor, projection_upper: torch.Tensor | None,
                       width_lower: float, width_upper: float | None, width_peak: float,
                       weighting_lower: float = 0.5, weighting_upper: float = 0.5, proj_meta=None,
                       sigma_theta: float = 0, sigma_r: float = 0, sigma_z: float = 0,
                       N_sigmas: int = 3, return_scatter_variance_estimate: bool = False
                       ) -> torch.Tensor | Sequence:
    """Triple (or dual, with no upper window) energy-window scatter estimate,
    optionally smoothed, and optionally the operator x -> F x * diag * F x of
    its variance (utils/scatter.py:52-99)."""
    lower = _lib.require_device_tensor(projection_lower, "projection_lower")
    dt = lower.dtype
    upper = None
    if projection_upper is not None:
        upper = _lib.require_device_tensor(projection_upper, "projection_upper", dt)
        if upper.shape != lower.shape:
            raise ValueError(f"window shapes differ: {tuple(lower.shape)} vs {tuple(upper.shape)}")
    width_upper = 1 if width_upper is None else width_upper
    smooth = (sigma_r > 0) + (sigma_theta > 0) + (sigma_z > 0)
    if smooth and proj_meta is None:
        raise ValueError("proj_meta must be provided if scatter is to be smoothed")
    estimate = torch.empty_like(lower)
    diag = torch.empty_like(lower) if return_scatter_variance_estimate else None
    _lib.check(_lib.lib().spect_ew_scatter(
        _lib.dtype_code(dt), lower.data_ptr(), _lib.ptr(upper), estimate.data_ptr(), _lib.ptr(diag),
        lower.numel(), float(width_lower), float(width_upper), float(width_peak),
        float(weighting_lower), float(weighting_upper), _lib.stream_ptr(lower.device)))
    if smooth:
        estimate = get_smoothed_scatter(estimate, proj_meta, sigma_theta, sigma_r, sigma_z, N_sigmas)
    if not return_scatter_variance_estimate:
        return estimate
    if smooth:
        def scatter_variance_estimate(x):
            xs = get_smoothed_scatter(x, proj_meta, sigma_theta, sigma_r, sigma_z, N_sigmas)
            return xs * diag * xs
    else:
        def scatter_variance_estimate(x):
            return x * diag * x
    return estimate, scatter_variance_estimate
