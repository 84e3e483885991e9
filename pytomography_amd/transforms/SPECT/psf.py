"""Depth-dependent collimator PSF (transforms/SPECT/psf.py:14-231).

The per-plane taps are built on the host as the reference builds its Conv1d
weights (psf.py:62-101, 123-142, 170-197), for the 'gaussian' and 'square'
shapes; the fused HIP kernels stage them per plane.  ``forward``/``backward``
blur one angle's padded, rotated object with the same taps (HIP kernel) for
the per-angle projector path and for callers that use the transform on its
own.  A ``psf_operator`` network (psf.py:42-60, 144-151) is called as the
reference calls it; it runs only on the per-angle path.
"""
from __future__ import annotations

import numpy as np
import torch

import pytomography_amd
from pytomography_amd.utils import compute_pad_size, get_distance
from ..transform import Transform


def gaussian_taps(sigmas: np.ndarray, kernel_size: int, dtype: torch.dtype,
                  shape: str = "gaussian") -> torch.Tensor:
    """Normalised taps per plane, [len(sigmas), kernel_size], in ``dtype`` from
    float64 sigmas, divided by the per-plane sum: exp(-x^2/(2 sigma^2 + delta))
    like get_1D_PSF_layer_gaussian (psf.py:62-83), or with shape 'square'
    max(0, 1 - |x|/sigma) like get_1D_PSF_layer_square (psf.py:85-101)."""
    n = len(sigmas)
    x = torch.arange(-int(kernel_size // 2), int(kernel_size // 2) + 1).unsqueeze(0).unsqueeze(0)
    x = x.repeat((n, 1, 1))
    s = torch.tensor(np.asarray(sigmas)).to(dtype).reshape((n, 1, 1))
    if shape == "square":
        kern = 1 - torch.abs(x) / s
        kern[kern < 0] = 0
    else:
        kern = torch.exp(-x ** 2 / (2 * s ** 2 + pytomography_amd.delta))
    kern = kern / kern.sum(axis=-1).unsqueeze(-1)
    return kern.to(dtype)[:, 0, :]


class _OperatorLayer:
    """PSF2D (psf.py:42-60): a ``psf_operator`` network called with the
    kernel grid and the plane distances of one detector radius."""

    def __init__(self, psf_operator, distances, kernel_size: int, dr) -> None:
        self.psf_operator = psf_operator
        dev = pytomography_amd.device
        self.distances = torch.tensor(distances).to(pytomography_amd.dtype).to(dev)
        x = torch.arange(-(kernel_size - 1) / 2, (kernel_size + 1) / 2, 1).to(dev) * dr[0]
        y = torch.arange(-(kernel_size - 1) / 2, (kernel_size + 1) / 2, 1).to(dev) * dr[1]
        self.xv, self.yv = torch.meshgrid(x, y, indexing="xy")

    @torch.no_grad()
    def __call__(self, input):
        return self.psf_operator(input, self.xv, self.yv, self.distances, normalize=True)


class SPECTPSFTransform(Transform):
    """obj2obj PSF transform configured from ``SPECTPSFMeta`` ('gaussian' or
    'square' shape, '2D' or '1D' kernels) or from a ``psf_operator``."""

    def __init__(self, psf_meta=None, psf_operator=None, assume_padded: bool = True) -> None:
        super().__init__()
        if (psf_meta is None) == (psf_operator is None):
            raise ValueError("give exactly one of psf_meta and psf_operator")
        if psf_meta is not None and psf_meta.shape not in ("gaussian", "square"):
            raise NotImplementedError(f"PSF shape {psf_meta.shape!r} is not supported ('gaussian', 'square')")
        if not assume_padded:
            raise NotImplementedError("assume_padded=False is not supported by the kernels")
        self.psf_meta = psf_meta
        self.psf_operator = psf_operator
        self.assume_padded = assume_padded

    @property
    def fusable(self) -> bool:
        """Whether the fused projector kernels can apply this PSF (taps from
        ``psf_meta``; a ``psf_operator`` network runs on the per-angle path)."""
        return self.psf_operator is None

    def configure(self, object_meta, proj_meta) -> None:
        super().configure(object_meta, proj_meta)
        if proj_meta.radii is None:
            raise ValueError("PSF modelling needs proj_meta.radii")
        self._tables = {}
        self._frames = None
        self.layers = {}
        if self.psf_operator is not None:  # _configure_manual_net (psf.py:144-151)
            dim = object_meta.shape[0] + 2 * compute_pad_size(object_meta.shape[0])
            for radius in np.unique(np.asarray(proj_meta.radii, dtype=np.float64)):
                distances = get_distance(dim, radius, object_meta.dx)
                self.layers[float(radius)] = _OperatorLayer(self.psf_operator, distances,
                                                            object_meta.shape[0] - 1, object_meta.dr)

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
        if self.psf_meta is None:
            raise ValueError("a psf_operator PSF has no taps")
        key = (float(radius), dtype)
        if key not in self._tables:
            kr = self._compute_kernel_size(radius, axis=0)
            kz = self._compute_kernel_size(radius, axis=2)
            sigma = self._get_sigma(radius)
            shape = self.psf_meta.shape
            taps_r = gaussian_taps(sigma / self.object_meta.dx, kr, dtype, shape)
            taps_z = (gaussian_taps(sigma / self.object_meta.dz, kz, dtype, shape)
                      if self.psf_meta.kernel_dimensions == "2D" else None)
            self._tables[key] = (taps_r, taps_z)
        return self._tables[key]

    def _rotated_frames(self):
        if self._frames is None:
            from pytomography_amd.projectors.SPECT.plan import RotatedFrames
            self._frames = RotatedFrames(self.object_meta, self.proj_meta, psf=self)
        return self._frames

    @torch.no_grad()
    def forward(self, object_i: torch.Tensor, ang_idx) -> torch.Tensor:
        """psf.py:199-214: every plane of ``object_i`` [P, P, Lz] (padded and
        rotated to angle ``ang_idx``) blurred with the taps of its distance to
        the detector at that angle's radius, along y then z (HIP kernel).
        With a vector of angles, ``object_i`` is [n, P, P, Lz]."""
        if self.psf_operator is not None:
            radii = np.asarray(self.proj_meta.radii, dtype=np.float64).reshape(-1)
            radius = float(radii[0] if radii.size == 1 else radii[int(ang_idx)])
            return self.layers[radius](object_i)
        single = object_i.dim() == 3
        out = self._rotated_frames().blur(object_i.unsqueeze(0) if single else object_i, ang_idx)
        return out[0] if single else out

    @torch.no_grad()
    def backward(self, object_i: torch.Tensor, ang_idx) -> torch.Tensor:
        """psf.py:216-231: the reference applies the same (symmetric) blur."""
        return self.forward(object_i, ang_idx)
