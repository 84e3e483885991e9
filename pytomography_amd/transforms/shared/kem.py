"""Kernel-EM basis transform f = K alpha (transforms/shared/kem.py:9-146).

The per-voxel kernel K[i,j,k](v) is built once at ``configure`` with the
reference's host-side tensor expressions (support kernels may be arbitrary
user callables), then kept resident in device memory as [size^3][Lx*Ly*Lz];
applying K (forward, and backward since K is symmetric) is one HIP kernel,
``spect_kem_apply``.
"""
from __future__ import annotations

import numpy as np
import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.transforms import Transform
from pytomography_amd.utils.misc import get_object_nearest_neighbour


class KEMTransform(Transform):
    def __init__(self, support_objects, support_kernels=None, support_kernels_params=None,
                 distance_kernel=None, distance_kernel_params=None, size: int = 5,
                 top_N: int | None = None, kernel_on_gpu: bool = False) -> None:
        super().__init__()
        self.support_objects = support_objects
        if support_kernels is None:
            self.support_kernels = [lambda obj_f, obj_j, sigma: torch.exp(-(obj_f - obj_j) ** 2 / (2 * sigma ** 2))
                                    for _ in range(len(support_objects))]
        else:
            self.support_kernels = support_kernels
        if support_kernels_params is None:
            self.support_kernel_params = [[1] for _ in range(len(support_objects))]
        else:
            self.support_kernel_params = support_kernels_params
        if distance_kernel is None:
            self.distance_kernel = lambda d, sigma: np.exp(-d ** 2 / (2 * sigma ** 2))
        else:
            self.distance_kernel = distance_kernel
        self.distance_kernel_params = [1] if distance_kernel_params is None else distance_kernel_params
        if size < 1 or size % 2 == 0:
            raise ValueError("size must be odd")
        self.size = size
        self.idx_max = int((size - 1) / 2)
        self.idxs = torch.arange(-self.idx_max, self.idx_max + 1)
        self.top_N = top_N
        # the kernel always lives in device memory: the HIP stencil reads it there
        self.kernel_on_gpu = kernel_on_gpu

    def compute_kernel(self) -> None:
        """kem.py:64-96: product of the support kernels over neighbours, optional
        top-N selection, distance weighting, normalisation over the taps."""
        shape = self.support_objects[0].shape
        kernel = torch.ones((self.size, self.size, self.size, *shape)).to(pytomography_amd.dtype)
        support = [so.cpu() for so in self.support_objects]
        for i in self.idxs:
            for j in self.idxs:
                for k in self.idxs:
                    component = 1
                    for l in range(len(support)):
                        so = support[l]
                        nb = get_object_nearest_neighbour(so, (int(i), int(j), int(k)))
                        component *= self.support_kernels[l](so, nb, *self.support_kernel_params[l])
                    kernel[i + self.idx_max, j + self.idx_max, k + self.idx_max] = component
        if self.top_N is not None:
            flat = kernel.reshape((self.size ** 3, *shape))
            keep = torch.argsort(torch.argsort(flat, dim=0), dim=0) >= self.size ** 3 - self.top_N
            kernel *= keep.reshape((self.size, self.size, self.size, *shape))
        xv, yv, zv = torch.meshgrid([self.object_meta.dx * self.idxs, self.object_meta.dy * self.idxs,
                                     self.object_meta.dz * self.idxs], indexing="ij")
        d = torch.sqrt(xv ** 2 + yv ** 2 + zv ** 2)
        kernel *= torch.as_tensor(self.distance_kernel(d, *self.distance_kernel_params)).unsqueeze(-1).unsqueeze(-1).unsqueeze(-1)
        kernel /= kernel.sum(dim=(0, 1, 2))
        self.kernel = kernel.to(pytomography_amd.device).contiguous()

    def configure(self, object_meta, proj_meta) -> None:
        super().configure(object_meta, proj_meta)
        self.compute_kernel()

    @torch.no_grad()
    def forward(self, object: torch.Tensor) -> torch.Tensor:
        """K alpha (kem.py:113-131)."""
        if object.dim() != 3 or tuple(object.shape) != tuple(self.kernel.shape[3:]):
            raise ValueError(f"object shape {tuple(object.shape)} != kernel support "
                             f"{tuple(self.kernel.shape[3:])}")
        dt = self.kernel.dtype
        f = _lib.require_device_tensor(object, "object", dt)
        out = torch.empty_like(f)
        nx, ny, nz = f.shape
        _lib.check(_lib.lib().spect_kem_apply(
            _lib.dtype_code(dt), self.kernel.data_ptr(), f.data_ptr(), out.data_ptr(), nx, ny, nz,
            self.size, _lib.stream_ptr(f.device)))
        return out

    @torch.no_grad()
    def backward(self, object: torch.Tensor) -> torch.Tensor:
        """K^T alpha = K alpha (kem.py:133-146)."""
        return self.forward(object)
