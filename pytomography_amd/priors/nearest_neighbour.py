"""Relative difference prior (priors/nearest_neighbour.py:13-220).

The 26-neighbour gradient runs as one stencil kernel (replicate-clamped
neighbours, inverse-Euclidean weights) instead of 26 pad/roll/crop passes.
"""
from __future__ import annotations

import numpy as np
import torch

import pytomography_amd
from pytomography_amd import _lib
from .prior import Prior


class NeighbourWeight:
    def set_object_meta(self, object_meta) -> None:
        self.object_meta = object_meta

    def __call__(self, coords):
        raise NotImplementedError


class EuclideanNeighbourWeight(NeighbourWeight):
    """w_rs = dx / |(i dx, j dy, k dz)| (nearest_neighbour.py:207-220)."""

    def __call__(self, coords):
        i, j, k = coords
        m = self.object_meta
        return m.dx / np.sqrt((m.dx * i) ** 2 + (m.dy * j) ** 2 + (m.dz * k) ** 2)


class NearestNeighbourPrior(Prior):
    def __init__(self, beta: float, weight: NeighbourWeight | None = None, **kwargs) -> None:
        super().__init__(beta)
        self.weight = EuclideanNeighbourWeight() if weight is None else weight
        if not isinstance(self.weight, EuclideanNeighbourWeight):
            raise NotImplementedError("only EuclideanNeighbourWeight is on the accelerated path")
        self.__dict__.update(kwargs)

    def set_object_meta(self, object_meta) -> None:
        self.weight.set_object_meta(object_meta)
        self.object_meta = object_meta
        self.FOV_scale = torch.ones(self.object_meta.shape).to(pytomography_amd.device)


class RelativeDifferencePrior(NearestNeighbourPrior):
    """phi0 = (fr-fs)^2/(fr+fs+gamma|fr-fs|+delta);
    phi1 = (fr-fs)(gamma|fr-fs|+3fs+fr+2delta)/(fr+fs+gamma|fr-fs|+delta)^2
    (nearest_neighbour.py:156-183)."""

    def __init__(self, beta: float, weight: NeighbourWeight | None = None, gamma: float = 1,
                 delta: float | None = None) -> None:
        if delta is None:
            delta = pytomography_amd.delta
        super().__init__(beta, weight=weight, gamma=gamma, delta=delta)

    def _fov_factor(self):
        """The FOV scale, or None when it is all ones (then the multiply is
        skipped).  The all-ones test syncs with the device, so its answer is
        kept until the tensor is replaced or modified."""
        fov = getattr(self, "FOV_scale", None)
        if fov is None:
            return None
        if getattr(self, "_fov_ref", None) is not fov or self._fov_version != fov._version:
            self._fov_ref, self._fov_version = fov, fov._version
            self._fov_is_one = bool(torch.all(fov == 1))
        return None if self._fov_is_one else fov

    def gradient(self, object: torch.Tensor, beta_scale: float,
                 beta: float | None = None) -> torch.Tensor:
        """beta * beta_scale * sum_s w_rs phi1(f_r, f_s) * FOV_scale;
        ``beta`` overrides the prior's beta."""
        dt = torch.float64 if object.dtype == torch.float64 else torch.float32
        f = _lib.require_device_tensor(object, "object", dt)
        out = torch.empty_like(f)
        m = self.object_meta
        nx, ny, nz = f.shape
        _lib.check(_lib.lib().spect_rdp_gradient(
            _lib.dtype_code(dt), f.data_ptr(), out.data_ptr(), nx, ny, nz, float(m.dx), float(m.dy),
            float(m.dz), float(self.beta if beta is None else beta), float(self.gamma),
            float(self.delta), float(beta_scale), _lib.stream_ptr(f.device)))
        fov = self._fov_factor()
        if fov is not None:
            out = out * fov
        return out

    def hessian_vjp(self, object: torch.Tensor, u: torch.Tensor, beta_scale: float,
                    beta: float | None = None) -> torch.Tensor:
        """J^T u with J the Jacobian of ``gradient(object, beta_scale)``, from
        phi2_1 / phi2_2 (nearest_neighbour.py:179-183) over the
        replicate-clamped neighbours; ``beta`` overrides the prior's beta."""
        dt = torch.float64 if object.dtype == torch.float64 else torch.float32
        f = _lib.require_device_tensor(object, "object", dt)
        uu = _lib.require_device_tensor(u, "u", dt)
        if uu.shape != f.shape:
            raise ValueError(f"u shape {tuple(uu.shape)} != object shape {tuple(f.shape)}")
        fov = self._fov_factor()
        if fov is not None:
            uu = (uu * fov).contiguous()
        out = torch.empty_like(f)
        m = self.object_meta
        nx, ny, nz = f.shape
        _lib.check(_lib.lib().spect_rdp_hessian_vjp(
            _lib.dtype_code(dt), f.data_ptr(), uu.data_ptr(), out.data_ptr(), nx, ny, nz,
            float(m.dx), float(m.dy), float(m.dz), float(self.beta if beta is None else beta),
            float(self.gamma), float(self.delta), float(beta_scale), _lib.stream_ptr(f.device)))
        return out

    def hessian_jvp(self, object: torch.Tensor, u: torch.Tensor, beta_scale: float,
                    beta: float | None = None) -> torch.Tensor:
        """The derivative of order 2 applied to u, as NearestNeighbourPrior.__call__(2)
        (nearest_neighbour.py:105-108): beta*beta_scale*[u sum_s w phi2_1 + sum_s w phi2_2 u_s]
        times the FOV scale."""
        dt = torch.float64 if object.dtype == torch.float64 else torch.float32
        f = _lib.require_device_tensor(object, "object", dt)
        uu = _lib.require_device_tensor(u, "u", dt)
        if uu.shape != f.shape:
            raise ValueError(f"u shape {tuple(uu.shape)} != object shape {tuple(f.shape)}")
        out = torch.empty_like(f)
        m = self.object_meta
        nx, ny, nz = f.shape
        _lib.check(_lib.lib().spect_rdp_hessian_jvp(
            _lib.dtype_code(dt), f.data_ptr(), uu.data_ptr(), out.data_ptr(), nx, ny, nz,
            float(m.dx), float(m.dy), float(m.dz), float(self.beta if beta is None else beta),
            float(self.gamma), float(self.delta), float(beta_scale), _lib.stream_ptr(f.device)))
        fov = self._fov_factor()
        if fov is not None:
            out = out * fov
        return out

    def value(self, object: torch.Tensor) -> float:
        """sum_r sum_s w_rs phi0(f_r, f_s) * beta (derivative order 0)."""
        f = object
        tot = torch.zeros((), dtype=torch.float64, device=f.device)
        p = torch.nn.functional.pad(f.unsqueeze(0).unsqueeze(0), (1, 1, 1, 1, 1, 1), mode="replicate")[0, 0]
        nx, ny, nz = f.shape
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    if i == 0 and j == 0 and k == 0:
                        continue
                    fs = p[1 + i:1 + i + nx, 1 + j:1 + j + ny, 1 + k:1 + k + nz]
                    w = self.weight((i, j, k))
                    d = f - fs
                    tot = tot + (w * d ** 2 / (f + fs + self.gamma * d.abs() + self.delta)).sum()
        return float(self.beta * tot)

    def __call__(self, derivative_order: int = 0):
        if derivative_order == 0:
            return self.value(self.object)
        if derivative_order == 1:
            return self.gradient(self.object, self.beta_scale_factor)
        if derivative_order == 2:
            f, scale = self.object, self.beta_scale_factor
            return lambda input: self.hessian_jvp(f, input, scale)
        raise NotImplementedError("Prior not implemented for derivative order >2")
