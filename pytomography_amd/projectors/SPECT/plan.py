"""Device plans of the SPECT kernels and whole rotated-frame volumes.

``SpectPlan`` owns one ``spect_plan`` (trig, linspace, activity and PSF
tables on one device); ``plan_geometry`` builds its ``spect_geom`` from the
reference's metadata (spect_metadata.py:8-64, psf.py:123-142).
``RotatedFrames`` runs the per-angle volumes of the reference's projector
loop (spect_system_matrix.py:145-156, 187-198) on the HIP kernels: the padded
object rotated to each angle, the back rotation and sum, and the attenuation
and PSF transforms applied to such volumes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

import pytomography_amd
from pytomography_amd import _lib


def _as_ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


class SpectPlan:
    """Owns one ``spect_plan`` (device tables) for one device."""

    def __init__(self, geom_args: dict, device: torch.device):
        self.lib = _lib.lib()
        self._keep = geom_args["_keep"]  # host arrays referenced by the geom struct
        g = _lib.SpectGeom(**{k: v for k, v in geom_args.items() if k != "_keep"})
        handle = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(self.lib.spect_plan_create(ctypes.byref(g), ctypes.byref(handle)))
        self.handle = handle
        dims = [ctypes.c_int32() for _ in range(4)]
        _lib.check(self.lib.spect_plan_dims(handle, *[ctypes.byref(d) for d in dims]))
        self.padded, self.pad, self.psf_half, self.chunks = (d.value for d in dims)

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                self.lib.spect_plan_destroy(self.handle)
                self.handle = None
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


def plan_geometry(object_meta, proj_meta, psf=None, chunk_planes: int = 0,
                  sched_rounds: int = 0) -> dict:
    """``spect_geom`` fields (plus ``_keep``, the host arrays they point to)
    for an object/projection geometry and an optional ``SPECTPSFTransform``:
    angles in float32 and float64, and the PSF taps of every distinct detector
    radius stacked to one odd width (smaller kernels centred in zero taps)."""
    nx, ny, nz = (int(s) for s in object_meta.shape)
    angles = proj_meta.angles.detach().cpu()
    a32 = np.ascontiguousarray(angles.to(torch.float32).numpy())
    a64 = np.ascontiguousarray(angles.to(torch.float64).numpy())
    keep = [a32, a64]
    args = dict(nx=nx, ny=ny, nz=nz, n_angles=len(a32), dx=float(object_meta.dx),
                dy=float(object_meta.dy), dz=float(object_meta.dz),
                angles_deg=_as_ptr(a32, ctypes.c_float),
                angles_deg64=_as_ptr(a64, ctypes.c_double),
                psf_kr=0, psf_kz=0, n_radii=0, chunk_planes=int(chunk_planes),
                sched_rounds=int(sched_rounds))
    if psf is not None:
        radii = np.asarray(proj_meta.radii, dtype=np.float64).reshape(-1)
        if radii.size == 1:
            radii = np.full(len(a32), radii.item())
        uniq = np.unique(radii)
        ridx = np.ascontiguousarray(np.searchsorted(uniq, radii).astype(np.int32))
        t32 = [psf.radius_taps(r, torch.float32) for r in uniq]
        t64 = [psf.radius_taps(r, torch.float64) for r in uniq]
        kr = max(t[0].shape[1] for t in t32)
        two_d = t32[0][1] is not None
        kz = max(t[1].shape[1] for t in t32) if two_d else 0

        def stack(tabs, k, which, npdt):
            out = []
            for t in tabs:
                w = t[which].numpy().astype(npdt)
                off = (k - w.shape[1]) // 2  # centre smaller kernels in zero taps
                full = np.zeros((w.shape[0], k), dtype=npdt)
                full[:, off:off + w.shape[1]] = w
                out.append(full)
            return np.ascontiguousarray(np.stack(out))

        tr32 = stack(t32, kr, 0, np.float32)
        tr64 = stack(t64, kr, 0, np.float64)
        keep += [ridx, tr32, tr64]
        args.update(psf_kr=kr, n_radii=len(uniq), radius_index=_as_ptr(ridx, ctypes.c_int32),
                    taps_r=_as_ptr(tr32, ctypes.c_float), taps_r64=_as_ptr(tr64, ctypes.c_double))
        if two_d:
            tz32 = stack(t32, kz, 1, np.float32)
            tz64 = stack(t64, kz, 1, np.float64)
            keep += [tz32, tz64]
            args.update(psf_kz=kz, taps_z=_as_ptr(tz32, ctypes.c_float),
                        taps_z64=_as_ptr(tz64, ctypes.c_double))
    args["_keep"] = keep
    return args


class RotatedFrames:
    """Whole rotated-frame volumes [n, P, P, Lz] of one geometry (P the padded
    width; axis 1 runs along the rays towards the detector at +x).

    Args:
        object_meta, proj_meta: the geometry (as configured on the transforms).
        psf: a ``SPECTPSFTransform`` whose taps ``psf_volume`` applies.
    """

    def __init__(self, object_meta, proj_meta, psf=None) -> None:
        self.object_meta = object_meta
        self.proj_meta = proj_meta
        self.psf = psf
        self.shape = tuple(int(s) for s in object_meta.shape)
        self._plans: dict = {}

    def plan(self, device: torch.device) -> SpectPlan:
        key = (device.type, device.index)
        if key not in self._plans:
            self._plans[key] = SpectPlan(plan_geometry(self.object_meta, self.proj_meta, self.psf), device)
        return self._plans[key]

    @staticmethod
    def _device(t: torch.Tensor) -> torch.device:
        return t.device if t.device.type == "cuda" else pytomography_amd.device

    @staticmethod
    def angle_index(angle_indices, device) -> torch.Tensor:
        """Angle indices (a tensor, a sequence or one integer) as a device int32 vector."""
        idx = torch.as_tensor(angle_indices).reshape(-1)
        return idx.to(device=device, dtype=torch.int32).contiguous()

    @staticmethod
    def _dtype(t: torch.Tensor) -> torch.dtype:
        return torch.float64 if t.dtype == torch.float64 else torch.float32

    def _frames_call(self, name: str, src: torch.Tensor, angle_indices, src_name: str) -> torch.Tensor:
        device = self._device(src)
        dt = self._dtype(src)
        x = _lib.require_device_tensor(src.to(device), src_name, dt)
        plan = self.plan(device)
        idx = self.angle_index(angle_indices, device)
        n = idx.numel()
        P, nz = plan.padded, self.shape[2]
        out = torch.empty((n, P, P, nz), dtype=dt, device=device)
        _lib.check(getattr(plan.lib, name)(plan.handle, _lib.dtype_code(dt), x.data_ptr(), idx.data_ptr(),
                                           n, out.data_ptr(), _lib.stream_ptr(device)))
        return out

    def rotate(self, object: torch.Tensor, angle_indices) -> torch.Tensor:
        """RotationTransform.backward(pad_object(object), 270 - beta) per angle
        (spect_system_matrix.py:149-151): [n, P, P, Lz]."""
        if tuple(object.shape) != self.shape:
            raise ValueError(f"object shape {tuple(object.shape)} != {self.shape}")
        return self._frames_call("spect_rotate_object", object, angle_indices, "object")

    def attenuation(self, attenuation_map: torch.Tensor, angle_indices) -> torch.Tensor:
        """get_prob_of_detection_matrix(rotate_detector_z(pad_object(mu), beta), dx)
        per angle (attenuation.py:11-21, 88-91): [n, P, P, Lz]."""
        if tuple(attenuation_map.shape) != self.shape:
            raise ValueError(f"attenuation map shape {tuple(attenuation_map.shape)} != {self.shape}")
        return self._frames_call("spect_att_volume", attenuation_map, angle_indices, "attenuation_map")

    def blur(self, volumes: torch.Tensor, angle_indices) -> torch.Tensor:
        """SPECTPSFTransform.forward of each volume (psf.py:199-214): [n, P, P, Lz]."""
        if self.psf is None:
            raise ValueError("RotatedFrames was built without a PSF")
        plan = self.plan(self._device(volumes))
        if tuple(volumes.shape[-3:]) != (plan.padded, plan.padded, self.shape[2]):
            raise ValueError(f"volume shape {tuple(volumes.shape)} is not [n, P, P, Lz] with P = {plan.padded}")
        if plan.psf_half == 0:  # every kernel a single centre tap of weight 1
            return volumes.reshape(-1, plan.padded, plan.padded, self.shape[2]).clone()
        return self._frames_call("spect_psf_volume", volumes, angle_indices, "volumes")

    def rotate_back(self, volumes: torch.Tensor, angle_indices, out: torch.Tensor | None = None) -> torch.Tensor:
        """unpad_object(sum_a RotationTransform.forward(volumes[a], 270 - beta_a))
        (spect_system_matrix.py:194-200); added to ``out`` when given."""
        device = self._device(volumes)
        dt = self._dtype(volumes)
        plan = self.plan(device)
        v = _lib.require_device_tensor(volumes.to(device), "volumes", dt)
        idx = self.angle_index(angle_indices, device)
        n = idx.numel()
        if tuple(v.shape) != (n, plan.padded, plan.padded, self.shape[2]):
            raise ValueError(f"volume shape {tuple(v.shape)} != {(n, plan.padded, plan.padded, self.shape[2])}")
        mode = _lib.OUT_ADD
        if out is None:
            out = torch.empty(self.shape, dtype=dt, device=device)
            mode = _lib.OUT_STORE
        elif (out.device != device or out.dtype != dt or tuple(out.shape) != self.shape
              or not out.is_contiguous() or out.data_ptr() % 16):
            raise ValueError(f"out must be a contiguous, 16-byte aligned {dt} tensor of shape "
                             f"{self.shape} on {device}")
        _lib.check(plan.lib.spect_rotate_back(plan.handle, _lib.dtype_code(dt), v.data_ptr(), idx.data_ptr(),
                                              n, mode, out.data_ptr(), _lib.stream_ptr(device)))
        return out
