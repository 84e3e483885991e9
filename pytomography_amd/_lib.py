"""ctypes binding of libspect_hip.so (declared in include/spect_hip.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded, every projector call raises ``SpectLibraryError``.  There
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspect_hip.so")

SPECT_F32, SPECT_F64 = 0, 1
ATT_FIELD = 0x100  # OR-ed into dtype: att_table is an attenuation field
BP_REFERENCE, BP_ADJOINT = 0, 1
PAD_CIRCULAR, PAD_REPLICATE, PAD_ZEROS = 0, 1, 2
OUT_STORE, OUT_ADD, OUT_OSEM = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 128 << 20  # SPECT_NOALLOC_SCRATCH_BYTES

EXPORTED = (
    "spect_last_error", "spect_version", "spect_plan_create", "spect_plan_destroy",
    "spect_plan_dims", "spect_workspace_bytes", "spect_schedule_info",
    "spect_att_table_bytes", "spect_att_table_build",
    "spect_att_field_bytes", "spect_att_field_build",
    "spect_forward", "spect_forward_noalloc", "spect_backward", "spect_backward_transpose",
    "spect_osem_subiteration",
    "spect_grad_workspace_bytes", "spect_forward_vjp",
    "spect_batch_create", "spect_batch_run", "spect_batch_destroy",
    "spect_poisson_ratio", "spect_rdp_gradient", "spect_rdp_hessian_vjp", "spect_rdp_hessian_jvp",
    "spect_bsrem_update", "spect_ew_scatter", "spect_smooth_axis",
    "spect_kem_apply", "spect_dvf_warp",
    "spect_rotate_object", "spect_rotate_back", "spect_att_volume", "spect_psf_volume",
)


class SpectLibraryError(RuntimeError):
    """libspect_hip.so is missing, failed to load, or a call returned an error."""


class SpectGeom(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("n_angles", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("angles_deg", ctypes.POINTER(ctypes.c_float)),
        ("angles_deg64", ctypes.POINTER(ctypes.c_double)),
        ("psf_kr", ctypes.c_int32), ("psf_kz", ctypes.c_int32),
        ("n_radii", ctypes.c_int32),
        ("radius_index", ctypes.POINTER(ctypes.c_int32)),
        ("taps_r", ctypes.POINTER(ctypes.c_float)),
        ("taps_z", ctypes.POINTER(ctypes.c_float)),
        ("taps_r64", ctypes.POINTER(ctypes.c_double)),
        ("taps_z64", ctypes.POINTER(ctypes.c_double)),
        ("chunk_planes", ctypes.c_int32),
        ("sched_rounds", ctypes.c_int32),
    ]


class SpectBatchProblem(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_void_p), ("f", ctypes.c_void_p), ("mu", ctypes.c_void_p),
        ("att_table", ctypes.c_void_p), ("att_field", ctypes.c_int32), ("n_subsets", ctypes.c_int32),
        ("angle_idx", ctypes.POINTER(ctypes.c_void_p)), ("counts", ctypes.POINTER(ctypes.c_int32)),
        ("g", ctypes.POINTER(ctypes.c_void_p)), ("s", ctypes.POINTER(ctypes.c_void_p)),
        ("norm", ctypes.POINTER(ctypes.c_void_p)), ("pred", ctypes.POINTER(ctypes.c_void_p)),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
    ]


_lock = threading.Lock()
_lib = None


def _declare(lib):
    vp, i32, i64, sz, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t, ctypes.c_double
    ci = ctypes.c_int
    lib.spect_last_error.restype = ctypes.c_char_p
    lib.spect_last_error.argtypes = []
    lib.spect_version.restype = ci
    lib.spect_plan_create.restype = ci
    lib.spect_plan_create.argtypes = [ctypes.POINTER(SpectGeom), ctypes.POINTER(vp)]
    lib.spect_plan_destroy.restype = ci
    lib.spect_plan_destroy.argtypes = [vp]
    lib.spect_plan_dims.restype = ci
    lib.spect_plan_dims.argtypes = [vp] + [ctypes.POINTER(i32)] * 4
    lib.spect_workspace_bytes.restype = sz
    lib.spect_workspace_bytes.argtypes = [vp, ci, ci]
    lib.spect_schedule_info.restype = ci
    lib.spect_schedule_info.argtypes = [vp, ci, ci, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.spect_att_table_bytes.restype = sz
    lib.spect_att_table_bytes.argtypes = [vp]
    lib.spect_att_table_build.restype = ci
    lib.spect_att_table_build.argtypes = [vp, ci, vp, vp, vp]
    lib.spect_att_field_bytes.restype = sz
    lib.spect_att_field_bytes.argtypes = [vp, ci]
    lib.spect_att_field_build.restype = ci
    lib.spect_att_field_build.argtypes = [vp, ci, vp, vp, vp]
    lib.spect_forward.restype = ci
    lib.spect_forward.argtypes = [vp, ci, vp, vp, vp, vp, ci, vp, vp, sz, vp]
    lib.spect_forward_noalloc.restype = ci
    lib.spect_forward_noalloc.argtypes = [vp, ci, vp, vp, vp, vp, ci, vp, vp]
    lib.spect_backward.restype = ci
    lib.spect_backward.argtypes = [vp, ci, ci, ci, vp, vp, vp, vp, ci, vp, vp, vp, dbl, dbl, vp, sz, vp]
    lib.spect_osem_subiteration.restype = ci
    lib.spect_osem_subiteration.argtypes = [vp, ci, vp, vp, vp, vp, ci, vp, vp, vp, vp, dbl, dbl, vp, sz, vp]
    lib.spect_grad_workspace_bytes.restype = sz
    lib.spect_grad_workspace_bytes.argtypes = [vp, ci, ci]
    lib.spect_backward_transpose.restype = ci
    lib.spect_backward_transpose.argtypes = [vp, ci, vp, vp, vp, vp, ci, vp, vp, sz, vp]
    lib.spect_forward_vjp.restype = ci
    lib.spect_forward_vjp.argtypes = [vp, ci, vp, vp, vp, vp, ci, vp, vp, vp, sz, vp]
    lib.spect_batch_create.restype = ci
    lib.spect_batch_create.argtypes = [ctypes.POINTER(SpectBatchProblem), ci, ci, dbl, dbl,
                                       ctypes.POINTER(vp)]
    lib.spect_batch_run.restype = ci
    lib.spect_batch_run.argtypes = [vp, ci, ci, vp]
    lib.spect_batch_destroy.restype = ci
    lib.spect_batch_destroy.argtypes = [vp]
    lib.spect_poisson_ratio.restype = ci
    lib.spect_poisson_ratio.argtypes = [ci, vp, vp, vp, vp, vp, i64, dbl, vp]
    lib.spect_rdp_gradient.restype = ci
    lib.spect_rdp_gradient.argtypes = [ci, vp, vp, i32, i32, i32, dbl, dbl, dbl, dbl, dbl, dbl, dbl, vp]
    lib.spect_rdp_hessian_jvp.restype = ci
    lib.spect_rdp_hessian_jvp.argtypes = [ci, vp, vp, vp, i32, i32, i32, dbl, dbl, dbl, dbl, dbl,
                                          dbl, dbl, vp]
    lib.spect_rdp_hessian_vjp.restype = ci
    lib.spect_rdp_hessian_vjp.argtypes = [ci, vp, vp, vp, i32, i32, i32, dbl, dbl, dbl, dbl, dbl,
                                          dbl, dbl, vp]
    lib.spect_ew_scatter.restype = ci
    lib.spect_ew_scatter.argtypes = [ci, vp, vp, vp, vp, i64, dbl, dbl, dbl, dbl, dbl, vp]
    lib.spect_smooth_axis.restype = ci
    lib.spect_smooth_axis.argtypes = [ci, vp, vp, i32, i32, i32, i32, vp, i32, i32, vp]
    lib.spect_kem_apply.restype = ci
    lib.spect_kem_apply.argtypes = [ci, vp, vp, vp, i32, i32, i32, i32, vp]
    lib.spect_dvf_warp.restype = ci
    lib.spect_dvf_warp.argtypes = [ci, vp, vp, vp, i32, i32, i32, vp]
    lib.spect_bsrem_update.restype = ci
    lib.spect_bsrem_update.argtypes = [ci, vp, vp, vp, vp, vp, i64, dbl, dbl, dbl, dbl, vp]
    for name in ("spect_rotate_object", "spect_att_volume", "spect_psf_volume"):
        getattr(lib, name).restype = ci
        getattr(lib, name).argtypes = [vp, ci, vp, vp, ci, vp, vp]
    lib.spect_rotate_back.restype = ci
    lib.spect_rotate_back.argtypes = [vp, ci, vp, vp, ci, ci, vp, vp]


def lib():
    """The loaded library (raises SpectLibraryError when it is unavailable)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise SpectLibraryError(
                    f"{LIB_PATH} is not built; run `python -m pytomography_amd.build` "
                    "(the HIP kernels are the only compute path)")
            try:
                handle = ctypes.CDLL(LIB_PATH)
            except OSError as exc:  # pragma: no cover - depends on the machine
                raise SpectLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
            _declare(handle)
            _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().spect_last_error().decode(errors="replace")
        raise SpectLibraryError(f"libspect_hip error {rc}: {msg}")


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return SPECT_F32
    if dt == torch.float64:
        return SPECT_F64
    raise SpectLibraryError(f"unsupported dtype {dt}: the kernels run in float32 or float64")


def stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_device_tensor(t: torch.Tensor, name: str, dtype: torch.dtype | None = None) -> torch.Tensor:
    """Validate a tensor handed to the kernels: on a HIP device, contiguous,
    16-byte aligned (a misaligned view is copied)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise SpectLibraryError(
            f"{name} is on {t.device}; the SPECT kernels need a ROCm/HIP device tensor")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()
    return t


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()
