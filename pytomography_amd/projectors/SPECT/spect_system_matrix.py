"""SPECTSystemMatrix: rotate+sum SPECT projector on the HIP kernels.

Drop-in for ``projectors/SPECT/spect_system_matrix.py:16-201``: same
constructor, ``forward(object, subset_idx)``, ``backward(proj, subset_idx)``,
subset bookkeeping and sensitivity image.  With obj2obj transforms
``SPECTAttenuationTransform`` followed by ``SPECTPSFTransform`` (either may be
absent; the order the reference's examples and tests use,
``tests/test_multi_bed_recon.py:68-72``) and no proj2proj transform, each call
is a handful of stream-ordered fused kernel launches (no per-angle Python
loop, no host syncs).  Any other transforms run the reference's per-angle loop
over whole rotated frames, with the rotation, attenuation and PSF on HIP
kernels (``fused`` is False).

Extras beyond the reference:
  * ``backward_adjoint(proj, subset_idx)`` -- the exact transpose of
    ``forward`` (the reference ``backward`` gathers at -theta instead);
  * ``compute_normalization_factor`` runs in float64 on the GPU and is
    rounded to the working dtype;
  * ``osem_step`` -- forward, Poisson ratio and back projection fused with
    the OSEM update for ``OSEM``.
"""
from __future__ import annotations

import ctypes
import inspect

import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
from pytomography_amd.utils import compute_pad_size, pad_proj, unpad_proj
from ..system_matrix import SystemMatrix
from .plan import RotatedFrames, SpectPlan, plan_geometry


def _accepts_angles(transform) -> bool:
    """Whether a proj2proj transform's forward takes the projections' angle
    indices after the projections (CutOffTransform's signature,
    transforms/SPECT/cutoff.py:9-20) or the projections alone (the call of
    spect_system_matrix.py:155-156)."""
    try:
        params = list(inspect.signature(transform.forward).parameters.values())
    except (TypeError, ValueError):
        return False
    positional = [p for p in params if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
    return len(positional) >= 2 or any(p.kind == p.VAR_POSITIONAL for p in params)


class SPECTSystemMatrix(SystemMatrix):
    """System matrix for SPECT imaging using the rotate+sum technique.

    Args:
        obj2obj_transforms: ``[SPECTAttenuationTransform?, SPECTPSFTransform?]``.
        proj2proj_transforms: must be empty (no proj2proj transform is on
            the accelerated path).
        object_meta: ``SPECTObjectMeta``.
        proj_meta: ``SPECTProjMeta``.
        object_initial_based_on_camera_path: not supported (must be False).
        sensitivity_dtype: dtype of the H^T 1 computation (float64 default).
        chunk_planes: planes per uniform work chunk (0: library default, 16).
        sched_rounds: the forward projection and the back projection's field
            kernel, without attenuation or with the attenuation field, split
            their visited planes evenly over this many rounds of the blocks the
            GPU holds at once; 0 (default) chooses per call: one round when the
            call's pieces of at most 32 planes fit the resident blocks (for the
            forward also at most 19 planes per block on average), uniform chunks
            otherwise (large problems); -1 always uses uniform chunks of
            ``chunk_planes`` planes.
        attenuation_field: keep the attenuation factor of every rotated-frame
            voxel of every angle resident in device memory (built once per
            attenuation map) instead of resampling mu in every projection.
            None: on for float32 when it needs at most ``ATT_FIELD_AUTO_BYTES`` and at
            most ``ATT_FIELD_AUTO_FRACTION`` of the device's memory (e.g. 19 GB at
            256^3 x 180 angles).
    """

    ATT_FIELD_AUTO_BYTES = 64 << 30
    ATT_FIELD_AUTO_FRACTION = 0.25

    def __init__(self, obj2obj_transforms, proj2proj_transforms, object_meta, proj_meta,
                 object_initial_based_on_camera_path: bool = False,
                 sensitivity_dtype: torch.dtype = torch.float64, chunk_planes: int = 0,
                 sched_rounds: int = 0,
                 attenuation_field: bool | None = None) -> None:
        super().__init__(object_meta, proj_meta, obj2obj_transforms, proj2proj_transforms)
        if object_initial_based_on_camera_path:
            raise NotImplementedError("object_initial_based_on_camera_path is not supported")
        self.object_initial_based_on_camera_path = False
        self.sensitivity_dtype = sensitivity_dtype
        self._classify_transforms()
        shape = tuple(int(s) for s in object_meta.shape)
        if shape[0] != shape[1]:
            raise NotImplementedError("the kernels need a square transaxial grid (Lx == Ly)")
        if tuple(proj_meta.shape[1:]) != (shape[1], shape[2]):
            raise ValueError(f"projection shape {tuple(proj_meta.shape[1:])} must be (Ly, Lz) = "
                             f"{(shape[1], shape[2])}")
        self._shape = shape
        self.chunk_planes = int(chunk_planes)
        self.sched_rounds = int(sched_rounds)
        self.attenuation_field = attenuation_field
        self._plans: dict = {}
        self._ws: dict = {}
        self._idx32: dict = {}
        self._mu_cache: dict = {}
        self._att_tables: dict = {}
        self.subset_indices_array = None

    # ------------------------------------------------------------------ set-up
    def _classify_transforms(self) -> None:
        """Fused path: no proj2proj transform and obj2obj transforms
        [SPECTAttenuationTransform?, SPECTPSFTransform?] in that order (a
        Gaussian or square PSF from ``psf_meta``).  Anything else runs the
        reference's per-angle loop over whole rotated frames."""
        ts = self.obj2obj_transforms
        kinds = ["att" if isinstance(t, SPECTAttenuationTransform)
                 else "psf" if isinstance(t, SPECTPSFTransform) and t.fusable else "other" for t in ts]
        self.fused = not self.proj2proj_transforms and kinds in ([], ["att"], ["psf"], ["att", "psf"])
        self._att = ts[kinds.index("att")] if self.fused and "att" in kinds else None
        self._psf = ts[kinds.index("psf")] if self.fused and "psf" in kinds else None
        self._frames = None if self.fused else RotatedFrames(self.object_meta, self.proj_meta)
        self._proj2proj_takes_angles = [_accepts_angles(t) for t in self.proj2proj_transforms]

    def _require_fused(self, what: str) -> None:
        if not self.fused:
            raise NotImplementedError(
                f"{what} needs the fused transforms (SPECTAttenuationTransform then SPECTPSFTransform, "
                "no proj2proj transform); this system matrix runs the per-angle path")

    def _geom_args(self) -> dict:
        return plan_geometry(self.object_meta, self.proj_meta, self._psf, self.chunk_planes,
                             self.sched_rounds)

    def _plan(self, device: torch.device) -> SpectPlan:
        key = (device.type, device.index)
        if key not in self._plans:
            self._plans[key] = SpectPlan(self._geom_args(), device)
        return self._plans[key]

    def _workspace(self, device, dt: torch.dtype, n: int):
        plan = self._plan(device)
        nbytes = int(plan.lib.spect_workspace_bytes(plan.handle, _lib.dtype_code(dt), n))
        key = (device.type, device.index)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            self._ws[key] = ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        return ws, ws.numel()

    def schedule(self, subset_idx: int | None = None, dtype: torch.dtype | None = None) -> dict:
        """The plane schedule the fused kernels run for this subset's angles
        (``spect_schedule_info``): ``{"forward": ..., "field": ...}`` with
        "balanced" (pieces of the visited planes over one round of resident
        blocks) or "uniform" (``chunk_planes``-plane chunks); "forward" is the
        forward projection without attenuation or with the attenuation field,
        "field" the back projection's field kernel ("none" without attenuation
        and PSF, where the back projection gathers the projection directly)."""
        self._require_fused("schedule")
        device = pytomography_amd.device
        dt = dtype or (pytomography_amd.dtype if pytomography_amd.dtype == torch.float64
                       else torch.float32)
        plan = self._plan(device)
        n = int(self._angle_index(subset_idx, device).numel())
        fwd, fld = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(plan.lib.spect_schedule_info(plan.handle, _lib.dtype_code(dt), n,
                                                ctypes.byref(fwd), ctypes.byref(fld)))
        name = lambda v: "balanced" if v.value else "uniform"
        direct = self._att is None and self._psf is None
        return {"forward": name(fwd), "field": "none" if direct else name(fld)}

    def _angle_index(self, subset_idx, device):
        key = (subset_idx, device.type, device.index)
        if key not in self._idx32:
            if subset_idx is None:
                idx = torch.arange(self.proj_meta.num_projections)
            else:
                idx = self.subset_indices_array[subset_idx]
            self._idx32[key] = idx.to(device=device, dtype=torch.int32).contiguous()
        return self._idx32[key]

    def _mu(self, device, dt):
        if self._att is None:
            return None
        key = (device.type, device.index, dt)
        mu = self._att.attenuation_map
        cached = self._mu_cache.get(key)
        if cached is None or cached[0] is not mu or cached[2] != mu._version:
            self._mu_cache[key] = (mu, _lib.require_device_tensor(mu.detach().to(device), "attenuation_map", dt),
                                   mu._version)
            self._att_tables.pop(key, None)
        return self._mu_cache[key][1]

    def _use_field(self, device, dt) -> bool:
        if self.attenuation_field is not None:
            return bool(self.attenuation_field)
        if dt != torch.float32:
            return False
        plan = self._plan(device)
        need = int(plan.lib.spect_att_field_bytes(plan.handle, _lib.dtype_code(dt)))
        total = torch.cuda.get_device_properties(device).total_memory
        return need <= min(self.ATT_FIELD_AUTO_BYTES, self.ATT_FIELD_AUTO_FRACTION * total)

    def _att_table(self, device, dt):
        """Attenuation data of the current mu, built once per map: the FP64
        suffix-sum table, or the attenuation field (see ``attenuation_field``)."""
        mu = self._mu(device, dt)
        if mu is None:
            return None
        key = (device.type, device.index, dt)
        tab = self._att_tables.get(key)
        if tab is None:
            plan = self._plan(device)
            if self._use_field(device, dt):
                nbytes = int(plan.lib.spect_att_field_bytes(plan.handle, _lib.dtype_code(dt)))
                tab = torch.empty(nbytes // (torch.finfo(dt).bits // 8), dtype=dt, device=device)
                _lib.check(plan.lib.spect_att_field_build(plan.handle, _lib.dtype_code(dt), mu.data_ptr(),
                                                          tab.data_ptr(), _lib.stream_ptr(device)))
                tab.att_field = True
            else:
                nbytes = int(plan.lib.spect_att_table_bytes(plan.handle))
                tab = torch.empty(nbytes // 8, dtype=torch.float64, device=device)
                _lib.check(plan.lib.spect_att_table_build(plan.handle, _lib.dtype_code(dt), mu.data_ptr(),
                                                          tab.data_ptr(), _lib.stream_ptr(device)))
                tab.att_field = False
            self._att_tables[key] = tab
        return tab

    @staticmethod
    def _code(dt, tab) -> int:
        """dtype argument of the projection entry points for attenuation data ``tab``."""
        code = _lib.dtype_code(dt)
        return code | _lib.ATT_FIELD if tab is not None and tab.att_field else code

    @staticmethod
    def _compute_dtype(t: torch.Tensor) -> torch.dtype:
        return torch.float64 if t.dtype == torch.float64 else torch.float32

    # --------------------------------------------------------------- subsets
    def set_n_subsets(self, n_subsets: int) -> None:
        """Interleaved subsets ``arange(A)[k::n_subsets]`` (spect_system_matrix.py:58-71)."""
        indices = torch.arange(self.proj_meta.shape[0]).to(torch.long).to(pytomography_amd.device)
        self.subset_indices_array = [indices[i::n_subsets] for i in range(n_subsets)]
        self._idx32 = {k: v for k, v in self._idx32.items() if k[0] is None}

    def get_projection_subset(self, projections: torch.Tensor, subset_idx: int) -> torch.Tensor:
        """spect_system_matrix.py:73-87."""
        return projections[..., self.subset_indices_array[subset_idx], :, :]

    def get_weighting_subset(self, subset_idx) -> float:
        """spect_system_matrix.py:89-104."""
        if subset_idx is None:
            return 1
        return len(self.subset_indices_array[subset_idx]) / self.proj_meta.num_projections

    def compute_normalization_factor(self, subset_idx=None) -> torch.Tensor:
        """H_m^T 1 (spect_system_matrix.py:106-118), computed in
        ``sensitivity_dtype`` on the GPU and returned in the working dtype."""
        device = pytomography_amd.device
        n = self.proj_meta.num_projections if subset_idx is None else len(self.subset_indices_array[subset_idx])
        ones = torch.ones((n, *self.proj_meta.shape[1:]), dtype=self.sensitivity_dtype, device=device)
        if not self.fused:
            return self._backward_frames(ones, subset_idx).to(pytomography_amd.dtype)
        out = self._backward(ones, subset_idx, _lib.BP_REFERENCE)
        return out.to(pytomography_amd.dtype)

    # ------------------------------------------------------------ projections
    def forward(self, object: torch.Tensor, subset_idx: int | None = None) -> torch.Tensor:
        """H_m f: projections [n, Ly, Lz] (spect_system_matrix.py:120-157)."""
        if not self.fused:
            return self._forward_frames(object, subset_idx)
        return self._project(object, subset_idx, "spect_forward")

    # ------------------------------------------------------- per-angle path
    def _angle_list(self, subset_idx):
        """The angle indices of a call as the reference's loop sees them (a
        long tensor on the device, spect_system_matrix.py:134-137)."""
        if subset_idx is None:
            return torch.arange(self.proj_meta.num_projections).to(pytomography_amd.device)
        return self.subset_indices_array[subset_idx]

    def _frame_chunk(self, plan, dt) -> int:
        """Angles whose rotated frames are held at once (about 1 GiB)."""
        per = plan.padded * plan.padded * self._shape[2] * (8 if dt == torch.float64 else 4)
        return max(1, (1 << 30) // per)

    def _proj2proj(self, proj, angles, backward: bool):
        ts = self.proj2proj_transforms
        order = range(len(ts) - 1, -1, -1) if backward else range(len(ts))
        for k in order:
            fn = ts[k].backward if backward else ts[k].forward
            proj = fn(proj, angles) if self._proj2proj_takes_angles[k] else fn(proj)
        return proj

    def _forward_frames(self, object: torch.Tensor, subset_idx) -> torch.Tensor:
        """spect_system_matrix.py:145-157: every angle's padded object rotated
        (HIP), the obj2obj transforms in order, the sum along the rays, then
        the proj2proj transforms on the padded projections."""
        device = pytomography_amd.device if object.device.type != "cuda" else object.device
        dt = self._compute_dtype(object)
        f = _lib.require_device_tensor(object.to(device), "object", dt)
        if tuple(f.shape) != self._shape:
            raise ValueError(f"object shape {tuple(f.shape)} != {self._shape}")
        angles = self._angle_list(subset_idx)
        n = len(angles)
        plan = self._frames.plan(device)
        proj = torch.empty((n, plan.padded, self._shape[2]), dtype=dt, device=device)
        chunk = self._frame_chunk(plan, dt)
        for c0 in range(0, n, chunk):
            frames = self._frames.rotate(f, angles[c0:c0 + chunk])
            for k in range(frames.shape[0]):
                object_i = frames[k]
                for transform in self.obj2obj_transforms:
                    object_i = transform.forward(object_i, angles[c0 + k])
                proj[c0 + k] = object_i.sum(axis=0)
            del frames
        proj = self._proj2proj(proj, angles, backward=False)
        return unpad_proj(proj)

    def _backward_frames(self, proj: torch.Tensor, subset_idx) -> torch.Tensor:
        """spect_system_matrix.py:178-201: the proj2proj transforms in reverse
        on the padded projections, each angle's projection smeared along the
        rays inside the back-projection box (object columns only), the
        obj2obj transforms in reverse, then the rotation and sum (HIP)."""
        device = pytomography_amd.device if proj.device.type != "cuda" else proj.device
        dt = self._compute_dtype(proj)
        angles = self._angle_list(subset_idx)
        n = len(angles)
        p = proj.to(device=device, dtype=dt)
        if tuple(p.shape) != (n, self._shape[1], self._shape[2]):
            raise ValueError(f"projection shape {tuple(p.shape)} != {(n, self._shape[1], self._shape[2])}")
        p = self._proj2proj(pad_proj(p), angles, backward=True)
        plan = self._frames.plan(device)
        P = plan.padded
        box = torch.zeros((P, 1), dtype=dt, device=device)  # pad_object(ones, 'back_project')
        box[plan.pad:plan.pad + self._shape[1]] = 1
        out = None
        chunk = self._frame_chunk(plan, dt)
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            frames = torch.empty((m, P, P, self._shape[2]), dtype=dt, device=device)
            for k in range(m):
                object_i = (p[c0 + k] * box).unsqueeze(0).expand(P, P, self._shape[2]).contiguous()
                for transform in self.obj2obj_transforms[::-1]:
                    object_i = transform.backward(object_i, angles[c0 + k])
                frames[k] = object_i
            out = self._frames.rotate_back(frames, angles[c0:c0 + m], out)
            del frames
        return out

    def backward_transpose(self, object: torch.Tensor, subset_idx: int | None = None) -> torch.Tensor:
        """B_m^T v, the transpose of the reference back projection ``backward``
        (spect_system_matrix.py:159-201): the forward chain with the rotation
        replaced by the transpose of the gather at 270-beta.  Used to
        differentiate through iterations that call ``backward``."""
        self._require_fused("backward_transpose")
        return self._project(object, subset_idx, "spect_backward_transpose")

    def _project(self, object: torch.Tensor, subset_idx, entry: str) -> torch.Tensor:
        device = pytomography_amd.device if object.device.type != "cuda" else object.device
        dt = self._compute_dtype(object)
        f = _lib.require_device_tensor(object.to(device), "object", dt)
        if tuple(f.shape) != self._shape:
            raise ValueError(f"object shape {tuple(f.shape)} != {self._shape}")
        plan = self._plan(device)
        idx = self._angle_index(subset_idx, device)
        n = idx.numel()
        mu = self._mu(device, dt)
        tab = self._att_table(device, dt)
        out = torch.empty((n, self._shape[1], self._shape[2]), dtype=dt, device=device)
        ws, nbytes = self._workspace(device, dt, n)
        _lib.check(getattr(plan.lib, entry)(plan.handle, self._code(dt, tab), f.data_ptr(),
                                            _lib.ptr(mu), _lib.ptr(tab), idx.data_ptr(), n,
                                            out.data_ptr(), ws.data_ptr(), nbytes,
                                            _lib.stream_ptr(device)))
        return out

    def _backward(self, proj, subset_idx, mode, out_mode=_lib.OUT_STORE, out=None, f=None,
                  norm=None, floor_value=0.0):
        device = pytomography_amd.device if proj.device.type != "cuda" else proj.device
        dt = self._compute_dtype(proj)
        p = _lib.require_device_tensor(proj.to(device), "proj", dt)
        plan = self._plan(device)
        idx = self._angle_index(subset_idx, device)
        n = idx.numel()
        if tuple(p.shape) != (n, self._shape[1], self._shape[2]):
            raise ValueError(f"projection shape {tuple(p.shape)} != {(n, self._shape[1], self._shape[2])}")
        mu = self._mu(device, dt)
        tab = self._att_table(device, dt)
        if out is None:
            out = torch.empty(self._shape, dtype=dt, device=device)
        ws, nbytes = self._workspace(device, dt, n)
        _lib.check(plan.lib.spect_backward(
            plan.handle, self._code(dt, tab), mode, out_mode, p.data_ptr(), _lib.ptr(mu),
            _lib.ptr(tab), idx.data_ptr(), n, out.data_ptr(), _lib.ptr(f), _lib.ptr(norm),
            float(pytomography_amd.delta), float(floor_value), ws.data_ptr(), nbytes,
            _lib.stream_ptr(device)))
        return out

    def backward(self, proj: torch.Tensor, subset_idx: int | None = None) -> torch.Tensor:
        """Reference back projection (spect_system_matrix.py:159-201): PSF,
        attenuation, then rotation by 270-beta (a bilinear gather)."""
        if not self.fused:
            return self._backward_frames(proj, subset_idx)
        return self._backward(proj, subset_idx, _lib.BP_REFERENCE)

    def backward_adjoint(self, proj: torch.Tensor, subset_idx: int | None = None) -> torch.Tensor:
        """Exact transpose H_m^T of ``forward``: <H f, y> = <f, H^T y>."""
        self._require_fused("backward_adjoint")
        return self._backward(proj, subset_idx, _lib.BP_ADJOINT)

    def osem_step(self, object: torch.Tensor, projections: torch.Tensor, additive: torch.Tensor | None,
                  norm: torch.Tensor, subset_idx, pred_out: torch.Tensor | None = None,
                  floor_value: float = 0.0, keep_prediction: bool = True) -> torch.Tensor | None:
        """One OSEM sub-iteration in place on ``object`` (a contiguous device
        tensor): pred = H_m f + s; r = g / (pred + delta);
        f += f/(n+delta) * (B_m r - n); values <= floor_value become floor_value.
        Returns pred (``projections_predicted``) unless keep_prediction is False."""
        self._require_fused("osem_step")
        device = object.device
        dt = object.dtype
        if object.device.type != "cuda" or not object.is_contiguous() or object.data_ptr() % 16:
            raise ValueError("osem_step updates a contiguous, aligned device tensor in place")
        plan = self._plan(device)
        idx = self._angle_index(subset_idx, device)
        n = idx.numel()
        g = _lib.require_device_tensor(projections, "projections", dt)
        s = None if additive is None else _lib.require_device_tensor(additive, "additive_term", dt)
        nrm = _lib.require_device_tensor(norm, "norm", dt)
        mu = self._mu(device, dt)
        tab = self._att_table(device, dt)
        if keep_prediction and pred_out is None:
            pred_out = torch.empty((n, self._shape[1], self._shape[2]), dtype=dt, device=device)
        ws, nbytes = self._workspace(device, dt, n)
        _lib.check(plan.lib.spect_osem_subiteration(
            plan.handle, self._code(dt, tab), object.data_ptr(), _lib.ptr(mu), _lib.ptr(tab),
            idx.data_ptr(), n, g.data_ptr(), _lib.ptr(s), nrm.data_ptr(),
            _lib.ptr(pred_out) if keep_prediction else None, float(pytomography_amd.delta),
            float(floor_value), ws.data_ptr(), nbytes, _lib.stream_ptr(device)))
        return pred_out

    # ------------------------------------------------- differentiable projector
    def forward_with_mu(self, object: torch.Tensor, mu: torch.Tensor | None,
                        subset_idx: int | None = None) -> torch.Tensor:
        """H(mu) f with an explicit attenuation map (rows rebuilt every call, no
        cache): the forward of the differentiable projector."""
        self._require_fused("forward_with_mu")
        device = object.device
        dt = self._compute_dtype(object)
        f = _lib.require_device_tensor(object, "object", dt)
        m = None if mu is None else _lib.require_device_tensor(mu.detach(), "mu", dt)
        if (m is None) != (self._att is None):
            raise ValueError("mu must be given exactly when the system matrix models attenuation")
        plan = self._plan(device)
        idx = self._angle_index(subset_idx, device)
        n = idx.numel()
        out = torch.empty((n, self._shape[1], self._shape[2]), dtype=dt, device=device)
        ws, nbytes = self._workspace(device, dt, n)
        _lib.check(plan.lib.spect_forward(plan.handle, _lib.dtype_code(dt), f.data_ptr(), _lib.ptr(m),
                                          None, idx.data_ptr(), n, out.data_ptr(), ws.data_ptr(),
                                          nbytes, _lib.stream_ptr(device)))
        return out

    def forward_vjp(self, grad_proj: torch.Tensor, object: torch.Tensor, mu: torch.Tensor | None,
                    subset_idx: int | None = None, need_mu: bool = True):
        """(H(mu)^T u, (d H(mu) f / d mu)^T u) for upstream gradient u = grad_proj."""
        self._require_fused("forward_vjp")
        device = object.device
        dt = self._compute_dtype(object)
        u = _lib.require_device_tensor(grad_proj, "grad_proj", dt)
        f = _lib.require_device_tensor(object.detach(), "object", dt)
        m = None if mu is None else _lib.require_device_tensor(mu.detach(), "mu", dt)
        plan = self._plan(device)
        idx = self._angle_index(subset_idx, device)
        n = idx.numel()
        gf = torch.empty(self._shape, dtype=dt, device=device)
        gm = torch.empty(self._shape, dtype=dt, device=device) if (need_mu and m is not None) else None
        nbytes = int(plan.lib.spect_grad_workspace_bytes(plan.handle, _lib.dtype_code(dt), n))
        key = ("grad", device.type, device.index)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            self._ws[key] = ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _lib.check(plan.lib.spect_forward_vjp(plan.handle, _lib.dtype_code(dt), u.data_ptr(),
                                              f.data_ptr(), _lib.ptr(m), idx.data_ptr(), n,
                                              gf.data_ptr(), _lib.ptr(gm), ws.data_ptr(), ws.numel(),
                                              _lib.stream_ptr(device)))
        return gf, gm

    def _get_object_initial(self, device=None):
        """Ones (spect_system_matrix.py:38-44)."""
        if device is None:
            device = pytomography_amd.device
        return torch.ones(self.object_meta.shape).to(device)
