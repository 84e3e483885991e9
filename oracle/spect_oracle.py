"""Torch-CPU restatement of PyTomography's SPECT rotate+sum path (the oracle).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this module, and only as the
checker.  The product package never calls it.

Every function follows one reference function op for op, in the same dtype
and the same order of operations, so that its float32 rounding behaviour is
the reference's.  Citations are relative to the PyTomography 3.3.2 source tree
(``src/pytomography/...``).

The reference delegates rotation to ``kornia.geometry.transform.rotate``
(kornia is not installed here; ``pyproject.toml:26`` pins only ``>=0.6.12``).
``kornia_rotate`` below restates kornia's published algorithm for that call:
3x3 matrix ``shift @ rot @ scale @ shift^-1`` built in the input dtype,
``normalize_homography`` to [-1, 1] coordinates, a 3x3 inverse, then torch's
own ``affine_grid`` + ``grid_sample(bilinear, zeros, align_corners=True)``.

Pinning (see tests/test_oracle.py): closed forms for padding, rotation by
0/90/180 degrees, linear ramps, constant fields, uniform-mu attenuation, PSF
taps by formula, scipy ``correlate1d`` for the blur, a numpy FP64 brute-force
dense system matrix on tiny geometry (``oracle/brute.py``), finite differences
for the prior and the likelihood gradient.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch
import torch.nn.functional as F

DELTA = 1e-11  # pytomography.delta for float32/float64 (__init__.py:21,24-32)


# --------------------------------------------------------------------------
# padding helpers -- utils/spatial.py:36-119
# --------------------------------------------------------------------------
def compute_pad_size(width: int) -> int:
    """utils/spatial.py:36-45."""
    return int(np.ceil((np.sqrt(2) * width - width) / 2))


def compute_pad_size_padded(width: int) -> int:
    """utils/spatial.py:47-62."""
    a = (np.sqrt(2) - 1) / 2
    if width % 2 == 0:
        width_old = int(2 * np.floor((width / 2) / (1 + 2 * a)))
    else:
        width_old = int(2 * np.floor(((width - 1) / 2) / (1 + 2 * a)))
    return int((width - width_old) / 2)


def pad_object(obj: torch.Tensor, mode: str = "constant") -> torch.Tensor:
    """utils/spatial.py:64-81 (pad from shape[-2] on x and y)."""
    p = compute_pad_size(obj.shape[-2])
    if mode == "back_project":
        obj = F.pad(obj.unsqueeze(0), [0, 0, 0, 0, p, p], mode="replicate").squeeze(0)
        return F.pad(obj, [0, 0, p, p], mode="constant")
    return F.pad(obj, [0, 0, p, p, p, p], mode=mode)


def unpad_object(obj: torch.Tensor) -> torch.Tensor:
    """utils/spatial.py:83-93."""
    p = compute_pad_size_padded(obj.shape[-2])
    return obj[p:-p, p:-p, :]


def pad_proj(proj: torch.Tensor) -> torch.Tensor:
    """utils/spatial.py:95-107."""
    p = compute_pad_size(proj.shape[-2])
    return F.pad(proj, [0, 0, p, p], mode="constant", value=0)


def unpad_proj(proj: torch.Tensor) -> torch.Tensor:
    """utils/spatial.py:109-119."""
    p = compute_pad_size_padded(proj.shape[-2])
    return proj[:, p:-p, :]


# --------------------------------------------------------------------------
# rotation -- kornia.geometry.transform.rotate as called at
# utils/spatial.py:29-33 and transforms/shared/spatial.py:36,53
# --------------------------------------------------------------------------
def _eye3(dt):
    return torch.eye(3, dtype=dt).unsqueeze(0)


def _normal_transform_pixel(h: int, w: int, dt) -> torch.Tensor:
    m = torch.tensor([[1.0, 0.0, -1.0], [0.0, 1.0, -1.0], [0.0, 0.0, 1.0]], dtype=dt)
    wd = 1e-14 if w == 1 else w - 1.0
    hd = 1e-14 if h == 1 else h - 1.0
    m[0, 0] = m[0, 0] * 2.0 / wd
    m[1, 1] = m[1, 1] * 2.0 / hd
    return m.unsqueeze(0)


def kornia_rotate_matrix(angle_deg: torch.Tensor, h: int, w: int) -> torch.Tensor:
    """src_norm <- dst_norm 2x3 matrix that kornia's ``rotate`` hands to
    ``affine_grid`` (rotation about ((W-1)/2, (H-1)/2), unit scale)."""
    dt = angle_deg.dtype
    angle = angle_deg.reshape(1)
    center = torch.tensor([[(w - 1) / 2, (h - 1) / 2]], dtype=dt)
    rad = angle * torch.tensor(math.pi, dtype=dt) / 180.0
    c, s = torch.cos(rad), torch.sin(rad)
    rot2 = torch.stack([c, s, -s, c], dim=-1).view(1, 2, 2)
    shift = _eye3(dt).clone()
    shift[:, :2, 2] = center
    shift_inv = _eye3(dt).clone()
    shift_inv[:, :2, 2] = -center
    scale = _eye3(dt).clone()
    rot = _eye3(dt).clone()
    rot[:, :2, :2] = rot2
    aff = (shift @ rot @ scale @ shift_inv)[:, :2, :]
    m3 = torch.cat([aff, torch.tensor([[[0.0, 0.0, 1.0]]], dtype=dt)], dim=1)
    src_n = _normal_transform_pixel(h, w, dt)
    dst_n = _normal_transform_pixel(h, w, dt)
    dst_norm_src_norm = dst_n @ (m3 @ torch.inverse(src_n))
    return torch.inverse(dst_norm_src_norm)[:, :2, :]


def kornia_rotate(t: torch.Tensor, angle_deg: torch.Tensor) -> torch.Tensor:
    """kornia ``rotate(t, angle)`` for t of shape [1, C, H, W]."""
    _, c, h, w = t.shape
    theta = kornia_rotate_matrix(angle_deg.to(t.dtype), h, w)
    grid = F.affine_grid(theta, [1, c, h, w], align_corners=True)
    return F.grid_sample(t, grid, mode="bilinear", padding_mode="zeros", align_corners=True)


def rotate_xyz(x: torch.Tensor, angle_deg: torch.Tensor) -> torch.Tensor:
    """``rotate(x.permute(2,0,1).unsqueeze(0), angle).squeeze().permute(1,2,0)``
    -- the shared body of RotationTransform.forward (transforms/shared/spatial.py:36)
    and rotate_detector_z (utils/spatial.py:31,33)."""
    out = kornia_rotate(x.permute(2, 0, 1).unsqueeze(0), angle_deg)
    return out[0].permute(1, 2, 0)


def rotate_detector_z(x: torch.Tensor, beta: torch.Tensor, negative: bool = False) -> torch.Tensor:
    """utils/spatial.py:8-34: phi = 270 - beta; rotate by -phi (or +phi)."""
    phi = 270 - beta
    return rotate_xyz(x, phi if negative else -phi)


# --------------------------------------------------------------------------
# attenuation -- utils/misc.py:8-18, transforms/SPECT/attenuation.py:11-115
# --------------------------------------------------------------------------
def rev_cumsum(x: torch.Tensor) -> torch.Tensor:
    """utils/misc.py:8-18 (CPU cumsum on float32 accumulates in float64)."""
    return torch.cumsum(x.flip(dims=(0,)), dim=0).flip(dims=(0,)) - x / 2


def prob_of_detection(mu_rot: torch.Tensor, dx: float) -> torch.Tensor:
    """attenuation.py:11-21."""
    return torch.exp(-rev_cumsum(mu_rot * dx))


def attenuation_factor(mu: torch.Tensor, beta: torch.Tensor, dx: float) -> torch.Tensor:
    """attenuation.py:87-91 (assume_padded=True)."""
    return prob_of_detection(rotate_detector_z(pad_object(mu), beta), dx)


# --------------------------------------------------------------------------
# PSF -- utils/misc.py:21-38, transforms/SPECT/psf.py:15-41,62-83,123-231
# --------------------------------------------------------------------------
def get_distance(lx: int, r: float, dx: float) -> np.ndarray:
    """utils/misc.py:21-38."""
    if lx % 2 == 0:
        d = r + (lx / 2 - np.arange(lx) - 1 / 2) * dx
    else:
        d = r + ((lx - 1) / 2 - np.arange(lx)) * dx
    d[d < 0] = 0
    return d


@dataclass
class PSFModel:
    """SPECTPSFMeta (metadata/SPECT/spect_metadata.py:66-94)."""
    sigma_fit_params: tuple
    sigma_fit: object = None  # default a*r+b (spect_metadata.py:79)
    kernel_dimensions: str = "2D"
    min_sigmas: float = 3
    shape: str = "gaussian"   # or "square"

    def fit(self, d):
        if self.sigma_fit is None:
            a, b = self.sigma_fit_params
            return a * d + b
        return self.sigma_fit(d, *self.sigma_fit_params)


def psf_sigma(psf: PSFModel, n: int, radius: float, dx: float) -> np.ndarray:
    """SPECTPSFTransform._get_sigma (psf.py:180-197), assume_padded=True."""
    dim = n + 2 * compute_pad_size(n)
    return psf.fit(get_distance(dim, radius, dx))


def psf_kernel_size(psf: PSFModel, n: int, radius: float, dx: float, d_axis: float) -> int:
    """SPECTPSFTransform._compute_kernel_size (psf.py:170-178)."""
    sigma_max = max(psf_sigma(psf, n, radius, dx)) / d_axis
    return int((np.ceil(sigma_max * psf.min_sigmas) * 2 + 1).astype(int))


def gaussian_taps(sigmas: np.ndarray, k: int, dtype=torch.float32) -> torch.Tensor:
    """get_1D_PSF_layer_gaussian (psf.py:62-83): taps [len(sigmas), k]."""
    n = len(sigmas)
    x = torch.arange(-int(k // 2), int(k // 2) + 1).unsqueeze(0).unsqueeze(0).repeat((n, 1, 1))
    s = torch.tensor(sigmas).to(dtype).reshape((n, 1, 1))
    kern = torch.exp(-x ** 2 / (2 * s ** 2 + DELTA))
    kern = kern / kern.sum(axis=-1).unsqueeze(-1)
    return kern.to(dtype)[:, 0, :]


def square_taps(sigmas: np.ndarray, k: int, dtype=torch.float32) -> torch.Tensor:
    """get_1D_PSF_layer_square (psf.py:85-101): taps [len(sigmas), k]."""
    n = len(sigmas)
    x = torch.arange(-int(k // 2), int(k // 2) + 1).unsqueeze(0).unsqueeze(0).repeat((n, 1, 1))
    s = torch.tensor(sigmas).to(dtype).reshape((n, 1, 1))
    kern = 1 - torch.abs(x) / s
    kern[kern < 0] = 0
    kern = kern / kern.sum(axis=-1).unsqueeze(-1)
    return kern.to(dtype)[:, 0, :]


def psf_layers(psf: PSFModel, shape, dr, radius: float, dtype=torch.float32):
    """_configure_simple_model (psf.py:123-142) for one radius: (taps_r, taps_z|None)."""
    n = shape[0]
    kr = psf_kernel_size(psf, n, radius, dr[0], dr[0])
    kz = psf_kernel_size(psf, n, radius, dr[0], dr[2])
    sig = psf_sigma(psf, n, radius, dr[0])
    taps = square_taps if psf.shape == "square" else gaussian_taps
    taps_r = taps(sig / dr[0], kr, dtype)
    taps_z = taps(sig / dr[2], kz, dtype) if psf.kernel_dimensions == "2D" else None
    return taps_r, taps_z


def psf_apply(x: torch.Tensor, taps_r: torch.Tensor, taps_z: torch.Tensor | None) -> torch.Tensor:
    """Seperable1DBlurNet.forward (psf.py:27-41): depthwise conv1d along y
    (channels = x planes), then along z."""
    p = taps_r.shape[0]
    out = F.conv1d(x.permute(2, 0, 1), taps_r.unsqueeze(1), padding="same", groups=p)
    if taps_z is not None:
        out = F.conv1d(out.permute(2, 1, 0), taps_z.unsqueeze(1), padding="same", groups=p)
        out = out.permute(2, 1, 0)
    return out.permute(1, 2, 0)


# --------------------------------------------------------------------------
# system matrix -- projectors/SPECT/spect_system_matrix.py:58-201
# --------------------------------------------------------------------------
@dataclass
class SPECTGeometry:
    """Object/projection metadata + transforms of one SPECT problem."""
    shape: tuple                      # (Lx, Ly, Lz)
    dr: tuple                         # voxel size (cm)
    angles: torch.Tensor              # beta, degrees, dtype of the run
    radii: np.ndarray | None = None   # cm per angle (PSF only)
    mu: torch.Tensor | None = None    # attenuation map (1/cm)
    psf: PSFModel | None = None
    dtype: torch.dtype = torch.float32
    _layers: dict = field(default_factory=dict, repr=False)

    @property
    def n_angles(self) -> int:
        return int(self.angles.shape[0])

    def layers(self, ang_idx: int):
        r = float(self.radii[ang_idx])
        if r not in self._layers:
            self._layers[r] = psf_layers(self.psf, self.shape, self.dr, r, self.dtype)
        return self._layers[r]

    def subsets(self, n_subsets: int):
        """set_n_subsets (spect_system_matrix.py:58-71)."""
        idx = torch.arange(self.n_angles, dtype=torch.long)
        return [idx[i::n_subsets] for i in range(n_subsets)]


def make_geometry(shape, dr, angles, radii=None, mu=None, psf=None, dtype=torch.float32):
    """SPECTProjMeta stores angles as a tensor of the global dtype (spect_metadata.py:50)."""
    ang = torch.tensor(np.asarray(angles, dtype=np.float64)).to(dtype)
    if mu is not None:
        mu = mu.to(dtype)
    return SPECTGeometry(tuple(shape), tuple(dr), ang, None if radii is None else np.asarray(radii),
                         mu, psf, dtype)


def _obj2obj_forward(g: SPECTGeometry, obj_i, ang_idx):
    """attenuation.forward (attenuation.py:72-92) then psf.forward (psf.py:199-214)."""
    if g.mu is not None:
        obj_i = obj_i * attenuation_factor(g.mu, g.angles[ang_idx], g.dr[0])
    if g.psf is not None:
        tr, tz = g.layers(ang_idx)
        obj_i = psf_apply(obj_i, tr, tz)
    return obj_i


def _obj2obj_backward(g: SPECTGeometry, obj_i, ang_idx):
    """Transforms in reverse order: psf.backward (psf.py:216-231) then
    attenuation.backward (attenuation.py:94-115)."""
    if g.psf is not None:
        tr, tz = g.layers(ang_idx)
        obj_i = psf_apply(obj_i, tr, tz)
    if g.mu is not None:
        obj_i = obj_i * attenuation_factor(g.mu, g.angles[ang_idx], g.dr[0])
    return obj_i


def _obj2obj_list(g: SPECTGeometry, obj2obj):
    """The obj2obj transforms of a call: the geometry's (attenuation then
    PSF) by default, else a list whose items are "att", "psf" or objects with
    ``forward(obj_i, ang_idx)`` / ``backward(obj_i, ang_idx)``."""
    if obj2obj is None:
        return (["att"] if g.mu is not None else []) + (["psf"] if g.psf is not None else [])
    return list(obj2obj)


def _apply_obj2obj(g: SPECTGeometry, t, obj_i, ang_idx: int, backward: bool):
    """attenuation.forward/backward (attenuation.py:72-115), psf.forward/
    backward (psf.py:199-231), or a caller's transform."""
    if t == "att":
        return obj_i * attenuation_factor(g.mu, g.angles[ang_idx], g.dr[0])
    if t == "psf":
        tr, tz = g.layers(ang_idx)
        return psf_apply(obj_i, tr, tz)
    return t.backward(obj_i, ang_idx) if backward else t.forward(obj_i, ang_idx)


def forward(g: SPECTGeometry, f: torch.Tensor, angle_indices=None, obj2obj=None,
            proj2proj=()) -> torch.Tensor:
    """SPECTSystemMatrix.forward (spect_system_matrix.py:120-157): per angle
    pad, rotate, obj2obj transforms in order, sum along x; then the proj2proj
    transforms (objects with ``forward(proj, angle_indices)``) on the padded
    projections; unpad."""
    if angle_indices is None:
        angle_indices = torch.arange(g.n_angles)
    ts = _obj2obj_list(g, obj2obj)
    f = f.to(g.dtype)
    p = compute_pad_size(g.shape[1])
    proj = torch.zeros((len(angle_indices), g.shape[1] + 2 * p, g.shape[2]), dtype=g.dtype)
    for i in range(len(angle_indices)):
        a = int(angle_indices[i])
        obj_i = pad_object(f)
        obj_i = rotate_xyz(obj_i, -(270 - g.angles[a]))       # RotationTransform.backward
        for t in ts:
            obj_i = _apply_obj2obj(g, t, obj_i, a, backward=False)
        proj[i] = obj_i.sum(axis=0)
    for t in proj2proj:
        proj = t.forward(proj, angle_indices)
    return unpad_proj(proj)


def backward(g: SPECTGeometry, proj: torch.Tensor, angle_indices=None, obj2obj=None,
             proj2proj=()) -> torch.Tensor:
    """SPECTSystemMatrix.backward (spect_system_matrix.py:159-201): gather at
    -theta (RotationTransform.forward with 270-beta), not the transpose; the
    proj2proj and obj2obj transforms' backward in reverse order."""
    if angle_indices is None:
        angle_indices = torch.arange(g.n_angles)
    ts = _obj2obj_list(g, obj2obj)
    box = pad_object(torch.ones(g.shape, dtype=g.dtype), mode="back_project")
    proj = pad_proj(proj.to(g.dtype))
    for t in list(proj2proj)[::-1]:
        proj = t.backward(proj, angle_indices)
    pp = compute_pad_size(g.shape[1])
    obj = torch.zeros((g.shape[0] + 2 * pp, g.shape[1] + 2 * pp, g.shape[2]), dtype=g.dtype)
    for i in range(len(angle_indices)):
        a = int(angle_indices[i])
        obj_i = proj[i].unsqueeze(0) * box
        for t in ts[::-1]:
            obj_i = _apply_obj2obj(g, t, obj_i, a, backward=True)
        obj_i = rotate_xyz(obj_i, 270 - g.angles[a])          # RotationTransform.forward
        obj = obj + obj_i
    return unpad_object(obj)


def average_prob_matrix(g: SPECTGeometry) -> torch.Tensor:
    """SPECTAttenuationTransform.compute_average_prob_matrix
    (attenuation.py:117-124): each angle's probability of detection rotated
    back (rotate_detector_z negative), averaged over the angles, unpadded."""
    mu = pad_object(g.mu)
    acc = torch.zeros(mu.shape, dtype=g.dtype)
    for a in range(g.n_angles):
        prob = prob_of_detection(rotate_detector_z(mu, g.angles[a]), g.dr[0])
        acc = acc + rotate_detector_z(prob, g.angles[a], negative=True)
    return unpad_object(acc / g.n_angles)


def forward_autograd(g: SPECTGeometry, f: torch.Tensor, mu: torch.Tensor | None,
                     angle_indices=None) -> torch.Tensor:
    """The same op sequence as ``forward`` with autograd left on (the reference
    wraps every transform in ``torch.no_grad``), differentiable in f and mu.
    This is the 'reference autograd gradient' of the differentiable projector."""
    if angle_indices is None:
        angle_indices = torch.arange(g.n_angles)
    p = compute_pad_size(g.shape[1])
    out = []
    for i in range(len(angle_indices)):
        a = int(angle_indices[i])
        obj_i = rotate_xyz(pad_object(f), -(270 - g.angles[a]))
        if mu is not None:
            obj_i = obj_i * attenuation_factor(mu, g.angles[a], g.dr[0])
        if g.psf is not None:
            tr, tz = g.layers(a)
            obj_i = psf_apply(obj_i, tr, tz)
        out.append(obj_i.sum(axis=0))
    proj = torch.stack(out)
    return proj[:, p:-p, :]


def exact_adjoint(g: SPECTGeometry, proj: torch.Tensor, angle_indices=None) -> torch.Tensor:
    """H^T proj for the forward operator above, via torch autograd (the
    vector-Jacobian product of the linear map f -> H f)."""
    f = torch.zeros(g.shape, dtype=g.dtype, requires_grad=True)
    y = forward_autograd(g, f, g.mu, angle_indices)
    (grad,) = torch.autograd.grad(y, f, grad_outputs=proj.to(g.dtype))
    return grad


def poisson_loglik(g: SPECTGeometry, f, mu, data, additive=None) -> torch.Tensor:
    """L(f, mu) = sum(g log(Hf + s) - (Hf + s)). The reference has no value
    function (likelihoods/poisson_log_likelihood.py:16-36 only gives the
    gradient H^T(g/(Hf+s)) - H^T 1, which this differentiates to)."""
    hf = forward_autograd(g, f, mu)
    if additive is not None:
        hf = hf + additive
    return (data * torch.log(hf) - hf).sum()


# --------------------------------------------------------------------------
# likelihood and algorithms -- likelihoods/*.py, algorithms/preconditioned_gradient_ascent.py
# --------------------------------------------------------------------------
def sensitivity(g: SPECTGeometry, subset_idx):
    """compute_normalization_factor (spect_system_matrix.py:106-118)."""
    ones = torch.ones((len(subset_idx), g.shape[1], g.shape[2]), dtype=g.dtype)
    return backward(g, ones, subset_idx)


def nearest_neighbour(obj: torch.Tensor, shifts) -> torch.Tensor:
    """utils/misc.py:40-57 (replicate padding, roll, crop)."""
    shift_max = max(np.abs(shifts))
    if shift_max == 0:
        return obj
    shifts = [-s for s in shifts]
    nb = F.pad(obj.unsqueeze(0), 6 * [int(shift_max)], mode="replicate").squeeze(0)
    nb = torch.roll(nb, shifts=shifts, dims=(0, 1, 2))
    m = int(shift_max)
    return nb[m:-m, m:-m, m:-m]


def rdp_gradient(f: torch.Tensor, beta: float, gamma: float, dr, beta_scale: float,
                 delta: float = DELTA) -> torch.Tensor:
    """RelativeDifferencePrior.phi1 (priors/nearest_neighbour.py:176-177) summed by
    NearestNeighbourPrior._pair_contribution (:44-75) with EuclideanNeighbourWeight
    (:207-220); valid points and FOV scale are all ones."""
    out = torch.zeros(f.shape, dtype=f.dtype)
    valid = torch.ones(f.shape, dtype=f.dtype)
    dx, dy, dz = dr
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                w = dx / np.sqrt((dx * i) ** 2 + (dy * j) ** 2 + (dz * k) ** 2)
                v_ijk = nearest_neighbour(valid, (i, j, k))
                nb = nearest_neighbour(f, (i, j, k))
                fr, fs = f, nb
                phi1 = ((fr - fs) * (gamma * torch.abs(fr - fs) + 3 * fs + fr + 2 * delta)) / \
                       ((fr + fs + gamma * torch.abs(fr - fs) + delta) ** 2)
                out = out + phi1 * w * v_ijk * valid
    return beta * beta_scale * out * torch.ones(f.shape, dtype=f.dtype)


def rdp_hessian_jvp(f: torch.Tensor, u: torch.Tensor, beta: float, gamma: float, dr,
                    beta_scale: float, delta: float = DELTA) -> torch.Tensor:
    """NearestNeighbourPrior.__call__(derivative_order=2) (nearest_neighbour.py:105-108)
    applied to u: diag * u + blur(u), diag = sum_s w phi2_1(f, f_s), blur = sum_s
    w phi2_2(f, f_s) u_s, with RelativeDifferencePrior.phi2_1 / phi2_2
    (:179-183); valid points and FOV scale are all ones."""
    diag = torch.zeros(f.shape, dtype=f.dtype)
    blur = torch.zeros(f.shape, dtype=f.dtype)
    dx, dy, dz = dr
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                w = dx / np.sqrt((dx * i) ** 2 + (dy * j) ** 2 + (dz * k) ** 2)
                fs = nearest_neighbour(f, (i, j, k))
                us = nearest_neighbour(u, (i, j, k))
                den = (gamma * torch.abs(f - fs) + f + fs + delta) ** 3
                diag = diag + (2 * (2 * fs + delta) ** 2) / den * w
                blur = blur + (-(2 * (2 * f + delta) * (2 * fs + delta)) / den) * us * w
    c = beta * beta_scale
    return c * diag * u + c * blur


def rdp_value(f: torch.Tensor, beta: float, gamma: float, dr, delta: float = DELTA) -> torch.Tensor:
    """Sum of RelativeDifferencePrior.phi0 (nearest_neighbour.py:173-174) pairs."""
    out = torch.zeros((), dtype=f.dtype)
    dx, dy, dz = dr
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                w = dx / np.sqrt((dx * i) ** 2 + (dy * j) ** 2 + (dz * k) ** 2)
                fs = nearest_neighbour(f, (i, j, k))
                out = out + (w * (f - fs) ** 2 / (f + fs + gamma * torch.abs(f - fs) + delta)).sum()
    return beta * out


def osem(g: SPECTGeometry, data, n_iters: int, n_subsets: int, additive=None,
         f0=None, callback=None):
    """OSEM (preconditioned_gradient_ascent.py:80-129,301-329)."""
    f = torch.ones(g.shape, dtype=g.dtype) if f0 is None else f0.clone().to(g.dtype)
    if additive is None:
        additive = torch.zeros(data.shape, dtype=g.dtype)
    subsets = g.subsets(n_subsets) if n_subsets > 1 else [torch.arange(g.n_angles)]
    norms = [sensitivity(g, s) for s in subsets]
    for j in range(n_iters):
        for k, sidx in enumerate(subsets):
            gk, sk = data[sidx], additive[sidx]
            pp = forward(g, f, sidx) + sk
            grad = backward(g, gk / (pp + DELTA), sidx) - norms[k]
            precond = f * (1 / (norms[k] + DELTA))
            f += precond * (grad - 0)
            f[f <= 0] = 0
            if callback is not None:
                callback(f, j, k, pp)
    return f


def bsrem(g: SPECTGeometry, data, n_iters: int, n_subsets: int, beta: float, gamma: float,
          additive=None, f0=None, relaxation=lambda n: 1, addition_after_iteration=1e-4):
    """BSREM + RDP (preconditioned_gradient_ascent.py:80-129,436-474)."""
    f = torch.ones(g.shape, dtype=g.dtype) if f0 is None else f0.clone().to(g.dtype)
    if additive is None:
        additive = torch.zeros(data.shape, dtype=g.dtype)
    subsets = g.subsets(n_subsets) if n_subsets > 1 else [torch.arange(g.n_angles)]
    norms = [sensitivity(g, s) for s in subsets]
    norm_sum = torch.stack(norms).sum(axis=0)
    for j in range(n_iters):
        for k, sidx in enumerate(subsets):
            wk = len(sidx) / g.n_angles if n_subsets > 1 else 1
            prior_grad = rdp_gradient(f.clone(), beta, gamma, g.dr, wk)
            gk, sk = data[sidx], additive[sidx]
            pp = forward(g, f, sidx) + sk
            grad = backward(g, gk / (pp + DELTA), sidx) - norms[k]
            d = relaxation(j) / (wk * norm_sum + DELTA)
            precond = f * d
            f += precond * (grad - prior_grad)
            f[f <= addition_after_iteration] = addition_after_iteration
    return f


def bsrem_unrolled(g: SPECTGeometry, data, n_iters: int, n_subsets: int, beta, gamma: float,
                   additive=None, f0=None, relaxation=lambda n: 1, addition_after_iteration=1e-4):
    """``bsrem`` written without in-place updates so that torch autograd can
    differentiate x_K with respect to ``data``, ``f0`` and a tensor ``beta``;
    the clamp ``f[f <= floor] = floor`` becomes ``where`` (same values)."""
    f = torch.ones(g.shape, dtype=g.dtype) if f0 is None else f0
    if additive is None:
        additive = torch.zeros(data.shape, dtype=g.dtype)
    subsets = g.subsets(n_subsets) if n_subsets > 1 else [torch.arange(g.n_angles)]
    with torch.no_grad():
        norms = [sensitivity(g, s) for s in subsets]
        norm_sum = torch.stack(norms).sum(axis=0)
    floor = torch.tensor(addition_after_iteration, dtype=g.dtype)
    for j in range(n_iters):
        for k, sidx in enumerate(subsets):
            wk = len(sidx) / g.n_angles if n_subsets > 1 else 1
            prior_grad = rdp_gradient(f, beta, gamma, g.dr, wk)
            pp = forward(g, f, sidx) + additive[sidx]
            grad = backward(g, data[sidx] / (pp + DELTA), sidx) - norms[k]
            d = relaxation(j) / (wk * norm_sum + DELTA)
            f = f + (f * d) * (grad - prior_grad)
            f = torch.where(f <= floor, floor, f)
    return f


def gradient_ascent(g: SPECTGeometry, data, n_iters: int, n_subsets: int, method: str,
                    beta=None, gamma: float = 2.0, additive=None, f0=None,
                    relaxation=lambda n: 1, floor: float = 0.0, callback=None):
    """PreconditionedGradientAscentAlgorithm.__call__ (preconditioned_gradient_ascent.py:80-129)
    with the preconditioner of ``method``: "osem" (:301-329), "osmaposl" (:331-362),
    "rbiem" (:364-398), "rbimap" (:400-434) or "bsrem" (:436-474), and an RDP prior
    when ``beta`` is not None.  Written without in-place updates (the clamp is a
    ``where``) so that autograd can differentiate the result.  ``callback(f, pp, j, k)``
    sees the iterate before sub-iteration (j, k) and its H_m f + s."""
    f = torch.ones(g.shape, dtype=g.dtype) if f0 is None else f0
    if additive is None:
        additive = torch.zeros(data.shape, dtype=g.dtype)
    subsets = g.subsets(n_subsets) if n_subsets > 1 else [torch.arange(g.n_angles)]
    with torch.no_grad():
        norms = [sensitivity(g, s) for s in subsets]
        norm_sum = torch.stack(norms).sum(axis=0) if n_subsets > 1 else norms[0]
    floor_t = torch.tensor(floor, dtype=g.dtype)
    for j in range(n_iters):
        for k, sidx in enumerate(subsets):
            wk = len(sidx) / g.n_angles if n_subsets > 1 else 1
            pg = 0 if beta is None else rdp_gradient(f, beta, gamma, g.dr, wk)
            pp = forward(g, f, sidx) + additive[sidx]
            if callback is not None:
                callback(f.detach().clone(), pp.detach().clone(), j, k)
            grad = backward(g, data[sidx] / (pp + DELTA), sidx) - norms[k]
            nm = norms[k]
            if method == "osem":
                c = f * (1 / (nm + DELTA))
            elif method == "osmaposl":
                c = f / (nm + pg + DELTA)
            elif method == "rbiem":
                rm = torch.max(nm / (norm_sum + DELTA))
                c = f / (norm_sum * rm + DELTA)
            elif method == "rbimap":
                rm = torch.max((nm + pg) / (norm_sum + pg + DELTA))
                c = f / (norm_sum * rm + pg + DELTA)
            elif method == "bsrem":
                c = f * (relaxation(j) / (wk * norm_sum + DELTA))
            else:
                raise ValueError(method)
            f = f + c * (grad - pg)
            f = torch.where(f <= floor_t, floor_t, f)
    return f


def uncertainty(g: SPECTGeometry, data, objects, fps, mask, n_subsets: int, factor,
                beta=None, gamma: float = 2.0, additive_variance=None):
    """LinearPreconditionedGradientAscentAlgorithm.compute_uncertainty
    (preconditioned_gradient_ascent.py:170-214) with _compute_Q (:216-247) and
    _compute_B (:249-299), for the last iterate.  ``objects[n]`` is the iterate
    before sub-iteration n (``objects[-1]`` the result), ``fps[n]`` its H_m f + s,
    ``factor(m)`` the linear preconditioner factor D of subset m.  Returns
    (absolute uncertainty, V) with V the (1 or 2, *data.shape) sensitivity."""
    subsets = g.subsets(n_subsets) if n_subsets > 1 else [torch.arange(g.n_angles)]
    K = len(objects) - 1
    q = mask.clone().to(g.dtype)
    V = torch.zeros((2 if additive_variance is not None else 1, *data.shape), dtype=g.dtype)
    for n in range(K, 0, -1):
        m = (n - 1) % n_subsets
        sidx = subsets[m]
        cur, fp = objects[n - 1], fps[n - 1]
        x = q * cur * factor(m)
        hx = forward(g, x, sidx)
        V[0, sidx] += hx / (fp + DELTA)
        if additive_variance is not None:
            V[1, sidx] += -hx * data[sidx] / (fp + DELTA) ** 2
        if n > 1:
            wk = len(sidx) / g.n_angles if n_subsets > 1 else 1
            out = -backward(g, forward(g, x, sidx) * data[sidx] / (fp ** 2 + DELTA), sidx)
            if beta is not None:
                out = out - rdp_hessian_jvp(cur, x, beta, gamma, g.dr, wk)
            q = out + (objects[n] / (cur + DELTA)) * q
    abs2 = torch.sum(V[0] * data * V[0])
    if additive_variance is not None:
        abs2 = abs2 + torch.sum(additive_variance(V[1]))
    return float(torch.sqrt(abs2)), V


# --------------------------------------------------------------------------
# energy-window scatter estimate -- utils/scatter.py:10-99
# --------------------------------------------------------------------------
def gaussian_kernel_1d(sigma: float, kernel_size: int, dtype=torch.float32) -> torch.Tensor:
    """get_1d_gaussian_kernel's weights (utils/misc.py:78-94)."""
    x = torch.arange(-int(kernel_size // 2), int(kernel_size // 2) + 1)
    k = torch.exp((-x ** 2 / (2 * sigma ** 2)).to(dtype))
    return k / k.sum()


def smoothed_scatter(scatter: torch.Tensor, dtheta: float, dr: float, dz: float,
                     sigma_theta: float = 0, sigma_r: float = 0, sigma_z: float = 0,
                     N_sigmas: int = 3, delta: float = DELTA) -> torch.Tensor:
    """get_smoothed_scatter (utils/scatter.py:10-50): per axis a Conv1d
    (padding 'same', circular in theta, replicate in r and z)."""
    dS = [dtheta, dr, dz]
    sigmas = [sigma_theta, sigma_r, sigma_z]
    ksize = [2 * int(np.ceil(N_sigmas * s / d)) + 1 for s, d in zip(sigmas, dS)]
    modes = ["circular", "replicate", "replicate"]
    for i in range(3):
        if sigmas[i] > delta:
            k = gaussian_kernel_1d(sigmas[i] / dS[i], ksize[i], scatter.dtype)
            layer = torch.nn.Conv1d(1, 1, ksize[i], padding="same", padding_mode=modes[i],
                                    bias=False).to(scatter.dtype)
            layer.weight.data = k.reshape(1, 1, -1)
            with torch.no_grad():
                x = scatter.swapaxes(i, 2)
                x = layer(x.flatten(end_dim=-2).unsqueeze(1)).reshape(x.shape)
                scatter = x.swapaxes(i, 2)
    return scatter


def ew_scatter(lower, upper, wl, wu, wp, kl=0.5, ku=0.5, dtheta=None, dr=None, dz=None,
               sigma_theta=0, sigma_r=0, sigma_z=0, N_sigmas=3):
    """compute_EW_scatter (utils/scatter.py:52-99); returns (estimate,
    variance diagonal, variance operator)."""
    upper = 0 if upper is None else upper
    wu = 1 if wu is None else wu
    est = (lower / wl * kl + upper / wu * ku) * wp
    smooth = (sigma_r > 0) + (sigma_theta > 0) + (sigma_z > 0)
    sm = lambda x: smoothed_scatter(x, dtheta, dr, dz, sigma_theta, sigma_r, sigma_z, N_sigmas)
    if smooth:
        est = sm(est)
    diag = (wp / wl * kl) ** 2 * lower + (wp / wu * ku) ** 2 * upper
    if smooth:
        op = lambda x: sm(x) * diag * sm(x)
    else:
        op = lambda x: x * diag * x
    return est, diag, op


# --------------------------------------------------------------------------
# wrappers of the projector -- transforms/shared/kem.py, motion.py,
# projectors/system_matrix.py:117-221
# --------------------------------------------------------------------------
def kem_kernel(support_objects, dr, size=5, sigmas=None, distance_sigma=1.0, top_N=None,
               dtype=torch.float32):
    """KEMTransform.compute_kernel (kem.py:64-96) with the default Gaussian
    support and distance kernels; returns [size, size, size, *shape]."""
    sigmas = [1.0] * len(support_objects) if sigmas is None else sigmas
    m = (size - 1) // 2
    idxs = torch.arange(-m, m + 1)
    shape = support_objects[0].shape
    K = torch.ones((size, size, size, *shape)).to(dtype)
    for i in idxs:
        for j in idxs:
            for k in idxs:
                comp = 1
                for so, sg in zip(support_objects, sigmas):
                    nb = nearest_neighbour(so, (int(i), int(j), int(k)))
                    comp *= torch.exp(-(so - nb) ** 2 / (2 * sg ** 2))
                K[i + m, j + m, k + m] = comp
    if top_N is not None:
        flat = K.reshape((size ** 3, *shape))
        K *= (torch.argsort(torch.argsort(flat, dim=0), dim=0) >= size ** 3 - top_N).reshape(K.shape)
    xv, yv, zv = torch.meshgrid([dr[0] * idxs, dr[1] * idxs, dr[2] * idxs], indexing="ij")
    d = torch.sqrt(xv ** 2 + yv ** 2 + zv ** 2)
    K *= np.exp(-d ** 2 / (2 * distance_sigma ** 2)).unsqueeze(-1).unsqueeze(-1).unsqueeze(-1)
    K /= K.sum(dim=(0, 1, 2))
    return K


def kem_apply(K: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
    """KEMTransform.forward (kem.py:113-131): sum_s K[s] * f(clamp(v + s))."""
    size = K.shape[0]
    m = (size - 1) // 2
    out = torch.zeros(f.shape, dtype=f.dtype)
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            for k in range(-m, m + 1):
                out += K[i + m, j + m, k + m] * nearest_neighbour(f, (i, j, k))
    return out


def dvf_warp(f: torch.Tensor, dvf: torch.Tensor) -> torch.Tensor:
    """DVFMotionTransform._apply_dvf (motion.py:40-78) with vol_ratio 1."""
    dims = torch.tensor(f.shape)
    old = torch.stack(torch.meshgrid(torch.arange(f.shape[0]), torch.arange(f.shape[1]),
                                     torch.arange(f.shape[2]), indexing="ij")).permute((1, 2, 3, 0)).to(f.dtype)
    new = old + dvf
    new = 2 / (dims - 1) * new - 1
    return F.grid_sample(f.unsqueeze(0).unsqueeze(0), new.to(f.dtype).unsqueeze(0).flip(dims=[-1]),
                         align_corners=True).squeeze() * 1
