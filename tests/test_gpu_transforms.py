"""GPU: the per-angle projector path and the stand-alone SPECT transforms.

A system matrix whose transforms are not [attenuation, PSF] (a caller's
obj2obj transform, a proj2proj transform, another order, a PSF network)
runs the reference's per-angle loop over whole rotated frames
(spect_system_matrix.py:145-156, 187-198) on the HIP rotation, attenuation
and PSF kernels.  Checked against the oracle run with the same transform
objects, and against the fused kernels where both apply.  Bar (FP32):
relative L2 <= 1e-5 per projection and per back projection.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import spect_oracle as so
from pytomography_amd.transforms import Transform
from tests.helpers import oracle_geometry, problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


class DepthWeight(Transform):
    """A caller's diagonal obj2obj transform: plane x of the rotated frame
    times 1 + 0.2 cos(0.15 x + angle index)."""

    def _w(self, obj_i, ang_idx):
        x = torch.arange(obj_i.shape[0], dtype=torch.float64)
        w = 1 + 0.2 * torch.cos(0.15 * x + float(ang_idx))
        return w.to(device=obj_i.device, dtype=obj_i.dtype).reshape(-1, 1, 1)

    def forward(self, obj_i, ang_idx):
        return obj_i * self._w(obj_i, ang_idx)

    def backward(self, obj_i, ang_idx):
        return obj_i * self._w(obj_i, ang_idx)


class ShiftRays(Transform):
    """A caller's non-diagonal obj2obj transform: every ray's samples moved
    one plane towards the detector; backward is its transpose."""

    def forward(self, obj_i, ang_idx):
        out = torch.zeros_like(obj_i)
        out[1:] = obj_i[:-1]
        return out

    def backward(self, obj_i, ang_idx):
        out = torch.zeros_like(obj_i)
        out[:-1] = obj_i[1:]
        return out


class Identity(Transform):
    def forward(self, obj_i, ang_idx):
        return obj_i

    def backward(self, obj_i, ang_idx):
        return obj_i


class OracleCutOff:
    """The oracle's restatement of CutOffTransform (cutoff.py:5-20) on the
    padded projections of the call's angles."""

    def __init__(self, mask):
        self.mask = so.pad_proj(mask)

    def forward(self, proj, angle_indices):
        return proj * self.mask[torch.as_tensor(angle_indices).long()]

    backward = forward


def _mask(pa, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand((len(pa["angles"]), pa["n"], pa["nz"]), generator=g) > 0.2).float()


def test_caller_transforms_and_cutoff_match_oracle(cuda):
    """[attenuation, DepthWeight, PSF, ShiftRays] with a CutOffTransform:
    forward (all angles and a subset) and back projection against the oracle
    running the same transforms."""
    from pytomography_amd.transforms.SPECT import CutOffTransform
    pa = problem_arrays(32, 12, seed=41, att=True, psf=True)
    mask = _mask(pa, 42)
    dw, sh = DepthWeight(), ShiftRays()
    sm = product_system_matrix(pa, obj2obj=["att", dw, "psf", sh], proj2proj=[CutOffTransform(mask.to(cuda))])
    assert not sm.fused
    g = oracle_geometry(pa)
    ts, p2p = ["att", dw, "psf", sh], [OracleCutOff(mask)]
    f = torch.from_numpy(pa["f"])
    out = sm.forward(f.to(cuda)).cpu()
    ref = so.forward(g, f, obj2obj=ts, proj2proj=p2p)
    errs = [rel_l2(out[k], ref[k]) for k in range(len(pa["angles"])) if float(ref[k].norm()) > 0]
    sm.set_n_subsets(3)
    sub = sm.subset_indices_array[1].cpu()
    out_s = sm.forward(f.to(cuda), 1).cpu()
    ref_s = so.forward(g, f, sub, obj2obj=ts, proj2proj=p2p)
    y = torch.rand(ref.shape, generator=torch.Generator().manual_seed(43))
    bp = sm.backward(y.to(cuda)).cpu()
    bp_ref = so.backward(g, y, obj2obj=ts, proj2proj=p2p)
    e_sub, e_bp = rel_l2(out_s, ref_s), rel_l2(bp, bp_ref)
    print(f"caller transforms: forward max rel-L2 {max(errs):.1e}, subset {e_sub:.1e}, backward {e_bp:.1e}")
    assert max(errs) <= TOL and e_sub <= TOL and e_bp <= TOL


@pytest.mark.parametrize("att,psf", [(True, True), (True, False), (False, True)])
def test_per_angle_path_matches_fused_path(cuda, att, psf):
    """The same attenuation and PSF with an identity transform in between
    (which forces the per-angle path) against the fused kernels: forward,
    back projection and sensitivity image."""
    pa = problem_arrays(40, 16, seed=44, att=att, psf=psf)
    fused = product_system_matrix(pa)
    names = (["att"] if att else []) + (["psf"] if psf else [])
    loop = product_system_matrix(pa, obj2obj=names[:1] + [Identity()] + names[1:])
    assert fused.fused and not loop.fused
    f = torch.from_numpy(pa["f"]).to(cuda)
    e_f = rel_l2(loop.forward(f), fused.forward(f))
    y = torch.rand((16, 40, 40), generator=torch.Generator().manual_seed(45)).to(cuda)
    e_b = rel_l2(loop.backward(y), fused.backward(y))
    e_n = rel_l2(loop.compute_normalization_factor(), fused.compute_normalization_factor())
    print(f"att {att} psf {psf}: per-angle vs fused forward {e_f:.1e}, backward {e_b:.1e}, norm {e_n:.1e}")
    assert e_f <= TOL and e_b <= TOL and e_n <= TOL


def test_psf_before_attenuation_runs_per_angle(cuda):
    """The reverse order [PSF, attenuation] (not fusable) against the oracle."""
    pa = problem_arrays(28, 10, seed=46, att=True, psf=True)
    sm = product_system_matrix(pa, obj2obj=["psf", "att"])
    assert not sm.fused
    g = oracle_geometry(pa)
    f = torch.from_numpy(pa["f"])
    e_f = rel_l2(sm.forward(f.to(cuda)).cpu(), so.forward(g, f, obj2obj=["psf", "att"]))
    y = torch.rand((10, 28, 28), generator=torch.Generator().manual_seed(47))
    e_b = rel_l2(sm.backward(y.to(cuda)).cpu(), so.backward(g, y, obj2obj=["psf", "att"]))
    print(f"PSF then attenuation: forward {e_f:.1e}, backward {e_b:.1e}")
    assert e_f <= TOL and e_b <= TOL


@pytest.mark.parametrize("shape,dims", [("gaussian", "2D"), ("square", "2D"), ("gaussian", "1D")])
def test_psf_transform_forward_matches_oracle(cuda, shape, dims):
    """SPECTPSFTransform.forward/backward on one angle's rotated frame (and
    the fused projector with that PSF) against the oracle's layers."""
    pa = problem_arrays(30, 8, seed=48, psf=True)
    pa["radii"] = np.linspace(22.0, 30.0, 8)
    sm = product_system_matrix(pa, psf_shape=shape, psf_dims=dims)
    psf = sm.obj2obj_transforms[0]
    g = oracle_geometry(pa, psf_shape=shape, psf_dims=dims)
    P = 30 + 2 * so.compute_pad_size(30)
    vol = torch.rand((P, P, 30), generator=torch.Generator().manual_seed(49))
    errs = []
    for a in (0, 5):
        tr, tz = g.layers(a)
        ref = so.psf_apply(vol, tr, tz)
        errs.append(rel_l2(psf.forward(vol.to(cuda), torch.tensor(a, device=cuda)).cpu(), ref))
        errs.append(rel_l2(psf.backward(vol.to(cuda), a).cpu(), ref))
    f = torch.from_numpy(pa["f"])
    e_proj = rel_l2(sm.forward(f.to(cuda)).cpu(), so.forward(g, f))
    print(f"PSF {shape} {dims}: transform max rel-L2 {max(errs):.1e}, fused projector {e_proj:.1e}")
    assert max(errs) <= TOL and e_proj <= TOL


def test_attenuation_transform_matches_oracle(cuda):
    """SPECTAttenuationTransform.forward/backward on rotated frames and
    compute_average_prob_matrix against the oracle."""
    pa = problem_arrays(30, 10, seed=50, att=True)
    sm = product_system_matrix(pa)
    att = sm.obj2obj_transforms[0]
    g = oracle_geometry(pa)
    P = 30 + 2 * so.compute_pad_size(30)
    vol = torch.rand((P, P, 30), generator=torch.Generator().manual_seed(51))
    errs = []
    for a in (0, 3, 7):
        ref = vol * so.attenuation_factor(g.mu, g.angles[a], g.dr[0])
        errs.append(rel_l2(att.forward(vol.to(cuda), torch.tensor(a, device=cuda)).cpu(), ref))
        errs.append(rel_l2(att.backward(vol.to(cuda), a).cpu(), ref))
    e_avg = rel_l2(att.compute_average_prob_matrix().cpu(), so.average_prob_matrix(g))
    print(f"attenuation transform: max rel-L2 {max(errs):.1e}, average prob matrix {e_avg:.1e}")
    assert max(errs) <= TOL and e_avg <= TOL


def test_rotated_frames_match_oracle(cuda):
    """RotatedFrames.rotate (the padded object rotated to each angle) and
    rotate_back (rotation by 270 - beta and sum, unpadded)."""
    from pytomography_amd.projectors.SPECT.plan import RotatedFrames
    pa = problem_arrays(26, 9, seed=52)
    sm = product_system_matrix(pa)
    frames = RotatedFrames(sm.object_meta, sm.proj_meta)
    g = oracle_geometry(pa)
    f = torch.from_numpy(pa["f"])
    idx = [0, 4, 8]
    vols = frames.rotate(f.to(cuda), idx).cpu()
    e_rot = max(rel_l2(vols[k], so.rotate_xyz(so.pad_object(f), -(270 - g.angles[a])))
                for k, a in enumerate(idx))
    v = torch.rand(vols.shape, generator=torch.Generator().manual_seed(53))
    back = frames.rotate_back(v.to(cuda), idx).cpu()
    ref = sum(so.rotate_xyz(v[k], 270 - g.angles[a]) for k, a in enumerate(idx))
    e_back = rel_l2(back, so.unpad_object(ref))
    acc = frames.rotate_back(v.to(cuda), idx, out=back.to(cuda)).cpu()
    e_acc = rel_l2(acc, 2 * so.unpad_object(ref))
    print(f"rotated frames: rotate max rel-L2 {e_rot:.1e}, rotate back {e_back:.1e}, accumulate {e_acc:.1e}")
    assert e_rot <= TOL and e_back <= TOL and e_acc <= TOL


def test_osem_on_per_angle_path_matches_oracle(cuda):
    """OSEM (2 iterations x 3 subsets) with an identity transform among the
    obj2obj transforms runs the algorithm's unfused branch on the per-angle
    projector; the iterate matches the oracle's OSEM."""
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    pa = problem_arrays(24, 12, seed=54, att=True, psf=True)
    sm = product_system_matrix(pa, obj2obj=["att", Identity(), "psf"])
    g = oracle_geometry(pa)
    f = torch.from_numpy(pa["f"])
    data = torch.poisson(so.forward(g, f) * 20, generator=torch.Generator().manual_seed(55))
    lk = PoissonLogLikelihood(sm, data.to(cuda))
    x = OSEM(lk)(n_iters=2, n_subsets=3).cpu()
    ref = so.osem(g, data, 2, 3)
    e = rel_l2(x, ref)
    print(f"OSEM on the per-angle path: rel-L2 {e:.1e}")
    assert e <= TOL


def test_fused_only_entry_points_refuse_per_angle_matrices(cuda):
    pa = problem_arrays(16, 4, seed=56, att=True)
    sm = product_system_matrix(pa, obj2obj=["att", Identity()])
    f = torch.from_numpy(pa["f"]).to(cuda)
    with pytest.raises(NotImplementedError, match="per-angle"):
        sm.backward_adjoint(sm.forward(f))
    with pytest.raises(NotImplementedError, match="per-angle"):
        sm.backward_transpose(f)
