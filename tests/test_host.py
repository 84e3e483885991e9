"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, argument validation fails loudly, and the host mirror of the
reference interface (metadata, PSF taps, padding, subsets) matches the oracle."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest
import torch

from oracle import spect_oracle as so

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "spect_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char \*)\s*(spect_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from pytomography_amd import _lib
    lib = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTED)
    assert lib.spect_version() >= 2


def test_c_abi_rejects_bad_arguments_without_touching_the_gpu():
    from pytomography_amd import _lib
    lib = _lib.lib()
    angles = (ctypes.c_float * 4)(0, 90, 180, 270)
    g = _lib.SpectGeom(nx=8, ny=9, nz=4, n_angles=4, dx=0.48, dy=0.48, dz=0.48,
                       angles_deg=ctypes.cast(angles, ctypes.POINTER(ctypes.c_float)))
    plan = ctypes.c_void_p()
    assert lib.spect_plan_create(ctypes.byref(g), ctypes.byref(plan)) == -1
    assert b"nx must equal ny" in lib.spect_last_error()
    g.ny = 8
    g.psf_kr = 4
    assert lib.spect_plan_create(ctypes.byref(g), ctypes.byref(plan)) == -1
    assert b"odd" in lib.spect_last_error()
    assert lib.spect_forward(None, 0, None, None, None, None, 1, None, None, 0, None) == -1
    assert lib.spect_workspace_bytes(None, 0, 4) == 0
    fwd, fld = ctypes.c_int32(), ctypes.c_int32()
    assert lib.spect_schedule_info(None, 0, 4, ctypes.byref(fwd), ctypes.byref(fld)) == -1
    assert lib.spect_poisson_ratio(0, None, None, None, None, None, 10, 1e-11, None) == -1
    # argument checks of the auxiliary entry points run before any launch
    fake = ctypes.c_void_p(0x1000)
    fake2 = ctypes.c_void_p(0x2000)
    assert lib.spect_smooth_axis(0, fake, fake2, 4, 4, 4, 1, fake, 4, 1, None) == -1
    assert b"odd" in lib.spect_last_error()
    assert lib.spect_smooth_axis(0, fake, fake, 4, 4, 4, 1, fake2, 5, 1, None) == -1
    assert lib.spect_smooth_axis(0, fake, fake2, 4, 4, 4, 3, fake, 5, 1, None) == -1
    assert lib.spect_smooth_axis(0, fake, fake2, 4, 4, 4, 1, fake, 5, 7, None) == -1
    assert lib.spect_ew_scatter(0, fake, None, fake2, None, 10, 0.0, 1.0, 1.0, 0.5, 0.5, None) == -1
    assert b"non-zero" in lib.spect_last_error()
    assert lib.spect_kem_apply(0, fake, fake2, fake, 4, 4, 4, 4, None) == -1
    assert lib.spect_kem_apply(0, fake, fake2, fake2, 4, 4, 4, 3, None) == -1
    assert lib.spect_dvf_warp(0, fake, fake2, fake, 1, 4, 4, None) == -1
    assert lib.spect_rdp_hessian_jvp(0, fake, fake2, fake, 0, 4, 4, 1, 1, 1, 0.1, 2, 1e-11, 1, None) == -1
    assert lib.spect_ew_scatter(7, fake, None, fake2, None, 10, 1.0, 1.0, 1.0, 0.5, 0.5, None) == -1
    # the rotated-frame entry points check the plan before anything else
    for name in ("spect_rotate_object", "spect_att_volume", "spect_psf_volume"):
        assert getattr(lib, name)(None, 0, fake, fake2, 1, fake, None) == -1
        assert b"null plan" in lib.spect_last_error()
    assert lib.spect_rotate_back(None, 0, fake, fake2, 1, 0, fake, None) == -1


def test_no_cpu_fallback():
    import pytomography_amd
    from pytomography_amd import _lib
    from tests.helpers import problem_arrays, product_system_matrix
    pa = problem_arrays(16, 4, seed=0)
    sm = product_system_matrix(pa, device="cpu")
    with pytest.raises(_lib.SpectLibraryError):
        sm.forward(torch.from_numpy(pa["f"]))
    pytomography_amd.set_device("cuda" if torch.cuda.is_available() else "cpu")


def test_transform_orders_and_shapes():
    """[PSF, attenuation] runs the per-angle path (fused is False); a map of
    the wrong shape is refused."""
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix
    om = SPECTObjectMeta((0.48,) * 3, (16, 16, 8))
    pm = SPECTProjMeta((16, 8), (0.48, 0.48), [0.0, 90.0], radii=[25.0, 25.0])
    att = SPECTAttenuationTransform(torch.zeros((16, 16, 8)))
    psf = SPECTPSFTransform(SPECTPSFMeta((0.03, 0.1)))
    assert not SPECTSystemMatrix([psf, att], [], om, pm).fused
    assert SPECTSystemMatrix([att, psf], [], om, pm).fused
    with pytest.raises(ValueError):
        SPECTSystemMatrix([SPECTAttenuationTransform(torch.zeros((8, 8, 8)))], [], om, pm)


def test_metadata_matches_reference_fields():
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta
    om = SPECTObjectMeta((0.48, 0.48, 0.48), (128, 128, 128))
    assert om.pad_size == 27 and om.padded_shape == (182, 182, 128)
    pm = SPECTProjMeta((128, 128), (0.48, 0.48), np.arange(0, 360, 3.0), radii=[25.0] * 120)
    assert pm.shape == (120, 128, 128) and pm.padded_shape == (120, 182, 128)
    assert pm.angles.dtype == torch.float32 and pm.num_projections == 120


def test_psf_taps_match_oracle_bitwise():
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTPSFTransform
    om = SPECTObjectMeta((0.48, 0.48, 0.48), (64, 64, 64))
    pm = SPECTProjMeta((64, 64), (0.48, 0.48), [0.0], radii=[25.0])
    t = SPECTPSFTransform(SPECTPSFMeta((0.03, 0.1)))
    t.configure(om, pm)
    tr, tz = t.radius_taps(25.0, torch.float32)
    ref_r, ref_z = so.psf_layers(so.PSFModel((0.03, 0.1)), (64, 64, 64), (0.48,) * 3, 25.0)
    assert tr.shape == (92, 21) and torch.equal(tr, ref_r) and torch.equal(tz, ref_z)


def test_padding_helpers_match_oracle():
    from pytomography_amd.utils import (compute_pad_size, compute_pad_size_padded, pad_object,
                                        unpad_object, pad_proj, unpad_proj)
    for n in (5, 64, 128, 191):
        assert compute_pad_size(n) == so.compute_pad_size(n)
        assert compute_pad_size_padded(n + 2 * compute_pad_size(n)) == compute_pad_size(n)
    x = torch.rand((10, 10, 3))
    assert torch.equal(pad_object(x), so.pad_object(x))
    assert torch.equal(pad_object(x, "back_project"), so.pad_object(x, "back_project"))
    assert torch.equal(unpad_object(pad_object(x)), x)
    p = torch.rand((4, 10, 3))
    assert torch.equal(unpad_proj(pad_proj(p)), p)


def test_subsets_and_weights():
    from tests.helpers import problem_arrays, product_system_matrix
    import pytomography_amd
    pa = problem_arrays(16, 12, seed=0)
    sm = product_system_matrix(pa, device="cpu")
    sm.set_n_subsets(4)
    ref = so.make_geometry((16, 16, 16), (0.48,) * 3, pa["angles"]).subsets(4)
    for k in range(4):
        assert torch.equal(sm.subset_indices_array[k].cpu(), ref[k])
        assert sm.get_weighting_subset(k) == 3 / 12
    assert sm.get_weighting_subset(None) == 1
    proj = torch.arange(12 * 2 * 2).reshape(12, 2, 2)
    assert torch.equal(sm.get_projection_subset(proj, 1), proj[1::4])
    pytomography_amd.set_device("cuda" if torch.cuda.is_available() else "cpu")


def test_synthetic_inputs_are_seeded():
    from pytomography_amd import synthetic
    a = synthetic.ellipsoid_phantom(32, 3)
    b = synthetic.ellipsoid_phantom(32, 3)
    assert np.array_equal(a, b) and a.max() > 0.1 and a.min() == 0
    mu = synthetic.attenuation_map(32, 3)
    assert set(np.unique(mu)) <= {0.0, np.float32(0.15), np.float32(0.25)}


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says so."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_transform_classification_and_proj2proj_signatures():
    """Fused path only for [attenuation?, PSF?] in that order without
    proj2proj; a CutOffTransform's forward takes the call's angle indices."""
    import pytomography_amd
    from pytomography_amd.projectors.SPECT.spect_system_matrix import _accepts_angles
    from pytomography_amd.transforms import Transform
    from pytomography_amd.transforms.SPECT import CutOffTransform, SPECTPSFTransform
    from tests.helpers import problem_arrays, product_system_matrix

    class Ident(Transform):
        def forward(self, x, a):
            return x

        def backward(self, x, a):
            return x

    class ProjOnly:
        def forward(self, proj):
            return proj

    pa = problem_arrays(12, 4, seed=1, att=True, psf=True)
    try:
        assert product_system_matrix(pa, device="cpu").fused
        assert product_system_matrix(pa, device="cpu", obj2obj=["att"]).fused
        assert not product_system_matrix(pa, device="cpu", obj2obj=["psf", "att"]).fused
        assert not product_system_matrix(pa, device="cpu", obj2obj=["att", Ident(), "psf"]).fused
        mask = torch.ones((4, 12, 12))
        assert not product_system_matrix(pa, device="cpu", proj2proj=[CutOffTransform(mask)]).fused
    finally:
        pytomography_amd.set_device("cuda" if torch.cuda.is_available() else "cpu")
    assert _accepts_angles(CutOffTransform(torch.ones((4, 12, 12))))
    assert not _accepts_angles(ProjOnly())
    with pytest.raises(ValueError):
        SPECTPSFTransform()


def test_cutoff_transform_pads_and_follows_subset_angles():
    """cutoff.py:5-20: the mask is zero-padded along Lr like the projections;
    for a subset's projections the rows of its angles are used."""
    from pytomography_amd.transforms.SPECT import CutOffTransform
    from pytomography_amd.utils import compute_pad_size
    mask = torch.rand((6, 10, 5), generator=torch.Generator().manual_seed(2))
    t = CutOffTransform(mask)
    p = compute_pad_size(10)
    assert t.mask.shape == (6, 10 + 2 * p, 5)
    assert float(t.mask[:, :p].abs().max()) == 0.0
    proj = torch.ones((6, 10 + 2 * p, 5))
    assert torch.equal(t.forward(proj), t.mask)
    sub = torch.tensor([1, 4])
    assert torch.equal(t.backward(proj[:2], sub), t.mask[sub])


def test_tuner_picks_the_fastest_candidate_and_caches_it(tmp_path, monkeypatch):
    """batch._tune (behind autotune_chunk_planes and autotune_schedule): the
    candidate with the smallest measured time wins, tuple candidates survive
    the JSON cache, and a cached key is answered without timing again."""
    from pytomography_amd import batch
    monkeypatch.setattr(batch, "_CACHE_PATH", str(tmp_path / "autotune.json"))
    monkeypatch.setattr(batch, "_cache", None)
    monkeypatch.setattr(batch.torch.cuda, "get_device_name", lambda device=None: "test device")
    times = {(0, 0): 3.0, (0, 1): 2.0, (16, -1): 1.0, (8, -1): 1.5}
    timed = []

    def seconds(cand, device):
        timed.append(cand)
        return times[cand]

    assert batch._tune("sched|k", "cpu", tuple(times), seconds) == (16, -1)
    assert timed == list(times)

    def no_timing(cand, device):
        raise AssertionError("a cached key must not be timed again")

    assert batch._tune("sched|k", "cpu", tuple(times), no_timing) == (16, -1)
    monkeypatch.setattr(batch, "_cache", None)  # read back from the file
    assert batch._tune("sched|k", "cpu", tuple(times), no_timing) == (16, -1)
    assert batch._tune("batch|k", "cpu", (8, 16, 32), lambda c, d: abs(c - 16)) == 16
