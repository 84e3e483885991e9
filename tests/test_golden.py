"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the numpy float64 brute force) against the oracle (CPU) and against the
HIP path through the C ABI (GPU, float64 and float32)."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest
import torch

from oracle import spect_oracle as so

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _load(path):
    d = np.load(path)  # allow_pickle=False: plain arrays only
    return {k: d[k] for k in d.files}


def _oracle_geometry(d, dtype):
    psf = so.PSFModel((0.03, 0.1)) if bool(d["psf"]) else None
    mu = torch.from_numpy(d["mu"]) if "mu" in d else None
    return so.make_geometry(d["f"].shape, (float(d["dx"]),) * 3, d["angles"],
                            radii=np.full(len(d["angles"]), float(d["radius"])), mu=mu, psf=psf,
                            dtype=dtype)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(path):
    d = _load(path)
    g = _oracle_geometry(d, torch.float64)
    if "taps" in d:  # the fixture's closed-form taps are the oracle's taps
        assert np.allclose(g.layers(0)[0].numpy(), d["taps"], rtol=1e-12, atol=0)
    scale = np.abs(d["proj"]).max()
    assert np.abs(so.forward(g, torch.from_numpy(d["f"])).numpy() - d["proj"]).max() < 1e-12 * scale
    bscale = np.abs(d["backward"]).max()
    assert np.abs(so.backward(g, torch.from_numpy(d["y"])).numpy() - d["backward"]).max() < 1e-12 * bscale
    ascale = np.abs(d["adjoint"]).max()
    assert np.abs(so.exact_adjoint(g, torch.from_numpy(d["y"])).numpy() - d["adjoint"]).max() < 1e-12 * ascale


def _product(d, device):
    import pytomography_amd
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix
    pytomography_amd.set_device(device)
    shape = d["f"].shape
    dx = float(d["dx"])
    om = SPECTObjectMeta((dx,) * 3, shape)
    pm = SPECTProjMeta((shape[1], shape[2]), (dx, dx), list(d["angles"]),
                       radii=[float(d["radius"])] * len(d["angles"]))
    tr = []
    if "mu" in d:
        tr.append(SPECTAttenuationTransform(torch.from_numpy(d["mu"]).to(device)))
    if bool(d["psf"]):
        tr.append(SPECTPSFTransform(SPECTPSFMeta((0.03, 0.1))))
    return SPECTSystemMatrix(tr, [], om, pm)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_hip_reproduces_golden(path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = _load(path)
    dev = torch.device("cuda")
    sm = _product(d, dev)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-6)):
        f = torch.from_numpy(d["f"]).to(dev, dt)
        y = torch.from_numpy(d["y"]).to(dev, dt)
        if "mu" in d:
            sm._att.attenuation_map = torch.from_numpy(d["mu"]).to(dev, dt)
        proj = sm.forward(f).double().cpu().numpy()
        bp = sm.backward(y).double().cpu().numpy()
        adj = sm.backward_adjoint(y).double().cpu().numpy()
        e = [np.abs(a - d[k]).max() / np.abs(d[k]).max()
             for a, k in ((proj, "proj"), (bp, "backward"), (adj, "adjoint"))]
        print(f"{os.path.basename(path)} {dt}: forward {e[0]:.1e} backward {e[1]:.1e} adjoint {e[2]:.1e}")
        assert max(e) < tol
