"""Shared builders: one SPECT problem as (product system matrix, oracle geometry)."""
from __future__ import annotations

import numpy as np
import torch

from oracle import spect_oracle as so
from pytomography_amd import synthetic


def problem_arrays(n, n_angles, seed=0, nz=None, att=False, psf=False):
    nz = n if nz is None else nz
    f = synthetic.ellipsoid_phantom(n, seed, nz=nz)
    mu = synthetic.attenuation_map(n, seed, nz=nz) if att else None
    return dict(n=n, nz=nz, angles=synthetic.angles(n_angles), f=f, mu=mu, psf=psf,
                radii=np.full(n_angles, synthetic.RADIUS))


def oracle_geometry(pa, dtype=torch.float32, psf_shape="gaussian", psf_dims="2D"):
    psf = (so.PSFModel(synthetic.PSF_PARAMS, kernel_dimensions=psf_dims, shape=psf_shape,
                       min_sigmas=1 if psf_shape == "square" else 3) if pa["psf"] else None)
    mu = None if pa["mu"] is None else torch.from_numpy(pa["mu"])
    return so.make_geometry((pa["n"], pa["n"], pa["nz"]), (synthetic.DX,) * 3, pa["angles"],
                            radii=pa["radii"], mu=mu, psf=psf, dtype=dtype)


def product_system_matrix(pa, device="cuda", chunk_planes=0, attenuation_field=None,
                          obj2obj=None, proj2proj=(), psf_shape="gaussian", psf_dims="2D",
                          sched_rounds=0):
    """The product system matrix of a problem: obj2obj transforms
    [attenuation?, PSF?] by default, else ``obj2obj`` whose items are "att",
    "psf" or transform objects, in that order."""
    import pytomography_amd
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix
    pytomography_amd.set_device(device)
    om = SPECTObjectMeta((synthetic.DX,) * 3, (pa["n"], pa["n"], pa["nz"]))
    pm = SPECTProjMeta((pa["n"], pa["nz"]), (synthetic.DX, synthetic.DX), list(pa["angles"]),
                       radii=list(pa["radii"]))
    if obj2obj is None:
        obj2obj = (["att"] if pa["mu"] is not None else []) + (["psf"] if pa["psf"] else [])
    tr = []
    for t in obj2obj:
        if t == "att":
            tr.append(SPECTAttenuationTransform(torch.from_numpy(pa["mu"]).to(device)))
        elif t == "psf":
            tr.append(SPECTPSFTransform(SPECTPSFMeta(synthetic.PSF_PARAMS, kernel_dimensions=psf_dims,
                                                     shape=psf_shape)))
        else:
            tr.append(t)
    return SPECTSystemMatrix(tr, list(proj2proj), om, pm, chunk_planes=chunk_planes,
                             attenuation_field=attenuation_field, sched_rounds=sched_rounds)


def rel_l2(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))
