"""Writes the golden fixtures of the SPECT path from oracle/brute.py (numpy
float64 brute force written from the mathematics).

    python tests/golden/make_golden.py

Each .npz holds the inputs (object, attenuation map, angles, PSF taps,
projection data) and the outputs (forward projection, reference back
projection, exact adjoint H^T y from the dense matrix).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import brute  # noqa: E402


def gaussian_taps(n_planes, radius, dx, a=0.03, b=0.1, min_sigmas=3):
    """sigma(d) = a d + b per padded plane, truncated at min_sigmas sigma,
    normalised per plane (transforms/SPECT/psf.py:62-83,170-197)."""
    i = np.arange(n_planes)
    d = radius + (n_planes / 2 - i - 0.5) * dx if n_planes % 2 == 0 else radius + ((n_planes - 1) / 2 - i) * dx
    d = np.maximum(d, 0)
    s = (a * d + b) / dx
    k = int(np.ceil(s.max() * min_sigmas) * 2 + 1)
    t = np.arange(-(k // 2), k // 2 + 1)
    w = np.exp(-t[None, :] ** 2 / (2 * s[:, None] ** 2 + 1e-11))
    return w / w.sum(1, keepdims=True)


def make(name, shape, angles, att, psf, seed, radius=6.0, dx=0.48):
    rng = np.random.default_rng(seed)
    f = rng.random(shape)
    mu = rng.random(shape) * 0.2 if att else None
    p = brute.pad_size(shape[1])
    taps = None
    if psf:
        w = gaussian_taps(shape[0] + 2 * p, radius, dx)
        taps = [(w, w)] * len(angles)
    proj = brute.forward(f, angles, dx, mu, taps)
    y = rng.random(proj.shape)
    bp = brute.backward_ref(y, angles, dx, shape, mu, taps)
    H = brute.dense_matrix(lambda e: brute.forward(e, angles, dx, mu, taps), shape)
    adj = (H.T @ y.ravel()).reshape(shape)
    out = dict(f=f, angles=np.asarray(angles, dtype=np.float64), proj=proj, y=y, backward=bp,
               adjoint=adj, dx=dx, radius=radius, psf=psf)
    if mu is not None:
        out["mu"] = mu
    if taps is not None:
        out["taps"] = taps[0][0]
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print("wrote", name, {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    make("plain_8x8x4_a6", (8, 8, 4), [0.0, 33.0, 90.0, 141.0, 222.0, 300.0], False, False, 11)
    make("att_psf_8x8x4_a6", (8, 8, 4), [0.0, 33.0, 90.0, 141.0, 222.0, 300.0], True, True, 12)
    make("att_psf_10x10x6_a5", (10, 10, 6), [7.5, 95.0, 180.0, 257.0, 333.0], True, True, 13)
