"""Numpy float64 brute force of the SPECT path for tiny geometry (test infrastructure only).

Written from the mathematics, not from the torch op sequence, so that it pins
``oracle/spect_oracle.py`` independently:

* rotation: output pixel (i, j) (row = padded x index, column = padded y
  index) samples the input at column ``c(j-cc) - s(i-cc) + cc`` and row
  ``s(j-cc) + c(i-cc) + cc`` with bilinear weights and zero outside the grid
  (kornia ``rotate``, align_corners=True, as called at utils/spatial.py:31 and
  transforms/shared/spatial.py:36,53);
* attenuation: a(i) = exp(-(sum_{i' >= i} mu_i' dx - mu_i dx / 2))
  (utils/misc.py:8-18, transforms/SPECT/attenuation.py:11-21);
* PSF: zero-padded correlation along y then z with per-plane taps
  (transforms/SPECT/psf.py:27-41);
* forward = sum over x (spect_system_matrix.py:145-157).

Only loops over voxels -- intended for volumes of a few hundred voxels.
"""
from __future__ import annotations

import math

import numpy as np


def pad_size(n: int) -> int:
    return int(np.ceil((np.sqrt(2) * n - n) / 2))


def rotate_plane(img: np.ndarray, theta_deg: float) -> np.ndarray:
    """Bilinear rotation of a 2D [H, W] image about its centre, zeros outside."""
    h, w = img.shape
    a = math.radians(theta_deg)
    c, s = math.cos(a), math.sin(a)
    cx, cy = (w - 1) / 2, (h - 1) / 2
    out = np.zeros_like(img, dtype=np.float64)
    for i in range(h):
        for j in range(w):
            xs = c * (j - cx) - s * (i - cy) + cx
            ys = s * (j - cx) + c * (i - cy) + cy
            x0, y0 = math.floor(xs), math.floor(ys)
            tx, ty = xs - x0, ys - y0
            acc = 0.0
            for dy, wy in ((0, 1 - ty), (1, ty)):
                for dx, wx in ((0, 1 - tx), (1, tx)):
                    yy, xx = y0 + dy, x0 + dx
                    if 0 <= yy < h and 0 <= xx < w:
                        acc += wy * wx * img[yy, xx]
            out[i, j] = acc
    return out


def rotate_volume(vol: np.ndarray, theta_deg: float) -> np.ndarray:
    """Rotate each z-slice of [X, Y, Z] (rows = X, columns = Y)."""
    return np.stack([rotate_plane(vol[:, :, z], theta_deg) for z in range(vol.shape[2])], axis=2)


def pad_xy(vol: np.ndarray) -> np.ndarray:
    p = pad_size(vol.shape[1])
    return np.pad(vol, ((p, p), (p, p), (0, 0)))


def attenuation(mu_rot: np.ndarray, dx: float) -> np.ndarray:
    out = np.zeros_like(mu_rot, dtype=np.float64)
    px = mu_rot.shape[0]
    for i in range(px):
        s = mu_rot[i:].sum(axis=0) * dx - mu_rot[i] * dx / 2
        out[i] = np.exp(-s)
    return out


def correlate_same(v: np.ndarray, taps: np.ndarray, axis: int) -> np.ndarray:
    k = len(taps)
    h = k // 2
    out = np.zeros_like(v, dtype=np.float64)
    n = v.shape[axis]
    vv = np.moveaxis(v, axis, 0)
    oo = np.moveaxis(out, axis, 0)
    for o in range(n):
        for t in range(k):
            src = o + t - h
            if 0 <= src < n:
                oo[o] += taps[t] * vv[src]
    return out


def psf(vol: np.ndarray, taps_r: np.ndarray, taps_z: np.ndarray | None) -> np.ndarray:
    """Per x-plane i: correlate along y with taps_r[i], then along z with taps_z[i]."""
    out = np.zeros_like(vol, dtype=np.float64)
    for i in range(vol.shape[0]):
        plane = correlate_same(vol[i], taps_r[i], axis=0)
        if taps_z is not None:
            plane = correlate_same(plane, taps_z[i], axis=1)
        out[i] = plane
    return out


def forward(f, angles, dx, mu=None, taps=None):
    """H f for every beta in ``angles``; taps = per-angle (taps_r, taps_z)."""
    p = pad_size(f.shape[1])
    n = f.shape[1]
    fp = pad_xy(f)
    mp = None if mu is None else pad_xy(mu)
    out = []
    for a, beta in enumerate(angles):
        v = rotate_volume(fp, beta - 270.0)
        if mp is not None:
            v = v * attenuation(rotate_volume(mp, beta - 270.0), dx)
        if taps is not None:
            v = psf(v, taps[a][0], taps[a][1])
        out.append(v.sum(axis=0)[p:p + n])
    return np.stack(out)


def backward_ref(proj, angles, dx, shape, mu=None, taps=None):
    """Reference back projection: smear along x, PSF, attenuation, rotate by
    270 - beta (a gather), sum, crop."""
    p = pad_size(shape[1])
    n = shape[1]
    px = n + 2 * p
    mp = None if mu is None else pad_xy(mu)
    acc = np.zeros((px, px, shape[2]))
    for a, beta in enumerate(angles):
        row = np.zeros((px, shape[2]))
        row[p:p + n] = proj[a]
        v = np.broadcast_to(row, (px, px, shape[2])).astype(np.float64).copy()
        if taps is not None:
            v = psf(v, taps[a][0], taps[a][1])
        if mp is not None:
            v = v * attenuation(rotate_volume(mp, beta - 270.0), dx)
        acc += rotate_volume(v, 270.0 - beta)
    return acc[p:p + n, p:p + n]


def dense_matrix(op, in_shape):
    """Columns op(e_k) of a linear operator on arrays of ``in_shape``."""
    size = int(np.prod(in_shape))
    cols = []
    for k in range(size):
        e = np.zeros(size)
        e[k] = 1.0
        cols.append(np.asarray(op(e.reshape(in_shape))).ravel())
    return np.stack(cols, axis=1)
