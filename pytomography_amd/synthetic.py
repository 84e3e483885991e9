"""Seeded synthetic SPECT problems (phantom, attenuation map, geometry).

There is no dataset on this path: benchmarks and tests build their inputs
from these generators.

* Activity: a sum of ``n`` axis-aligned ellipsoids with semi-axes uniform in
  [5, 40] voxels (capped at 0.45 N), activity uniform in [0.1, 1.0], centres
  drawn so that each ellipsoid lies inside the cylinder inscribed in the
  transaxial square; zero background.
* Attenuation: a body ellipsoid of water (0.15/cm, semi-axes 0.42/0.36/0.45 N)
  and two bone ellipsoids (0.25/cm) inside it that override the body value.
* Geometry: 4.8 mm voxels, angles beta_k = k * 360 / A degrees, radius 25 cm,
  PSF sigma(d) = 0.03 d + 0.1 cm truncated at 3 sigma.
"""
from __future__ import annotations

import numpy as np
import torch


def _ellipsoid_mask(n: int, nz: int, centre, semi) -> np.ndarray:
    x = np.arange(n)[:, None, None]
    y = np.arange(n)[None, :, None]
    z = np.arange(nz)[None, None, :]
    return (((x - centre[0]) / semi[0]) ** 2 + ((y - centre[1]) / semi[1]) ** 2
            + ((z - centre[2]) / semi[2]) ** 2) <= 1.0


def ellipsoid_phantom(n: int, seed: int, n_ellipsoids: int = 6, nz: int | None = None) -> np.ndarray:
    nz = n if nz is None else nz
    rng = np.random.default_rng(seed)
    out = np.zeros((n, n, nz), dtype=np.float32)
    c0 = (n - 1) / 2
    rmax = n / 2 - 1
    cap = 0.45 * n
    for _ in range(n_ellipsoids):
        semi = np.minimum(rng.uniform(5, 40, size=3), cap)
        act = rng.uniform(0.1, 1.0)
        reach = max(semi[0], semi[1])
        room = max(rmax - reach, 0.0)
        rad = room * np.sqrt(rng.uniform(0, 1))
        phi = rng.uniform(0, 2 * np.pi)
        cz_lo, cz_hi = semi[2], nz - 1 - semi[2]
        cz = rng.uniform(cz_lo, cz_hi) if cz_hi > cz_lo else (nz - 1) / 2
        centre = (c0 + rad * np.cos(phi), c0 + rad * np.sin(phi), cz)
        out[_ellipsoid_mask(n, nz, centre, semi)] += np.float32(act)
    return out


def attenuation_map(n: int, seed: int, nz: int | None = None) -> np.ndarray:
    nz = n if nz is None else nz
    rng = np.random.default_rng(seed + 1000)
    c = (n - 1) / 2
    mu = np.zeros((n, n, nz), dtype=np.float32)
    body = (0.42 * n, 0.36 * n, 0.45 * nz)
    mu[_ellipsoid_mask(n, nz, (c, c, (nz - 1) / 2), body)] = 0.15
    for _ in range(2):
        semi = rng.uniform(0.05, 0.12, size=3) * np.array([n, n, nz])
        off = rng.uniform(-0.2, 0.2, size=2) * n
        centre = (c + off[0], c + off[1], rng.uniform(0.3, 0.7) * (nz - 1))
        mu[_ellipsoid_mask(n, nz, centre, semi)] = 0.25
    return mu


def angles(n_angles: int) -> np.ndarray:
    return np.arange(n_angles, dtype=np.float64) * (360.0 / n_angles)


DX = 0.48          # cm
RADIUS = 25.0      # cm
PSF_PARAMS = (0.03, 0.1)


def poisson_data(expected: torch.Tensor, seed: int) -> torch.Tensor:
    """Poisson counts with mean ``expected``, drawn on the host (seeded CPU
    generator) and returned on ``expected``'s device and dtype.

    The rates are checked first: a NaN, infinite or negative rate raises
    ``ValueError`` naming the offending count, instead of reaching a sampler
    whose device build aborts on it."""
    host = expected.detach().to("cpu", torch.float64)
    bad_finite = int((~torch.isfinite(host)).sum())
    bad_sign = int((host < 0).sum())
    if bad_finite or bad_sign:
        raise ValueError(f"Poisson rates must be finite and >= 0: {bad_finite} non-finite and "
                         f"{bad_sign} negative of {host.numel()}")
    gen = torch.Generator(device="cpu").manual_seed(seed)
    return torch.poisson(host, generator=gen).to(device=expected.device, dtype=expected.dtype)
