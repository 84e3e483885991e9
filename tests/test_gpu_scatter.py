"""Energy-window scatter estimate (utils/scatter.py:10-99) on the HIP kernels
against the oracle's restatement (pinned against scipy and closed forms)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import pytomography_amd
from oracle import spect_oracle as so
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _windows(shape, seed):
    gen = torch.Generator().manual_seed(seed)
    lo = torch.poisson(torch.full(shape, 30.0, dtype=torch.float64), generator=gen)
    up = torch.poisson(torch.full(shape, 8.0, dtype=torch.float64), generator=gen)
    return lo, up


def _proj_meta(cuda, n_angles, nr, nz, dt):
    from pytomography_amd.metadata.SPECT import SPECTProjMeta
    old = pytomography_amd.dtype
    pytomography_amd.set_dtype(dt)
    try:
        return SPECTProjMeta((nr, nz), (0.48, 0.5), list(np.arange(n_angles) * 360.0 / n_angles))
    finally:
        pytomography_amd.set_dtype(old)
        torch.set_default_dtype(torch.float32)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_tew_smoothed_matches_oracle(cuda, dt):
    from pytomography_amd.utils import compute_EW_scatter
    A, nr, nz = 60, 64, 48
    lo, up = _windows((A, nr, nz), 51)
    pm = _proj_meta(cuda, A, nr, nz, dt)
    dtheta = float(torch.diff(pm.angles)[0].item())
    kw = dict(sigma_theta=8.0, sigma_r=1.2, sigma_z=0.9)
    est, var_op = compute_EW_scatter(lo.to(cuda, dt), up.to(cuda, dt), 6.0, 6.0, 28.0, 0.5, 0.5,
                                     proj_meta=pm, return_scatter_variance_estimate=True, **kw)
    ref_est, _, ref_op = so.ew_scatter(lo.to(dt), up.to(dt), 6.0, 6.0, 28.0, 0.5, 0.5,
                                       dtheta=dtheta, dr=0.48, dz=0.5, **kw)
    x = torch.rand((A, nr, nz), generator=torch.Generator().manual_seed(52), dtype=dt)
    e1 = rel_l2(est, ref_est)
    e2 = rel_l2(var_op(x.to(cuda)), ref_op(x))
    print(f"TEW smoothed {dt}: estimate rel-L2 {e1:.1e}, variance operator rel-L2 {e2:.1e}")
    tol = 1e-6 if dt == torch.float32 else 1e-14
    assert e1 <= tol and e2 <= tol


def test_dew_unsmoothed_is_bit_exact(cuda):
    from pytomography_amd.utils import compute_EW_scatter
    lo, _ = _windows((30, 32, 40), 53)
    lo = lo.float()
    est, var_op = compute_EW_scatter(lo.to(cuda), None, 7.3, None, 21.1, 0.7,
                                     return_scatter_variance_estimate=True)
    ref_est, ref_diag, _ = so.ew_scatter(lo, None, 7.3, None, 21.1, 0.7)
    assert torch.equal(est.cpu(), ref_est)
    x = torch.rand(lo.shape, generator=torch.Generator().manual_seed(54))
    assert torch.equal(var_op(x.to(cuda)).cpu(), x * ref_diag * x)


def test_tew_unsmoothed_is_bit_exact_and_errors(cuda):
    from pytomography_amd.utils import compute_EW_scatter
    lo, up = _windows((30, 32, 40), 55)
    lo, up = lo.float(), up.float()
    est = compute_EW_scatter(lo.to(cuda), up.to(cuda), 5.5, 4.5, 20.8)
    ref_est, _, _ = so.ew_scatter(lo, up, 5.5, 4.5, 20.8)
    assert torch.equal(est.cpu(), ref_est)
    with pytest.raises(ValueError):
        compute_EW_scatter(lo.to(cuda), up.to(cuda), 5.5, 4.5, 20.8, sigma_r=1.0)
    with pytest.raises(ValueError):
        compute_EW_scatter(lo.to(cuda), up[:, :16].to(cuda), 5.5, 4.5, 20.8)
