"""Differentiable projector: gradients of the Poisson log-likelihood with
respect to the activity f and the attenuation map mu.

* against the oracle's torch-autograd gradient (the op-for-op restatement of
  the reference path with autograd left on), in float64 and float32;
* against float64 central finite differences along 16 seeded unit
  directions with step 1e-3 at 128^3 x 120 angles (bar: 1e-4 relative).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pytomography_amd import synthetic
from oracle import spect_oracle as so
from tests.helpers import oracle_geometry, problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _data(sm, f_true, mu, scale=100.0):
    with torch.no_grad():
        hf = sm.forward_with_mu(f_true, mu) * scale
    s = torch.full_like(hf, 0.05 * float(hf.mean()))
    g = synthetic.poisson_data(hf + s, 3)
    return g, s


@pytest.mark.parametrize("psf", [False, True])
def test_gradient_matches_oracle_autograd(cuda, psf):
    from pytomography_amd.autograd import poisson_log_likelihood
    pa = problem_arrays(24, 12, seed=21, att=True, psf=psf)
    sm = product_system_matrix(pa)
    f_true = torch.from_numpy(pa["f"]).double().to(cuda)
    mu0 = torch.from_numpy(pa["mu"]).double().to(cuda)
    g, s = _data(sm, f_true, mu0)
    f = torch.full_like(f_true, 0.4) + 0.1 * torch.rand(f_true.shape, generator=torch.Generator().manual_seed(4), dtype=torch.float64).to(cuda)
    results = {}
    for dt in (torch.float64, torch.float32):
        fv = f.to(dt).clone().requires_grad_(True)
        mv = mu0.to(dt).clone().requires_grad_(True)
        L = poisson_log_likelihood(fv, mv, sm, g.to(dt), s.to(dt))
        L.backward()
        results[dt] = (float(L), fv.grad.double().cpu(), mv.grad.double().cpu())
    geo = oracle_geometry(pa, torch.float64)
    fo = f.cpu().clone().requires_grad_(True)
    mo = mu0.cpu().clone().requires_grad_(True)
    Lo = so.poisson_loglik(geo, fo, mo, g.cpu(), s.cpu())
    Lo.backward()
    for dt, tol in ((torch.float64, 1e-10), (torch.float32, 1e-5)):
        L, gf, gm = results[dt]
        ef, em = rel_l2(gf, fo.grad), rel_l2(gm, mo.grad)
        print(f"psf={psf} {dt}: L rel {abs(L - float(Lo)) / abs(float(Lo)):.1e} "
              f"grad_f rel-L2 {ef:.2e} grad_mu rel-L2 {em:.2e}")
        assert ef <= tol and em <= tol


@pytest.mark.parametrize("psf", [False, True])
def test_gradient_finite_differences_fp64(cuda, psf):
    """128^3 x 120 angles, water/bone mu, 16 seeded unit directions, step 1e-3."""
    from pytomography_amd.autograd import poisson_log_likelihood
    pa = problem_arrays(128, 120, seed=0, att=True, psf=psf)
    sm = product_system_matrix(pa)
    f_true = torch.from_numpy(pa["f"]).double().to(cuda)
    mu = torch.from_numpy(pa["mu"]).double().to(cuda)
    g, s = _data(sm, f_true, mu, scale=1e3)
    f0 = (0.5 * torch.ones_like(f_true)).requires_grad_(True)
    mu0 = mu.clone().requires_grad_(True)
    L = poisson_log_likelihood(f0, mu0, sm, g, s)
    L.backward()
    gf, gm = f0.grad.detach(), mu0.grad.detach()

    def loss(f, m):
        with torch.no_grad():
            return float(poisson_log_likelihood(f, m, sm, g, s))

    gen = torch.Generator(device="cpu").manual_seed(123)
    h = 1e-3
    worst = 0.0
    for k in range(16):
        for which in ("f", "mu"):
            d = torch.randn(f_true.shape, generator=gen, dtype=torch.float64).to(cuda)
            d /= d.norm()
            if which == "f":
                fd = (loss(f0.detach() + h * d, mu) - loss(f0.detach() - h * d, mu)) / (2 * h)
                an = float((gf * d).sum())
            else:
                fd = (loss(f0.detach(), mu + h * d) - loss(f0.detach(), mu - h * d)) / (2 * h)
                an = float((gm * d).sum())
            err = abs(fd - an) / max(abs(fd), 1e-30)
            worst = max(worst, err)
            if k < 2:
                print(f"psf={psf} dir {k} {which}: FD {fd:.10e} analytic {an:.10e} rel {err:.2e}")
    print(f"psf={psf}: worst relative FD error over 16 directions x (f, mu): {worst:.2e}")
    assert worst <= 1e-4


def test_fp32_gradient_error_vs_fp64(cuda):
    from pytomography_amd.autograd import poisson_log_likelihood
    pa = problem_arrays(64, 60, seed=5, att=True, psf=True)
    sm = product_system_matrix(pa)
    f_true = torch.from_numpy(pa["f"]).double().to(cuda)
    mu = torch.from_numpy(pa["mu"]).double().to(cuda)
    g, s = _data(sm, f_true, mu)
    grads = {}
    for dt in (torch.float64, torch.float32):
        f = (0.5 * torch.ones_like(f_true)).to(dt).requires_grad_(True)
        m = mu.to(dt).clone().requires_grad_(True)
        poisson_log_likelihood(f, m, sm, g.to(dt), s.to(dt)).backward()
        grads[dt] = (f.grad, m.grad)
    ef = rel_l2(grads[torch.float32][0], grads[torch.float64][0])
    em = rel_l2(grads[torch.float32][1], grads[torch.float64][1])
    print(f"float32 gradient vs float64: f {ef:.2e}, mu {em:.2e}")
    assert ef <= 1e-5 and em <= 1e-4
