"""OSMAPOSL, RBI-EM, RBI-MAP, the prior's derivative of order 2 and the
uncertainty of a masked sum, on the accelerated projectors, against the
oracle's restatements (pinned in tests/test_oracle.py)."""
from __future__ import annotations

import pytest
import torch

import pytomography_amd
from oracle import spect_oracle as so
from tests.helpers import oracle_geometry, problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


@pytest.fixture
def float64_run():
    old = pytomography_amd.dtype
    pytomography_amd.set_dtype(torch.float64)
    yield
    pytomography_amd.set_dtype(old)
    torch.set_default_dtype(torch.float32)


def test_rdp_hessian_jvp_kernel_matches_oracle(cuda):
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta
    from pytomography_amd.priors import RelativeDifferencePrior
    gen = torch.Generator().manual_seed(41)
    f = torch.rand((20, 18, 16), generator=gen) + 0.05
    f[3:7, 4:9, 2:5] = 0.0
    u = torch.rand(f.shape, generator=gen) - 0.5
    dr = (0.48, 0.5, 0.52)
    prior = RelativeDifferencePrior(beta=0.1, gamma=2.0)
    prior.set_object_meta(SPECTObjectMeta(dr, f.shape))
    for dt in (torch.float32, torch.float64):
        ref = so.rdp_hessian_jvp(f.to(dt), u.to(dt), 0.1, 2.0, dr, 0.25)
        out = prior.hessian_jvp(f.to(cuda, dt), u.to(cuda, dt), beta_scale=0.25).cpu()
        prior.set_object(f.to(cuda, dt))
        prior.set_beta_scale(0.25)
        via_call = prior(derivative_order=2)(u.to(cuda, dt)).cpu()
        e, e2 = rel_l2(out, ref), rel_l2(via_call, ref)
        print(f"RDP order-2 operator {dt}: rel-L2 {e:.2e} (via __call__ {e2:.2e})")
        tol = 1e-6 if dt == torch.float32 else 1e-14
        assert e <= tol and e2 <= tol


def _problem(cuda, n=24, na=12, seed=43):
    pa = problem_arrays(n, na, seed=seed, att=True, psf=True)
    g32 = oracle_geometry(pa, torch.float32)
    f_true = torch.from_numpy(pa["f"])
    expected = so.forward(g32, f_true) * 100
    s = torch.full_like(expected, 0.05 * float(expected.mean()))
    data = torch.poisson((expected + s).double(), generator=torch.Generator().manual_seed(seed + 1)).float()
    return pa, data, s


@pytest.mark.parametrize("name", ["OSMAPOSL", "RBIEM", "RBIMAP"])
def test_map_algorithms_match_oracle(cuda, name):
    import pytomography_amd.algorithms as algs
    from pytomography_amd.callbacks import Callback
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    from pytomography_amd.priors import RelativeDifferencePrior
    iters, subsets, beta = 2, 3, 0.05
    pa, data, s = _problem(cuda)
    g32 = oracle_geometry(pa, torch.float32)
    use_prior = name != "RBIEM"
    ref = []
    so.gradient_ascent(g32, data, iters, subsets, name.lower(), beta=beta if use_prior else None,
                       additive=s, callback=lambda f, pp, j, k: ref.append(f))

    class Keep(Callback):
        def __init__(self):
            self.iterates = []

        def run(self, obj, n_iter, n_subset):
            self.iterates.append(obj.detach().cpu().clone())
            return obj

    sm = product_system_matrix(pa)
    lk = PoissonLogLikelihood(sm, data.to(cuda), additive_term=s.to(cuda))
    prior = RelativeDifferencePrior(beta=beta, gamma=2.0) if use_prior else None
    cb = Keep()
    out = getattr(algs, name)(lk, prior=prior)(iters, subsets, callback=cb)
    fin = so.gradient_ascent(g32, data, iters, subsets, name.lower(), beta=beta if use_prior else None,
                             additive=s)
    errs = [rel_l2(a, b) for a, b in zip(cb.iterates[:-1], ref[1:])] + [rel_l2(out, fin)]
    print(f"{name} per-iterate rel-L2: " + " ".join(f"{e:.1e}" for e in errs))
    assert len(errs) == iters * subsets and max(errs) <= 1e-5


@pytest.mark.parametrize("method", ["osem", "bsrem"])
def test_compute_uncertainty_matches_oracle(cuda, float64_run, method):
    from pytomography_amd.algorithms import BSREM, OSEM
    from pytomography_amd.callbacks import DataStorageCallback
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    from pytomography_amd.priors import RelativeDifferencePrior
    iters, subsets, beta = 2, 2, 0.1
    pa, data, s = _problem(cuda, n=16, na=8, seed=45)
    data, s = data.double(), s.double()
    g64 = oracle_geometry(pa, torch.float64)
    objects, fps = [], []
    fK = so.gradient_ascent(g64, data, iters, subsets, method,
                            beta=beta if method == "bsrem" else None, additive=s,
                            floor=1e-4 if method == "bsrem" else 0.0,
                            callback=lambda f, pp, j, k: (objects.append(f), fps.append(pp)))
    objects.append(fK)
    norms = [so.sensitivity(g64, x) for x in g64.subsets(subsets)]
    N = torch.stack(norms).sum(0)
    w = 1 / subsets
    factor = (lambda m: 1 / (norms[m] + so.DELTA)) if method == "osem" else \
        (lambda m: 1 / (w * N + so.DELTA))
    mask = torch.zeros(g64.shape, dtype=torch.float64)
    mask[4:12, 5:11, 6:10] = 1
    var = lambda v: 0.5 * v ** 2
    ref_abs, _ = so.uncertainty(g64, data, objects, fps, mask, subsets, factor,
                                beta=beta if method == "bsrem" else None, additive_variance=var)
    ref_pct = ref_abs / float((fK * mask).sum()) * 100

    sm = product_system_matrix(pa)
    lk = PoissonLogLikelihood(sm, data.to(cuda), additive_term=s.to(cuda),
                              additive_term_variance_estimate=var)
    if method == "osem":
        alg = OSEM(lk)
    else:
        alg = BSREM(lk, prior=RelativeDifferencePrior(beta=beta, gamma=2.0))
    cb = DataStorageCallback(lk, alg.object_prediction)
    alg(iters, subsets, callback=cb)
    unc, pct = alg.compute_uncertainty(mask.to(cuda), cb, return_pct=True, include_additive_term=True)
    print(f"{method} uncertainty: {unc:.10g} vs oracle {ref_abs:.10g} "
          f"({abs(unc - ref_abs) / ref_abs:.1e}), pct {pct:.6g} vs {ref_pct:.6g}")
    assert abs(unc - ref_abs) <= 1e-9 * ref_abs and abs(pct - ref_pct) <= 1e-9 * ref_pct
