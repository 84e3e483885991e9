"""GPU parity at full size against the CPU oracle: the 128^3 object with 120
angles, attenuation (water body, two bone ellipsoids) and the depth-dependent
PSF (K = 29), one 15-angle subset of 8 -- the launches the OSEM benchmark
times; the same 128^3 x 120 geometry with neither attenuation nor PSF (the
reference's CPU-runnable case, subset 5 of 8); and the 192^3 object with 180
angles (K = 37, the unrolled BSREM geometry), one 9-angle subset of 20:
forward projection, reference back projection, and one fused OSEM
sub-iteration (forward, Poisson ratio with the additive term, back projection
and update) from a non-uniform iterate.

Bar (FP32): relative L2 <= 1e-5 per projection and for the back projection and
the updated iterate; each result also reports its error against the float64
oracle."""
from __future__ import annotations

import pytest
import torch

from oracle import spect_oracle as so
from tests.helpers import oracle_geometry, problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
# (side, angles, subsets, subset tested, attenuation + PSF)
SIZES = [(128, 120, 8, 3, True), (128, 120, 8, 5, False), (192, 180, 20, 7, True)]


@pytest.fixture(scope="module", params=SIZES,
                ids=lambda p: f"{p[0]}^3x{p[1]}" + ("_att_psf" if p[4] else "_plain"))
def problem(request):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    n, na, n_subsets, subset, att_psf = request.param
    print(f"[fullsize] building {n}^3 x {na} att+psf={att_psf}", flush=True)
    pa = problem_arrays(n, na, seed=0, att=att_psf, psf=att_psf)
    sm = product_system_matrix(pa)
    sm.set_n_subsets(n_subsets)
    g32 = oracle_geometry(pa, torch.float32)
    g64 = oracle_geometry(pa, torch.float64)
    sidx = g32.subsets(n_subsets)[subset]
    return pa, sm, g32, g64, sidx, n_subsets, subset


def test_headline_subset_forward_and_back_projection(problem):
    pa, sm, g32, g64, sidx, _, subset = problem
    dev = torch.device("cuda")
    f = torch.from_numpy(pa["f"])
    if pa["mu"] is not None:
        assert sm._att_table(dev, torch.float32).att_field  # the benchmark's attenuation mode
    kind = "att+PSF" if pa["psf"] else "plain"
    tag = f"{pa['n']}^3 x {len(pa['angles'])} {kind} (H {sm._plan(dev).psf_half}) subset {subset}"
    out = sm.forward(f.to(dev), subset).cpu()
    ref = so.forward(g32, f, sidx)
    ref64 = so.forward(g64, f.double(), sidx)
    errs = [rel_l2(out[k], ref[k]) for k in range(len(sidx))]
    print(f"{tag}: forward rel-L2 max {max(errs):.2e} "
          f"(vs fp64 {rel_l2(out, ref64):.2e})")
    assert max(errs) <= TOL
    y = torch.rand(out.shape, generator=torch.Generator().manual_seed(11))
    b = sm.backward(y.to(dev), subset).cpu()
    rb = so.backward(g32, y, sidx)
    rb64 = so.backward(g64, y.double(), sidx)
    e = rel_l2(b, rb)
    print(f"{tag}: reference back projection rel-L2 {e:.2e} "
          f"(vs fp64 {rel_l2(b, rb64):.2e})")
    assert e <= TOL


def test_headline_osem_subiteration(problem):
    """x <- x + x / (n + d) * (H^T(g / (Hx + s + d)) - n), floored at 0, with
    n = H^T 1 of the subset (preconditioned_gradient_ascent.py:121-122,329)."""
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    pa, sm, g32, g64, sidx, n_subsets, subset = problem
    dev = torch.device("cuda")
    gen = torch.Generator().manual_seed(12)
    f = torch.from_numpy(pa["f"])
    expected = torch.zeros((len(pa["angles"]), pa["n"], pa["nz"]))
    expected[sidx] = so.forward(g32, f, sidx)  # only this subset's data is read
    s = torch.full_like(expected, 0.05 * float(expected[sidx].mean()))
    data = torch.poisson(expected * 50 + s, generator=gen)
    x0 = 0.5 + torch.rand(f.shape, generator=gen)
    lk = PoissonLogLikelihood(sm, data.to(dev), additive_term=s.to(dev))
    out = OSEM(lk, object_initial=x0.to(dev))(1, n_subsets, n_subset_specific=subset).cpu()
    norm = so.sensitivity(g32, sidx)
    pp = so.forward(g32, x0, sidx)
    ratio = data[sidx] / (pp + s[sidx] + so.DELTA)
    ref = x0 + x0 * (1 / (norm + so.DELTA)) * (so.backward(g32, ratio, sidx) - norm)
    ref[ref <= 0] = 0
    e = rel_l2(out, ref)
    print(f"{pa['n']}^3 x {len(pa['angles'])} subset {subset}: OSEM sub-iteration rel-L2 {e:.2e}")
    assert e <= TOL
