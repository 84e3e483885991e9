"""GPU parity: HIP kernels (through the C ABI) against the CPU oracle.

Bar (FP32): relative L2 <= 1e-5 per projection and per OSEM iterate; each
result also reports its error against the float64 oracle.  Adjoint identity
<Hx, y> = <x, H^T y> to <= 1e-6 in FP64.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import spect_oracle as so
from tests.helpers import oracle_geometry, problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [
    dict(n=16, n_angles=8, att=False, psf=False),
    dict(n=24, n_angles=12, att=True, psf=False),
    dict(n=24, n_angles=12, att=False, psf=True),
    dict(n=32, n_angles=12, att=True, psf=True),
    dict(n=33, n_angles=10, att=True, psf=True, nz=20),
    # z extent not a multiple of 4: the element-wise (ragged) load paths
    dict(n=21, n_angles=9, att=True, psf=True, nz=13),
]


def _ids(c):
    return f"n{c['n']}a{c['n_angles']}{'_att' if c['att'] else ''}{'_psf' if c['psf'] else ''}"


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _per_projection_rel(a, b):
    return [rel_l2(a[k], b[k]) for k in range(a.shape[0])]


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_forward_matches_oracle(cuda, case):
    pa = problem_arrays(case["n"], case["n_angles"], seed=1, nz=case.get("nz"), att=case["att"],
                        psf=case["psf"])
    sm = product_system_matrix(pa)
    f = torch.from_numpy(pa["f"])
    out = sm.forward(f.to(cuda)).cpu()
    ref32 = so.forward(oracle_geometry(pa, torch.float32), f)
    ref64 = so.forward(oracle_geometry(pa, torch.float64), f.double())
    errs = _per_projection_rel(out, ref32)
    print(f"[{_ids(case)}] forward rel-L2 max {max(errs):.2e}  vs fp64 {rel_l2(out, ref64):.2e}")
    assert max(errs) <= TOL


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_backward_reference_matches_oracle(cuda, case):
    pa = problem_arrays(case["n"], case["n_angles"], seed=2, nz=case.get("nz"), att=case["att"],
                        psf=case["psf"])
    sm = product_system_matrix(pa)
    gen = torch.Generator().manual_seed(3)
    y = torch.rand((case["n_angles"], case["n"], pa["nz"]), generator=gen)
    out = sm.backward(y.to(cuda)).cpu()
    ref32 = so.backward(oracle_geometry(pa, torch.float32), y)
    ref64 = so.backward(oracle_geometry(pa, torch.float64), y.double())
    e = rel_l2(out, ref32)
    print(f"[{_ids(case)}] backward rel-L2 {e:.2e}  vs fp64 {rel_l2(out, ref64):.2e}")
    assert e <= TOL


@pytest.mark.parametrize("case", CASES[:4], ids=_ids)
def test_backward_adjoint_matches_autograd(cuda, case):
    pa = problem_arrays(case["n"], case["n_angles"], seed=4, att=case["att"], psf=case["psf"])
    sm = product_system_matrix(pa)
    gen = torch.Generator().manual_seed(5)
    y = torch.rand((case["n_angles"], case["n"], case["n"]), generator=gen)
    out = sm.backward_adjoint(y.to(cuda)).cpu()
    ref = so.exact_adjoint(oracle_geometry(pa, torch.float64), y.double())
    e = rel_l2(out, ref)
    print(f"[{_ids(case)}] exact adjoint rel-L2 vs fp64 autograd {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("case", CASES[:4], ids=_ids)
def test_adjoint_identity_fp64(cuda, case):
    pa = problem_arrays(case["n"], case["n_angles"], seed=6, att=case["att"], psf=case["psf"])
    sm = product_system_matrix(pa)
    gen = torch.Generator().manual_seed(7)
    x = torch.rand((case["n"],) * 3, generator=gen, dtype=torch.float64).to(cuda)
    y = torch.rand((case["n_angles"], case["n"], case["n"]), generator=gen, dtype=torch.float64).to(cuda)
    hx = sm.forward(x)
    hty = sm.backward_adjoint(y)
    lhs = float((hx * y).sum())
    rhs = float((x * hty).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    ref_bp = sm.backward(y)
    rel_ref = abs(lhs - float((x * ref_bp).sum())) / abs(lhs)
    print(f"[{_ids(case)}] <Hx,y>={lhs:.12e} <x,HTy>={rhs:.12e} rel={rel:.2e}; "
          f"reference backward mismatch {rel_ref:.2e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("chunk_planes", [8, 24])
def test_noncircular_orbit_and_chunking(cuda, chunk_planes):
    """A different detector radius at every angle (a non-circular orbit: one
    PSF table per radius) and a non-default plane chunk, forward and
    reference back projection against the oracle."""
    pa = problem_arrays(24, 10, seed=11, nz=20, att=True, psf=True)
    pa["radii"] = 20.0 + 6.0 * np.abs(np.sin(np.deg2rad(pa["angles"])))
    sm = product_system_matrix(pa, chunk_planes=chunk_planes)
    geo = oracle_geometry(pa, torch.float32)
    f = torch.from_numpy(pa["f"])
    y = torch.rand((10, 24, 20), generator=torch.Generator().manual_seed(12))
    ef = max(_per_projection_rel(sm.forward(f.to(cuda)).cpu(), so.forward(geo, f)))
    eb = rel_l2(sm.backward(y.to(cuda)).cpu(), so.backward(geo, y))
    print(f"non-circular orbit, L={chunk_planes}: forward {ef:.2e}, backward {eb:.2e}")
    assert ef <= TOL and eb <= TOL


def test_ragged_z_adjoint_and_identity(cuda):
    """z extent 13 (not a multiple of 4): the exact adjoint's element-wise
    gather path against the oracle's autograd, and <Hx,y> = <x,H^T y> in FP64."""
    pa = problem_arrays(21, 9, seed=9, nz=13, att=True, psf=True)
    sm = product_system_matrix(pa)
    gen = torch.Generator().manual_seed(10)
    y = torch.rand((9, 21, 13), generator=gen, dtype=torch.float64)
    x = torch.rand((21, 21, 13), generator=gen, dtype=torch.float64)
    ref = so.exact_adjoint(oracle_geometry(pa, torch.float64), y)
    e32 = rel_l2(sm.backward_adjoint(y.float().to(cuda)).cpu(), ref)
    hx = sm.forward(x.to(cuda))
    hty = sm.backward_adjoint(y.to(cuda))
    lhs, rhs = float((hx * y.to(cuda)).sum()), float((x.to(cuda) * hty).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"ragged z: exact adjoint FP32 vs FP64 oracle {e32:.2e}; adjoint identity {rel:.2e}")
    assert e32 <= TOL and rel <= 1e-12


@pytest.mark.parametrize("nz", [24, 160])
def test_fp64_forward_matches_fp64_oracle(cuda, nz):
    """nz = 160 needs ceil(160 / RZ) z runs per block within the forward
    kernel's 512-thread launch bound."""
    pa = problem_arrays(24, 12, seed=8, nz=nz, att=True, psf=True)
    sm = product_system_matrix(pa)
    f = torch.from_numpy(pa["f"]).double()
    out = sm.forward(f.to(cuda)).cpu()
    ref = so.forward(oracle_geometry(pa, torch.float64), f)
    e = rel_l2(out, ref)
    print(f"fp64 forward rel-L2 vs fp64 oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("att,psf", [(False, False), (True, True)])
def test_osem_iterates_match_oracle(cuda, att, psf):
    import pytomography_amd
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    from pytomography_amd.callbacks import Callback

    n, na, iters, subsets = 24, 16, 2, 4
    pa = problem_arrays(n, na, seed=9, att=att, psf=psf)
    g32 = oracle_geometry(pa, torch.float32)
    f_true = torch.from_numpy(pa["f"])
    expected = so.forward(g32, f_true) * 100
    gen = torch.Generator().manual_seed(10)
    data = torch.poisson(expected.double(), generator=gen).float()
    additive = torch.full_like(data, 0.05 * float(expected.mean())) if att else None
    if additive is not None:
        data = data + torch.poisson(additive.double(), generator=gen).float()

    ref_iterates = []
    so.osem(g32, data, iters, subsets, additive=additive,
            callback=lambda f, j, k, pp: ref_iterates.append(f.clone()))

    class Keep(Callback):
        def __init__(self):
            self.iterates = []

        def run(self, obj, n_iter, n_subset):
            self.iterates.append(obj.detach().cpu().clone())
            return obj

    sm = product_system_matrix(pa)
    lk = PoissonLogLikelihood(sm, data.to(cuda),
                              additive_term=None if additive is None else additive.to(cuda))
    cb = Keep()
    OSEM(lk)(iters, subsets, callback=cb)
    errs = [rel_l2(a, b) for a, b in zip(cb.iterates, ref_iterates)]
    print(f"OSEM att={att} psf={psf} per-iterate rel-L2: " + " ".join(f"{e:.1e}" for e in errs))
    assert len(errs) == iters * subsets
    assert max(errs) <= TOL


def test_subset_forward_backward(cuda):
    pa = problem_arrays(24, 16, seed=11, att=True, psf=True)
    sm = product_system_matrix(pa)
    sm.set_n_subsets(4)
    g32 = oracle_geometry(pa, torch.float32)
    subsets = g32.subsets(4)
    f = torch.from_numpy(pa["f"])
    for k in range(4):
        out = sm.forward(f.to(cuda), k).cpu()
        ref = so.forward(g32, f, subsets[k])
        assert rel_l2(out, ref) <= TOL
        y = torch.rand(out.shape, generator=torch.Generator().manual_seed(k))
        b = sm.backward(y.to(cuda), k).cpu()
        assert rel_l2(b, so.backward(g32, y, subsets[k])) <= TOL


def test_rdp_gradient_kernel_matches_oracle(cuda):
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta
    from pytomography_amd.priors import RelativeDifferencePrior
    gen = torch.Generator().manual_seed(12)
    f = torch.rand((20, 18, 16), generator=gen) + 0.05
    f[3:7, 4:9, 2:5] = 0.0
    prior = RelativeDifferencePrior(beta=0.1, gamma=2.0)
    prior.set_object_meta(SPECTObjectMeta((0.48, 0.5, 0.52), f.shape))
    for dt in (torch.float32, torch.float64):
        out = prior.gradient(f.to(cuda, dt), beta_scale=0.125).cpu()
        ref = so.rdp_gradient(f.to(dt), 0.1, 2.0, (0.48, 0.5, 0.52), 0.125)
        e = rel_l2(out, ref)
        print(f"RDP gradient {dt}: rel-L2 {e:.2e}")
        assert e <= (1e-6 if dt == torch.float32 else 1e-14)


def test_bsrem_iterates_match_oracle(cuda):
    from pytomography_amd.algorithms import BSREM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    from pytomography_amd.priors import RelativeDifferencePrior
    from pytomography_amd.callbacks import Callback

    n, na, iters, subsets = 24, 16, 2, 4
    pa = problem_arrays(n, na, seed=13, att=True, psf=True)
    g32 = oracle_geometry(pa, torch.float32)
    f_true = torch.from_numpy(pa["f"])
    expected = so.forward(g32, f_true) * 100
    gen = torch.Generator().manual_seed(14)
    s = torch.full_like(expected, 0.05 * float(expected.mean()))
    data = torch.poisson((expected + s).double(), generator=gen).float()

    ref_iterates = []
    f_ref = torch.ones(g32.shape)
    subs = g32.subsets(subsets)
    norms = [so.sensitivity(g32, sb) for sb in subs]
    nsum = torch.stack(norms).sum(axis=0)
    for j in range(iters):
        for k, sidx in enumerate(subs):
            wk = len(sidx) / na
            pg = so.rdp_gradient(f_ref.clone(), 0.1, 2.0, g32.dr, wk)
            pp = so.forward(g32, f_ref, sidx) + s[sidx]
            grad = so.backward(g32, data[sidx] / (pp + so.DELTA), sidx) - norms[k]
            f_ref += f_ref * (1 / (wk * nsum + so.DELTA)) * (grad - pg)
            f_ref[f_ref <= 1e-4] = 1e-4
            ref_iterates.append(f_ref.clone())

    class Keep(Callback):
        def __init__(self):
            self.iterates = []

        def run(self, obj, n_iter, n_subset):
            self.iterates.append(obj.detach().cpu().clone())
            return obj

    sm = product_system_matrix(pa)
    lk = PoissonLogLikelihood(sm, data.to(cuda), additive_term=s.to(cuda))
    cb = Keep()
    BSREM(lk, prior=RelativeDifferencePrior(beta=0.1, gamma=2.0))(iters, subsets, callback=cb)
    errs = [rel_l2(a, b) for a, b in zip(cb.iterates, ref_iterates)]
    print("BSREM+RDP per-iterate rel-L2: " + " ".join(f"{e:.1e}" for e in errs))
    assert len(errs) == iters * subsets and max(errs) <= TOL


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_attenuation_field_matches_suffix_table(cuda, dt):
    """The resident attenuation field (SPECT_ATT_FIELD) and the FP64 suffix
    table give the same projections: both evaluate the same FP64 march."""
    pa = problem_arrays(33, 10, seed=15, nz=20, att=True, psf=True)
    a = product_system_matrix(pa, attenuation_field=True)
    b = product_system_matrix(pa, attenuation_field=False)
    a.set_n_subsets(2)
    b.set_n_subsets(2)
    gen = torch.Generator().manual_seed(16)
    f = torch.rand((33, 33, 20), generator=gen, dtype=dt).to(cuda)
    y = torch.rand((5, 33, 20), generator=gen, dtype=dt).to(cuda)
    pairs = [
        ("forward", a.forward(f, 1), b.forward(f, 1)),
        ("backward", a.backward(y, 1), b.backward(y, 1)),
        ("backward_adjoint", a.backward_adjoint(y, 1), b.backward_adjoint(y, 1)),
        ("backward_transpose", a.backward_transpose(f, 1), b.backward_transpose(f, 1)),
    ]
    assert a._att_table(f.device, dt).att_field and not b._att_table(f.device, dt).att_field
    for name, x, z in pairs:
        e = rel_l2(x, z)
        print(f"{dt} {name}: field vs suffix table rel-L2 {e:.1e}, max |diff| {float((x - z).abs().max()):.1e}")
        assert e <= (1e-6 if dt == torch.float32 else 1e-14)


def test_second_order_likelihood_operators(cuda):
    """compute_gradient_ff / gf / sf (poisson_log_likelihood.py:38-114) against
    the same compositions of the oracle's forward and reference backward."""
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    pa = problem_arrays(24, 12, seed=17, att=True, psf=True)
    sm = product_system_matrix(pa)
    sm.set_n_subsets(3)
    geo = oracle_geometry(pa, torch.float32)
    gen = torch.Generator().manual_seed(18)
    f = torch.from_numpy(pa["f"]) + 0.1
    g = torch.poisson(so.forward(geo, f) * 20, generator=gen)
    x = torch.rand(f.shape, generator=gen)
    lk = PoissonLogLikelihood(sm, g.to(cuda))
    lk._set_n_subsets(3)
    sub = geo.subsets(3)[1]
    fp = so.forward(geo, f, sub)
    hx = so.forward(geo, x, sub)
    ref = {
        "ff": -so.backward(geo, hx * g[sub] / (fp ** 2 + so.DELTA), sub),
        "gf": hx / (fp + so.DELTA),
        "sf": -hx * g[sub] / (fp + so.DELTA) ** 2,
    }
    for name in ("ff", "gf", "sf"):
        op = getattr(lk, f"compute_gradient_{name}")(f.to(cuda), subset_idx=1)
        e = rel_l2(op(x.to(cuda)), ref[name])
        print(f"compute_gradient_{name}: rel-L2 vs oracle {e:.2e}")
        assert e <= TOL


@pytest.mark.parametrize("n", [64, 128, 192])
def test_quarter_turns_keep_projections_non_negative(cuda, n):
    """At beta = 0, 90, 180, 270 degrees many samples land within an ulp of an
    integer pixel position.  Bilinear weights are then still >= 0, so a
    non-negative object projects and back-projects to non-negative values
    (the fractional part of a sample position must be computed from the
    rounded position; fused with the product that forms it, it could dip to
    about -6e-6 and make some projections of the 128^3 object negative)."""
    pa = problem_arrays(n, 4, seed=5, att=False, psf=False)
    print(f"[quarter turns] {n}^3, angles {list(pa['angles'])}", flush=True)
    sm = product_system_matrix(pa, device=cuda)
    f = torch.from_numpy(pa["f"])
    out = sm.forward(f.to(cuda)).cpu()
    ref = so.forward(oracle_geometry(pa, torch.float32), f)
    neg = int((out < 0).sum())
    print(f"{n}^3 quarter turns: {neg} negative projections, min {float(out.min()):.3e}, "
          f"rel-L2 vs oracle {rel_l2(out, ref):.2e}")
    assert neg == 0
    assert rel_l2(out, ref) <= TOL
    b = sm.backward(out.clamp_min(0).to(cuda)).cpu()
    assert int((b < 0).sum()) == 0
