"""Batched OSEM (many uneven problems per launch, hipGraph capture) against
the one-at-a-time OSEM of each problem, and graph replays against eager runs."""
from __future__ import annotations

import pytest
import torch

from pytomography_amd import synthetic
from tests.helpers import problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu

SPECS = [(24, 12, 0), (32, 16, 1), (24, 16, 2), (32, 12, 3), (40, 12, 4), (24, 12, 5)]


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _problems(cuda, chunk_planes=0):
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    out = []
    for n, na, seed in SPECS:
        pa = problem_arrays(n, na, seed=seed, att=True, psf=True)
        sm = product_system_matrix(pa, device=cuda, chunk_planes=chunk_planes)
        f_true = torch.from_numpy(pa["f"]).to(cuda)
        exp = sm.forward(f_true) * 100
        s = torch.full_like(exp, 0.05 * float(exp.mean()))
        g = synthetic.poisson_data(exp + s, seed)
        out.append(PoissonLogLikelihood(sm, g, additive_term=s))
    return out


def test_batched_equals_one_at_a_time(cuda):
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.batch import BatchedOSEM
    iters, subsets = 2, 4
    ref = [OSEM(lk)(iters, subsets).clone() for lk in _problems(cuda)]
    batch = BatchedOSEM(_problems(cuda), subsets, use_graph=False)
    out = batch.run(iters)
    torch.cuda.synchronize()
    for (n, na, _), a, b in zip(SPECS, out, ref):
        e = rel_l2(a, b)
        print(f"{n}^3 x {na}: batched vs one-at-a-time rel-L2 {e:.1e}, max abs {float((a - b).abs().max()):.1e}")
        assert e <= 1e-6


def test_graph_replays_match_eager(cuda):
    from pytomography_amd.batch import BatchedOSEM
    subsets = 4
    graph = BatchedOSEM(_problems(cuda), subsets, use_graph=True)
    eager = BatchedOSEM(_problems(cuda), subsets, use_graph=False)
    for replay in range(3):
        graph.run(1)
        torch.cuda.synchronize()
        x = torch.ones(1000, device=cuda).sum()  # unrelated eager work between replays
        eager.run(1)
        torch.cuda.synchronize()
        assert float(x) == 1000
        for a, b in zip(graph.objects, eager.objects):
            assert torch.equal(a, b), f"replay {replay} differs from the eager run"
        print(f"replay {replay}: graph == eager for {len(SPECS)} problems")


def test_autotuner_returns_candidate(cuda):
    from pytomography_amd.batch import autotune_chunk_planes, CHUNK_CANDIDATES
    pa = problem_arrays(32, 16, seed=0, att=True, psf=True)
    best = autotune_chunk_planes(lambda c: product_system_matrix(pa, device=cuda, chunk_planes=c),
                                 32, 16, 4, key="test_32_16_4", device=cuda)
    assert best in CHUNK_CANDIDATES


def test_schedule_autotuner_returns_candidate(cuda):
    """The schedule autotuner returns one of its candidates, and every
    candidate schedule (balanced pieces over 1..4 rounds, uniform chunks of
    8..32 planes) projects and back-projects as the default one does: the
    schedules differ only in how the planes' partial sums are grouped."""
    from pytomography_amd.batch import autotune_schedule, SCHEDULE_CANDIDATES
    pa = problem_arrays(32, 16, seed=0, att=True, psf=True)
    make = lambda c, r: product_system_matrix(pa, device=cuda, chunk_planes=c, sched_rounds=r)
    best = autotune_schedule(make, 32, 16, 4, key="test_32_16_4", device=cuda)
    assert best in SCHEDULE_CANDIDATES
    f = torch.from_numpy(pa["f"]).to(cuda)
    ref = make(0, 0)
    p_ref = ref.forward(f)
    b_ref = ref.backward(p_ref)
    for c, r in SCHEDULE_CANDIDATES:
        sm = make(c, r)
        p = sm.forward(f)
        b = sm.backward(p_ref)
        print(f"chunk {c} rounds {r}: forward {rel_l2(p, p_ref):.2e} backward {rel_l2(b, b_ref):.2e}")
        assert rel_l2(p, p_ref) < 1e-6
        assert rel_l2(b, b_ref) < 1e-6


def test_batched_problems_sharing_a_system_matrix(cuda):
    """Several datasets on one scanner geometry: the likelihoods share one
    SPECTSystemMatrix (and its plan and attenuation data); each batched
    problem must still get its own workspace, so the batch equals the
    one-at-a-time OSEM of every dataset."""
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.batch import BatchedOSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    pa = problem_arrays(32, 16, seed=7, att=True, psf=True)
    sm = product_system_matrix(pa, device=cuda)
    f_true = torch.from_numpy(pa["f"]).to(cuda)
    lks = []
    for k, scale in enumerate((50.0, 100.0, 200.0)):
        exp = sm.forward(f_true) * scale
        s = torch.full_like(exp, 0.05 * float(exp.mean()))
        lks.append(PoissonLogLikelihood(sm, synthetic.poisson_data(exp + s, 100 + k), additive_term=s))
    iters, subsets = 2, 4
    ref = [OSEM(lk)(iters, subsets).clone() for lk in lks]
    out = BatchedOSEM(lks, subsets, use_graph=False).run(iters)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(out, ref)):
        e = rel_l2(a, b)
        print(f"shared system matrix, dataset {k}: batched vs one-at-a-time rel-L2 {e:.1e}")
        assert e <= 1e-6
    # the datasets differ, so the reconstructions must too
    assert rel_l2(out[0], out[2]) > 1e-3


def test_batch_rejects_shared_workspace(cuda):
    """Two problems handed the same workspace (or iterate) are refused."""
    import ctypes
    from pytomography_amd import _lib
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    pa = problem_arrays(24, 12, seed=8, att=True, psf=True)
    sm = product_system_matrix(pa, device=cuda)
    exp = sm.forward(torch.from_numpy(pa["f"]).to(cuda)) * 50
    lk = PoissonLogLikelihood(sm, synthetic.poisson_data(exp, 9))
    lk._set_n_subsets(2)
    lib = _lib.lib()
    plan = sm._plan(cuda)
    idx = [sm._angle_index(k, cuda) for k in range(2)]
    g = [lk._get_projection_subset(lk.projections, k) for k in range(2)]
    norm = [lk._get_normBP(k) for k in range(2)]
    nbytes = int(lib.spect_workspace_bytes(plan.handle, _lib.SPECT_F32, max(t.numel() for t in idx)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    fs = [torch.ones(sm._shape, device=cuda) for _ in range(2)]

    def vp(ts):
        return (ctypes.c_void_p * 2)(*[t.data_ptr() for t in ts])

    counts = (ctypes.c_int32 * 2)(*[int(t.numel()) for t in idx])
    arrs = [vp(idx), vp(g), vp(norm)]
    probs = (_lib.SpectBatchProblem * 2)()
    for i in range(2):
        probs[i] = _lib.SpectBatchProblem(
            plan=plan.handle, f=fs[i].data_ptr(), mu=None, att_table=None, att_field=0,
            n_subsets=2, angle_idx=arrs[0], counts=counts, g=arrs[1], s=None, norm=arrs[2],
            pred=None, workspace=ws.data_ptr(), workspace_bytes=nbytes)
    handle = ctypes.c_void_p()
    rc = lib.spect_batch_create(probs, 2, _lib.SPECT_F32, 1e-11, 0.0, ctypes.byref(handle))
    assert rc == -1 and b"overlap" in lib.spect_last_error()
    assert not handle.value


# A reduced config-4 mix that still spans every side N in {64..192}, every
# angle count in {60, 90, 120} and every PSF width of that range (K = 21..37,
# i.e. the batch kernels' H = 10..18 instantiations), attenuation + PSF.
MIX = [(64, 60, 11), (96, 90, 12), (128, 120, 13), (160, 60, 14), (192, 90, 15)]


def _mix_likelihoods(cuda):
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    out = []
    for n, na, seed in MIX:
        print(f"[mix] building {n}^3 x {na}", flush=True)
        pa = problem_arrays(n, na, seed=seed, att=True, psf=True)
        sm = product_system_matrix(pa, device=cuda)
        exp = sm.forward(torch.from_numpy(pa["f"]).to(cuda)) * 1e3
        s = torch.full_like(exp, 0.05 * float(exp.mean()))
        out.append((pa, PoissonLogLikelihood(sm, synthetic.poisson_data(exp + s, seed), additive_term=s)))
    return out


def test_uneven_mix_batched_vs_one_at_a_time_and_oracle(cuda):
    """Every problem of the mix: the batched hipGraph run (one iteration per
    replay, 4 x 8) matches its one-at-a-time OSEM to 1e-6 rel-L2 (the same
    per-plane arithmetic; the one-at-a-time forward adds its partial ray sums
    in the balanced schedule's pieces, the batch in uniform chunks, so the
    float32 sums are rounded in a different order); the 64^3 x 60 problem also
    matches the CPU oracle after every iteration (rel-L2 1e-5)."""
    from oracle import spect_oracle as so
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.batch import BatchedOSEM
    from tests.helpers import oracle_geometry
    iters, subsets = 4, 8
    probs = _mix_likelihoods(cuda)
    H = [lk.system_matrix._plan(cuda).psf_half for _, lk in probs]
    assert sorted(set(H)) == [10, 12, 14, 16, 18], H
    ref = []
    for (pa, lk) in probs:
        ref.append(OSEM(lk)(iters, subsets).clone())
    batch = BatchedOSEM([lk for _, lk in probs], subsets, use_graph=True)
    iterates = []
    for _ in range(iters):
        batch.run(1)
        iterates.append([o.clone() for o in batch.objects])
    torch.cuda.synchronize()
    for (n, na, _), a, b in zip(MIX, iterates[-1], ref):
        e = rel_l2(a, b)
        print(f"{n}^3 x {na}: batched vs one-at-a-time rel-L2 {e:.1e}, "
              f"max abs {float((a - b).abs().max()):.1e}")
        assert e <= 1e-6
    pa, lk = probs[0]
    g32 = oracle_geometry(pa, torch.float32)
    data = lk.projections.cpu()
    add = lk.additive_term.cpu()
    states = []
    so.osem(g32, data, iters, subsets, additive=add,
            callback=lambda f, j, k, pp: states.append(f.clone()) if k == subsets - 1 else None)
    for j in range(iters):
        e = rel_l2(iterates[j][0], states[j])
        print(f"64^3 x 60 iteration {j + 1}: batched vs oracle rel-L2 {e:.2e}")
        assert e <= 1e-5
