"""Batched OSEM (many uneven problems per launch, hipGraph capture) against
the one-at-a-time OSEM of each problem, and graph replays against eager runs."""
from __future__ import annotations

import pytest
import torch

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
        g = torch.poisson(exp + s, generator=torch.Generator(device=cuda).manual_seed(seed))
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
