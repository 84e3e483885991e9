"""The plane schedule the library picks per call (sched_rounds 0): the
balanced schedule when the call's pieces of at most 32 planes fit one round of
resident blocks, uniform chunks otherwise.  The automatic choice must give the
same bits as the forced schedule it picks -- forward and back projection of an
attenuation-field + PSF problem -- at a size where the schedule fits (96^3 x
60, 8 subsets) and at one where it does not (160^3 x 120, 8 subsets); parity of
both schedules with the oracle is covered at 128^3 and 192^3 by
test_gpu_fullsize.py and at small sizes by test_gpu_parity.py."""
from __future__ import annotations

import pytest
import torch

from tests.helpers import problem_arrays, product_system_matrix, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,n_angles,picked", [(96, 60, 1), (160, 120, -1)],
                         ids=["96^3x60_balanced", "160^3x120_uniform"])
def test_automatic_schedule_matches_the_one_it_picks(n, n_angles, picked):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"[schedule] {n}^3 x {n_angles}, expecting sched_rounds {picked}", flush=True)
    pa = problem_arrays(n, n_angles, seed=3, att=True, psf=True)
    f = torch.from_numpy(pa["f"]).cuda()
    out = {}
    for rounds in (0, 1, -1):
        sm = product_system_matrix(pa, sched_rounds=rounds)
        sm.set_n_subsets(8)
        # spect_schedule_info reports the choice the launches make
        want = "balanced" if (picked if rounds == 0 else rounds) == 1 else "uniform"
        assert sm.schedule(2) == {"forward": want, "field": want}
        p = sm.forward(f, subset_idx=2)
        b = sm.backward(p, subset_idx=2)
        out[rounds] = (p, b)
        torch.cuda.synchronize()
        del sm
    other = 1 if picked == -1 else -1
    e_p = rel_l2(out[0][0], out[other][0])
    e_b = rel_l2(out[0][1], out[other][1])
    print(f"auto vs sched_rounds {other}: forward {e_p:.1e}, back projection {e_b:.1e}")
    assert torch.equal(out[0][0], out[picked][0])
    assert torch.equal(out[0][1], out[picked][1])
    # the two schedules group the planes' partial sums differently
    assert e_p < 1e-6 and e_b < 1e-6
