"""spect_forward_noalloc (the allocation-free forward projection) against
spect_forward through the C ABI: bit-identical projections on two calls in a
row, the second on a smaller problem, with the attenuation rows in the
library's scratch (mu without a table) and with the attenuation field; and a
call above the scratch limit is refused before anything is launched."""
from __future__ import annotations

import ctypes

import pytest
import torch

from tests.helpers import problem_arrays, product_system_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _both(sm, cuda, f, subset, use_table):
    """(spect_forward, spect_forward_noalloc) projections of f for one subset."""
    from pytomography_amd import _lib
    lib = _lib.lib()
    plan = sm._plan(cuda)
    idx = sm._angle_index(subset, cuda)
    n = idx.numel()
    mu = sm._mu(cuda, torch.float32)
    tab = sm._att_table(cuda, torch.float32) if use_table else None
    code = sm._code(torch.float32, tab)
    ws_bytes = int(lib.spect_workspace_bytes(plan.handle, _lib.SPECT_F32, n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=cuda)
    out = []
    for noalloc in (False, True):
        proj = torch.full((n, sm._shape[1], sm._shape[2]), float("nan"), device=cuda)
        st = _lib.stream_ptr(cuda)
        if noalloc:
            rc = lib.spect_forward_noalloc(plan.handle, code, f.data_ptr(), _lib.ptr(mu), _lib.ptr(tab),
                                           idx.data_ptr(), n, proj.data_ptr(), st)
        else:
            rc = lib.spect_forward(plan.handle, code, f.data_ptr(), _lib.ptr(mu), _lib.ptr(tab),
                                   idx.data_ptr(), n, proj.data_ptr(), ws.data_ptr(), ws_bytes, st)
        _lib.check(rc)
        out.append(proj)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("use_table", [False, True], ids=["mu_rows_in_scratch", "att_field"])
def test_noalloc_forward_equals_forward(cuda, use_table):
    # first call: 64^3 x 32 angles, attenuation + PSF, one 16-angle subset;
    # second call: a smaller problem, 24^3 x 12 angles, one 6-angle subset
    for n, na, subsets, subset, seed in ((64, 32, 2, 1, 3), (24, 12, 2, 0, 4)):
        print(f"[noalloc] {n}^3 x {na}, subset {subset} of {subsets}, table={use_table}", flush=True)
        pa = problem_arrays(n, na, seed=seed, att=True, psf=True)
        sm = product_system_matrix(pa, device=cuda, attenuation_field=True if use_table else None)
        sm.set_n_subsets(subsets)
        f = torch.from_numpy(pa["f"]).to(cuda)
        a, b = _both(sm, cuda, f, subset, use_table)
        assert torch.isfinite(a).all()
        assert torch.equal(a, b), f"{n}^3: noalloc differs by {float((a - b).abs().max()):.3e}"
        print(f"[noalloc] {n}^3 x {na}: identical ({a.numel()} values, max {float(a.max()):.3f})")


def test_noalloc_refuses_above_the_limit(cuda):
    from pytomography_amd import _lib
    lib = _lib.lib()
    pa = problem_arrays(128, 120, seed=0, att=False, psf=True)
    sm = product_system_matrix(pa, device=cuda)
    plan = sm._plan(cuda)
    idx = sm._angle_index(None, cuda)      # all 120 angles: ~220 MB of partials
    f = torch.zeros(sm._shape, device=cuda)
    proj = torch.zeros((120, 128, 128), device=cuda)
    rc = lib.spect_forward_noalloc(plan.handle, _lib.SPECT_F32, f.data_ptr(), None, None,
                                   idx.data_ptr(), 120, proj.data_ptr(), _lib.stream_ptr(cuda))
    assert rc == -3
    msg = lib.spect_last_error().decode()
    assert "scratch" in msg and str(_lib.NOALLOC_SCRATCH_BYTES) in msg
    torch.cuda.synchronize()
    assert float(proj.abs().sum()) == 0.0   # nothing was launched
