"""Batched OSEM over many independent SPECT reconstructions, and the
autotuners of the kernels' plane-chunk length and work schedule.

The reference reconstructs one problem at a time through the Python subset
loop of ``PreconditionedGradientAscentAlgorithm.__call__``
(algorithms/preconditioned_gradient_ascent.py:103-122).  ``BatchedOSEM``
runs OSEM for many problems of uneven size and angle count at once: each
sub-iteration launches every kernel once for all problems that share a PSF
width and z extent (``spect_batch_*`` in include/spect_hip.h), and the whole
schedule can be captured into one hipGraph and replayed.  Each problem's
arithmetic is that of ``OSEM`` on its own ``SPECTSystemMatrix``.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.likelihoods import PoissonLogLikelihood


class BatchedOSEM:
    """OSEM for a list of ``PoissonLogLikelihood`` objects on ``SPECTSystemMatrix``
    system matrices, all with the same subset count."""

    def __init__(self, likelihoods: list[PoissonLogLikelihood], n_subsets: int,
                 object_initial: list[torch.Tensor] | None = None, use_graph: bool = True):
        if not likelihoods:
            raise ValueError("no problems")
        self.likelihoods = likelihoods
        self.n_subsets = n_subsets
        self.use_graph = use_graph
        self._keep = []
        dt = pytomography_amd.dtype if pytomography_amd.dtype == torch.float64 else torch.float32
        self.dtype = dt
        self.objects = []
        probs = (_lib.SpectBatchProblem * len(likelihoods))()
        lib = _lib.lib()
        for i, lk in enumerate(likelihoods):
            sm = lk.system_matrix
            sm._require_fused("batched OSEM")
            device = lk.projections.device
            if device.type != "cuda":
                raise _lib.SpectLibraryError("batched OSEM needs device tensors")
            lk._set_n_subsets(n_subsets)
            f = (object_initial[i] if object_initial is not None else sm._get_object_initial(device))
            f = f.to(device=device, dtype=dt).contiguous().clone()
            self.objects.append(f)
            plan = sm._plan(device)
            idx = [sm._angle_index(k if n_subsets > 1 else None, device) for k in range(n_subsets)]
            counts = (ctypes.c_int32 * n_subsets)(*[int(t.numel()) for t in idx])
            sub = (lambda t, k: lk._get_projection_subset(t, k if n_subsets > 1 else None))
            g = [_lib.require_device_tensor(sub(lk.projections, k), "projections", dt).contiguous()
                 for k in range(n_subsets)]
            s = ([_lib.require_device_tensor(sub(lk.additive_term, k), "additive_term", dt).contiguous()
                  for k in range(n_subsets)] if lk.exists_additive_term else None)
            norm = [_lib.require_device_tensor(lk._get_normBP(k if n_subsets > 1 else None), "norm", dt)
                    for k in range(n_subsets)]
            pred = [torch.empty_like(gk) for gk in g]
            mu = sm._mu(device, dt)
            tab = sm._att_table(device, dt)
            # each problem gets a workspace of its own: problems may share a
            # system matrix (one scanner geometry, several datasets), and the
            # jobs of one launch run concurrently
            nbytes = int(lib.spect_workspace_bytes(plan.handle, _lib.dtype_code(dt),
                                                   max(int(t.numel()) for t in idx)))
            ws = torch.empty(nbytes, dtype=torch.uint8, device=device)

            def ptrs(ts):
                return (ctypes.c_void_p * n_subsets)(*[None if t is None else t.data_ptr() for t in ts])

            arrs = dict(idx=ptrs(idx), g=ptrs(g), s=ptrs(s) if s is not None else None,
                        norm=ptrs(norm), pred=ptrs(pred))
            self._keep.append((idx, counts, g, s, norm, pred, mu, tab, ws, arrs, plan))
            probs[i] = _lib.SpectBatchProblem(
                plan=plan.handle, f=f.data_ptr(), mu=_lib.ptr(mu), att_table=_lib.ptr(tab),
                att_field=1 if tab is not None and tab.att_field else 0,
                n_subsets=n_subsets, angle_idx=arrs["idx"], counts=counts, g=arrs["g"],
                s=arrs["s"], norm=arrs["norm"], pred=arrs["pred"], workspace=ws.data_ptr(),
                workspace_bytes=nbytes)
            lk.projections_predicted_subsets = pred
        handle = ctypes.c_void_p()
        _lib.check(lib.spect_batch_create(probs, len(likelihoods), _lib.dtype_code(dt),
                                          float(pytomography_amd.delta), 0.0, ctypes.byref(handle)))
        self.handle = handle
        self.lib = lib

    def run(self, n_iters: int, use_graph: bool | None = None) -> list[torch.Tensor]:
        """Runs n_iters OSEM iterations of every problem (in place on ``objects``)."""
        g = self.use_graph if use_graph is None else use_graph
        device = self.objects[0].device
        _lib.check(self.lib.spect_batch_run(self.handle, int(n_iters), 1 if g else 0,
                                            _lib.stream_ptr(device)))
        return self.objects

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                self.lib.spect_batch_destroy(self.handle)
                self.handle = None
        except Exception:  # pragma: no cover
            pass


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------
_CACHE_PATH = os.path.join(os.path.expanduser("~"), ".cache", "pytomography_amd", "autotune.json")
_cache: dict | None = None
CHUNK_CANDIDATES = (4, 8, 12, 16, 24, 32)


def _load_cache() -> dict:
    global _cache
    if _cache is None:
        try:
            with open(_CACHE_PATH) as fh:
                _cache = json.load(fh)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _store_cache() -> None:
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        with open(_CACHE_PATH, "w") as fh:
            json.dump(_cache, fh, indent=1)
    except OSError:
        pass


def _osem_step_seconds(sm, n: int, n_subsets: int, device, reps: int) -> float:
    """Mean seconds of one OSEM sub-iteration (subset 0) on unit data, after
    one untimed call (attenuation data, plan, workspace)."""
    sm.set_n_subsets(n_subsets)
    a_s = len(sm.subset_indices_array[0])
    nz = sm.object_meta.shape[2]
    f = torch.ones((n, n, nz), device=device)
    g = torch.ones((a_s, n, nz), device=device)
    norm = torch.ones_like(f)
    sm.osem_step(f, g, None, norm, 0, keep_prediction=False)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(reps):
        sm.osem_step(f, g, None, norm, 0, keep_prediction=False)
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / reps


def _tune(key: str, device, candidates, seconds):
    device = device or pytomography_amd.device
    cache = _load_cache()
    ck = f"{torch.cuda.get_device_name(device)}|{key}"
    if ck in cache:
        v = cache[ck]
        return tuple(v) if isinstance(v, list) else int(v)
    best, best_t = None, float("inf")
    for cand in candidates:
        t = seconds(cand, device)
        if t < best_t:
            best, best_t = cand, t
    cache[ck] = list(best) if isinstance(best, tuple) else best
    _store_cache()
    return best


def _batch_iteration_seconds(sm, n: int, n_angles: int, n_subsets: int, device, reps: int) -> float:
    """Mean seconds of one OSEM iteration of a one-problem ``BatchedOSEM``
    (the batched kernels, eager launches) on unit data, after one untimed
    iteration."""
    nz = sm.object_meta.shape[2]
    g = torch.ones((n_angles, n, nz), device=device)
    batch = BatchedOSEM([PoissonLogLikelihood(sm, g)], n_subsets, use_graph=False)
    batch.run(1)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    batch.run(reps)
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / reps


def autotune_chunk_planes(make_system_matrix, n: int, n_angles: int, n_subsets: int, key: str,
                          device=None, candidates=CHUNK_CANDIDATES, reps: int = 3) -> int:
    """Plane-chunk length (planes per work block) of the batched kernels
    (``BatchedOSEM``, which cut every problem into uniform chunks) with the
    fastest OSEM iteration for this problem class, timed on a one-problem
    batch on the device.

    ``make_system_matrix(chunk_planes)`` builds a system matrix of the class.
    Results are cached per device name and key."""
    return _tune("batch|" + key, device, candidates,
                 lambda c, dev: _batch_iteration_seconds(make_system_matrix(c), n, n_angles,
                                                         n_subsets, dev, reps))


# (chunk_planes, sched_rounds): the library's per-call choice, the balanced
# schedule forced over 1, 2 or 4 rounds of resident blocks, and uniform chunks
# of 8, 16 or 32 planes
SCHEDULE_CANDIDATES = ((0, 0), (0, 1), (0, 2), (0, 4), (8, -1), (16, -1), (32, -1))


def autotune_schedule(make_system_matrix, n: int, n_angles: int, n_subsets: int, key: str,
                      device=None, candidates=SCHEDULE_CANDIDATES, reps: int = 3) -> tuple[int, int]:
    """Work schedule ``(chunk_planes, sched_rounds)`` of the projector kernels
    with the fastest OSEM sub-iteration for this problem class, measured on the
    device: how the (angle, column tile) pairs' visited planes are cut into
    blocks -- balanced pieces over a number of rounds of resident blocks, or
    uniform chunks of a plane count.

    ``make_system_matrix(chunk_planes, sched_rounds)`` builds a system matrix
    of the class.  Results are cached per device name and key."""
    return _tune("sched|" + key, device, candidates,
                 lambda c, dev: _osem_step_seconds(make_system_matrix(*c), n, n_subsets, dev, reps))
