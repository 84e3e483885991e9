#!/usr/bin/env python3
"""OSEM iterations/s of the SPECT system-matrix path on one MI355X.

Workload (default): 128^3 object at 4.8 mm, 120 angles over 360 degrees,
128 x 128 detector, attenuation (water body + 2 bone ellipsoids) and the
depth-dependent Gaussian PSF (sigma = 0.03 d + 0.1 cm, 3 sigma), radius 25
cm, Poisson data + 5 % constant scatter, OSEM with 8 subsets in float32.
One step = one full OSEM iteration (8 sub-iterations = one forward and one
back projection over all 120 angles plus 8 updates), inputs resident in HBM.

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (timed with HIP events on its stream) and a CPU baseline (the oracle's
reference-equivalent torch-CPU path on a bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload att_psf|plain|batch|unrolled|aux|schedules]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "OSEM iter/s at 128³×120 ang att+PSF: plain vs batched/adjoint; % HBM peak"


def build_problem(n: int, n_angles: int, att: bool, psf: bool, seed: int, device):
    import pytomography_amd
    from pytomography_amd import synthetic
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix

    pytomography_amd.set_device(device)
    f_true = torch.from_numpy(synthetic.ellipsoid_phantom(n, seed))
    mu = torch.from_numpy(synthetic.attenuation_map(n, seed)) if att else None
    angles = synthetic.angles(n_angles)
    om = SPECTObjectMeta((synthetic.DX,) * 3, (n, n, n))
    pm = SPECTProjMeta((n, n), (synthetic.DX, synthetic.DX), list(angles),
                       radii=[synthetic.RADIUS] * n_angles)
    tr = []
    if att:
        tr.append(SPECTAttenuationTransform(mu.to(device)))
    if psf:
        tr.append(SPECTPSFTransform(SPECTPSFMeta(synthetic.PSF_PARAMS)))
    sm = SPECTSystemMatrix(tr, [], om, pm)
    expected = sm.forward(f_true.to(device)) * 1e3
    # counts drawn on the host from the checked rates (synthetic.poisson_data)
    if att:
        scatter = torch.full_like(expected, 0.05 * float(expected.mean()))
        data = synthetic.poisson_data(expected + scatter, seed)
    else:
        scatter = None
        data = synthetic.poisson_data(expected, seed)
    return dict(sm=sm, f_true=f_true, mu=mu, angles=angles, data=data, scatter=scatter)


def algorithmic_bytes(n: int, a_s: int, att: bool):
    """Per-launch algorithmic HBM bytes (fp32, each unpadded operand streamed
    once per angle): forward reads f (+ mu) and writes the projection;
    back projection reads the ratio projection (+ mu), read-modify-writes the
    accumulator, and the fused OSEM update reads f, b, n and writes f."""
    n3, n2 = 4 * n ** 3, 4 * n ** 2
    fwd = a_s * (n3 + (n3 if att else 0) + n2)
    bwd = a_s * (n2 + (n3 if att else 0) + 2 * n3)
    upd = 4 * n3
    ratio = a_s * 4 * n2
    return fwd, bwd, upd, ratio


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob: dict, n_subsets: int, att: bool, psf: bool, reps: int = 3):
    """The oracle (op-for-op torch-CPU restatement of the reference path,
    including its per-angle loop and per-call mu rotation) run on the same
    workload as the GPU: the same counts, additive term and attenuation map,
    each subset with its own sensitivity image (computed first, untimed).
    ``reps`` full OSEM iterations (all subsets) are timed one by one from the
    uniform image; the baseline is 1 / median iteration time."""
    from oracle import spect_oracle as so
    from pytomography_amd import synthetic

    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "64"))))
    torch.set_num_threads(threads)
    n = int(prob["f_true"].shape[0])
    n_angles = len(prob["angles"])
    g = so.make_geometry((n, n, n), (synthetic.DX,) * 3, prob["angles"],
                         radii=np.full(n_angles, synthetic.RADIUS),
                         mu=prob["mu"] if att else None,
                         psf=so.PSFModel(synthetic.PSF_PARAMS) if psf else None)
    data = prob["data"].detach().cpu().float()
    additive = (prob["scatter"].detach().cpu().float() if prob["scatter"] is not None
                else torch.zeros_like(data))
    subsets = g.subsets(n_subsets)
    t0 = time.perf_counter()
    norms = [so.sensitivity(g, sidx) for sidx in subsets]
    sens_s = time.perf_counter() - t0
    f = torch.ones((n, n, n))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for k, sidx in enumerate(subsets):
            pp = so.forward(g, f, sidx) + additive[sidx]
            b = so.backward(g, data[sidx] / (pp + so.DELTA), sidx)
            f = f + f * (1 / (norms[k] + so.DELTA)) * (b - norms[k])
            f[f <= 0] = 0
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return dict(value=1.0 / med, unit="OSEM it/s", cores=threads, kind="port",
                cpu_model=cpu_model(), iteration_s=times, sensitivity_s=sens_s,
                sample=f"median of {reps} full OSEM iterations ({n_subsets} subsets of "
                       f"{len(subsets[0])} angles: forward + additive term + reference back "
                       f"projection + update) of the same {n}^3 x {n_angles} workload and data, "
                       f"after the {n_subsets} sensitivity images ({sens_s:.1f} s, untimed)")


def bench_batch(args, device):
    """Config 4: 32 problems, N in {64..192}, A in {60, 90, 120}, attenuation +
    PSF, OSEM 4 x 8 -- batched hipGraph vs the one-at-a-time GPU loop."""
    import pytomography_amd
    from pytomography_amd import synthetic
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.batch import BatchedOSEM, autotune_chunk_planes
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix

    pytomography_amd.set_device(device)
    rng = np.random.default_rng(2024)
    specs = [(int(rng.choice([64, 96, 128, 160, 192])), int(rng.choice([60, 90, 120])), s)
             for s in range(args.batch)]
    iters, subsets = 4, 8

    def system_matrix(n, na, seed, chunk):
        mu = torch.from_numpy(synthetic.attenuation_map(n, seed)).to(device)
        om = SPECTObjectMeta((synthetic.DX,) * 3, (n, n, n))
        pm = SPECTProjMeta((n, n), (synthetic.DX, synthetic.DX), list(synthetic.angles(na)),
                           radii=[synthetic.RADIUS] * na)
        return SPECTSystemMatrix([SPECTAttenuationTransform(mu),
                                  SPECTPSFTransform(SPECTPSFMeta(synthetic.PSF_PARAMS))], [], om, pm,
                                 chunk_planes=chunk)

    t0 = time.perf_counter()
    tuned = {}
    for n, na, _ in specs:
        if (n, na) not in tuned:
            tuned[(n, na)] = autotune_chunk_planes(lambda c, n=n, na=na: system_matrix(n, na, 0, c),
                                                   n, na, subsets, key=f"att_psf_{n}_{na}_{subsets}",
                                                   device=device)
    tune_s = time.perf_counter() - t0

    def likelihoods():
        out = []
        for n, na, seed in specs:
            sm = system_matrix(n, na, seed, tuned[(n, na)])
            f_true = torch.from_numpy(synthetic.ellipsoid_phantom(n, seed)).to(device)
            exp = sm.forward(f_true) * 1e3
            s = torch.full_like(exp, 0.05 * float(exp.mean()))
            g = synthetic.poisson_data(exp + s, seed)
            out.append(PoissonLogLikelihood(sm, g, additive_term=s))
        return out

    # one-at-a-time GPU loop (same kernels, one problem after another), each
    # problem timed on its own after one untimed sub-iteration (first-call
    # allocations), restarted from its initial object
    lks = likelihoods()
    for lk in lks:
        lk._set_n_subsets(subsets)
    per_problem_s = []
    torch.cuda.synchronize()
    for lk in lks:
        alg = OSEM(lk)
        alg._set_n_subsets = lambda k: None
        f0 = alg.object_prediction.clone()
        alg._fused_subiteration(0, 0)
        alg.object_prediction = f0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            for k in range(subsets):
                alg._fused_subiteration(0, k)
        torch.cuda.synchronize()
        per_problem_s.append(time.perf_counter() - t0)
    single_s = sum(per_problem_s)
    del lks

    batch = BatchedOSEM(likelihoods(), subsets, use_graph=True)
    batch.run(iters)                      # capture + first replay (warm-up)
    torch.cuda.synchronize()
    reps = max(3, args.steps)
    runs_s = []
    for _ in range(reps):
        t0 = time.perf_counter()
        batch.run(iters)
        torch.cuda.synchronize()
        runs_s.append(time.perf_counter() - t0)
    batch_s = float(np.median(runs_s))
    n_prob = len(specs)
    # Ceiling: the work of a problem is N^3 x A per iteration; the best rate
    # any problem reaches alone (the large ones fill the GPU) prices the whole
    # mix, so single_s / ideal_s bounds what any batching can gain.
    work = [n ** 3 * na for n, na, _ in specs]
    rates = [w / t for w, t in zip(work, per_problem_s)]
    ideal_s = sum(work) / max(rates)
    by_size = {}
    for (n, na, _), r in zip(specs, rates):
        by_size.setdefault(f"{n}x{na}", []).append(r / max(rates))
    out = {
        "metric": METRIC,
        "value": n_prob * iters / batch_s,
        "unit": "problem OSEM it/s (batched, 32 uneven problems)",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": 1e3 * batch_s,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, Poisson counts)",
        "config": {
            "workload": f"{n_prob} OSEM reconstructions, N in {{64..192}}, A in {{60,90,120}}, "
                        f"attenuation + PSF, {iters} iterations x {subsets} subsets, one hipGraph",
            "problems": [[n, na] for n, na, _ in specs],
            "chunk_planes": {f"{n}x{na}": c for (n, na), c in tuned.items()},
            "autotune_s": tune_s,
            "one_at_a_time_s": single_s,
            "batched_graph_s": batch_s,
            "batched_graph_runs_s": runs_s,
            "speedup_vs_one_at_a_time": single_s / batch_s,
            "work_weighted_ceiling_speedup": single_s / ideal_s,
            "solo_rate_vs_best": {k: float(np.mean(v)) for k, v in sorted(by_size.items())},
            "schedule": "per sub-iteration one forward, reduce, field and gather launch per group "
                        "of problems with equal PSF width and z extent; groups are parallel "
                        "branches of one hipGraph replayed per run",
        },
    }
    print(json.dumps(out))


def bench_unrolled(args, device):
    """Config 5: BSREM + RDP (beta 0.1, gamma 2) unrolled for 10 iterations at
    192^3 x 180 angles with attenuation + PSF; loss ||x_10 - x_true||^2
    backpropagated to the projections, the initial object and beta, with
    only the iterates checkpointed.  float32 run timed and compared with the
    same run in float64; each is preceded by an untimed one-iteration run."""
    import pytomography_amd
    from pytomography_amd.priors import RelativeDifferencePrior
    from pytomography_amd.unrolled import UnrolledBSREM

    pytomography_amd.set_device(device)
    n, na, iters, subsets = args.n, args.angles, args.iters, args.subsets
    prob = build_problem(n, na, True, True, seed=0, device=device)
    sm = prob["sm"]
    sm._ws.clear()               # drop the all-angle workspace of the data simulation
    torch.cuda.empty_cache()
    f_true = prob["f_true"].to(device)
    res = {}
    for dt in (torch.float32, torch.float64):
        un = UnrolledBSREM(sm, RelativeDifferencePrior(beta=0.1, gamma=2.0),
                           additive_term=prob["scatter"].to(dt), n_iters=iters, n_subsets=subsets)
        g = prob["data"].to(dt).detach().clone().requires_grad_(True)
        x0 = torch.ones(sm._shape, dtype=dt, device=device).requires_grad_(True)
        beta = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        un._norms(dt, device)                      # sensitivities, one-off
        sm._att_table(device, dt)                  # attenuation data of the map, one-off
        # warm-up: one unrolled iteration forward and reverse, untimed (loads every
        # kernel's code object and grows the caching allocator once)
        warm = UnrolledBSREM(sm, RelativeDifferencePrior(beta=0.1, gamma=2.0),
                             additive_term=prob["scatter"].to(dt), n_iters=1, n_subsets=subsets)
        gw = prob["data"].to(dt).detach().clone().requires_grad_(True)
        bw = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        ((warm(gw, x0.detach().clone().requires_grad_(True), bw) - f_true.to(dt)) ** 2).sum().backward()
        del warm, gw, bw
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(device)
        base = torch.cuda.memory_allocated(device)
        t0 = time.perf_counter()
        xk = un(g, x0, beta)
        loss = ((xk - f_true.to(dt)) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        after_fwd = torch.cuda.memory_allocated(device)
        peak_fwd = torch.cuda.max_memory_allocated(device)
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res[dt] = dict(fwd_s=t1 - t0, bwd_s=t2 - t1, loss=float(loss.detach()),
                       peak=torch.cuda.max_memory_allocated(device), base=base,
                       after_fwd=after_fwd, peak_fwd=peak_fwd,
                       g=g.grad.detach(), x0=x0.grad.detach(), beta=float(beta.grad))
        print(f"[unrolled {dt}] forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s, "
              f"resident before {base / 2**30:.2f} GiB, after forward {after_fwd / 2**30:.2f} GiB, "
              f"peak forward {peak_fwd / 2**30:.2f} GiB, peak {res[dt]['peak'] / 2**30:.2f} GiB",
              file=sys.stderr, flush=True)
        del un, g, x0, beta, xk, loss

    def rel(a, b):
        return float((a.double() - b.double()).norm() / b.double().norm())
    r32, r64 = res[torch.float32], res[torch.float64]
    step_s = r32["fwd_s"] + r32["bwd_s"]
    out = {
        "metric": METRIC,
        "value": iters / step_s,
        "unit": "unrolled BSREM iterations/s (forward + reverse)",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * step_s,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, water/bone mu, Poisson counts)",
        "config": {
            "workload": (f"unrolled BSREM+RDP {n}^3 x {na} angles, attenuation + PSF, "
                         f"{iters} iterations x {subsets} subsets, loss ||x_K - x_true||^2"),
            "forward_s": r32["fwd_s"], "reverse_s": r32["bwd_s"],
            "peak_hbm_gib_f32": r32["peak"] / 2**30,
            "resident_before_gib_f32": r32["base"] / 2**30,
            "checkpoints_gib_f32": (r32["after_fwd"] - r32["base"]) / 2**30,
            "peak_hbm_above_inputs_gib_f32": (r32["peak"] - r32["base"]) / 2**30,
            "peak_hbm_gib_f64": r64["peak"] / 2**30,
            "f64_forward_s": r64["fwd_s"], "f64_reverse_s": r64["bwd_s"],
            "loss_rel_err_f32_vs_f64": abs(r32["loss"] - r64["loss"]) / abs(r64["loss"]),
            "grad_g_rel_l2_f32_vs_f64": rel(r32["g"], r64["g"]),
            "grad_x0_rel_l2_f32_vs_f64": rel(r32["x0"], r64["x0"]),
            "grad_beta_rel_err_f32_vs_f64": abs(r32["beta"] - r64["beta"]) / abs(r64["beta"]),
            "grad_beta_f64": r64["beta"],
        },
    }
    print(json.dumps(out))


def bench_aux(args, device):
    """The HBM-streaming kernels beside the projector at 128^3 / 120 x 128 x 128:
    RDP gradient and its order-2 operators, the BSREM update, the Poisson
    ratio, the energy-window scatter estimate and its smoothing, the KEM basis
    (size 5) and the DVF warp.  Each is timed with HIP events around a graph
    replay of ``steps`` launches; achieved = algorithmic bytes / time."""
    import pytomography_amd
    from pytomography_amd import _lib
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta
    from pytomography_amd.priors import RelativeDifferencePrior
    from pytomography_amd.transforms.shared import DVFMotionTransform, KEMTransform
    from pytomography_amd.utils.scatter import _smooth_axis, get_1d_gaussian_taps

    pytomography_amd.set_device(device)
    n, na = args.n, args.angles
    gen = torch.Generator(device=device).manual_seed(0)
    f = torch.rand((n, n, n), generator=gen, device=device) + 0.05
    u = torch.rand((n, n, n), generator=gen, device=device)
    b = torch.rand_like(f)
    nrm = torch.rand_like(f) + 1
    nsum = torch.rand_like(f) + 8
    proj = [torch.rand((na, n, n), generator=gen, device=device) + 0.1 for _ in range(3)]
    prior = RelativeDifferencePrior(beta=0.1, gamma=2.0)
    prior.set_object_meta(SPECTObjectMeta((0.48,) * 3, (n, n, n)))
    dvf = 2 * torch.rand((n, n, n, 3), generator=gen, device=device) - 1
    warp = DVFMotionTransform(dvf, -dvf)
    kem = KEMTransform([f.cpu()], size=5)
    kem.configure(SPECTObjectMeta((0.48,) * 3, (n, n, n)), None)
    taps = get_1d_gaussian_taps(1.3, 9).to(device)
    stream = torch.cuda.current_stream(device)
    N3, P2 = n ** 3, na * n * n

    def bsrem_update():
        _lib.check(_lib.lib().spect_bsrem_update(
            _lib.SPECT_F32, u.data_ptr(), b.data_ptr(), nrm.data_ptr(), nsum.data_ptr(), b.data_ptr(),
            u.numel(), 1.0, 0.125, 1e-11, 1e-4, _lib.stream_ptr(device)))

    def ew():
        out, var = torch.empty_like(proj[0]), torch.empty_like(proj[0])
        _lib.check(_lib.lib().spect_ew_scatter(
            _lib.SPECT_F32, proj[0].data_ptr(), proj[1].data_ptr(), out.data_ptr(), var.data_ptr(),
            proj[0].numel(), 6.0, 6.0, 28.0, 0.5, 0.5, _lib.stream_ptr(device)))

    def ratio():
        pred, r = torch.empty_like(proj[0]), torch.empty_like(proj[0])
        _lib.check(_lib.lib().spect_poisson_ratio(
            _lib.SPECT_F32, proj[0].data_ptr(), proj[1].data_ptr(), proj[2].data_ptr(), pred.data_ptr(),
            r.data_ptr(), proj[0].numel(), 1e-11, _lib.stream_ptr(device)))

    cases = [
        ("spect_rdp_gradient", lambda: prior.gradient(f, 0.125), 8 * N3),
        ("spect_rdp_hessian_jvp", lambda: prior.hessian_jvp(f, u, 0.125), 12 * N3),
        ("spect_rdp_hessian_vjp", lambda: prior.hessian_vjp(f, u, 0.125), 12 * N3),
        ("spect_bsrem_update", bsrem_update, 24 * N3),
        ("spect_poisson_ratio", ratio, 20 * P2),
        ("spect_ew_scatter", ew, 16 * P2),
        ("spect_smooth_axis (theta, circular)", lambda: _smooth_axis(proj[0], 0, taps, "circular"), 8 * P2),
        ("spect_smooth_axis (r, replicate)", lambda: _smooth_axis(proj[0], 1, taps, "replicate"), 8 * P2),
        ("spect_smooth_axis (z, replicate)", lambda: _smooth_axis(proj[0], 2, taps, "replicate"), 8 * P2),
        ("spect_kem_apply (size 5)", lambda: kem.forward(f), (125 + 2) * 4 * N3),
        ("spect_dvf_warp", lambda: warp.forward(f), 20 * N3),
    ]
    rows = []
    for name, fn, nbytes in cases:
        # the launches are captured in a hipGraph and replayed, so the time is the
        # kernels' and not the Python call overhead between them
        side = torch.cuda.Stream(device)
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            for _ in range(args.warmup):
                fn()
        stream.wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(args.steps):
                fn()
        graph.replay()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        del graph
        us = e0.elapsed_time(e1) * 1e3 / args.steps
        gbs = nbytes / (us * 1e-6) / 1e9
        rows.append({"kernel": name, "us": us, "algorithmic_bytes": nbytes, "achieved_GBs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS})
        print(f"[aux] {name}: {us:.1f} us, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f} % of HBM peak)",
              file=sys.stderr, flush=True)
    fracs = np.array([r["hbm_frac"] for r in rows])
    out = {
        "metric": METRIC, "value": float(np.exp(np.log(fracs).mean())),
        "unit": "geometric-mean fraction of HBM peak over the auxiliary kernels",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(sum(r["us"] for r in rows) / 1e3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded uniform random volumes and projections)",
        "config": {"workload": f"auxiliary HBM-streaming kernels, {n}^3 objects, {na} x {n} x {n} projections",
                   "kernels": rows},
    }
    print(json.dumps(out))


def bench_schedules(args, device):
    """One att+PSF OSEM sub-iteration (subset 0 of 8) per problem size, under
    the library's per-call schedule choice (sched_rounds 0), the balanced
    schedule forced to one round (1) and uniform 16-plane chunks (-1); HIP
    events over ``steps`` back-to-back sub-iterations after ``warmup``."""
    import pytomography_amd
    from pytomography_amd import synthetic
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.transforms.SPECT import SPECTAttenuationTransform, SPECTPSFTransform
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix

    pytomography_amd.set_device(device)
    sizes = [(64, 60), (96, 120), (128, 60), (128, 120), (160, 60), (160, 120), (192, 60), (192, 120)]
    modes = {"auto": 0, "balanced": 1, "uniform": -1}
    rows = []
    for n, na in sizes:
        mu = torch.from_numpy(synthetic.attenuation_map(n, 0)).to(device)
        om = SPECTObjectMeta((synthetic.DX,) * 3, (n, n, n))
        pm = SPECTProjMeta((n, n), (synthetic.DX, synthetic.DX), list(synthetic.angles(na)),
                           radii=[synthetic.RADIUS] * na)
        row = {"n": n, "angles": na}
        for name, rounds in modes.items():
            print(f"[bench] schedules: {n}^3 x {na}, {name}", file=sys.stderr, flush=True)
            sm = SPECTSystemMatrix([SPECTAttenuationTransform(mu),
                                    SPECTPSFTransform(SPECTPSFMeta(synthetic.PSF_PARAMS))], [], om, pm,
                                   sched_rounds=rounds)
            sm.set_n_subsets(8)
            a_s = len(sm.subset_indices_array[0])
            f = torch.rand((n, n, n), device=device) + 0.5
            g = torch.rand((a_s, n, n), device=device) + 0.5
            norm = torch.ones_like(f)
            for _ in range(max(1, args.warmup)):
                sm.osem_step(f, g, None, norm, 0, keep_prediction=False)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                sm.osem_step(f, g, None, norm, 0, keep_prediction=False)
            e1.record()
            torch.cuda.synchronize()
            row[f"{name}_us"] = 1e3 * e0.elapsed_time(e1) / args.steps
            del sm, f, g, norm
        rows.append(row)
        del mu
        torch.cuda.empty_cache()
    best = [min(r["balanced_us"], r["uniform_us"]) for r in rows]
    out = {
        "metric": "att+PSF OSEM sub-iteration time per problem size and plane schedule",
        "value": float(np.mean([r["auto_us"] / b for r, b in zip(rows, best)])),
        "unit": "mean auto / best-of-two sub-iteration time",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (uniform random object and projections, seeded mu)",
        "config": {"workload": "one OSEM sub-iteration (subset 0 of 8, attenuation field + PSF) "
                               "per size: N in {64..192}, A in {60, 120}",
                   "sizes": rows},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["att_psf", "plain", "batch", "unrolled", "aux",
                                           "schedules"],
                    default="att_psf")
    ap.add_argument("--iters", type=int, default=10, help="iterations of the unrolled workload")
    ap.add_argument("--batch", type=int, default=32, help="problems of the batch workload")
    ap.add_argument("--n", type=int, default=None, help="object size (128; 192 for unrolled)")
    ap.add_argument("--angles", type=int, default=None, help="angles (120; 180 for unrolled)")
    ap.add_argument("--subsets", type=int, default=None, help="subsets (8; 4 for unrolled)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-leg", action="store_true",
                    help="skip the no-attenuation, no-PSF leg of the default workload")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    args = ap.parse_args()
    unrolled = args.workload == "unrolled"
    args.n = args.n or (192 if unrolled else 128)
    args.angles = args.angles or (180 if unrolled else 120)
    args.subsets = args.subsets or (4 if unrolled else 8)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for {max(args.gpus, world)}): the "
              "projectors run on a single MI355X; sharding across GPUs is not implemented.",
              file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)
    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    if args.workload == "batch":
        bench_batch(args, device)
        return
    if args.workload == "unrolled":
        bench_unrolled(args, device)
        return
    if args.workload == "aux":
        bench_aux(args, device)
        return
    if args.workload == "schedules":
        bench_schedules(args, device)
        return
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=device).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    att = psf = args.workload == "att_psf"
    n, na, ns = args.n, args.angles, args.subsets
    head = measure_osem(n, na, ns, att, psf, device, args.steps, args.warmup,
                        cpu=not args.no_cpu_baseline, tag=args.workload)
    plain = None
    if att and not args.no_plain_leg:
        # the same geometry without attenuation and PSF (the reference's
        # CPU-runnable case), timed the same way with its own CPU baseline
        plain = measure_osem(n, na, ns, False, False, device, args.steps, args.warmup,
                             cpu=not args.no_cpu_baseline, tag="plain")
    a_s = na // ns
    out = {
        "metric": METRIC,
        "value": head["its"],
        "unit": "OSEM it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / head["its"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts drawn on the host)",
        "config": {
            "workload": (f"SPECT OSEM {n}^3 x {na} angles, "
                         + ("attenuation + PSF" if att else "no attenuation, no PSF")
                         + f", {ns} subsets"),
            "n": n, "angles": na, "subsets": ns, "attenuation": att, "psf": psf,
            **head["config"],
        },
        "roofline": {
            "bound": "hbm", "kernel": head["kernel"], "achieved": head["achieved"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": head["achieved"] / HBM_PEAK_GBS,
            "traffic": head["traffic"], "algorithmic_bytes_per_launch": head["bytes_launch"],
            "launch_ms": head["launch_ms"],
            "algorithmic_bytes_unit": f"15 angles x (8 N^3 + 4 N^2) bytes, N = {n}" if att else
                                      f"{a_s} angles x (4 N^3 + 4 N^2) bytes, N = {n}",
        },
        "cpu_baseline": head["cpu"],
    }
    if plain is not None:
        out["config"]["plain_leg"] = {
            "workload": f"SPECT OSEM {n}^3 x {na} angles, no attenuation, no PSF, {ns} subsets",
            "value": plain["its"], "unit": "OSEM it/s",
            "roofline_frac": plain["achieved"] / HBM_PEAK_GBS, "kernel": plain["kernel"],
            "launch_ms": plain["launch_ms"], **plain["config"],
            "cpu_baseline": plain["cpu"],
        }
    print(json.dumps(out))


def measure_osem(n: int, na: int, ns: int, att: bool, psf: bool, device, steps: int, warmup: int,
                 cpu: bool, tag: str) -> dict:
    """One OSEM workload: throughput over ``steps`` iterations after ``warmup``,
    the per-launch device time of the forward projection and of the fused
    back projection + update (HIP events on the launch stream), the float32
    iterate against the float64 kernels, and the CPU baseline."""
    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood
    import pytomography_amd

    print(f"[bench] {tag}: {n}^3 x {na}, att {att}, psf {psf}: building the problem",
          file=sys.stderr, flush=True)
    prob = build_problem(n, na, att, psf, seed=0, device=device)
    sm = prob["sm"]
    lk = PoissonLogLikelihood(sm, prob["data"], additive_term=prob["scatter"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lk._set_n_subsets(ns)           # FP64 sensitivity images H_m^T 1, one-off
    torch.cuda.synchronize()
    sens_s = time.perf_counter() - t0
    alg = OSEM(lk)
    alg._set_n_subsets = lambda k: None  # sensitivities already cached

    def osem_iteration():
        for k in range(ns):
            alg._fused_subiteration(0, k)

    for _ in range(warmup):
        osem_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        osem_iteration()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    its = steps / elapsed
    f32_iterate = alg.object_prediction.detach().clone()
    n_done = warmup + steps
    print(f"[bench] {tag}: {its:.1f} OSEM it/s", file=sys.stderr, flush=True)

    # Per-entry-point timing with HIP events on the launch stream: each entry
    # point is launched for every subset, REPS times back to back between two
    # events, so the device never waits for the host between launches and
    # the mean is the launch's device time.  The back projections update a
    # copy of the iterate (mode 0: reference, out_mode 2: fused OSEM update).
    REPS = 4
    stream = torch.cuda.current_stream(device)
    f = alg.object_prediction
    ratios, norms = [], []
    for k in range(ns):
        g = lk._get_projection_subset(lk.projections, k)
        s = lk._get_projection_subset(lk.additive_term, k)
        _, ratio = lk.poisson_ratio(g, sm.forward(f, k), s)
        ratios.append(ratio)
        norms.append(lk._get_normBP(k))
    f_tmp = f.clone()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(REPS):
        for k in range(ns):
            sm.forward(f, k)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ev[2].record(stream)
    for _ in range(REPS):
        for k in range(ns):
            sm._backward(ratios[k], k, 0, 2, out=f_tmp, f=f_tmp, norm=norms[k])
    ev[3].record(stream)
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / (REPS * ns)
    bwd_ms = ev[2].elapsed_time(ev[3]) / (REPS * ns)
    del f_tmp, ratios
    a_s = na // ns
    fwd_b, bwd_b, upd_b, ratio_b = algorithmic_bytes(n, a_s, att)
    # Dominant kernel: the forward projection (its projector kernel + the
    # partials' reduce), timed above with HIP events on the stream it is
    # launched on; which pair runs follows the plane schedule of the call
    kern = (f"fwd_sched_kernel + reduce_sched_kernel (spect_forward, {a_s} angles)"
            if sm.schedule(0, torch.float32)["forward"] == "balanced"
            else f"fwd_kernel + fwd_reduce_kernel (spect_forward, {a_s} angles)")
    achieved = fwd_b / (fwd_ms * 1e-3) / 1e9
    iter_bytes = ns * (fwd_b + bwd_b + upd_b + ratio_b)

    # The same number of iterations with the float64 kernels (untimed): the
    # float32 iterate's error against an FP64 run of the same path.
    pytomography_amd.set_dtype(torch.float64)
    try:
        lk64 = PoissonLogLikelihood(sm, prob["data"].double(), additive_term=prob["scatter"])
        lk64._set_n_subsets(ns)
        alg64 = OSEM(lk64)
        alg64._set_n_subsets = lambda k: None
        for _ in range(n_done):
            for k in range(ns):
                alg64._fused_subiteration(0, k)
        ref64 = alg64.object_prediction
        fp64_err = float((f32_iterate.double() - ref64).norm() / ref64.norm())
        del lk64, alg64, ref64
    finally:
        pytomography_amd.set_dtype(torch.float32)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            entry = json.load(fh).get(f"{tag}_{n}_{na}_{ns}")
        if entry:
            traffic = entry.get("spect_forward")

    cpu_res = None
    if cpu:
        print(f"[bench] {tag}: CPU baseline (oracle on the host)", file=sys.stderr, flush=True)
        cpu_res = cpu_baseline(prob, ns, att, psf)
        print(f"[bench] {tag}: CPU {cpu_res['value']:.4f} OSEM it/s", file=sys.stderr, flush=True)
    # the attenuation data the projectors read: the resident attenuation
    # field (factors of every visited rotated-frame voxel of every angle,
    # built once per mu map) or the FP64 suffix table
    att_data = None
    if att:
        from pytomography_amd import _lib
        plan = sm._plan(device)
        field = sm._use_field(device, torch.float32)
        nbytes = int(plan.lib.spect_att_field_bytes(plan.handle, _lib.SPECT_F32) if field
                     else plan.lib.spect_att_table_bytes(plan.handle))
        att_data = {"mode": "attenuation field" if field else "FP64 suffix table",
                    "resident_bytes": nbytes,
                    "note": ("read as a coalesced stream instead of resampling mu per call; the "
                             "forward's measured HBM bytes include it (traffic vs algorithmic bytes "
                             "in the roofline entry)")}
    # the plane schedule the timed kernels ran (spect_schedule_info)
    schedule = sm.schedule(0, torch.float32)
    del alg, lk, prob, sm
    torch.cuda.empty_cache()
    return dict(
        its=its, kernel=kern, achieved=achieved, bytes_launch=fwd_b, launch_ms=fwd_ms,
        traffic=traffic, cpu=cpu_res,
        config={
            "projections_per_s": its * na,
            "iteration_algorithmic_bytes": iter_bytes,
            "iteration_hbm_frac": iter_bytes * its / (HBM_PEAK_GBS * 1e9),
            "sensitivity_fp64_s": sens_s,
            "iterate_rel_l2_f32_vs_f64": fp64_err,
            "iterations_compared": n_done,
            "forward_ms_per_subset": fwd_ms,
            "backward_update_ms_per_subset": bwd_ms,
            "schedule": schedule,
            **({"attenuation_data": att_data} if att_data else {}),
        })


if __name__ == "__main__":
    main()
