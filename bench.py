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

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain|att_psf]
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
    gen = torch.Generator(device=device).manual_seed(seed)
    if att:
        scatter = torch.full_like(expected, 0.05 * float(expected.mean()))
        data = torch.poisson(expected + scatter, generator=gen)
    else:
        scatter = None
        data = torch.poisson(expected, generator=gen)
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


def cpu_baseline(n: int, n_angles: int, att: bool, psf: bool, seed: int, budget_s: float):
    """Oracle (op-for-op torch-CPU restatement of the reference path) timed on
    whole OSEM sub-iterations until ``budget_s`` is spent; returns it/s."""
    from oracle import spect_oracle as so
    from pytomography_amd import synthetic

    torch.set_num_threads(max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "64")))))
    mu = torch.from_numpy(synthetic.attenuation_map(n, seed)) if att else None
    g = so.make_geometry((n, n, n), (synthetic.DX,) * 3, synthetic.angles(n_angles),
                         radii=np.full(n_angles, synthetic.RADIUS), mu=mu,
                         psf=so.PSFModel(synthetic.PSF_PARAMS) if psf else None)
    subsets = g.subsets(8)
    f = torch.ones((n, n, n))
    gen = torch.Generator().manual_seed(seed)
    data = torch.poisson(torch.full((n_angles, n, n), 50.0), generator=gen)
    norm = so.sensitivity(g, subsets[0])
    done, t0 = 0, time.perf_counter()
    while True:
        sidx = subsets[done % 8]
        pp = so.forward(g, f, sidx)
        b = so.backward(g, data[sidx] / (pp + so.DELTA), sidx)
        f = f + f * (1 / (norm + so.DELTA)) * (b - norm)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 8:
            break
    per_sub = el / done
    return dict(value=1.0 / (8 * per_sub), unit="OSEM it/s", cores=torch.get_num_threads(),
                kind="port",
                sample=f"{done} OSEM sub-iteration(s) of {len(subsets[0])} angles (forward + reference "
                       f"back projection + update) of the same {n}^3 x {n_angles} workload, "
                       f"{el:.1f} s, scaled x8 to one iteration")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["att_psf", "plain"], default="att_psf")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--angles", type=int, default=120)
    ap.add_argument("--subsets", type=int, default=8)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for {max(args.gpus, world)}): this build "
              "targets a single MI355X; the multi-GPU sharding is described in DESIGN.md.",
              file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)
    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    att = psf = args.workload == "att_psf"
    n, na, ns = args.n, args.angles, args.subsets

    from pytomography_amd.algorithms import OSEM
    from pytomography_amd.likelihoods import PoissonLogLikelihood

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=device).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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

    for _ in range(args.warmup):
        osem_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        osem_iteration()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    its = args.steps / elapsed

    # Per-entry-point timing with HIP events on the launch stream (one more iteration).
    stream = torch.cuda.current_stream(device)
    f = alg.object_prediction
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(ns)]
    for k in range(ns):
        g = lk._get_projection_subset(lk.projections, k)
        s = lk._get_projection_subset(lk.additive_term, k)
        nrm = lk._get_normBP(k)
        ev[k][0].record(stream)
        hf = sm.forward(f, k)
        ev[k][1].record(stream)
        _, ratio = lk.poisson_ratio(g, hf, s)
        sm._backward(ratio, k, 0, 2, out=f, f=f, norm=nrm)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    fwd_ms = float(np.mean([ev[k][0].elapsed_time(ev[k][1]) for k in range(ns)]))
    bwd_ms = float(np.mean([ev[k][1].elapsed_time(ev[k][2]) for k in range(ns)]))
    a_s = na // ns
    fwd_b, bwd_b, upd_b, ratio_b = algorithmic_bytes(n, a_s, att)
    # Dominant kernel: the forward projection (fwd_kernel + its chunk reduce),
    # timed above with HIP events on the stream it is launched on.
    kern, bytes_launch, ms = "fwd_kernel + fwd_reduce_kernel (spect_forward, 15 angles)", fwd_b, fwd_ms
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    iter_bytes = ns * (fwd_b + bwd_b + upd_b + ratio_b)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            entry = json.load(fh).get(f"{args.workload}_{n}_{na}_{ns}")
        if entry:
            traffic = entry.get("spect_forward")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, na, att, psf, seed=0, budget_s=args.cpu_budget)

    out = {
        "metric": METRIC,
        "value": its,
        "unit": "OSEM it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts)",
        "config": {
            "workload": (f"SPECT OSEM {n}^3 x {na} angles, "
                         + ("attenuation + PSF" if att else "no attenuation, no PSF")
                         + f", {ns} subsets"),
            "n": n, "angles": na, "subsets": ns, "attenuation": att, "psf": psf,
            "projections_per_s": its * na,
            "iteration_algorithmic_bytes": iter_bytes,
            "iteration_hbm_frac": iter_bytes * its / (HBM_PEAK_GBS * 1e9),
            "sensitivity_fp64_s": sens_s,
            "forward_ms_per_subset": fwd_ms,
            "backward_update_ms_per_subset": bwd_ms,
        },
        "roofline": {
            "bound": "hbm", "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_launch, "launch_ms": ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
