"""The forward projector's block bodies (csrc/forward_body.hpp) built for the
host with g++ and a thread-per-lane stand-in for the HIP runtime
(tests/cpu_build), under AddressSanitizer and UBSan, against the oracle.

The GPU parity tests check the compiled kernels; this checks the kernel
source itself on the CPU: index arithmetic, lane coverage of every phase,
barriers and the chunk reduce, with dynamic LDS and the chunk partials
pre-filled with NaN so that a read of an entry no lane wrote would show.
The cases include whole-tile geometries with several phase-A passes (the
no-PSF layout of a 128^3 object), ragged z (element loads) and PSF widths."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

from oracle import spect_oracle as so
from tests.helpers import oracle_geometry, problem_arrays, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_build", "fwd_cpu.cpp")


@pytest.fixture(scope="module")
def fwd_cpu(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("cpu_build") / "fwd_cpu")
    cmd = [gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-pthread", "-Wno-attributes", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "tests", "cpu_build"), "-I", os.path.join(ROOT, "pytomography_amd", "csrc"),
           SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "barrier" in r.stderr:
        pytest.skip("the host compiler lacks C++20 <barrier>")
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def _psf_tables(pa):
    """Taps exactly as the product hands them to spect_plan_create."""
    import pytomography_amd
    from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta
    from pytomography_amd.projectors.SPECT import SPECTSystemMatrix
    from pytomography_amd import synthetic
    from pytomography_amd.transforms.SPECT import SPECTPSFTransform
    old = pytomography_amd.device
    pytomography_amd.set_device("cpu")
    try:
        om = SPECTObjectMeta((synthetic.DX,) * 3, (pa["n"], pa["n"], pa["nz"]))
        pm = SPECTProjMeta((pa["n"], pa["nz"]), (synthetic.DX, synthetic.DX), list(pa["angles"]),
                           radii=list(pa["radii"]))
        sm = SPECTSystemMatrix([SPECTPSFTransform(SPECTPSFMeta(synthetic.PSF_PARAMS))], [], om, pm)
        args = sm._geom_args()
    finally:
        pytomography_amd.set_device(old)
    keep = args["_keep"]
    ridx, tr32 = keep[2], keep[3]
    tz32 = keep[5] if args["psf_kz"] > 0 else None
    return args["psf_kr"], args["psf_kz"], args["n_radii"], ridx, tr32, tz32


def _run(fwd_cpu, tmp_path, pa, sel, chunk_planes=16):
    d = tmp_path
    n, nz, A = pa["n"], pa["nz"], len(pa["angles"])
    kr = kz = n_radii = 0
    if pa["psf"]:
        kr, kz, n_radii, ridx, tr, tz = _psf_tables(pa)
        ridx.astype(np.int32).tofile(d / "ridx.bin")
        tr.astype(np.float32).tofile(d / "taps_r.bin")
        if tz is not None:
            tz.astype(np.float32).tofile(d / "taps_z.bin")
    (d / "meta.txt").write_text(f"{n} {nz} {A} {chunk_planes} {kr} {kz} {n_radii} {len(sel)}\n")
    np.ascontiguousarray(pa["f"], dtype=np.float32).tofile(d / "f.bin")
    np.asarray(pa["angles"], dtype=np.float32).tofile(d / "angles.bin")
    np.asarray(sel, dtype=np.int32).tofile(d / "sel.bin")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([fwd_cpu, str(d)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    stats = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
    proj = torch.from_numpy(np.fromfile(d / "proj.bin", dtype=np.float32).reshape(len(sel), n, nz))
    return proj, stats


CASES = [
    # whole 32-column tiles, 4 phase-A passes per lane: the no-PSF layout of 128^3
    ("plain_whole_tiles", dict(n=64, n_angles=60), [0, 7, 15, 22, 30, 37, 45, 59]),
    ("plain_ragged", dict(n=21, n_angles=9, nz=13), list(range(9))),
    ("psf", dict(n=24, n_angles=12, psf=True), [0, 3, 5, 11]),
    ("psf_ragged", dict(n=21, n_angles=9, psf=True, nz=13), [1, 4, 8]),
]


@pytest.mark.parametrize("name,spec,sel", CASES, ids=[c[0] for c in CASES])
def test_forward_body_on_cpu_matches_oracle(fwd_cpu, tmp_path, name, spec, sel):
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=5, **spec)
    proj, stats = _run(fwd_cpu, tmp_path, pa, sel)
    ref = so.forward(oracle_geometry(pa), torch.from_numpy(pa["f"]), torch.tensor(sel))
    errs = [rel_l2(proj[k], ref[k]) for k in range(len(sel))]
    print(f"{name}: H {stats['H']}, {stats['blocks']} blocks x {stats['threads']} lanes, "
          f"max rel-L2 {max(errs):.1e}")
    assert stats["nan_partial"] == 0
    assert bool(torch.isfinite(proj).all()) and bool((proj >= 0).all())
    assert max(errs) <= 1e-5  # the per-projection bar of the GPU parity tests


@pytest.mark.skipif(os.environ.get("SPECT_CPU_BUILD_FULL") != "1",
                    reason="about 10 minutes: set SPECT_CPU_BUILD_FULL=1")
def test_forward_body_on_cpu_plain_128(fwd_cpu, tmp_path):
    """The no-PSF, no-attenuation forward of the 128^3 ellipsoid phantom (seed
    0) over all 120 angles, the data-simulation call of the plain OSEM
    workload, run from the kernel source on the CPU."""
    from pytomography_amd import synthetic
    n, A = 128, 120
    pa = dict(n=n, nz=n, angles=synthetic.angles(A), f=synthetic.ellipsoid_phantom(n, 0), mu=None, psf=False,
              radii=np.full(A, synthetic.RADIUS))
    proj, stats = _run(fwd_cpu, tmp_path, pa, list(range(A)))
    ref = so.forward(oracle_geometry(pa), torch.from_numpy(pa["f"]))
    errs = [rel_l2(proj[k], ref[k]) for k in range(A)]
    print(f"plain 128^3 x 120: {stats['blocks']} blocks x {stats['threads']} lanes, max rel-L2 "
          f"{max(errs):.1e}, max {float(proj.max()):.1f}")
    assert stats["nan_partial"] == 0
    assert bool(torch.isfinite(proj).all()) and bool((proj >= 0).all())
    assert max(errs) <= 1e-5
