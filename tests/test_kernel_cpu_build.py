"""The projector's block bodies (csrc/forward_body.hpp, csrc/backward_body.hpp)
built for the host with g++ and a thread-per-lane stand-in for the HIP
runtime (tests/cpu_build), under AddressSanitizer and UBSan, against the
oracle.

The GPU parity tests check the compiled kernels; this checks the kernel
source itself on the CPU: index arithmetic, lane coverage of every phase,
barriers, the chunk reduce and the back projection's field/gather split.
Dynamic LDS, the forward's chunk partials and the back projection's field W
start as NaN, so a read of an entry no lane wrote -- also one taken with
weight zero -- would show in the output.  The cases include whole-tile
geometries with several phase-A passes (the no-PSF layout of a 128^3 object),
ragged z (element loads) and PSF widths."""
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


@pytest.fixture(scope="module")
def cpu_build(tmp_path_factory):
    """Paths of the forward (fwd_cpu) and back projection (bwd_cpu) builds."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out_dir = tmp_path_factory.mktemp("cpu_build")
    exes = {}
    for name in ("fwd_cpu", "bwd_cpu"):
        exe = str(out_dir / name)
        cmd = [gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               "-fno-omit-frame-pointer", "-pthread", "-Wno-attributes", "-Wno-unknown-pragmas",
               "-I", os.path.join(ROOT, "tests", "cpu_build"),
               "-I", os.path.join(ROOT, "pytomography_amd", "csrc"),
               os.path.join(ROOT, "tests", "cpu_build", name + ".cpp"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0 and "barrier" in r.stderr:
            pytest.skip("the host compiler lacks C++20 <barrier>")
        assert r.returncode == 0, r.stderr[-4000:]
        exes[name] = exe
    return exes


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


def _write_problem(d, pa, sel, chunk_planes=16, att_mode=0):
    """att_mode 0: no attenuation, 1: attenuation field, 2: FP64 suffix table."""
    n, nz, A = pa["n"], pa["nz"], len(pa["angles"])
    kr = kz = n_radii = 0
    if pa["psf"]:
        kr, kz, n_radii, ridx, tr, tz = _psf_tables(pa)
        ridx.astype(np.int32).tofile(d / "ridx.bin")
        tr.astype(np.float32).tofile(d / "taps_r.bin")
        if tz is not None:
            tz.astype(np.float32).tofile(d / "taps_z.bin")
    if att_mode:
        np.ascontiguousarray(pa["mu"], dtype=np.float32).tofile(d / "mu.bin")
    from pytomography_amd import synthetic
    (d / "meta.txt").write_text(f"{n} {nz} {A} {chunk_planes} {kr} {kz} {n_radii} {len(sel)} {att_mode} "
                                f"{synthetic.DX}\n")
    np.asarray(pa["angles"], dtype=np.float32).tofile(d / "angles.bin")
    np.asarray(sel, dtype=np.int32).tofile(d / "sel.bin")


def _exec(exe, d, extra=()):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe, str(d), *extra], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    words = r.stdout.split()
    return dict(zip(words[0::2], map(int, words[1::2])))


def _run(exe, tmp_path, pa, sel, chunk_planes=16, att_mode=0, trb=False, obj=None, sched_blocks=None):
    """Forward of pa["f"] (or ``obj``) over the angles ``sel`` by the CPU
    build; with ``trb`` the B^T mode of the same body, with ``sched_blocks``
    the balanced schedule over that many blocks."""
    _write_problem(tmp_path, pa, sel, chunk_planes, att_mode)
    np.ascontiguousarray(pa["f"] if obj is None else obj, dtype=np.float32).tofile(tmp_path / "f.bin")
    extra = ["trb"] if trb else (["sched", str(sched_blocks)] if sched_blocks else [])
    stats = _exec(exe, tmp_path, extra)
    proj = np.fromfile(tmp_path / "proj.bin", dtype=np.float32).reshape(len(sel), pa["n"], pa["nz"])
    return torch.from_numpy(proj), stats


CASES = [
    # name, problem, angles, attenuation mode (1: field, 2: FP64 suffix table)
    # whole 32-column tiles, 4 phase-A passes per lane: the no-PSF layout of 128^3
    ("plain_whole_tiles", dict(n=64, n_angles=60), [0, 7, 15, 22, 30, 37, 45, 59], 0),
    ("plain_ragged", dict(n=21, n_angles=9, nz=13), list(range(9)), 0),
    ("psf", dict(n=24, n_angles=12, psf=True), [0, 3, 5, 11], 0),
    ("psf_ragged", dict(n=21, n_angles=9, psf=True, nz=13), [1, 4, 8], 0),
    ("att_field_psf", dict(n=24, n_angles=12, att=True, psf=True), [0, 2, 7, 10], 1),
    ("att_table_psf", dict(n=24, n_angles=12, att=True, psf=True), [1, 5, 11], 2),
    ("att_field_ragged", dict(n=21, n_angles=9, att=True, psf=True, nz=13), [0, 3, 6], 1),
    ("att_table_whole_tiles", dict(n=32, n_angles=16, att=True), [0, 4, 9, 15], 2),
]


@pytest.mark.parametrize("name,spec,sel,att_mode", CASES, ids=[c[0] for c in CASES])
def test_forward_body_on_cpu_matches_oracle(cpu_build, tmp_path, name, spec, sel, att_mode):
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=5, **spec)
    proj, stats = _run(cpu_build["fwd_cpu"], tmp_path, pa, sel, att_mode=att_mode)
    ref = so.forward(oracle_geometry(pa), torch.from_numpy(pa["f"]), torch.tensor(sel))
    errs = [rel_l2(proj[k], ref[k]) for k in range(len(sel))]
    print(f"{name}: H {stats['H']}, {stats['blocks']} blocks x {stats['threads']} lanes, "
          f"max rel-L2 {max(errs):.1e}")
    assert stats["nan_partial"] == 0
    assert bool(torch.isfinite(proj).all()) and bool((proj >= 0).all())
    assert max(errs) <= 1e-5  # the per-projection bar of the GPU parity tests


def _sched_cases(counts):
    """(case, resident blocks) pairs of the balanced-schedule tests: every
    block count on the attenuation-field + PSF case, one count per other case
    (rotating through the counts), so each count runs on several layouts
    without running every pairing."""
    sched = [c for c in CASES if c[3] != 2]
    out = []
    for k, c in enumerate(sched):
        for b in (counts if c[0] == "att_field_psf" else [counts[k % len(counts)]]):
            out.append(pytest.param(*c, b, id=f"{c[0]}-{b}"))
    return out


@pytest.mark.parametrize("name,spec,sel,att_mode,blocks", _sched_cases([1, 5, 37]))
def test_balanced_schedule_on_cpu_matches_oracle(cpu_build, tmp_path, name, spec, sel, att_mode, blocks):
    """The balanced schedule (every block evaluating the schedule and
    locating its piece, blocks dealt to pieces in the XCD order, the piece
    reduce) sized for 1, 5 and 37 resident blocks: pairs split into several
    pieces, grids larger than the resident blocks.  Every slot the reduce
    reads is written (the partials start as NaN) and the projections match
    the oracle."""
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=5, **spec)
    proj, stats = _run(cpu_build["fwd_cpu"], tmp_path, pa, sel, att_mode=att_mode, sched_blocks=blocks)
    ref = so.forward(oracle_geometry(pa), torch.from_numpy(pa["f"]), torch.tensor(sel))
    errs = [rel_l2(proj[k], ref[k]) for k in range(len(sel))]
    print(f"{name}, {blocks} blocks: {stats['slots']} pieces, max rel-L2 {max(errs):.1e}")
    assert stats["nan_partial"] == 0
    assert bool(torch.isfinite(proj).all()) and bool((proj >= 0).all())
    assert max(errs) <= 1e-5


@pytest.mark.skipif(os.environ.get("SPECT_CPU_BUILD_FULL") != "1",
                    reason="about 10 minutes: set SPECT_CPU_BUILD_FULL=1")
def test_forward_body_on_cpu_plain_128(cpu_build, tmp_path):
    """The no-PSF, no-attenuation forward of the 128^3 ellipsoid phantom (seed
    0) over all 120 angles, the data-simulation call of the plain OSEM
    workload, run from the kernel source on the CPU."""
    from pytomography_amd import synthetic
    n, A = 128, 120
    pa = dict(n=n, nz=n, angles=synthetic.angles(A), f=synthetic.ellipsoid_phantom(n, 0), mu=None, psf=False,
              radii=np.full(A, synthetic.RADIUS))
    proj, stats = _run(cpu_build["fwd_cpu"], tmp_path, pa, list(range(A)))
    ref = so.forward(oracle_geometry(pa), torch.from_numpy(pa["f"]))
    errs = [rel_l2(proj[k], ref[k]) for k in range(A)]
    print(f"plain 128^3 x 120: {stats['blocks']} blocks x {stats['threads']} lanes, max rel-L2 "
          f"{max(errs):.1e}, max {float(proj.max()):.1f}")
    assert stats["nan_partial"] == 0
    assert bool(torch.isfinite(proj).all()) and bool((proj >= 0).all())
    assert max(errs) <= 1e-5


@pytest.mark.parametrize("name,spec,sel,att_mode", CASES, ids=[c[0] for c in CASES])
def test_transpose_of_back_projection_on_cpu_matches_oracle(cpu_build, tmp_path, name, spec, sel, att_mode):
    """The forward body's B^T mode (transpose of the reference back
    projection's gather) against the oracle's autograd VJP of its reference
    back projection, in float64."""
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=9, **spec)
    v = torch.rand((pa["n"], pa["n"], pa["nz"]), generator=torch.Generator().manual_seed(10), dtype=torch.float64)
    out, stats = _run(cpu_build["fwd_cpu"], tmp_path, pa, sel, att_mode=att_mode, trb=True, obj=v.numpy())
    geo = oracle_geometry(pa, torch.float64)
    y = torch.zeros((len(sel), pa["n"], pa["nz"]), dtype=torch.float64, requires_grad=True)
    (ref,) = torch.autograd.grad(so.backward(geo, y, torch.tensor(sel)), y, v)
    errs = [rel_l2(out[k], ref[k]) for k in range(len(sel))]
    print(f"{name}: B^T, H {stats['H']}, max rel-L2 {max(errs):.1e}")
    assert stats["nan_partial"] == 0
    assert max(errs) <= 1e-5


@pytest.mark.parametrize("name,spec,sel,att_mode", CASES, ids=[c[0] for c in CASES])
def test_backward_bodies_on_cpu_match_oracle(cpu_build, tmp_path, name, spec, sel, att_mode):
    """Field + gather (reference back projection, mode 0) and field + exact
    adjoint gather (mode 1) of seeded projections, against the oracle's
    reference backward and its autograd transpose of the forward."""
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=6, **spec)
    _write_problem(tmp_path, pa, sel, att_mode=att_mode)
    y = torch.rand((len(sel), pa["n"], pa["nz"]), generator=torch.Generator().manual_seed(7))
    y.numpy().astype(np.float32).tofile(tmp_path / "p.bin")
    stats = _exec(cpu_build["bwd_cpu"], tmp_path)
    shape = (pa["n"], pa["n"], pa["nz"])
    ref_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_ref.bin", dtype=np.float32).reshape(shape))
    adj_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_adj.bin", dtype=np.float32).reshape(shape))
    g = oracle_geometry(pa)
    ref = so.backward(g, y, torch.tensor(sel))
    adj = so.exact_adjoint(g, y, torch.tensor(sel))
    e_ref, e_adj = rel_l2(ref_cpu, ref), rel_l2(adj_cpu, adj)
    print(f"{name}: H {stats['H']}, W written {stats['w_written']} of {stats['w_total']}, "
          f"reference rel-L2 {e_ref:.1e}, adjoint rel-L2 {e_adj:.1e}")
    assert bool(torch.isfinite(ref_cpu).all()) and bool(torch.isfinite(adj_cpu).all())
    assert e_ref <= 1e-5 and e_adj <= 1e-5
    if (tmp_path / "bp_direct.bin").exists():
        # no attenuation, no PSF: the gather of the projection itself
        direct = torch.from_numpy(np.fromfile(tmp_path / "bp_direct.bin", dtype=np.float32).reshape(shape))
        e_dir = rel_l2(direct, ref)
        print(f"{name}: direct reference back projection rel-L2 {e_dir:.1e}")
        assert bool(torch.isfinite(direct).all()) and e_dir <= 1e-5


@pytest.mark.parametrize("name,spec,sel,att_mode,blocks", _sched_cases([1, 6, 41]))
def test_balanced_field_schedule_on_cpu_matches_oracle(cpu_build, tmp_path, name, spec, sel, att_mode,
                                                       blocks):
    """The back-projection field from the balanced schedule (the forward's
    plane ranges split into pieces for 1, 6 and 41 blocks), then both
    gathers: every W entry a gather reads is written (W starts as NaN) and
    the reference back projection and exact adjoint match the oracle."""
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=6, **spec)
    _write_problem(tmp_path, pa, sel, att_mode=att_mode)
    y = torch.rand((len(sel), pa["n"], pa["nz"]), generator=torch.Generator().manual_seed(7))
    y.numpy().astype(np.float32).tofile(tmp_path / "p.bin")
    stats = _exec(cpu_build["bwd_cpu"], tmp_path, ["sched", str(blocks)])
    shape = (pa["n"], pa["n"], pa["nz"])
    ref_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_ref.bin", dtype=np.float32).reshape(shape))
    adj_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_adj.bin", dtype=np.float32).reshape(shape))
    g = oracle_geometry(pa)
    e_ref = rel_l2(ref_cpu, so.backward(g, y, torch.tensor(sel)))
    e_adj = rel_l2(adj_cpu, so.exact_adjoint(g, y, torch.tensor(sel)))
    print(f"{name}, {blocks} blocks: {stats['pieces']} pieces, reference rel-L2 {e_ref:.1e}, "
          f"adjoint rel-L2 {e_adj:.1e}")
    assert bool(torch.isfinite(ref_cpu).all()) and bool(torch.isfinite(adj_cpu).all())
    assert e_ref <= 1e-5 and e_adj <= 1e-5


@pytest.mark.skipif(os.environ.get("SPECT_CPU_BUILD_FULL") != "1",
                    reason="about 20 minutes: set SPECT_CPU_BUILD_FULL=1")
def test_backward_bodies_on_cpu_plain_128(cpu_build, tmp_path):
    """Reference back projection and exact adjoint of seeded 120 x 128 x 128
    projections (no PSF, no attenuation) from the kernel source on the CPU."""
    from pytomography_amd import synthetic
    n, A = 128, 120
    pa = dict(n=n, nz=n, angles=synthetic.angles(A), f=None, mu=None, psf=False,
              radii=np.full(A, synthetic.RADIUS))
    sel = list(range(A))
    _write_problem(tmp_path, pa, sel)
    y = torch.rand((A, n, n), generator=torch.Generator().manual_seed(8))
    y.numpy().astype(np.float32).tofile(tmp_path / "p.bin")
    stats = _exec(cpu_build["bwd_cpu"], tmp_path)
    ref_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_ref.bin", dtype=np.float32).reshape(n, n, n))
    adj_cpu = torch.from_numpy(np.fromfile(tmp_path / "bp_adj.bin", dtype=np.float32).reshape(n, n, n))
    g = oracle_geometry(pa)
    e_ref = rel_l2(ref_cpu, so.backward(g, y))
    e_adj = rel_l2(adj_cpu, so.exact_adjoint(g, y))
    print(f"plain 128^3 x 120 back projection: W written {stats['w_written']} of {stats['w_total']}, "
          f"reference rel-L2 {e_ref:.1e}, adjoint rel-L2 {e_adj:.1e}")
    assert bool(torch.isfinite(ref_cpu).all()) and bool(torch.isfinite(adj_cpu).all())
    assert e_ref <= 1e-5 and e_adj <= 1e-5


GRAD_CASES = [c for c in CASES if c[3]]


@pytest.mark.parametrize("name,spec,sel,att_mode", GRAD_CASES, ids=[c[0] for c in GRAD_CASES])
def test_projector_vjp_on_cpu_matches_oracle_autograd(cpu_build, tmp_path, name, spec, sel, att_mode):
    """The differentiable projector's VJP (field with the mu-gradient parts,
    prefix of chunk totals, gather mode 2) against torch autograd through the
    oracle's forward, in float64: gradients w.r.t. the activity and mu."""
    spec = dict(spec)
    pa = problem_arrays(spec.pop("n"), spec.pop("n_angles"), seed=13, **spec)
    _write_problem(tmp_path, pa, sel, att_mode=att_mode)
    gen = torch.Generator().manual_seed(14)
    u = torch.rand((len(sel), pa["n"], pa["nz"]), generator=gen, dtype=torch.float64)
    f = torch.from_numpy(pa["f"]).double() + 0.1 * torch.rand(pa["f"].shape, generator=gen, dtype=torch.float64)
    u.numpy().astype(np.float32).tofile(tmp_path / "p.bin")
    f.numpy().astype(np.float32).tofile(tmp_path / "f.bin")
    _exec(cpu_build["bwd_cpu"], tmp_path, ["grad"])
    shape = (pa["n"], pa["n"], pa["nz"])
    gf = torch.from_numpy(np.fromfile(tmp_path / "grad_f.bin", dtype=np.float32).reshape(shape))
    gm = torch.from_numpy(np.fromfile(tmp_path / "grad_mu.bin", dtype=np.float32).reshape(shape))
    geo = oracle_geometry(pa, torch.float64)
    fr = f.clone().requires_grad_(True)
    mr = torch.from_numpy(pa["mu"]).double().requires_grad_(True)
    y = so.forward_autograd(geo, fr, mr, torch.tensor(sel))
    rf, rm = torch.autograd.grad(y, (fr, mr), u)
    ef, em = rel_l2(gf, rf), rel_l2(gm, rm)
    print(f"{name}: VJP rel-L2 d/df {ef:.1e}, d/dmu {em:.1e}")
    assert bool(torch.isfinite(gf).all()) and bool(torch.isfinite(gm).all())
    assert ef <= 1e-5 and em <= 1e-5
