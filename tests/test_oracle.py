"""Pins of the CPU oracle (no GPU).

The reference's own test suite pins no number on this path (its single test
needs external DICOM data), so the oracle is pinned by closed forms, by the
numpy float64 brute force in oracle/brute.py (written from the mathematics,
not from the torch op sequence) and by invariants chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails at least one test.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import brute
from oracle import spect_oracle as so

pytestmark = pytest.mark.filterwarnings("ignore::UserWarning")


# ----------------------------------------------------------------- padding
def test_pad_size_closed_form():
    for n in (1, 2, 16, 64, 128, 192, 1000):
        assert so.compute_pad_size(n) == math.ceil((math.sqrt(2) * n - n) / 2)
    # Appendix-B values of the benchmark geometries
    assert [so.compute_pad_size(n) for n in (64, 96, 128, 160, 192)] == [14, 20, 27, 34, 40]


def test_pad_size_roundtrip():
    for n in range(2, 4001):
        p = so.compute_pad_size(n)
        assert so.compute_pad_size_padded(n + 2 * p) == p


def test_back_project_box():
    box = so.pad_object(torch.ones((6, 6, 3)), mode="back_project")
    p = so.compute_pad_size(6)
    assert box.shape == (6 + 2 * p, 6 + 2 * p, 3)
    assert torch.all(box[:, p:p + 6] == 1) and torch.all(box[:, :p] == 0) and torch.all(box[:, p + 6:] == 0)


# ----------------------------------------------------------------- rotation
def _rand(shape, seed=0, dtype=torch.float64):
    return torch.rand(shape, generator=torch.Generator().manual_seed(seed), dtype=dtype)


def test_rotation_zero_is_identity():
    x = _rand((9, 9, 4))
    out = so.rotate_xyz(x, torch.tensor(0.0, dtype=torch.float64))
    assert torch.allclose(out, x, atol=1e-12)


@pytest.mark.parametrize("theta,k", [(90.0, 1), (180.0, 2), (270.0, 3)])
def test_rotation_quarter_turns_are_permutations(theta, k):
    x = _rand((9, 9, 3), seed=1)
    out = so.rotate_xyz(x, torch.tensor(theta, dtype=torch.float64))
    # rows = x, columns = y; kornia's positive angle turns the image counter-clockwise
    ref = torch.rot90(x, k=k, dims=(0, 1))
    assert torch.allclose(out, ref, atol=1e-10)


def test_rotation_matches_brute_force_nonsquare():
    x = _rand((5, 7, 3), seed=2)
    for theta in (13.0, -37.0, 123.0):
        out = so.rotate_xyz(x, torch.tensor(theta, dtype=torch.float64)).numpy()
        ref = brute.rotate_volume(x.numpy(), theta)
        assert np.abs(out - ref).max() < 1e-12


def test_rotation_reproduces_linear_ramps_and_constants():
    n = 21
    i = torch.arange(n, dtype=torch.float64)
    ramp = (0.3 * i[:, None] - 0.7 * i[None, :] + 2.0).unsqueeze(-1).repeat(1, 1, 2)
    theta = 29.0
    out = so.rotate_xyz(ramp, torch.tensor(theta, dtype=torch.float64))
    c, s = math.cos(math.radians(theta)), math.sin(math.radians(theta))
    cc = (n - 1) / 2
    r = i[:, None].expand(n, n)
    q = i[None, :].expand(n, n)
    col = c * (q - cc) - s * (r - cc) + cc
    row = s * (q - cc) + c * (r - cc) + cc
    expect = 0.3 * row - 0.7 * col + 2.0
    inside = (row >= 0) & (row <= n - 1) & (col >= 0) & (col <= n - 1)
    assert torch.allclose(out[..., 0][inside], expect[inside], atol=1e-10)
    const = torch.ones((n, n, 2), dtype=torch.float64)
    outc = so.rotate_xyz(const, torch.tensor(theta, dtype=torch.float64))[..., 0]
    disk = (r - cc) ** 2 + (q - cc) ** 2 <= (cc - 1) ** 2
    assert torch.allclose(outc[disk], torch.ones_like(outc[disk]), atol=1e-12)


def test_rotation_float32_matches_float64_closely():
    x = _rand((30, 30, 4), seed=3)
    a = so.rotate_xyz(x.float(), torch.tensor(-57.0))
    b = so.rotate_xyz(x, torch.tensor(-57.0, dtype=torch.float64))
    assert float((a.double() - b).abs().max()) < 1e-5


# -------------------------------------------------------------- attenuation
def test_cumsum_is_float64_accumulated():
    x = torch.rand((182, 7, 5), generator=torch.Generator().manual_seed(4)) * 0.07
    c32 = torch.cumsum(x, 0)
    c64 = torch.cumsum(x.double(), 0).float()
    assert torch.equal(c32, c64)


def test_uniform_mu_attenuation_closed_form():
    p, mu, dx = 17, 0.15, 0.48
    m = torch.full((p, 3, 2), mu, dtype=torch.float64)
    a = so.prob_of_detection(m, dx)
    i = torch.arange(p, dtype=torch.float64)
    expect = torch.exp(-mu * dx * (p - i - 0.5))
    assert torch.allclose(a[:, 1, 1], expect, rtol=1e-14)


def test_attenuation_matches_brute_force():
    m = _rand((11, 4, 3), seed=5) * 0.3
    a = so.prob_of_detection(m, 0.48).numpy()
    assert np.abs(a - brute.attenuation(m.numpy(), 0.48)).max() < 1e-14


# ----------------------------------------------------------------------- PSF
def test_psf_kernel_sizes_appendix_b():
    psf = so.PSFModel((0.03, 0.1))
    ks = [so.psf_kernel_size(psf, n, 25.0, 0.48, 0.48) for n in (64, 96, 128, 160, 192)]
    assert ks == [21, 25, 29, 33, 37]


def test_psf_sigma_and_taps_closed_form():
    psf = so.PSFModel((0.03, 0.1))
    n, r, dx = 128, 25.0, 0.48
    sig = so.psf_sigma(psf, n, r, dx)
    p = n + 2 * so.compute_pad_size(n)
    i = np.arange(p)
    d = np.maximum(r + (p / 2 - i - 0.5) * dx, 0)
    assert np.allclose(sig, 0.03 * d + 0.1, rtol=0, atol=1e-15)
    taps = so.gaussian_taps(sig / dx, 29, torch.float64).numpy()
    t = np.arange(-14, 15)
    s = sig[:, None] / dx
    ref = np.exp(-t ** 2 / (2 * s ** 2 + 1e-11))
    ref /= ref.sum(1, keepdims=True)
    assert np.allclose(taps, ref, rtol=1e-13, atol=0)
    assert np.allclose(taps, taps[:, ::-1])


def test_psf_blur_matches_scipy_and_is_self_adjoint():
    from scipy.ndimage import correlate1d
    rng = np.random.default_rng(6)
    x = rng.random((5, 9, 7))
    taps_r = torch.from_numpy(rng.random((5, 5)))
    taps_z = torch.from_numpy(rng.random((5, 3)))
    taps_r = (taps_r + taps_r.flip(1)) / 2
    taps_z = (taps_z + taps_z.flip(1)) / 2
    out = so.psf_apply(torch.from_numpy(x), taps_r, taps_z).numpy()
    ref = np.stack([correlate1d(correlate1d(x[i], taps_r[i].numpy(), axis=0, mode="constant"),
                                taps_z[i].numpy(), axis=1, mode="constant") for i in range(5)])
    assert np.abs(out - ref).max() < 1e-13
    assert np.abs(out - brute.psf(x, taps_r.numpy(), taps_z.numpy())).max() < 1e-13
    y = rng.random(x.shape)
    lhs = (out * y).sum()
    rhs = (x * so.psf_apply(torch.from_numpy(y), taps_r, taps_z).numpy()).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_psf_delta_gives_kernel():
    taps = so.gaussian_taps(np.full(3, 1.7), 9, torch.float64)
    x = torch.zeros((3, 11, 11), dtype=torch.float64)
    x[:, 5, 5] = 1
    out = so.psf_apply(x, taps, taps)
    assert torch.allclose(out[1, 1:10, 1:10], torch.outer(taps[1], taps[1]), atol=1e-15)


# --------------------------------------------------------- system matrix
def _tiny(att=True, psf=True):
    rng = np.random.default_rng(7)
    shape = (8, 8, 4)
    f = rng.random(shape)
    mu = rng.random(shape) * 0.2 if att else None
    angles = [0.0, 33.0, 90.0, 141.0, 222.0, 300.0]
    g = so.make_geometry(shape, (0.48,) * 3, angles, radii=[6.0] * 6,
                         mu=None if mu is None else torch.from_numpy(mu),
                         psf=so.PSFModel((0.03, 0.1)) if psf else None, dtype=torch.float64)
    taps = None
    if psf:
        taps = [tuple(None if t is None else t.numpy() for t in g.layers(a)) for a in range(6)]
    return g, f, mu, angles, taps


@pytest.mark.parametrize("att,psf", [(False, False), (True, False), (False, True), (True, True)])
def test_forward_backward_match_brute_force(att, psf):
    g, f, mu, angles, taps = _tiny(att, psf)
    out = so.forward(g, torch.from_numpy(f)).numpy()
    ref = brute.forward(f, angles, 0.48, mu, taps)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()
    y = np.random.default_rng(8).random(out.shape)
    bo = so.backward(g, torch.from_numpy(y)).numpy()
    bb = brute.backward_ref(y, angles, 0.48, g.shape, mu, taps)
    assert np.abs(bo - bb).max() < 1e-13 * np.abs(bb).max()


def test_exact_adjoint_is_dense_transpose_and_reference_backward_is_not():
    g, f, mu, angles, taps = _tiny(True, True)
    H = brute.dense_matrix(lambda e: brute.forward(e, angles, 0.48, mu, taps), g.shape)
    out = so.forward(g, torch.from_numpy(f)).numpy().ravel()
    assert np.abs(out - H @ f.ravel()).max() < 1e-12
    y = np.random.default_rng(9).random((6, 8, 4))
    adj = so.exact_adjoint(g, torch.from_numpy(y)).numpy().ravel()
    assert np.abs(adj - H.T @ y.ravel()).max() < 1e-12
    ref_bp = so.backward(g, torch.from_numpy(y)).numpy().ravel()
    assert np.abs(ref_bp - H.T @ y.ravel()).max() > 1e-4  # the reference back projection is not H^T


def test_osem_one_subset_is_mlem_on_dense_matrix():
    g, f, mu, angles, taps = _tiny(False, False)
    H = brute.dense_matrix(lambda e: brute.forward(e, angles, 0.48, None, None), g.shape)
    B = brute.dense_matrix(lambda p: brute.backward_ref(p, angles, 0.48, g.shape, None, None),
                           (6, 8, 4))
    data = H @ (f.ravel() * 3)
    x = np.ones(f.size)
    sens = B @ np.ones(data.size)
    for _ in range(3):
        x = x + x / (sens + 1e-11) * (B @ (data / (H @ x + 1e-11)) - sens)
        x[x <= 0] = 0
    out = so.osem(g, torch.from_numpy(data.reshape(6, 8, 4)), 3, 1).numpy().ravel()
    assert np.abs(out - x).max() < 1e-10 * np.abs(x).max()


# ----------------------------------------------------------------- prior
def test_rdp_constant_image_has_zero_gradient():
    f = torch.full((6, 6, 6), 2.5, dtype=torch.float64)
    assert torch.all(so.rdp_gradient(f, 0.1, 2.0, (0.48,) * 3, 1.0) == 0)


def test_rdp_gradient_is_half_fd_of_value_in_interior():
    f = _rand((7, 7, 7), seed=10) + 0.5
    beta, gamma = 0.3, 2.0
    grad = so.rdp_gradient(f, beta, gamma, (0.48,) * 3, 1.0)
    h = 1e-6
    for idx in [(3, 3, 3), (2, 4, 3), (4, 2, 2)]:
        fp, fm = f.clone(), f.clone()
        fp[idx] += h
        fm[idx] -= h
        fd = (so.rdp_value(fp, beta, gamma, (0.48,) * 3) - so.rdp_value(fm, beta, gamma, (0.48,) * 3)) / (2 * h)
        assert abs(float(fd) - 2 * float(grad[idx])) < 1e-6 * abs(float(fd)) + 1e-9


def test_rdp_gradient_jacobian_is_symmetric_fd():
    """The VJP of the prior gradient (the RDP Hessian applied to u) from
    autograd agrees with central differences of the gradient."""
    f = _rand((6, 5, 7), seed=11) + 0.2
    u = _rand((6, 5, 7), seed=12)
    d = _rand((6, 5, 7), seed=13)
    fr = f.clone().requires_grad_(True)
    r = so.rdp_gradient(fr, 0.2, 2.0, (0.48, 0.5, 0.52), 0.5)
    (jtu,) = torch.autograd.grad(r, fr, u)
    h = 1e-6
    jd = (so.rdp_gradient(f + h * d, 0.2, 2.0, (0.48, 0.5, 0.52), 0.5)
          - so.rdp_gradient(f - h * d, 0.2, 2.0, (0.48, 0.5, 0.52), 0.5)) / (2 * h)
    # <u, J d> = <J^T u, d>
    assert abs(float((u * jd).sum()) - float((jtu * d).sum())) < 1e-7 * abs(float((u * jd).sum()))


# ------------------------------------------------------- unrolled BSREM
def _bsrem_problem():
    from tests.helpers import oracle_geometry, problem_arrays
    pa = problem_arrays(10, 6, seed=3, att=True, psf=True)
    g = oracle_geometry(pa, torch.float64)
    f_true = torch.from_numpy(pa["f"]).double()
    expected = so.forward(g, f_true) * 50
    s = torch.full_like(expected, 0.05 * float(expected.mean()))
    data = torch.poisson(expected + s, generator=torch.Generator().manual_seed(4))
    return g, f_true, data, s


def _dense_plain(angles, shape=(8, 8, 4)):
    """Dense forward H and reference back projection B (numpy brute force)."""
    H = brute.dense_matrix(lambda e: brute.forward(e, angles, 0.48), shape)
    B = brute.dense_matrix(lambda p: brute.backward_ref(p, angles, 0.48, shape),
                           (len(angles), shape[1], shape[2]))
    return H, B


def _em_counts(H, shape, seed):
    rng = np.random.default_rng(seed)
    f_true = rng.random(shape) + 0.2
    expected = H @ f_true.ravel() * 50
    s = np.full(expected.shape, 0.05 * expected.mean())
    data = rng.poisson(expected + s).astype(np.float64)
    x0 = np.ones(int(np.prod(shape)))
    x0[::7] = 1e-6   # below BSREM's floor: the first update must lift them to 1e-4
    return data, s, x0


def test_bsrem_one_subset_without_prior_is_floored_mlem():
    """BSREM (preconditioned_gradient_ascent.py:436-474) with one subset and
    beta = 0: w = 1 and sum_m n_m = n, so D = 1/(n + delta) and each step is
    the MLEM step (written here on the dense brute-force matrices), followed
    by the 1e-4 floor (:452, :122)."""
    shape = (8, 8, 4)
    angles = [0.0, 33.0, 90.0, 141.0, 222.0, 300.0]
    H, B = _dense_plain(angles, shape)
    data, s, x0 = _em_counts(H, shape, seed=31)
    x = x0.copy()
    sens = B @ np.ones(data.size)
    floored = 0
    for _ in range(4):
        x = x + x / (sens + 1e-11) * (B @ (data / (H @ x + s + 1e-11)) - sens)
        floored += int((x <= 1e-4).sum())
        x[x <= 1e-4] = 1e-4
    g = so.make_geometry(shape, (0.48,) * 3, angles, radii=[6.0] * 6, dtype=torch.float64)
    out = so.bsrem(g, torch.from_numpy(data.reshape(6, 8, 4)), 4, 1, 0.0, 2.0,
                   additive=torch.from_numpy(s.reshape(6, 8, 4)),
                   f0=torch.from_numpy(x0.reshape(shape))).numpy().ravel()
    assert floored > 0   # the floor acted
    assert np.abs(out - x).max() < 1e-10 * np.abs(x).max()


def test_bsrem_subset_weight_scales_the_sensitivity_sum():
    """Two subsets {0, 180} and {90, 270} degrees without attenuation or PSF
    have equal sensitivities n_1 = n_2 (quarter turns), so BSREM's
    w_k * (n_1 + n_2) with w_k = |subset| / A = 1/2 is n_k and BSREM without
    a prior steps exactly as OSEM (floored at 1e-4): per subset
    x <- x + x/(n_k + delta) * (B_k(g_k / (H_k x + s_k)) - n_k), on the dense
    brute-force matrices.  A missing or inverted w_k changes every step."""
    shape = (8, 8, 4)
    angles = [0.0, 90.0, 180.0, 270.0]
    H, B = _dense_plain(angles, shape)
    data, s, x0 = _em_counts(H, shape, seed=32)
    rows = np.arange(data.size).reshape(4, -1)
    subsets = [rows[[0, 2]].ravel(), rows[[1, 3]].ravel()]
    sens = [B[:, k] @ np.ones(len(k)) for k in subsets]
    assert np.abs(sens[0] - sens[1]).max() < 1e-12
    x = x0.copy()
    for _ in range(3):
        for k, (sk, nk) in enumerate(zip(subsets, sens)):
            x = x + x / (nk + 1e-11) * (B[:, sk] @ (data[sk] / (H[sk] @ x + s[sk] + 1e-11)) - nk)
            x[x <= 1e-4] = 1e-4
    g = so.make_geometry(shape, (0.48,) * 3, angles, radii=[6.0] * 4, dtype=torch.float64)
    out = so.bsrem(g, torch.from_numpy(data.reshape(4, 8, 4)), 3, 2, 0.0, 2.0,
                   additive=torch.from_numpy(s.reshape(4, 8, 4)),
                   f0=torch.from_numpy(x0.reshape(shape))).numpy().ravel()
    assert np.abs(out - x).max() < 1e-10 * np.abs(x).max()


def test_bsrem_converges_to_the_penalised_likelihood_kkt_point():
    """With a diminishing relaxation BSREM converges to the maximiser of
    L(f) - V(f) over f >= floor (Ahn & Fessler 2003): with three subsets of
    the six angles and the RDP (beta 0.5), the limit satisfies
    f * (sum_k [B_k(g_k/(H_k f + s_k)) - B_k 1] - grad V(f)) = 0, i.e. the
    subset gradients add up to the full one and the prior, scaled by w_k per
    subset (preconditioned_gradient_ascent.py:115), adds up to beta * phi1
    once.  B and H are the dense brute-force matrices; grad V is the RDP
    gradient at beta_scale 1 (pinned separately by finite differences).  The
    residual shrinks like the relaxation; a prior counted once per subset
    instead leaves a residual of the size of the prior term itself."""
    shape = (8, 8, 4)
    angles = [0.0, 33.0, 90.0, 141.0, 222.0, 300.0]
    H, B = _dense_plain(angles, shape)
    rng = np.random.default_rng(33)
    f_true = rng.random(shape) + 0.2
    expected = H @ f_true.ravel() * 50
    s = np.full(expected.shape, 0.05 * expected.mean())
    data = rng.poisson(expected + s).astype(np.float64)
    beta, gamma = 0.5, 2.0
    g = so.make_geometry(shape, (0.48,) * 3, angles, radii=[6.0] * 6, dtype=torch.float64)
    f = so.bsrem(g, torch.from_numpy(data.reshape(6, 8, 4)), 150, 3, beta, gamma,
                 additive=torch.from_numpy(s.reshape(6, 8, 4)),
                 relaxation=lambda j: 1.0 / (1 + j / 4))
    x = f.numpy().ravel()
    assert x.min() > 1e-3     # interior point: the floor is not active
    grad_l = B @ (data / (H @ x + s + 1e-11)) - B @ np.ones(data.size)
    grad_v = so.rdp_gradient(f.clone(), beta, gamma, g.dr, 1.0).numpy().ravel()
    scale = np.linalg.norm(x * grad_l)
    kkt = np.linalg.norm(x * (grad_l - grad_v)) / scale
    wrong = np.linalg.norm(x * (grad_l - 3 * grad_v)) / scale
    assert np.linalg.norm(x * grad_v) > 0.5 * scale   # the prior term is not negligible
    assert kkt < 0.04, kkt
    assert wrong > 1.0, wrong


def test_unrolled_bsrem_values_equal_inplace_bsrem():
    g, f_true, data, s = _bsrem_problem()
    a = so.bsrem(g, data, 2, 3, 0.1, 2.0, additive=s)
    b = so.bsrem_unrolled(g, data, 2, 3, 0.1, 2.0, additive=s)
    assert torch.equal(a, b)


def test_unrolled_bsrem_gradient_matches_finite_differences():
    g, f_true, data, s = _bsrem_problem()
    x0 = torch.ones(g.shape, dtype=torch.float64) + 0.1 * _rand(g.shape, seed=5)

    def loss(dat, f0, beta):
        x = so.bsrem_unrolled(g, dat, 2, 3, beta, 2.0, additive=s, f0=f0)
        return ((x - f_true) ** 2).sum()

    dv, fv = data.clone().requires_grad_(True), x0.clone().requires_grad_(True)
    bv = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    loss(dv, fv, bv).backward()
    h = 1e-5
    dd = _rand(data.shape, seed=6)
    df = _rand(g.shape, seed=7)
    with torch.no_grad():
        fd_d = (loss(data + h * dd, x0, 0.1) - loss(data - h * dd, x0, 0.1)) / (2 * h)
        fd_f = (loss(data, x0 + h * df, 0.1) - loss(data, x0 - h * df, 0.1)) / (2 * h)
        fd_b = (loss(data, x0, 0.1 + h) - loss(data, x0, 0.1 - h)) / (2 * h)
    for fd, an in ((fd_d, (dv.grad * dd).sum()), (fd_f, (fv.grad * df).sum()), (fd_b, bv.grad)):
        assert abs(float(fd) - float(an)) <= 1e-5 * abs(float(fd)), (float(fd), float(an))


# --------------------------------------------- prior derivative of order 2
@pytest.mark.parametrize("gamma", [0.0, 2.0])
def test_rdp_hessian_jvp_is_jacobian_of_gradient(gamma):
    """phi2_1 / phi2_2 are the exact partials of phi1, and the replicated border
    neighbours move with the voxel, so the order-2 operator applied to u is the
    Jacobian of the prior gradient times u (autograd forward-mode), everywhere."""
    f = _rand((6, 5, 7), seed=21) + 0.2
    f[1:3, 1:3, 2:4] = 0.0
    u = _rand((6, 5, 7), seed=22) - 0.5
    dr = (0.48, 0.5, 0.52)
    _, jvp = torch.autograd.functional.jvp(lambda x: so.rdp_gradient(x, 0.2, gamma, dr, 0.5), f, u)
    out = so.rdp_hessian_jvp(f, u, 0.2, gamma, dr, 0.5)
    assert torch.allclose(out, jvp, rtol=1e-11, atol=1e-11 * float(jvp.abs().max()))


# ------------------------------------------ preconditioned gradient ascent
def _small_problem(n_angles=6, quarter=False):
    from tests.helpers import oracle_geometry, problem_arrays
    pa = problem_arrays(8, n_angles, seed=23, nz=4, att=True, psf=True)
    if quarter:
        pa["angles"] = np.array([0.0, 90.0, 180.0, 270.0])
    g = oracle_geometry(pa, torch.float64)
    f_true = torch.from_numpy(pa["f"]).double() + 0.05
    expected = so.forward(g, f_true) * 40
    s = torch.full_like(expected, 0.05 * float(expected.mean()))
    data = torch.poisson(expected + s, generator=torch.Generator().manual_seed(24)) + 1
    return g, data, s


def test_gradient_ascent_reduces_to_pinned_loops():
    g, data, s = _small_problem()
    osem = so.osem(g, data, 2, 3, additive=s)
    bsrem = so.bsrem(g, data, 2, 3, 0.1, 2.0, additive=s)
    run = lambda m, **kw: so.gradient_ascent(g, data, 2, 3, m, additive=s, **kw)
    assert torch.allclose(run("osem"), osem, rtol=1e-13, atol=0)
    assert torch.allclose(run("bsrem", beta=0.1, floor=1e-4), bsrem, rtol=1e-13, atol=0)
    # OSMAPOSL without a prior is OSEM; RBI-MAP without a prior is RBI-EM.
    assert torch.allclose(run("osmaposl"), osem, rtol=1e-13, atol=0)
    assert torch.allclose(run("rbimap"), run("rbiem"), rtol=1e-13, atol=0)
    # RBI-EM with one block: r = max(n / (n + delta)) = 1 - O(delta), i.e. MLEM.
    one = so.gradient_ascent(g, data, 2, 1, "rbiem", additive=s)
    assert torch.allclose(one, so.osem(g, data, 2, 1, additive=s), rtol=1e-9, atol=0)


def test_one_step_late_and_rbi_closed_forms():
    """One sub-iteration in the textbook forms: OSL-EM (Green 1990)
    f' = f B(g/p) / (n_m + grad V); RBI-EM (Byrne 1996)
    f' = f + f (B(g/p) - n_m) / (r_m N), r_m = max n_m / N."""
    g, data, s = _small_problem()
    sidx = g.subsets(3)[0]
    f0 = torch.ones(g.shape, dtype=torch.float64)
    norms = [so.sensitivity(g, x) for x in g.subsets(3)]
    N = torch.stack(norms).sum(0)
    ratio_bp = so.backward(g, data[sidx] / (so.forward(g, f0, sidx) + s[sidx] + so.DELTA), sidx)
    pg = so.rdp_gradient(f0 + 0, 0.3, 2.0, g.dr, len(sidx) / g.n_angles)
    got = []
    so.gradient_ascent(g, data, 1, 3, "osmaposl", beta=0.3, additive=s,
                       callback=lambda f, pp, j, k: got.append(f))
    assert len(got) == 3
    # got[1] is the iterate after sub-iteration 0
    assert torch.allclose(got[1], f0 * ratio_bp / (norms[0] + pg), rtol=1e-10, atol=0)
    got = []
    so.gradient_ascent(g, data, 1, 3, "rbiem", additive=s, callback=lambda f, pp, j, k: got.append(f))
    rm = float((norms[0] / N).max())
    assert torch.allclose(got[1], f0 + f0 * (ratio_bp - norms[0]) / (rm * N), rtol=1e-10, atol=0)


@pytest.mark.parametrize("method", ["osem", "bsrem"])
def test_uncertainty_recursion_is_autograd_sensitivity(method):
    """At quarter-turn angles the reference back projection is the exact
    transpose of the forward projection, so the Q/B recursion must equal the
    reverse-mode derivative of sum(mask * f_K) with respect to the projections
    (V[0]) and to the additive term (V[1])."""
    g, data, s = _small_problem(quarter=True)
    assert torch.allclose(so.backward(g, data), so.exact_adjoint(g, data), rtol=1e-12, atol=1e-12)
    n_sub, iters = 2, 3
    beta = 0.3 if method == "bsrem" else None
    floor = 1e-4 if method == "bsrem" else 0.0
    objects, fps = [], []
    dv, sv = data.clone().requires_grad_(True), s.clone().requires_grad_(True)
    fK = so.gradient_ascent(g, dv, iters, n_sub, method, beta=beta, additive=sv, floor=floor,
                            callback=lambda f, pp, j, k: (objects.append(f), fps.append(pp)))
    assert float(fK.min()) > floor
    objects.append(fK.detach().clone())
    mask = torch.zeros(g.shape, dtype=torch.float64)
    mask[2:6, 3:6, 1:3] = 1
    gd, gs = torch.autograd.grad((mask * fK).sum(), (dv, sv))
    norms = [so.sensitivity(g, x) for x in g.subsets(n_sub)]
    N = torch.stack(norms).sum(0)
    if method == "osem":
        factor = lambda m: 1 / (norms[m] + so.DELTA)
    else:
        factor = lambda m: 1 / (0.5 * N + so.DELTA)
    unc, V = so.uncertainty(g, data, objects, fps, mask, n_sub, factor, beta=beta,
                            additive_variance=lambda v: v ** 2)
    assert torch.allclose(V[0], gd, rtol=1e-8, atol=1e-10 * float(gd.abs().max()))
    assert torch.allclose(V[1], gs, rtol=1e-8, atol=1e-10 * float(gs.abs().max()))
    expected = math.sqrt(float((gd ** 2 * data).sum() + (gs ** 2).sum()))
    assert abs(unc - expected) <= 1e-8 * expected


# ------------------------------------------------- energy-window scatter
def test_ew_scatter_closed_form():
    lo = torch.tensor([[[10.0, 20.0]]], dtype=torch.float64)
    up = torch.tensor([[[4.0, 0.0]]], dtype=torch.float64)
    est, diag, op = so.ew_scatter(lo, up, 6.0, 4.0, 20.0, 0.5, 0.5)
    # TEW trapezoid: (C_l/w_l + C_u/w_u) * w_p / 2
    assert torch.allclose(est, torch.tensor([[[(10 / 6 + 4 / 4) * 10, (20 / 6) * 10]]], dtype=torch.float64))
    assert torch.allclose(diag, (20 / 6 * 0.5) ** 2 * lo + (20 / 4 * 0.5) ** 2 * up)
    x = torch.tensor([[[2.0, -1.0]]], dtype=torch.float64)
    assert torch.allclose(op(x), x * diag * x)
    # DEW: no upper window
    est, diag, _ = so.ew_scatter(lo, None, 6.0, None, 20.0, 1.0, 0.5)
    assert torch.equal(est, lo / 6.0 * 1.0 * 20.0)


@pytest.mark.parametrize("axis,mode", [(0, "wrap"), (1, "nearest"), (2, "nearest")])
def test_smoothed_scatter_matches_scipy_correlate(axis, mode):
    from scipy.ndimage import correlate1d
    x = _rand((12, 9, 7), seed=31)
    sig = [0.0, 0.0, 0.0]
    d = [3.0, 0.48, 0.5]
    sig[axis] = [7.0, 0.9, 0.6][axis]
    out = so.smoothed_scatter(x, d[0], d[1], d[2], sig[0], sig[1], sig[2])
    K = 2 * int(np.ceil(3 * sig[axis] / d[axis])) + 1
    w = so.gaussian_kernel_1d(sig[axis] / d[axis], K, torch.float64).numpy()
    ref = correlate1d(x.numpy(), w, axis=axis, mode=mode)
    assert np.abs(out.numpy() - ref).max() < 1e-12
    # normalised taps: a constant stays constant
    c = so.smoothed_scatter(torch.full((12, 9, 7), 3.0, dtype=torch.float64), d[0], d[1], d[2], *sig)
    assert torch.allclose(c, torch.full_like(c, 3.0), rtol=1e-14, atol=0)


def test_gaussian_taps_closed_form():
    # the exponent is an int64 tensor over a Python float: torch's default
    # (float32) dtype, as in the reference
    k = so.gaussian_kernel_1d(1.5, 7, torch.float64)
    x = np.arange(-3, 4)
    ref = np.exp(-x ** 2 / (2 * 1.5 ** 2))
    assert np.allclose(k.numpy(), ref / ref.sum(), rtol=2e-7, atol=0)
    assert torch.equal(k, k.flip(0))


# ------------------------------------------------ KEM basis and DVF warp
def test_kem_apply_matches_numpy_brute_force():
    f = _rand((4, 5, 3), seed=41)
    K = torch.rand((3, 3, 3, 4, 5, 3), generator=torch.Generator().manual_seed(42), dtype=torch.float64)
    out = so.kem_apply(K, f).numpy()
    fn, Kn = f.numpy(), K.numpy()
    ref = np.zeros_like(fn)
    for x in range(4):
        for y in range(5):
            for z in range(3):
                for i in range(-1, 2):
                    for j in range(-1, 2):
                        for k in range(-1, 2):
                            xs, ys, zs = min(max(x + i, 0), 3), min(max(y + j, 0), 4), min(max(z + k, 0), 2)
                            ref[x, y, z] += Kn[i + 1, j + 1, k + 1, x, y, z] * fn[xs, ys, zs]
    assert np.abs(out - ref).max() < 1e-13


def test_kem_kernel_normalised_and_uniform_support_is_distance_only():
    sup = torch.full((6, 6, 6), 2.0, dtype=torch.float64)
    K = so.kem_kernel([sup], (0.5, 0.5, 0.5), size=3, dtype=torch.float64)
    assert torch.allclose(K.sum(dim=(0, 1, 2)), torch.ones(6, 6, 6, dtype=torch.float64))
    # a constant support object leaves only the Gaussian distance weights
    x = np.arange(-1, 2) * 0.5
    d2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    w = np.exp(-d2 / 2)
    assert np.allclose(K[..., 2, 3, 1].numpy(), w / w.sum(), rtol=1e-6, atol=0)
    # top_N keeps exactly N taps per voxel
    sup2 = _rand((6, 6, 6), seed=43)
    K2 = so.kem_kernel([sup2], (0.5, 0.5, 0.5), size=3, top_N=5, dtype=torch.float64)
    assert torch.all((K2 > 0).sum(dim=(0, 1, 2)) == 5)


@pytest.fixture
def default_float64():
    """The reference's float64 mode sets torch's default dtype (__init__.py:24-32),
    which the DVF normalisation 2/(dims-1) is computed in."""
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def test_dvf_warp_closed_forms(default_float64):
    f = _rand((6, 7, 5), seed=44)
    zero = torch.zeros((6, 7, 5, 3), dtype=torch.float64)
    assert torch.allclose(so.dvf_warp(f, zero), f, rtol=1e-12, atol=1e-12)
    # integer displacement: a shift with zeros entering
    sh = zero.clone()
    sh[..., 0] = 1.0
    sh[..., 2] = -2.0
    out = so.dvf_warp(f, sh)
    ref = torch.zeros_like(f)
    ref[:-1, :, 2:] = f[1:, :, :-2]
    assert torch.allclose(out, ref, rtol=1e-12, atol=1e-12)
    # fractional displacement inside: trilinear interpolation of a linear ramp is exact
    ramp = torch.from_numpy(np.fromfunction(lambda x, y, z: 0.3 * x - 0.7 * y + 1.1 * z + 2, (6, 7, 5)))
    fr = zero.clone()
    fr[..., 0], fr[..., 1], fr[..., 2] = 0.25, 0.4, -0.35
    out = so.dvf_warp(ramp, fr)
    exp = ramp + 0.3 * 0.25 - 0.7 * 0.4 + 1.1 * -0.35
    assert torch.allclose(out[:5, :6, 1:], exp[:5, :6, 1:], rtol=1e-12, atol=1e-12)


# ------------------------------------------------ transform lists (per-angle path)
class _Scale:
    """A caller's transform: multiply by a constant (backward the same)."""

    def __init__(self, c):
        self.c = c

    def forward(self, obj_i, ang_idx):
        return obj_i * self.c

    backward = forward


class _MaskAngles:
    """proj2proj: zero the projections of some angles (a diagonal mask)."""

    def __init__(self, n_angles, zero):
        self.m = torch.ones(n_angles)
        self.m[list(zero)] = 0

    def forward(self, proj, angle_indices):
        return proj * self.m[torch.as_tensor(angle_indices).long()].reshape(-1, 1, 1)

    backward = forward


def _small_geometry(att=True, psf=True):
    n, a = 12, 6
    f = torch.zeros((n, n, n))
    f[4:8, 3:9, 2:10] = 1.0
    mu = torch.full((n, n, n), 0.12) if att else None
    return so.make_geometry((n, n, n), (0.4, 0.4, 0.4), np.arange(a) * 60.0, radii=np.full(a, 20.0),
                            mu=mu, psf=so.PSFModel((0.03, 0.1)) if psf else None), f


def test_transform_lists_reduce_to_the_default_chain():
    """The default obj2obj list is [attenuation, PSF]; an identity inserted
    anywhere changes nothing; a constant factor c scales forward and
    backward by c (linearity) wherever it sits."""
    g, f = _small_geometry()
    y = torch.rand((6, 12, 12), generator=torch.Generator().manual_seed(3))
    fw, bw = so.forward(g, f), so.backward(g, y)
    assert torch.equal(so.forward(g, f, obj2obj=["att", "psf"]), fw)
    assert torch.equal(so.backward(g, y, obj2obj=["att", _Scale(1.0), "psf"]), bw)
    for pos in (0, 1, 2):
        ts = ["att", "psf"]
        ts.insert(pos, _Scale(3.0))
        assert torch.allclose(so.forward(g, f, obj2obj=ts), 3 * fw, rtol=1e-6, atol=1e-6)
        assert torch.allclose(so.backward(g, y, obj2obj=ts), 3 * bw, rtol=1e-6, atol=1e-6)
    # order matters once the transforms do not commute: PSF then attenuation
    assert not torch.allclose(so.forward(g, f, obj2obj=["psf", "att"]), fw, rtol=1e-4)


def test_proj2proj_masks_forward_and_backward():
    """A diagonal proj2proj mask zeroes exactly its angles' projections in
    forward, and backward equals the default back projection of the masked
    projections."""
    g, f = _small_geometry()
    m = _MaskAngles(6, zero=(1, 4))
    out = so.forward(g, f, proj2proj=[m])
    ref = so.forward(g, f)
    assert float(out[[1, 4]].abs().max()) == 0.0
    assert torch.equal(out[[0, 2, 3, 5]], ref[[0, 2, 3, 5]])
    y = torch.rand((6, 12, 12), generator=torch.Generator().manual_seed(4))
    ym = y.clone()
    ym[[1, 4]] = 0
    assert torch.allclose(so.backward(g, y, proj2proj=[m]), so.backward(g, ym), rtol=1e-6, atol=1e-7)
    # a subset: the mask rows follow the call's angle indices
    sub = torch.tensor([1, 3])
    assert float(so.forward(g, f, sub, proj2proj=[m])[0].abs().max()) == 0.0


def test_average_prob_matrix_without_attenuation_is_one():
    """mu = 0: every probability of detection is 1, so the average over
    angles is 1 wherever all bilinear taps of the back rotation fall inside
    the padded frame (the inscribed disk of the object)."""
    n = 16
    g = so.make_geometry((n, n, 4), (0.4,) * 3, np.arange(8) * 45.0, mu=torch.zeros((n, n, 4)))
    avg = so.average_prob_matrix(g)
    yy, xx = np.meshgrid(np.arange(n) - (n - 1) / 2, np.arange(n) - (n - 1) / 2, indexing="ij")
    disk = torch.from_numpy(np.hypot(xx, yy) < n / 2 - 1)
    assert float((avg[disk] - 1).abs().max()) < 1e-6


def test_square_taps_closed_form():
    """get_1D_PSF_layer_square (psf.py:85-101): max(0, 1 - |x|/sigma),
    normalised; sigma = 2 gives [0, .5, 1, .5, 0] / 2."""
    t = so.square_taps(np.array([2.0, 1.0]), 5)
    assert torch.allclose(t[0], torch.tensor([0.0, 0.25, 0.5, 0.25, 0.0]))
    assert torch.allclose(t[1], torch.tensor([0.0, 0.0, 1.0, 0.0, 0.0]))
