"""Pins of the FP64 oracle (oracle/np64.py) and the torch restatement
(oracle/torch_ref.py): closed forms, library routines, invariants, and the
reference's own fixtures.  CPU only."""
import math
import os

import numpy as np
import pytest
import torch
from scipy import signal, stats

from oracle import np64, torch_ref
from tests import common as C
from gigalens_amd import psf as gpsf

RNG = np.random.default_rng(0)


# ---- EPL (tests/test_profiles.py:50-63 parameter points) -----------------
@pytest.mark.parametrize("prm", [(1.0, 2.0, 0.0, 0.0, 0.0, 0.0), (1.2, 2.2, -0.1, 0.1, 0.0, 0.0)])
def test_epl_series_vs_hyp2f1(prm):
    x, y = RNG.normal(size=(1, 10000)), RNG.normal(size=(1, 10000))
    a = np64.epl_deriv(x, y, *prm, niter=100)
    b = np64.epl_deriv_hyp2f1(x, y, *prm)
    assert np.allclose(a[0], b[0], rtol=1e-5, atol=1e-4)
    assert np.allclose(a[1], b[1], rtol=1e-5, atol=1e-4)
    assert np.abs(a[0] - b[0]).max() < 1e-12


def test_epl_gamma2_is_sie():
    x, y = RNG.normal(size=(1, 5000)), RNG.normal(size=(1, 5000))
    for e in [(1e-3, 1e-3), (0.1, -0.1), (0.3, 0.2)]:
        a = np64.epl_deriv(x, y, 1.2, 2.0, e[0], e[1], 0.05, -0.02, niter=100)
        b = np64.sie_deriv(x, y, 1.2, e[0], e[1], 0.05, -0.02)
        assert np.abs(a[0] - b[0]).max() < 1e-10 and np.abs(a[1] - b[1]).max() < 1e-10


def test_epl_round_is_power_law():
    x, y = RNG.normal(size=(1, 5000)), RNG.normal(size=(1, 5000))
    for g in (1.6, 2.0, 2.5):
        a = np64.epl_deriv(x, y, 1.3, g, 0.0, 0.0, 0.0, 0.0)
        r = np.hypot(x, y)
        amp = 1.3 ** (g - 1) * r ** (2 - g)
        assert np.allclose(a[0], amp * x / r, rtol=1e-12) and np.allclose(a[1], amp * y / r, rtol=1e-12)


def test_epl_series_length_rule():
    # log(1e-12)/log(0.1) + 2 = 14 -> n = 1..13
    assert np64.epl_series_length(0.1, 50) == 13
    assert np64.epl_series_length(0.99, 50) == 50
    assert np64.epl_series_length(0.0, 50) == 1
    assert np64.epl_series_length(1.0, 50) == 0
    # truncation error of the per-particle length stays below 1e-12 relative
    f = 0.37
    n = np64.epl_series_length(f, 50)
    assert f ** (n + 1) / (1 - f) < 1e-11


def test_sis_shear_closed_forms():
    x, y = RNG.normal(size=(1, 100)), RNG.normal(size=(1, 100))
    ax, ay = np64.sis_deriv(x, y, 1.2, 0.0, 0.0)
    r = np.hypot(x, y)
    assert np.allclose(np.hypot(ax, ay), 1.2)
    sx, sy = np64.shear_deriv(x, y, 0.1, -0.05)
    assert np.allclose(sx, 0.1 * x - 0.05 * y) and np.allclose(sy, -0.05 * x - 0.1 * y)


# ---- Sersic (tests/test_profiles.py:14-32) ---------------------------------
def test_sersic_known_answer_and_rotation():
    v = np64.sersic_ellipse_light(np.zeros((1, 1)), np.ones((1, 1)), 1.0, 2.0, 0.0, 0.0, 0.0, 0.0, Ie=5.0)
    assert math.isclose(float(v.flat[0]), 5.0)
    # independent construction: rotate the frame by phi, scale axes by sqrt(q), 1/sqrt(q)
    x, y = RNG.normal(size=(1, 200)), RNG.normal(size=(1, 200))
    e1, e2 = 0.12, -0.2
    phi = 0.5 * math.atan2(e2, e1)
    c = math.hypot(e1, e2)
    q = (1 - c) / (1 + c)
    R = np64.sersic_distance(x, y, 0.1, 0.2, e1, e2)[0]
    dx, dy = x - 0.1, y - 0.2
    rot = np.array([[math.cos(phi), math.sin(phi)], [-math.sin(phi), math.cos(phi)]])
    u, w = rot @ np.vstack([dx, dy])
    assert np.allclose(R, np.sqrt(q * u ** 2 + w ** 2 / q))


# ---- Shapelets (tests/test_profiles.py:35-47) ------------------------------
def test_hermite_functions_orthonormal():
    xs = np.linspace(-12, 12, 20001)
    dx = xs[1] - xs[0]
    H = np.stack([np64.hermite_phi(n, xs) for n in range(11)])
    G = H @ H.T * dx
    assert np.allclose(G, np.eye(11), atol=1e-10)


def test_shapelet_orders():
    o = np64.shapelet_orders(2)
    assert o == [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]
    assert len(np64.shapelet_orders(10)) == 66


def test_shapelet_table_vs_direct():
    x, y = RNG.normal(size=(1, 5, 5)), RNG.normal(size=(1, 5, 5))
    a = np64.shapelet_basis(x, y, 0.0, 0.0, 1.0, 5, interpolate=True)
    b = np64.shapelet_basis(x, y, 0.0, 0.0, 1.0, 5, interpolate=False)
    # the reference's own tolerance for both modes (test_profiles.py:47)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-4)
    # zero fill outside |u| > 5
    out = np64.shapelet_basis(np.full((1, 1, 1), 5.5), np.zeros((1, 1, 1)), 0.0, 0.0, 1.0, 3, interpolate=True)
    assert np.all(out == 0.0)


# ---- conv / pool -----------------------------------------------------------
def test_conv_pool_vs_scipy_and_fused_identity():
    img = RNG.normal(size=(2, 24, 24))
    k = RNG.random((5, 5))
    out = np64.conv_pool(img, k, 2)
    ref = np.stack([signal.convolve2d(im, k, mode="same") for im in img])
    ref = ref.reshape(2, 12, 2, 12, 2).mean(axis=(2, 4))
    assert np.allclose(out, ref)
    # pool(conv) == stride-ss correlation with flipped kernel convolved with the ss x ss box
    kf = signal.convolve2d(k[::-1, ::-1], np.ones((2, 2)) / 4)
    pb = (5 - 1) // 2
    pad = np.pad(img, ((0, 0), (pb, pb + 2), (pb, pb + 2)))
    fused = np.zeros((2, 12, 12))
    for I in range(12):
        for J in range(12):
            fused[:, I, J] = np.sum(pad[:, 2 * I:2 * I + 6, 2 * J:2 * J + 6] * kf, axis=(1, 2))
    assert np.allclose(fused, out)


# ---- likelihood, prior, bijectors ------------------------------------------
def test_forward_loglike_vs_scipy():
    im = RNG.random((3, 8, 8)) * 5 - 0.5
    obs = RNG.random((8, 8)) * 5
    ll, chi2 = np64.forward_loglike(im, obs, 0.2, 100.0)
    sig = np.sqrt(0.04 + np.clip(im, 0, None) / 100)
    ref = stats.norm.logpdf(obs[None], im, sig).sum(axis=(1, 2))
    assert np.allclose(ll, ref)
    assert np.allclose(chi2, np.mean(((im - obs) / sig) ** 2, axis=(1, 2)))


def test_prior_logpdf_and_bijectors():
    for kind, args in [("Normal", (0.0, 0.1)), ("LogNormal", (math.log(1.25), 0.25)), ("Uniform", (0.5, 8.0)),
                       ("TruncatedNormal", (2.0, 0.25, 1.2, 2.8))]:
        z = np.linspace(-2, 2, 9)
        x, fldj = np64.bijector_forward(kind, args, z)
        assert np.allclose(np64.bijector_inverse(kind, args, x), z)
        h = 1e-6
        xp, _ = np64.bijector_forward(kind, args, z + h)
        xm, _ = np64.bijector_forward(kind, args, z - h)
        assert np.allclose(fldj, np.log((xp - xm) / (2 * h)), atol=1e-6)
    assert np.isneginf(np64.prior_logpdf("Uniform", (0.0, 1.0), 1.0))
    assert math.isclose(float(np64.prior_logpdf("TruncatedNormal", (2.0, 0.25, 1.2, 2.8), 2.0)),
                        float(stats.truncnorm.logpdf(2.0, -3.2, 3.2, loc=2.0, scale=0.25)))


# ---- least squares (tf/simulator.py:157-165) --------------------------------
def test_lstsq_solve_is_weighted_least_squares():
    stacked = RNG.normal(size=(2, 6, 6, 4))
    obs = RNG.normal(size=(6, 6))
    err = 0.5 + RNG.random((6, 6))
    img, coeffs = np64.lstsq_solve(stacked, obs, err)
    for i in range(2):
        A = (stacked[i] / err[..., None]).reshape(-1, 4)
        b = (obs / err).reshape(-1)
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.allclose(coeffs[i], ref)


# ---- sub-grid PSF (lenstronomy subgrid_kernel restatement) ------------------
@pytest.mark.parametrize("ss", [2, 3, 4])
def test_subgrid_kernel_invariants(ss):
    k = np.load(os.path.join(C.GOLDEN, "psf.npy")).astype(np.float64)
    ks = np64.subgrid_kernel(k, ss)
    assert ks.shape[0] % 2 == 1 and ks.shape[0] in (13 * ss - 1, 13 * ss)
    assert math.isclose(ks.sum(), 1.0, rel_tol=1e-12)
    if ss % 2 == 0:
        back = np64._averaging_even_kernel(ks, ss)
    else:
        back = ks.reshape(13, ss, 13, ss).mean(axis=(1, 3))
    assert np.abs(back / back.sum() - k / k.sum()).max() < 1e-4
    # the package's implementation agrees with the oracle's
    assert np.allclose(gpsf.subgrid_kernel(k, ss), ks, atol=1e-12)
    # a symmetric PSF stays symmetric
    g = gpsf.gaussian(0.1, 0.065, 11).astype(np.float64)
    gs = np64.subgrid_kernel(g, ss)
    assert np.allclose(gs, gs[::-1, ::-1], atol=1e-12) and np.allclose(gs, gs.T, atol=1e-12)


def test_analytic_psfs():
    g = gpsf.gaussian(0.1, 0.065, 11)
    assert math.isclose(float(g.sum()), 1.0, rel_tol=1e-6) and g.argmax() == 60
    m = gpsf.moffat(0.08, 3.0, 0.04, 21)
    # half maximum at r = FWHM/2 (closed form of the Moffat profile)
    alpha = 0.08 / (2 * math.sqrt(2 ** (1 / 3.0) - 1))
    assert math.isclose((1 + 0.04 ** 2 / alpha ** 2) ** -3.0, 0.5, rel_tol=1e-12)
    assert math.isclose(float(m.sum()), 1.0, rel_tol=1e-6)


# ---- grid (simulator.py:91-125) + the reference fixtures -------------------
def test_grid_and_demo_fixture_statistics():
    from gigalens_amd import workloads as wl
    from gigalens_amd.simulator import SimulatorConfig, Camera
    cfg = SimulatorConfig(delta_pix=0.065, num_pix=60, supersample=2,
                          kernel=np.load(os.path.join(C.GOLDEN, "psf.npy")).astype(np.float32))
    cam = Camera(cfg)
    x, y = cam.coords()
    assert math.isclose(float(x.mean()), 0.0, abs_tol=1e-7) and math.isclose(float(y.mean()), 0.0, abs_tol=1e-7)
    assert x[0, 1] > x[0, 0] and y[1, 0] > y[0, 0]  # columns move in +x (RA), rows in +y
    assert math.isclose(float(x[0, 1] - x[0, 0]), 0.0325, rel_tol=1e-5)
    ocam = C.oracle_camera(cfg, cam)
    assert np.array_equal(ocam.x, x) and np.array_equal(ocam.y, y)
    phys = wl.epl_shear_model()
    lenses, ll, src = C.oracle_model_lists(phys, wl.DEMO_TRUTH)
    img = np64.simulate(ocam, lenses, ll, src)[0]
    obs = np.load(os.path.join(C.GOLDEN, "demo.npy"))
    ll_, chi2 = np64.forward_loglike(img[None], obs, 0.2, 100.0)
    # the reference's demo image is a noisy draw of the notebook truth: mean r^2 ~ 1
    assert abs(float(chi2[0]) - 1.0) < 3 * math.sqrt(2.0 / 3600) + 0.02
    # a transposed grid or an unflipped PSF would show up as a misfit
    lenses_t, ll_t, src_t = C.oracle_model_lists(phys, wl.DEMO_TRUTH)
    bad = np64.simulate(ocam, lenses_t, ll_t, src_t)[0].T
    assert float(np64.forward_loglike(bad[None], obs, 0.2, 100.0)[1][0]) > 1.5


# ---- torch restatement == numpy oracle -------------------------------------
def test_torch_ref_matches_np64():
    prob = C.epl_shear_problem(bs=3)
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SERSIC_ELLIPSE", {})]}
    sim = torch_ref.Sim(model, prob["ocam"], dtype=torch.float64)
    x = prob["x"].double()
    nested = C.nested_from_x(prob["prior"], x)
    img_t = sim.simulate(nested).numpy()
    # the torch restatement uses the batch-coupled series length (epl.py:41)
    img_o = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prob["prior"], x.numpy()),
                              per_particle_series=False)
    assert np.abs(img_t - img_o).max() < 1e-9 * np.abs(img_o).max()
    # per-particle vs batch-coupled series length: identical to 1e-10
    img_pp = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prob["prior"], x.numpy()))
    assert np.abs(img_pp - img_o).max() < 1e-9 * np.abs(img_o).max()
    # log_prob of the torch restatement == oracle log_prob
    op = [(p, k, a) for p, k, a in C.oracle_prior(prob["prior"])]
    nested_prior = C.oracle_nested_prior(prob["prior"])
    z = torch.as_tensor(torch_ref.x_to_z(nested_prior, x.numpy()))
    lp_t, chi_t = torch_ref.log_prob(sim, nested_prior, z, prob["obs"], 0.2, 100.0)
    lp_o, chi_o, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], z.numpy(), prob["obs"], 0.2, 100.0)
    assert np.allclose(lp_t.numpy(), lp_o, rtol=1e-10) and np.allclose(chi_t.numpy(), chi_o, rtol=1e-10)


def test_core_sersic_reductions():
    """Core-Sersic restatement (sersic.py:98-121): Rb -> 0 with alpha = 1/n gives
    exp(-1 - bn) times the Sersic ellipse (pinned above), and the core factor at R = Rb
    is 2^(gamma/alpha) relative to the core-less expression."""
    rng = np.random.default_rng(3)
    x, y = rng.normal(size=(2, 40, 40)) * 0.8
    Rs, n, e1, e2, cx, cy, Ie = 0.7, 2.5, 0.1, -0.05, 0.02, -0.03, 3.0
    bn = 1.9992 * n - 0.3271
    core = np64.core_sersic_light(x, y, Rs, n, 1e-40, 1.0 / n, 0.4, e1, e2, cx, cy, Ie)
    ser = np64.sersic_ellipse_light(x, y, Rs, n, e1, e2, cx, cy, Ie)
    np.testing.assert_allclose(core, ser * np.exp(-1.0 - bn), rtol=1e-10)
    # at R = Rb the core factor (1 + 1)^(gamma/alpha)
    Rb, al, gam = 0.3, 1.7, 0.6
    xs, ys = np.array([[cx + Rb]]), np.array([[cy]])
    R = np64.sersic_distance(xs, ys, cx, cy, 0.0, 0.0)[0, 0, 0]
    full = np64.core_sersic_light(xs, ys, Rs, n, Rb, al, gam, 0.0, 0.0, cx, cy, Ie)[0, 0, 0]
    nocore = Ie * np.exp(-bn * (R ** al + Rb ** al) / Rs ** al / (al * n) - 1.0)
    assert full == pytest.approx(nocore * 2.0 ** (gam / al), rel=1e-12)


@pytest.mark.parametrize("alpha", [0.5, 1.3, 2.0, 3.5])
def test_core_sersic_general_alpha(alpha):
    """Core-Sersic restatement at general alpha, from properties of the profile rather
    than its expression: with Rb -> 0 the log-profile is ln Ie - 1 - b (R/Rs)^alpha, so
    -(ln I - ln Ie + 1) has log-log slope alpha and equals b = bn / (alpha n) at R = Rs;
    inside the core (R << Rb) the logarithmic slope d ln I / d ln R tends to -gamma;
    outside it (R >> Rb) the core factor tends to 1 + (gamma/alpha) (Rb/R)^alpha."""
    # centre at the origin so R = x exactly (a shifted centre loses R's low digits at R << 1)
    Rs, n, gam, Ie, cx, cy = 0.7, 2.5, 0.6, 3.0, 0.0, 0.0
    bn = 1.9992 * n - 0.3271
    b = bn / (alpha * n)

    def light(R, Rb):
        xs, ys = np.array([[cx + R]]), np.array([[cy]])
        return float(np64.core_sersic_light(xs, ys, Rs, n, Rb, alpha, gam, 0.0, 0.0, cx, cy, Ie)[0, 0, 0])

    # Rb -> 0: log-log slope alpha (central differences in ln R) and the value at R = Rs
    for R in (0.1, 0.5, 1.3):
        h = 1e-4
        fp = -(math.log(light(R * math.exp(h), 1e-300)) - math.log(Ie) + 1.0)
        fm = -(math.log(light(R * math.exp(-h), 1e-300)) - math.log(Ie) + 1.0)
        assert (math.log(fp) - math.log(fm)) / (2 * h) == pytest.approx(alpha, rel=1e-7)
    assert -(math.log(light(Rs, 1e-300)) - math.log(Ie) + 1.0) == pytest.approx(b, rel=1e-12)
    # inner logarithmic slope -> -gamma
    Rb = 0.4
    R, h = 1e-12, 1e-4
    slope = (math.log(light(R * math.exp(h), Rb)) - math.log(light(R * math.exp(-h), Rb))) / (2 * h)
    assert slope == pytest.approx(-gam, rel=1e-4)
    # outer core factor
    R = 4.0
    ratio = light(R, Rb) / light(R, 1e-300)
    # the Rb^alpha term of the exponent is part of the core factor at finite Rb
    ratio *= math.exp(b * (Rb / Rs) ** alpha)
    x = (Rb / R) ** alpha
    assert ratio - 1.0 == pytest.approx(gam / alpha * x, rel=abs(gam / alpha - 1.0) * x + 1e-9)
