"""CPU checks of the device math (lens_math.h, host build) against the FP64
oracle, closed forms and FP64 finite differences.  No GPU needed."""
import ctypes
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

from oracle import np64

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "math_harness.hip")
F32P = ctypes.POINTER(ctypes.c_float)


@pytest.fixture(scope="module")
def mh():
    out = os.path.join(tempfile.mkdtemp(), "libmh.so")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", out, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    L = ctypes.CDLL(out)
    L.h_series_len.restype = ctypes.c_int
    L.h_series_len.argtypes = [ctypes.c_double, ctypes.c_int]
    return L


def f32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


def p(a):
    return a.ctypes.data_as(F32P)


def epl_alpha(mh, prm, x, y, niter=50):
    x, y = f32(x), f32(y)
    ax, ay = np.empty_like(x), np.empty_like(x)
    n = ctypes.c_int()
    mh.h_epl_alpha(p(f32(prm)), niter, len(x), p(x), p(y), p(ax), p(ay), ctypes.byref(n))
    return ax, ay, n.value


def test_series_length_rule(mh):
    for f in [0.0, 1e-7, 0.01, 0.1, 0.3, 0.5, 0.58, 0.7, 0.95, 1.0]:
        assert mh.h_series_len(f, 50) == np64.epl_series_length(f, 50), f
    assert mh.h_series_len(float("nan"), 50) == 0


@pytest.mark.parametrize("prm", [(1.0, 2.0, 0.0, 0.0, 0.0, 0.0), (1.2, 2.2, -0.1, 0.1, 0.0, 0.0),
                                 (1.1, 1.7, 0.3, -0.2, 0.05, -0.03), (0.9, 2.6, 0.05, 0.2, -0.1, 0.2)])
def test_epl_forward_vs_hyp2f1(mh, prm):
    """Clenshaw series vs the scipy 2F1 closed form (test_profiles.py:50-63 tolerances)."""
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=2000), rng.normal(size=2000)
    ax, ay, _ = epl_alpha(mh, prm, x, y, niter=100)
    rx, ry = np64.epl_deriv_hyp2f1(x.reshape(1, -1), y.reshape(1, -1), *prm)
    assert np.allclose(ax, rx[0, 0], rtol=1e-5, atol=1e-5)
    assert np.allclose(ay, ry[0, 0], rtol=1e-5, atol=1e-5)


def test_epl_forward_vs_series_oracle(mh):
    rng = np.random.default_rng(1)
    x, y = rng.normal(size=500) * 1.5, rng.normal(size=500) * 1.5
    for e in [(0.4, 0.3), (-0.45, 0.1), (0.0, 0.01)]:
        prm = (1.3, 2.1, e[0], e[1], 0.02, -0.04)
        ax, ay, N = epl_alpha(mh, prm, x, y, niter=50)
        rx, ry = np64.epl_deriv(x.reshape(1, -1), y.reshape(1, -1), *prm, n_terms=[N])
        assert np.abs(ax - rx[0, 0]).max() < 2e-5 * np.abs(rx).max()
        assert np.abs(ay - ry[0, 0]).max() < 2e-5 * np.abs(ry).max()


def _epl_contract(prm, x, y, Hx, Hy, n_terms):
    ax, ay = np64.epl_deriv(x.reshape(1, -1), y.reshape(1, -1), *prm, n_terms=[n_terms])
    return float(np.sum(Hx * ax[0, 0] + Hy * ay[0, 0]))


@pytest.mark.parametrize("prm", [(1.2, 2.2, -0.1, 0.1, 0.03, -0.02), (0.9, 1.8, 0.25, 0.15, -0.05, 0.1),
                                 (1.4, 2.5, 0.02, -0.3, 0.0, 0.0)])
def test_epl_adjoint_vs_fd(mh, prm):
    """H . d alpha / d(theta_E, gamma, e1, e2, cx, cy) vs FP64 central differences."""
    rng = np.random.default_rng(2)
    n = 3000
    x, y = rng.normal(size=n), rng.normal(size=n)
    Hx, Hy = rng.normal(size=n), rng.normal(size=n)
    g = np.zeros(6, np.float32)
    ax, ay = np.empty(n, np.float32), np.empty(n, np.float32)
    xs, ys, hxs, hys = f32(x), f32(y), f32(Hx), f32(Hy)
    mh.h_epl_grad(p(f32(prm)), 100, n, p(xs), p(ys), p(hxs), p(hys), p(g), p(ax), p(ay))
    N = epl_alpha(mh, prm, x, y, niter=100)[2]
    X, Y = xs.astype(np.float64), ys.astype(np.float64)
    fd = np.zeros(6)
    for j in range(6):
        h = 1e-5 * max(abs(prm[j]), 1.0)
        pp, pm = list(prm), list(prm)
        pp[j] += h
        pm[j] -= h
        fd[j] = (_epl_contract(pp, X, Y, hxs, hys, N) - _epl_contract(pm, X, Y, hxs, hys, N)) / (2 * h)
    scale = np.abs(fd).max()
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd) + 1e-4 * scale), (g, fd)


def test_epl_convergence_shear_identity(mh):
    """Closed form pin of the EPL spatial Jacobian used by the adjoint:
    d alpha/dx . (1,0) etc. from the centre adjoint equals -FD in x."""
    prm = (1.1, 2.3, 0.2, -0.1, 0.0, 0.0)
    rng = np.random.default_rng(3)
    n = 500
    x, y = rng.normal(size=n), rng.normal(size=n)
    # kappa = (2-t)/2 (b/R)^t: divergence of alpha equals 2 kappa
    h = 1e-6
    X, Y = x.reshape(1, -1), y.reshape(1, -1)
    axp, _ = np64.epl_deriv_hyp2f1(X + h, Y, *prm)
    axm, _ = np64.epl_deriv_hyp2f1(X - h, Y, *prm)
    _, ayp = np64.epl_deriv_hyp2f1(X, Y + h, *prm)
    _, aym = np64.epl_deriv_hyp2f1(X, Y - h, *prm)
    div = ((axp - axm) + (ayp - aym))[0, 0] / (2 * h)
    th, g, e1, e2 = prm[:4]
    phi = math.atan2(e2, e1) / 2
    c = math.hypot(e1, e2)
    q = (1 - c) / (1 + c)
    b = th * math.sqrt(q)
    t = g - 1
    u = x * math.cos(phi) + y * math.sin(phi)
    v = -x * math.sin(phi) + y * math.cos(phi)
    R = np.sqrt(q * q * u * u + v * v)
    kappa = (2 - t) / 2 * (b / R) ** t
    assert np.allclose(div, 2 * kappa, rtol=1e-5)


@pytest.mark.parametrize("ell", [1, 0])
def test_sersic_forward_and_adjoint(mh, ell):
    rng = np.random.default_rng(4)
    n = 2000
    x, y = rng.normal(size=n), rng.normal(size=n)
    if ell:
        prm = [0.8, 2.5, 0.09, 0.15, 0.1, -0.05, 499.0]
        names = ["R_sersic", "n_sersic", "e1", "e2", "center_x", "center_y", "Ie"]
    else:
        prm = [0.8, 2.5, 0.1, -0.05, 499.0]
        names = ["R_sersic", "n_sersic", "center_x", "center_y", "Ie"]
    gup = rng.normal(size=n)
    xs, ys, gs = f32(x), f32(y), f32(gup)
    I = np.empty(n, np.float32)
    gp = np.zeros(7, np.float32)
    gdx, gdy = np.empty(n, np.float32), np.empty(n, np.float32)
    mh.h_sersic(p(f32(prm)), ell, n, p(xs), p(ys), p(gs), p(I), p(gp), p(gdx), p(gdy))

    def light(pr, X, Y):
        kw = dict(zip(names, pr))
        fn = np64.sersic_ellipse_light if ell else np64.sersic_light
        return fn(X, Y, **kw)[0]
    X, Y = xs.astype(np.float64).reshape(1, -1), ys.astype(np.float64).reshape(1, -1)
    ref = light(prm, X, Y)[0]
    assert np.allclose(I, ref, rtol=2e-5)
    # known answer of tests/test_profiles.py:17-26: light(0, 1) = Ie at R_sersic=1, e=0
    if ell:
        ka = np64.sersic_ellipse_light(np.zeros((1, 1)), np.ones((1, 1)), 1.0, 2.0, 0.0, 0.0, 0.0, 0.0, Ie=5.0)
        assert math.isclose(float(ka.flat[0]), 5.0)
    fd = np.zeros(len(prm))
    for j in range(len(prm)):
        h = 1e-6 * max(abs(prm[j]), 1.0)
        pp, pm = list(prm), list(prm)
        pp[j] += h
        pm[j] -= h
        fd[j] = np.sum(gs * (light(pp, X, Y)[0] - light(pm, X, Y)[0])) / (2 * h)
    gmap = gp if ell else gp[[0, 1, 4, 5, 6]]
    scale = np.abs(fd).max()
    assert np.all(np.abs(gmap - fd) <= 1e-3 * np.abs(fd) + 1e-5 * scale), (gmap, fd)
    # position adjoint (used to chain a source into the lenses)
    h = 1e-6
    dIdx = (light(prm, X + h, Y)[0] - light(prm, X - h, Y)[0]) / (2 * h)
    dIdy = (light(prm, X, Y + h)[0] - light(prm, X, Y - h)[0]) / (2 * h)
    assert np.allclose(gdx, gs * dIdx, rtol=1e-3, atol=1e-4 * np.abs(gs * dIdx).max())
    assert np.allclose(gdy, gs * dIdy, rtol=1e-3, atol=1e-4 * np.abs(gs * dIdy).max())


def test_sis_forward_and_adjoint(mh):
    rng = np.random.default_rng(5)
    n = 1000
    x, y, Hx, Hy = (f32(rng.normal(size=n)) for _ in range(4))
    prm = [1.2, 0.05, -0.1]
    ax, ay, g = np.empty(n, np.float32), np.empty(n, np.float32), np.zeros(3, np.float32)
    mh.h_sis(p(f32(prm)), n, p(x), p(y), p(Hx), p(Hy), p(ax), p(ay), p(g))
    X, Y = x.astype(np.float64).reshape(1, -1), y.astype(np.float64).reshape(1, -1)
    rx, ry = np64.sis_deriv(X, Y, *prm)
    assert np.allclose(ax, rx[0, 0], rtol=1e-5) and np.allclose(ay, ry[0, 0], rtol=1e-5)
    fd = np.zeros(3)
    for j in range(3):
        pp, pm = list(prm), list(prm)
        pp[j] += 1e-6
        pm[j] -= 1e-6
        a1 = np64.sis_deriv(X, Y, *pp)
        a0 = np64.sis_deriv(X, Y, *pm)
        fd[j] = np.sum(Hx * (a1[0] - a0[0])[0, 0] + Hy * (a1[1] - a0[1])[0, 0]) / 2e-6
    assert np.allclose(g, fd, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("prm", [[0.8, 2.5, 0.12, 1.6, 0.4, 0.09, 0.15, 0.1, -0.05, 499.0],
                                 [0.5, 1.2, 0.05, 3.0, -0.3, -0.2, 0.05, -0.1, 0.2, 20.0]])
def test_core_sersic_forward_and_adjoint(mh, prm):
    """Core-Sersic (sersic.py:98-121): forward vs the FP64 restatement, the 10 parameter
    gradients and the position adjoint vs FP64 central differences."""
    names = ["R_sersic", "n_sersic", "Rb", "alpha", "gamma", "e1", "e2", "center_x", "center_y", "Ie"]
    rng = np.random.default_rng(9)
    n = 2000
    x, y = rng.normal(size=n), rng.normal(size=n)
    gup = rng.normal(size=n)
    xs, ys, gs = f32(x), f32(y), f32(gup)
    I = np.empty(n, np.float32)
    gp = np.zeros(10, np.float32)
    gdx, gdy = np.empty(n, np.float32), np.empty(n, np.float32)
    mh.h_core_sersic(p(f32(prm)), n, p(xs), p(ys), p(gs), p(I), p(gp), p(gdx), p(gdy))

    def light(pr, X, Y):
        return np64.core_sersic_light(X, Y, **dict(zip(names, pr)))[0]
    X, Y = xs.astype(np.float64).reshape(1, -1), ys.astype(np.float64).reshape(1, -1)
    ref = light(prm, X, Y)[0]
    assert np.allclose(I, ref, rtol=5e-6, atol=1e-6 * np.abs(ref).max())
    fd = np.zeros(len(prm))
    for j in range(len(prm)):
        h = 1e-6 * max(abs(prm[j]), 1.0)
        pp, pm = list(prm), list(prm)
        pp[j] += h
        pm[j] -= h
        fd[j] = np.sum(gs * (light(pp, X, Y)[0] - light(pm, X, Y)[0])) / (2 * h)
    scale = np.abs(fd).max()
    assert np.all(np.abs(gp - fd) <= 1e-3 * np.abs(fd) + 1e-5 * scale), (gp, fd)
    h = 1e-6
    dIdx = (light(prm, X + h, Y)[0] - light(prm, X - h, Y)[0]) / (2 * h)
    dIdy = (light(prm, X, Y + h)[0] - light(prm, X, Y - h)[0]) / (2 * h)
    assert np.allclose(gdx, gs * dIdx, rtol=1e-3, atol=1e-4 * np.abs(gs * dIdx).max())
    assert np.allclose(gdy, gs * dIdy, rtol=1e-3, atol=1e-4 * np.abs(gs * dIdy).max())


def _mass_check(mh, fn, oracle, prm, x, y, Hx, Hy, h, rtol_fwd, rtol_grad):
    n = len(x)
    ax, ay, g = np.empty(n, np.float32), np.empty(n, np.float32), np.zeros(len(prm), np.float32)
    fn(p(f32(prm)), n, p(x), p(y), p(Hx), p(Hy), p(ax), p(ay), p(g))
    X, Y = x.astype(np.float64), y.astype(np.float64)
    rx, ry = oracle(X, Y, *prm)
    rx, ry = rx.reshape(-1), ry.reshape(-1)
    scale = max(np.abs(rx).max(), np.abs(ry).max())
    assert np.abs(ax - rx).max() <= rtol_fwd * scale and np.abs(ay - ry).max() <= rtol_fwd * scale
    fd = np.zeros(len(prm))
    for j in range(len(prm)):
        pp, pm = list(prm), list(prm)
        step = h * max(abs(prm[j]), 0.1)
        pp[j] += step
        pm[j] -= step
        a1, a0 = oracle(X, Y, *pp), oracle(X, Y, *pm)
        fd[j] = np.sum(Hx * (a1[0] - a0[0]).reshape(-1) + Hy * (a1[1] - a0[1]).reshape(-1)) / (2 * step)
    assert np.allclose(g, fd, rtol=rtol_grad, atol=rtol_grad * np.abs(fd).max()), (g, fd)


@pytest.mark.parametrize("prm", [[1.0, 1e-3, 1e-3, 0.0, 0.0], [1.2, 0.1, -0.1, 0.0, 0.0],
                                 [0.9, -0.3, 0.2, 0.05, -0.1], [1.1, 0.02, 0.0, -0.2, 0.1]])
def test_sie_forward_and_adjoint(mh, prm):
    """SIE (sie.py:13-43): deflection vs the FP64 reference formula at the points of
    the reference's test_sie (test_profiles.py:82-95) and others, reverse-mode
    statistics vs FP64 central differences."""
    rng = np.random.default_rng(7)
    n = 2000
    x, y, Hx, Hy = (f32(rng.normal(size=n)) for _ in range(4))
    _mass_check(mh, mh.h_sie, np64.sie_deriv, prm, x, y, Hx, Hy, 1e-5, 2e-6, 1e-4)


@pytest.mark.parametrize("prm", [[1.0, 0.5, 5.0, 0.0, 0.0], [0.3, 0.05, 2.0, 0.1, -0.05],
                                 [2.0, 1.2, 40.0, -0.3, 0.2]])
def test_tnfw_forward_and_adjoint(mh, prm):
    """Truncated NFW (tnfw.py:14-58): the device form (regrouped for FP32) vs the
    oracle's literal FP64 expression, including points inside the 0.001 Rs clamp
    and around x = 1; reverse-mode statistics vs FP64 central differences."""
    rng = np.random.default_rng(8)
    n = 2000
    x, y = f32(rng.normal(size=n) * 1.5), f32(rng.normal(size=n) * 1.5)
    Rs, cx, cy = prm[0], prm[3], prm[4]
    ring = np.linspace(0, 2 * np.pi, 40)
    x[:40] = f32(cx + Rs * (1 + 0.03 * np.sin(5 * ring)) * np.cos(ring))
    y[:40] = f32(cy + Rs * (1 + 0.03 * np.sin(5 * ring)) * np.sin(ring))
    x[40:45] = f32(cx + Rs * np.array([1e-4, 5e-4, 2e-3, 1e-2, 0.05]))
    y[40:45] = f32(cy)
    Hx, Hy = f32(rng.normal(size=n)), f32(rng.normal(size=n))
    _mass_check(mh, mh.h_tnfw, np64.tnfw_deriv, prm, x, y, Hx, Hy, 1e-5, 5e-6, 1e-4)


@pytest.mark.parametrize("kind,args", [(1, (0.0, 0.1)), (2, (math.log(1.25), 0.25)), (3, (0.5, 8.0)),
                                       (4, (2.0, 0.25, 1.2, 2.8))])
@pytest.mark.parametrize("fldj_at_x", [0, 1])
def test_prior_bijector(mh, kind, args, fldj_at_x):
    """TFP-semantics prior log-densities (scipy) and bijector Jacobians (closed form + FD)."""
    name = {1: "Normal", 2: "LogNormal", 3: "Uniform", 4: "TruncatedNormal"}[kind]
    a = f32(list(args) + [0.0] * (8 - len(args)))
    if kind == 4:
        loc, sc, lo, hi = args
        a[4] = math.log(stats.norm.cdf((hi - loc) / sc) - stats.norm.cdf((lo - loc) / sc))
    out = np.zeros(6, np.float32)
    for z in [-1.3, -0.2, 0.0, 0.4, 1.7]:
        mh.h_prior(kind, p(a), ctypes.c_float(z), fldj_at_x, p(out))
        x, fl = np64.bijector_forward(name, args, z)
        assert math.isclose(out[0], float(x), rel_tol=1e-5, abs_tol=1e-6)
        lp = float(np64.prior_logpdf(name, args, float(x)))
        assert math.isclose(out[2], lp, rel_tol=1e-5, abs_tol=1e-5)
        if fldj_at_x:
            _, fl = np64.bijector_forward(name, args, float(x))
        assert math.isclose(out[4], float(fl), rel_tol=1e-5, abs_tol=1e-5)
        h = 1e-5
        xp, _ = np64.bijector_forward(name, args, z + h)
        xm, _ = np64.bijector_forward(name, args, z - h)
        assert math.isclose(out[1], (xp - xm) / (2 * h), rel_tol=1e-4, abs_tol=1e-6)
        dl = (np64.prior_logpdf(name, args, xp) - np64.prior_logpdf(name, args, xm)) / (xp - xm) if xp != xm else 0
        assert math.isclose(out[3], float(dl), rel_tol=1e-3, abs_tol=1e-4)
        if fldj_at_x:
            fp = np64.bijector_forward(name, args, float(xp))[1]
            fm = np64.bijector_forward(name, args, float(xm))[1]
        else:
            fp = np64.bijector_forward(name, args, z + h)[1]
            fm = np64.bijector_forward(name, args, z - h)[1]
        assert math.isclose(out[5], float(fp - fm) / (2 * h), rel_tol=1e-3, abs_tol=1e-4)


def test_likelihood_terms(mh):
    out = np.zeros(4, np.float32)
    for im, obs in [(3.0, 2.5), (-0.2, 0.1), (0.0, 0.3), (40.0, 44.0)]:
        mh.h_lik(0, ctypes.c_float(im), ctypes.c_float(obs), ctypes.c_float(0.04), ctypes.c_float(0.01), p(out))
        sig = math.sqrt(0.04 + max(im, 0) * 0.01)
        ll = stats.norm.logpdf(obs, im, sig) + 0.5 * math.log(2 * math.pi)
        assert math.isclose(out[0], ll, rel_tol=1e-5, abs_tol=1e-5)
        assert math.isclose(out[1], ((obs - im) / sig) ** 2, rel_tol=1e-5)
        h = 1e-4

        def f(m):
            s = math.sqrt(0.04 + max(m, 0) * 0.01)
            return stats.norm.logpdf(obs, m, s), ((obs - m) / s) ** 2
        if im != 0.0:
            assert math.isclose(out[2], (f(im + h)[0] - f(im - h)[0]) / (2 * h), rel_tol=1e-3, abs_tol=1e-3)
            assert math.isclose(out[3], (f(im + h)[1] - f(im - h)[1]) / (2 * h), rel_tol=1e-3, abs_tol=1e-3)
