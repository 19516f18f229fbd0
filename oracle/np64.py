"""FP64 NumPy restatement of the GIGA-Lens TensorFlow forward model (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in ``gigalens_amd`` imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

Every function cites the reference file:line whose semantics it restates
(paths relative to the reference repository root).  Arrays are float64
unless stated otherwise.  Batched parameters are 1-D arrays of length ``bs``;
images are ``(bs, H, W)``.

Pinning (what checks each function beyond retyping):
  * ``epl_deriv``      -- scipy ``hyp2f1`` closed form, q=1 power law, the
                          reference SIE formula at gamma=2 (tests/test_oracle_profiles.py)
  * ``shear_deriv``    -- closed form
  * ``sie_deriv``      -- equals ``epl_deriv`` (hyp2f1 form) at gamma = 2
  * ``tnfw_deriv``     -- the NFW closed form in the limit r_trunc -> inf, and
                          alpha R equal to the projected (cylinder) mass of the
                          truncated-NFW 3-D density integrated with scipy quad
                          (tests/test_oracle_tnfw.py)
  * ``sersic_*``       -- known answer light(0,1)=Ie (tests/test_profiles.py:17-26)
                          and an independent rotated-ellipse construction
  * ``core_sersic_light`` -- at Rb -> 0 and alpha = 1/n it is e^{-1-bn} times the
                          (pinned) Sersic ellipse; at R = Rb its core factor is
                          2^{gamma/alpha}; at general alpha the Rb -> 0 log-profile has
                          log-log slope alpha and value bn/(alpha n) at R = R_sersic, the
                          inner logarithmic slope tends to -gamma and the outer core
                          factor to 1 + (gamma/alpha)(Rb/R)^alpha (tests/test_oracle.py)
  * ``shapelet_*``     -- numpy.polynomial.hermite, orthonormality integrals
  * ``conv_pool``      -- scipy.signal.convolve2d + block mean
  * ``log_prob_*``     -- scipy.stats log-pdfs, closed-form log-Jacobians
  * ``lstsq``          -- numpy.linalg.pinv / lstsq
  * ``subgrid_kernel`` -- lenstronomy 1.9.3 (absent here): its re-sampling equals
                          scipy RegularGridInterpolator, its even-factor averaging the
                          sub-pixel area overlap, its even-ss output the directly solved
                          fixed point of "re-averaging reproduces the input", its odd-ss
                          output re-averages exactly; planted mistakes (nearest
                          up-sampling, shifted grid, no correction, whole-pixel
                          averaging) each fail a check (tests/test_oracle_subgrid.py).
                          lenstronomy's own floating-point output is not available.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

LN2PI_HALF = 0.5 * math.log(2.0 * math.pi)


# --------------------------------------------------------------------------
# Camera / grid
# --------------------------------------------------------------------------
def get_coords(supersample: int, num_pix: int, transform_pix2angle: np.ndarray):
    """Sub-pixel coordinate grid, ``src/gigalens/simulator.py:91-125``.

    lenstronomy ``PixelGrid`` (simulator.py:112-124) places pixel (i, j)
    (row i, column j) at ``M @ (j, i) + (ra0, dec0)`` with
    ``(ra0, dec0) = M @ (lo, lo)``, ``lo = -(n-1)/2``.  Returned as float32
    like the reference (simulator.py:121-124).
    """
    n = supersample * num_pix
    lo = np.arange(0, n, dtype=np.float32)
    lo = np.min(lo - np.mean(lo))
    M = np.asarray(transform_pix2angle, dtype=np.float32).astype(np.float64)
    ra0, dec0 = (M.astype(np.float32) @ np.array([lo, lo], dtype=np.float32)).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64))
    x = jj * M[0, 0] + ii * M[0, 1] + ra0
    y = jj * M[1, 0] + ii * M[1, 1] + dec0
    return float(ra0), float(dec0), x.astype(np.float32), y.astype(np.float32)


def _interp_matrix(x_in: np.ndarray, x_out: np.ndarray) -> np.ndarray:
    """Linear interpolation with constant (nearest) extrapolation, as a matrix."""
    L = np.zeros((len(x_out), len(x_in)))
    for r, xo in enumerate(x_out):
        if xo <= x_in[0]:
            L[r, 0] = 1.0
        elif xo >= x_in[-1]:
            L[r, -1] = 1.0
        else:
            k = int(np.searchsorted(x_in, xo) - 1)
            k = min(max(k, 0), len(x_in) - 2)
            a = (xo - x_in[k]) / (x_in[k + 1] - x_in[k])
            L[r, k] = 1.0 - a
            L[r, k + 1] = a
    return L


def _re_size_array(x_in, y_in, values, x_out, y_out):
    # bilinear interpolation of values[y, x] with nearest extrapolation
    return _interp_matrix(y_in, y_out) @ values @ _interp_matrix(x_in, x_out).T


def _averaging_even_kernel(kh: np.ndarray, ss: int) -> np.ndarray:
    """Down-sample an odd, centred sub-grid kernel by an even factor: sub-pixels
    on a coarse-pixel border are shared equally between the coarse pixels
    they straddle (lenstronomy 1.9.3 ``kernel_util`` behaviour)."""
    n_h = len(kh)
    n_low = int(round(n_h / ss + 0.5))
    if n_low % 2 == 0:
        n_low += 1
    n_high = int(n_low * ss - 1)
    if n_high == n_h:
        ke = kh
    else:
        s = (n_high - n_h) // 2
        ke = np.zeros((n_high, n_high))
        ke[s:s + n_h, s:s + n_h] = kh
    out = np.zeros((n_low, n_low))
    for i in range(ss - 1):
        for j in range(ss - 1):
            out += ke[i::ss, j::ss]
    i = ss - 1
    for j in range(ss - 1):
        out[1:, :] += ke[i::ss, j::ss] / 2
        out[:-1, :] += ke[i::ss, j::ss] / 2
    j = ss - 1
    for i in range(ss - 1):
        out[:, 1:] += ke[i::ss, j::ss] / 2
        out[:, :-1] += ke[i::ss, j::ss] / 2
    e = ke[ss - 1::ss, ss - 1::ss]
    out[1:, 1:] += e / 4
    out[:-1, 1:] += e / 4
    out[1:, :-1] += e / 4
    out[:-1, :-1] += e / 4
    return out


def subgrid_kernel(kernel: np.ndarray, subgrid_res: int, odd: bool = True, num_iter: int = 100) -> np.ndarray:
    """PSF on the supersampled grid, as called at ``src/gigalens/tf/simulator.py:47-50``.

    Restates lenstronomy 1.9.3 ``Util.kernel_util.subgrid_kernel`` (third-party,
    not vendored in the reference): bilinear up-sampling, then ``num_iter``
    corrections so that re-averaging to the coarse grid reproduces the input,
    normalised to unit sum.  Pinned by independent constructions
    (tests/test_oracle_subgrid.py); lenstronomy's own output is not available.
    """
    kernel = np.asarray(kernel, dtype=np.float64)
    ss = int(subgrid_res)
    if ss == 1:
        return kernel
    nx, ny = kernel.shape
    x_in = np.linspace(0.5 / nx, 1 - 0.5 / nx, nx)
    y_in = np.linspace(0.5 / nx, 1 - 0.5 / nx, ny)
    nxn, nyn = nx * ss, ny * ss
    if odd:
        if nxn % 2 == 0:
            nxn -= 1
        if nyn % 2 == 0:
            nyn -= 1
    x_out = np.linspace(0.5 / nxn, 1 - 0.5 / nxn, nxn)
    y_out = np.linspace(0.5 / nyn, 1 - 0.5 / nyn, nyn)
    k_in = kernel.copy()
    ks = _re_size_array(x_in, y_in, k_in, x_out, y_out)
    ks = ks / ks.sum()
    for _ in range(max(num_iter, 1)):
        if ss % 2 == 0:
            kp = _averaging_even_kernel(ks, ss)
        else:
            kp = ks.reshape(nx, ss, ny, ss).mean(3).mean(1)
        tmp = k_in + (kernel - kp)
        ks = _re_size_array(x_in, y_in, tmp, x_out, y_out)
        ks = ks / ks.sum()
        k_in = tmp
    if ss % 2 == 0:
        return ks
    kp = ks.reshape(nx, ss, ny, ss).mean(3).mean(1)
    kp = kp / kp.sum()
    dk = kp - kernel / kernel.sum()
    dks = np.kron(dk, np.ones((ss, ss))) / ss ** 2
    out = ks - dks
    return out / out.sum()


def gaussian_psf(fwhm: float, delta_pix: float, size: int) -> np.ndarray:
    """Pixel-sampled, unit-sum circular Gaussian PSF (kernel generator)."""
    sigma = fwhm / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    c = (size - 1) / 2.0
    r = (np.arange(size) - c) * delta_pix
    g = np.exp(-0.5 * (r[:, None] ** 2 + r[None, :] ** 2) / sigma ** 2)
    return g / g.sum()


def moffat_psf(fwhm: float, beta: float, delta_pix: float, size: int) -> np.ndarray:
    """Pixel-sampled, unit-sum circular Moffat PSF (kernel generator)."""
    alpha = fwhm / (2.0 * math.sqrt(2.0 ** (1.0 / beta) - 1.0))
    c = (size - 1) / 2.0
    r = (np.arange(size) - c) * delta_pix
    rr = r[:, None] ** 2 + r[None, :] ** 2
    m = (1.0 + rr / alpha ** 2) ** (-beta)
    return m / m.sum()


# --------------------------------------------------------------------------
# Mass profiles
# --------------------------------------------------------------------------
def _rotate(x, y, phi):
    # src/gigalens/tf/profiles/mass/epl.py:63-68
    c, s = np.cos(phi), np.sin(phi)
    return x * c + y * s, -x * s + y * c


def epl_series_length(f_max: np.ndarray, niter: int = 50) -> int:
    """Iterations the reference's while-loop runs (epl.py:41,51-58): ``n``
    starts at 1 and runs while ``n < log(1e-12)/log(max f) + 2``, at most
    ``niter`` times."""
    with np.errstate(divide="ignore", invalid="ignore"):
        lim = math.log(1e-12) / np.log(np.float64(f_max)) + 2.0
    if not np.isfinite(lim):
        lim = -np.inf if lim < 0 or np.isnan(lim) else lim
    if lim == -np.inf:
        return 0
    n = int(max(0, math.ceil(lim) - 1))
    return min(n, niter)


def epl_deriv(x, y, theta_E, gamma, e1, e2, center_x, center_y, niter: int = 50,
              n_terms: Optional[Sequence[int]] = None):
    """EPL deflection, ``src/gigalens/tf/profiles/mass/epl.py:23-61``.

    ``x, y``: (H, W) grids; parameters: (bs,).  Returns (bs, H, W) pairs.
    ``n_terms`` (per particle) overrides the series length; by default the
    reference's batch-coupled length (epl.py:41) is used.
    """
    th, g, e1, e2, cx, cy = [np.asarray(v, np.float64).reshape(-1, 1, 1)
                             for v in (theta_E, gamma, e1, e2, center_x, center_y)]
    phi = np.arctan2(e2, e1) / 2
    c = np.clip(np.sqrt(e1 ** 2 + e2 ** 2), 0, 1)
    q = (1 - c) / (1 + c)
    with np.errstate(divide="ignore", invalid="ignore"):
        theta_E_conv = th / np.sqrt((1.0 + q ** 2) / (2.0 * q))
    b = theta_E_conv * np.sqrt((1 + q ** 2) / 2)
    t = g - 1
    xx, yy = x[None].astype(np.float64) - cx, y[None].astype(np.float64) - cy
    xx, yy = _rotate(xx, yy, phi)
    R = np.clip(np.sqrt((q * xx) ** 2 + yy ** 2), 1e-10, 1e10)
    ang = np.arctan2(yy, q * xx)
    f = (1 - q) / (1 + q)
    Cs, Ss = np.cos(ang), np.sin(ang)
    Cs2, Ss2 = np.cos(2 * ang), np.sin(2 * ang)
    bs = th.shape[0]
    if n_terms is None:
        n_terms = [epl_series_length(np.max(f), niter)] * bs
    n_terms = np.asarray(n_terms).reshape(-1)
    fx, fy = Cs.copy(), Ss.copy()
    lx, ly = Cs.copy(), Ss.copy()
    for n in range(1, int(n_terms.max(initial=0)) + 1):
        active = (n <= n_terms).reshape(-1, 1, 1)
        pref = -f * (2 * n - (2 - t)) / (2 * n + (2 - t))
        nlx = pref * (Cs2 * lx - Ss2 * ly)
        nly = pref * (Ss2 * lx + Cs2 * ly)
        lx = np.where(active, nlx, lx)
        ly = np.where(active, nly, ly)
        fx = np.where(active, fx + lx, fx)
        fy = np.where(active, fy + ly, fy)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        pref = (2 * b) / (1 + q) * np.power(b / R, t - 1)
    fx, fy = fx * pref, fy * pref
    return _rotate(fx, fy, -phi)


def epl_series_length_per_particle(e1, e2, niter: int = 50) -> np.ndarray:
    """Per-particle series length: the reference's rule (epl.py:41) applied to
    each particle's own f instead of the batch maximum."""
    c = np.clip(np.sqrt(np.asarray(e1, np.float64) ** 2 + np.asarray(e2, np.float64) ** 2), 0, 1)
    q = (1 - c) / (1 + c)
    f = (1 - q) / (1 + q)
    return np.array([epl_series_length(fi, niter) for fi in np.atleast_1d(f)])


def epl_deriv_hyp2f1(x, y, theta_E, gamma, e1, e2, center_x, center_y):
    """Closed form (Tessore & Metcalf 2015) via scipy's Gauss hypergeometric
    function: alpha = 2b/(1+q) (b/R)^(t-1) e^{i phi} 2F1(1, t/2; 2-t/2; -f e^{2 i phi}).
    Independent of the series; used to pin ``epl_deriv``."""
    from scipy.special import hyp2f1
    th, g, e1, e2, cx, cy = [np.asarray(v, np.float64).reshape(-1, 1, 1)
                             for v in (theta_E, gamma, e1, e2, center_x, center_y)]
    phi = np.arctan2(e2, e1) / 2
    c = np.clip(np.sqrt(e1 ** 2 + e2 ** 2), 0, 1)
    q = (1 - c) / (1 + c)
    b = th * np.sqrt(q)
    t = g - 1
    xx, yy = _rotate(x[None] - cx, y[None] - cy, phi)
    R = np.sqrt((q * xx) ** 2 + yy ** 2)
    ang = np.arctan2(yy, q * xx)
    f = (1 - q) / (1 + q)
    z = -f * np.exp(2j * ang)
    F = np.exp(1j * ang) * hyp2f1(1.0, t / 2, 2 - t / 2, z)
    a = 2 * b / (1 + q) * (b / R) ** (t - 1) * F
    return _rotate(a.real, a.imag, -phi)


def sie_deriv(x, y, theta_E, e1, e2, center_x, center_y):
    """SIE deflection, ``src/gigalens/tf/profiles/mass/sie.py:13-43`` (s_scale=0)."""
    th, e1, e2, cx, cy = [np.asarray(v, np.float64).reshape(-1, 1, 1)
                          for v in (theta_E, e1, e2, center_x, center_y)]
    phi = np.arctan2(e2, e1) / 2
    c = np.minimum(np.sqrt(e1 ** 2 + e2 ** 2), 0.9999)
    q = (1 - c) / (1 + c)
    b = th / np.sqrt((1.0 + q ** 2) / (2.0 * q)) * np.sqrt((1 + q ** 2) / 2)
    s = 0.0
    xx, yy = _rotate(x[None] - cx, y[None] - cy, phi)
    psi = np.sqrt(q ** 2 * (s ** 2 + xx ** 2) + yy ** 2)
    sq = np.sqrt(1.0 - q ** 2)
    fx = b / sq * np.arctan(sq * xx / (psi + s))
    fy = b / sq * np.arctanh(sq * yy / (psi + q ** 2 * s))
    return _rotate(fx, fy, -phi)


def tnfw_F(x):
    """F(x) of ``src/gigalens/tf/profiles/mass/tnfw.py:40-58``: atanh/atan forms, 1 at x = 1."""
    x = np.asarray(x, np.float64)
    out = np.ones_like(x)
    lo, hi = x < 1, x > 1
    s = np.sqrt(1 - x[lo] ** 2)
    out[lo] = np.arctanh(s) / s
    s = np.sqrt(x[hi] ** 2 - 1)
    out[hi] = np.arctan(s) / s
    return out


def tnfw_deriv(x, y, Rs, alpha_Rs, r_trunc, center_x, center_y):
    """Truncated-NFW deflection, ``src/gigalens/tf/profiles/mass/tnfw.py:14-38``."""
    Rs, aR, rt, cx, cy = [np.asarray(v, np.float64).reshape(-1, 1, 1)
                          for v in (Rs, alpha_Rs, r_trunc, center_x, center_y)]
    rho0 = aR / (4.0 * Rs ** 2 * (1.0 + np.log(0.5)))
    dx, dy = x[None] - cx, y[None] - cy
    R = np.sqrt(dx ** 2 + dy ** 2)
    R = np.maximum(R, 0.001 * Rs)
    xx = R / Rs
    tau = rt / Rs
    L = np.log(xx / (tau + np.sqrt(tau ** 2 + xx ** 2)))
    F = tnfw_F(xx)
    gx = (tau ** 2) / (tau ** 2 + 1) ** 2 * (
        (tau ** 2 + 1 + 2 * (xx ** 2 - 1)) * F + tau * np.pi + (tau ** 2 - 1) * np.log(tau)
        + np.sqrt(tau ** 2 + xx ** 2) * (-np.pi + L * (tau ** 2 - 1) / tau))
    a = 4 * rho0 * Rs * gx / xx ** 2
    return a * dx, a * dy


def sis_deriv(x, y, theta_E, center_x, center_y):
    """SIS deflection, ``src/gigalens/tf/profiles/mass/sis.py:12-18``."""
    th, cx, cy = [np.asarray(v, np.float64).reshape(-1, 1, 1) for v in (theta_E, center_x, center_y)]
    dx, dy = x[None] - cx, y[None] - cy
    R = np.sqrt(dx ** 2 + dy ** 2)
    a = np.where(R == 0, 0.0, th / np.where(R == 0, 1.0, R))
    return a * dx, a * dy


def shear_deriv(x, y, gamma1, gamma2):
    """External shear, ``src/gigalens/tf/profiles/mass/shear.py:12-14``."""
    g1 = np.asarray(gamma1, np.float64).reshape(-1, 1, 1)
    g2 = np.asarray(gamma2, np.float64).reshape(-1, 1, 1)
    x, y = x[None].astype(np.float64), y[None].astype(np.float64)
    return g1 * x + g2 * y, g2 * x - g1 * y


# --------------------------------------------------------------------------
# Light profiles
# --------------------------------------------------------------------------
def _as_b(v):
    return np.asarray(v, np.float64).reshape(-1, 1, 1)


def sersic_distance(x, y, cx, cy, e1=None, e2=None):
    """Elliptical radius, ``src/gigalens/tf/profiles/light/sersic.py:33-59``.
    ``x, y`` may be (H, W) (shared grid) or (bs, H, W) (ray-traced)."""
    cx, cy = _as_b(cx), _as_b(cy)
    e1 = np.zeros_like(cx) if e1 is None else _as_b(e1)
    e2 = np.zeros_like(cx) if e2 is None else _as_b(e2)
    phi = np.arctan2(e2, e1) / 2
    c = np.sqrt(e1 ** 2 + e2 ** 2)
    q = (1 - c) / (1 + c)
    x = x if x.ndim == 3 else x[None]
    y = y if y.ndim == 3 else y[None]
    dx, dy = x - cx, y - cy
    cp, sp = np.cos(phi), np.sin(phi)
    xt1 = (cp * dx + sp * dy) * np.sqrt(q)
    xt2 = (-sp * dx + cp * dy) / np.sqrt(q)
    return np.sqrt(xt1 ** 2 + xt2 ** 2)


def sersic_ellipse_light(x, y, R_sersic, n_sersic, e1, e2, center_x, center_y, Ie=None):
    """``SersicEllipse.light``, ``src/gigalens/tf/profiles/light/sersic.py:66-72``.
    ``Ie=None`` is the least-squares (unit amplitude) mode."""
    R = sersic_distance(x, y, center_x, center_y, e1, e2)
    n = _as_b(n_sersic)
    bn = 1.9992 * n - 0.3271
    Ie = 1.0 if Ie is None else _as_b(Ie)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return Ie * np.exp(-bn * ((R / _as_b(R_sersic)) ** (1 / n) - 1.0))


def core_sersic_light(x, y, R_sersic, n_sersic, Rb, alpha, gamma, e1, e2, center_x, center_y, Ie=None):
    """``CoreSersic.light``, ``src/gigalens/tf/profiles/light/sersic.py:98-121``, as written there:
    ``R_sersic ** alpha ** 1.0`` is ``R_sersic ** alpha`` and the ``- 1.0`` sits outside the
    ``bn`` product."""
    R = sersic_distance(x, y, center_x, center_y, e1, e2)
    n, Rs, Rb, al, gam = (_as_b(v) for v in (n_sersic, R_sersic, Rb, alpha, gamma))
    bn = 1.9992 * n - 0.3271
    Ie = 1.0 if Ie is None else _as_b(Ie)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return (Ie * (1 + (Rb / R) ** al) ** (gam / al)
                * np.exp(-bn * ((R ** al + Rb ** al) / Rs ** al ** 1.0 / (al * n)) - 1.0))


def sersic_light(x, y, R_sersic, n_sersic, center_x, center_y, Ie=None):
    """Circular ``Sersic.light``, ``src/gigalens/tf/profiles/light/sersic.py:21-31``."""
    R = sersic_distance(x, y, center_x, center_y)
    n = _as_b(n_sersic)
    bn = 1.9992 * n - 0.3271
    Ie = 1.0 if Ie is None else _as_b(Ie)
    return Ie * np.exp(-bn * ((R / _as_b(R_sersic)) ** (1 / n) - 1.0))


def hermite_phi(n: int, x: np.ndarray) -> np.ndarray:
    """Orthonormal 1-D Hermite function phi_n(x) = H_n(x) e^{-x^2/2} / sqrt(2^n sqrt(pi) n!)
    (lenstronomy ``Shapelets.phi_n``, used at ``shapelets.py:39-40``)."""
    from numpy.polynomial.hermite import hermval
    coef = np.zeros(n + 1)
    coef[n] = 1.0
    pref = 1.0 / math.sqrt(2 ** n * math.sqrt(math.pi) * math.factorial(n))
    return pref * hermval(x, coef) * np.exp(-x ** 2 / 2)


def shapelet_orders(n_max: int) -> List[Tuple[int, int]]:
    """Basis order (n1, n2), ``src/gigalens/tf/profiles/light/shapelets.py:27-46``:
    (0,0),(1,0),(0,1),(2,0),(1,1),(0,2),..."""
    out = []
    n1 = n2 = 0
    for _ in range((n_max + 1) * (n_max + 2) // 2):
        out.append((n1, n2))
        if n1 == 0:
            n1, n2 = n2 + 1, 0
        else:
            n1, n2 = n1 - 1, n2 + 1
    return out


TABLE_N = 6000
TABLE_LO, TABLE_HI = -5.0, 5.0


def shapelet_table(n_max: int) -> np.ndarray:
    """(n_max+1, 6000) table of phi_n on linspace(-5, 5, 6000) stored as float32
    (shapelets.py:39-40,50-51)."""
    xs = np.linspace(TABLE_LO, TABLE_HI, TABLE_N)
    return np.stack([hermite_phi(n, xs) for n in range(n_max + 1)]).astype(np.float32)


def interp_regular(x: np.ndarray, table: np.ndarray) -> np.ndarray:
    """Linear interpolation on the regular [-5, 5] grid with zero fill outside
    (tfp.math.interp_regular_1d_grid with fill_value_below/above=0,
    shapelets.py:55-60).  ``table`` (..., 6000) -> (..., *x.shape)."""
    ny = table.shape[-1]
    xi = (ny - 1) * (x - TABLE_LO) / (TABLE_HI - TABLE_LO)
    xc = np.clip(xi, 0, ny - 1)
    lo = np.floor(xc)
    hi = np.minimum(lo + 1, ny - 1)
    lo = np.maximum(hi - 1, 0)
    a = xc - lo
    lo, hi = lo.astype(np.int64), hi.astype(np.int64)
    tb = table.astype(np.float64)
    out = (1 - a) * tb[..., lo] + a * tb[..., hi]
    out = np.where((xi < 0) | (xi > ny - 1), 0.0, out)
    return out


def shapelet_basis(x, y, center_x, center_y, beta, n_max: int, interpolate: bool = True):
    """Unit-amplitude basis images, ``src/gigalens/tf/profiles/light/shapelets.py:53-75``.
    Returns (n_layers, bs, H, W)."""
    x = x if x.ndim == 3 else x[None]
    y = y if y.ndim == 3 else y[None]
    u = (x - _as_b(center_x)) / _as_b(beta)
    v = (y - _as_b(center_y)) / _as_b(beta)
    orders = shapelet_orders(n_max)
    if interpolate:
        tab = shapelet_table(n_max)
        hx = interp_regular(u, tab)  # (n_max+1, bs, H, W)
        hy = interp_regular(v, tab)
        return np.stack([hx[a] * hy[b] for a, b in orders])
    fac = np.exp(-(u ** 2 + v ** 2) / 2)
    hx = [_hermite_norm(n, u) for n in range(n_max + 1)]
    hy = [_hermite_norm(n, v) for n in range(n_max + 1)]
    return np.stack([fac * hx[a] * hy[b] for a, b in orders])


def _hermite_norm(n, x):
    from numpy.polynomial.hermite import hermval
    coef = np.zeros(n + 1)
    coef[n] = 1.0
    return hermval(x, coef) / math.sqrt(2 ** n * math.sqrt(math.pi) * math.factorial(n))


def shapelets_light(x, y, center_x, center_y, beta, amps, n_max: int, interpolate: bool = True):
    """Shapelet light with amplitudes ``amps`` (n_layers, bs)."""
    B = shapelet_basis(x, y, center_x, center_y, beta, n_max, interpolate)
    return np.einsum("lbhw,lb->bhw", B, np.asarray(amps, np.float64).reshape(len(B), -1))


# --------------------------------------------------------------------------
# Simulator
# --------------------------------------------------------------------------
class Camera:
    """Grid + PSF state of ``LensSimulator.__init__``, ``src/gigalens/tf/simulator.py:11-55``."""

    def __init__(self, delta_pix: float, num_pix: int, supersample: int = 1, kernel=None,
                 transform_pix2angle=None, subgrid: Optional[np.ndarray] = None):
        self.ss = int(supersample)
        self.num_pix = int(num_pix)
        T = (np.eye(2, dtype=np.float32) * np.float32(delta_pix) if transform_pix2angle is None
             else np.asarray(transform_pix2angle, np.float32))
        self.conversion_factor = float(np.float32(np.linalg.det(T.astype(np.float64))))
        self.T_sub = (T / np.float32(self.ss)).astype(np.float32)
        _, _, self.x, self.y = get_coords(self.ss, self.num_pix, self.T_sub)
        if kernel is None:
            self.kernel = None
        else:
            k = subgrid_kernel(kernel, self.ss, odd=True) if subgrid is None else subgrid
            self.kernel = np.asarray(k, np.float32)

    @property
    def n_sub(self):
        return self.ss * self.num_pix


def conv_pool(img: np.ndarray, kernel: Optional[np.ndarray], ss: int) -> np.ndarray:
    """SAME true convolution with the sub-grid PSF then ss x ss block mean,
    ``src/gigalens/tf/simulator.py:83-96`` (conv2d with the flipped kernel
    == convolution).  ``img`` (..., H, W)."""
    if kernel is not None:
        from scipy.signal import convolve2d
        flat = img.reshape(-1, *img.shape[-2:])
        k = np.asarray(kernel, np.float64)
        kh = k.shape[0]
        # SAME zero padding; for even sizes TF pads (k-1)//2 before
        pb = (kh - 1) // 2
        out = np.empty_like(flat, dtype=np.float64)
        from scipy.signal import fftconvolve
        for i in range(flat.shape[0]):
            # FFT for large kernels (exact to FP64 rounding); direct otherwise
            full = (fftconvolve(flat[i].astype(np.float64), k, mode="full") if kh > 15
                    else convolve2d(flat[i].astype(np.float64), k, mode="full"))
            # full conv index = out index + (k-1) - pb_conv; true conv with flipped
            # correlation offset: correlation pad pb  -> conv offset kh-1-pb
            off = kh - 1 - pb
            out[i] = full[off:off + flat.shape[1], off:off + flat.shape[2]]
        img = out.reshape(img.shape)
    if ss != 1:
        H, W = img.shape[-2:]
        img = img.reshape(*img.shape[:-2], H // ss, ss, W // ss, ss).mean(axis=(-3, -1))
    return img


def beta_map(cam: Camera, lenses: Sequence[Tuple[str, Dict]]):
    """Ray-traced source-plane coordinates, ``src/gigalens/tf/simulator.py:57-63``."""
    X, Y = cam.x.astype(np.float64), cam.y.astype(np.float64)
    bx, by = X[None], Y[None]
    for kind, p in lenses:
        ax, ay = MASS[kind](cam.x, cam.y, **p)
        bx, by = bx - ax, by - ay
    return bx, by


def _epl(x, y, **p):
    return epl_deriv(x, y, **p)


MASS = {
    "EPL": _epl,
    "SHEAR": lambda x, y, **p: shear_deriv(x, y, **p),
    "SIS": lambda x, y, **p: sis_deriv(x, y, **p),
    "SIE": lambda x, y, **p: sie_deriv(x, y, **p),
    "TNFW": lambda x, y, **p: tnfw_deriv(x, y, **p),
}


def light_eval(kind: str, x, y, p: Dict, n_max=None, interpolate=True):
    if kind == "SERSIC_ELLIPSE":
        return sersic_ellipse_light(x, y, **p)
    if kind == "SERSIC":
        return sersic_light(x, y, **p)
    if kind == "CORE_SERSIC":
        return core_sersic_light(x, y, **p)
    if kind == "SHAPELETS":
        amps = np.stack([np.asarray(p[k], np.float64) for k in sorted(p) if k.startswith("amp")])
        return shapelets_light(x, y, p["center_x"], p["center_y"], p["beta"], amps, n_max, interpolate)
    raise KeyError(kind)


def simulate(cam: Camera, lenses, lens_light, source_light, no_deflection=False):
    """``LensSimulator.simulate``, ``src/gigalens/tf/simulator.py:65-97``.

    ``lenses``: list of (kind, params); lights: list of (kind, params[, n_max, interpolate]).
    Returns (bs, N, N) float64.
    """
    bx, by = beta_map(cam, lenses)
    if no_deflection:
        bx, by = cam.x[None].astype(np.float64), cam.y[None].astype(np.float64)
    img = 0.0
    for item in lens_light:
        kind, p = item[0], item[1]
        img = img + light_eval(kind, cam.x.astype(np.float64), cam.y.astype(np.float64), p, *item[2:])
    for item in source_light:
        kind, p = item[0], item[1]
        img = img + light_eval(kind, bx, by, p, *item[2:])
    img = np.where(np.isnan(img), 0.0, img)
    out = conv_pool(img, cam.kernel, cam.ss)
    return out * cam.conversion_factor


def lstsq_components(cam: Camera, lenses, lens_light, source_light, no_deflection=False):
    """Convolved, pooled unit-amplitude components, ``src/gigalens/tf/simulator.py:109-153``.
    Returns (bs, N, N, depth) (the ``return_stacked`` output)."""
    bx, by = beta_map(cam, lenses)
    if no_deflection:
        bx, by = cam.x[None].astype(np.float64), cam.y[None].astype(np.float64)
    comps = []
    X, Y = cam.x.astype(np.float64), cam.y.astype(np.float64)
    for item in lens_light:
        comps.extend(_unit_components(item, X, Y))
    for item in source_light:
        comps.extend(_unit_components(item, bx, by))
    bs = bx.shape[0]
    comps = [np.broadcast_to(c, (bs,) + c.shape[-2:]) for c in comps]
    img = np.stack(comps, axis=-1)  # (bs, H, W, D)
    img = np.where(np.isnan(img), 0.0, img)
    img = np.moveaxis(img, -1, 1)
    ret = conv_pool(img, cam.kernel, cam.ss)
    ret = np.moveaxis(ret, 1, -1)
    return np.where(np.isnan(ret), 0.0, ret)


def _unit_components(item, x, y):
    kind, p = item[0], item[1]
    if kind == "SERSIC_ELLIPSE":
        q = {k: v for k, v in p.items() if k != "Ie"}
        return [sersic_ellipse_light(x, y, **q)]
    if kind == "SERSIC":
        q = {k: v for k, v in p.items() if k != "Ie"}
        return [sersic_light(x, y, **q)]
    if kind == "SHAPELETS":
        n_max, interp = item[2], item[3]
        B = shapelet_basis(x, y, p["center_x"], p["center_y"], p["beta"], n_max, interp)
        return list(B)
    raise KeyError(kind)


def lstsq_solve(stacked: np.ndarray, observed_image: np.ndarray, err_map: np.ndarray, rcond=1e-6):
    """Weighted linear amplitudes, ``src/gigalens/tf/simulator.py:157-165``.
    Returns (image (bs, N, N), coeffs (bs, D))."""
    bs, H, W, D = stacked.shape
    Wt = 1.0 / np.asarray(err_map, np.float64)
    Yv = (np.asarray(observed_image, np.float64) * Wt).reshape(-1, 1)
    X = (stacked * Wt[None, :, :, None]).reshape(bs, -1, D)
    coeffs = np.empty((bs, D))
    for i in range(bs):
        A = X[i].T @ X[i]
        coeffs[i] = (np.linalg.pinv(A, rcond=rcond) @ X[i].T @ Yv)[:, 0]
    img = np.einsum("bhwd,bd->bhw", stacked, coeffs)
    return img, coeffs


# --------------------------------------------------------------------------
# Likelihood
# --------------------------------------------------------------------------
def forward_loglike(im_sim: np.ndarray, observed: np.ndarray, background_rms: float, exp_time: float):
    """Log-likelihood and mean r^2 of ``ForwardProbModel.log_prob``,
    ``src/gigalens/tf/model.py:77-86``: sigma from the *simulated* image."""
    im = np.asarray(im_sim, np.float64)
    obs = np.asarray(observed, np.float64)
    var = background_rms ** 2 + np.clip(im, 0, np.inf) / exp_time
    sig = np.sqrt(var)
    r = (obs - im) / sig
    ll = np.sum(-0.5 * r ** 2 - np.log(sig) - LN2PI_HALF, axis=(-2, -1))
    return ll, np.mean(r ** 2, axis=(-2, -1))


def backward_err_map(observed, background_rms, exp_time):
    """``BackwardProbModel`` error map from the observed image, ``src/gigalens/tf/model.py:113-115``."""
    obs = np.asarray(observed, np.float64)
    return np.sqrt(background_rms ** 2 + np.clip(obs, 0, np.inf) / exp_time)


def backward_loglike(im_sim, observed, err_map):
    """``src/gigalens/tf/model.py:116-118,159,161-163``."""
    im = np.asarray(im_sim, np.float64)
    obs = np.asarray(observed, np.float64)
    r = (im - obs) / err_map
    ll = np.sum(-0.5 * r ** 2 - np.log(err_map) - LN2PI_HALF, axis=(-2, -1))
    return ll, np.mean(r ** 2, axis=(-2, -1))


# --------------------------------------------------------------------------
# Priors and bijectors (TFP semantics)
# --------------------------------------------------------------------------
def prior_logpdf(kind: str, args: Sequence[float], x: np.ndarray) -> np.ndarray:
    """Log density of the four prior families the reference uses
    (tests/conftest.py:20-73): Normal(loc, scale), LogNormal(loc, scale),
    Uniform(low, high) (support [low, high)), TruncatedNormal(loc, scale, low, high)."""
    from scipy import stats
    x = np.asarray(x, np.float64)
    if kind == "Normal":
        return stats.norm.logpdf(x, args[0], args[1])
    if kind == "LogNormal":
        return stats.lognorm.logpdf(x, s=args[1], scale=math.exp(args[0]))
    if kind == "Uniform":
        lo, hi = args
        return np.where((x >= lo) & (x < hi), -math.log(hi - lo), -np.inf)
    if kind == "TruncatedNormal":
        loc, sc, lo, hi = args
        a, b = (lo - loc) / sc, (hi - loc) / sc
        return stats.truncnorm.logpdf(x, a, b, loc=loc, scale=sc)
    raise KeyError(kind)


def bijector_forward(kind: str, args, z):
    """Default event-space bijector (TFP ``experimental_default_event_space_bijector``,
    used at ``src/gigalens/tf/model.py:50``): Exp for LogNormal, Sigmoid(low, high)
    for Uniform / TruncatedNormal, Identity for Normal.  Returns (x, fldj)."""
    z = np.asarray(z, np.float64)
    if kind == "Normal":
        return z, np.zeros_like(z)
    if kind == "LogNormal":
        return np.exp(z), z
    if kind in ("Uniform", "TruncatedNormal"):
        lo, hi = args[-2], args[-1]
        s = 1.0 / (1.0 + np.exp(-z))
        x = lo + (hi - lo) * s
        fldj = math.log(hi - lo) - np.logaddexp(0, -z) - np.logaddexp(0, z)
        return x, fldj
    raise KeyError(kind)


def bijector_inverse(kind: str, args, x):
    x = np.asarray(x, np.float64)
    if kind == "Normal":
        return x
    if kind == "LogNormal":
        return np.log(x)
    lo, hi = args[-2], args[-1]
    s = (x - lo) / (hi - lo)
    return np.log(s) - np.log1p(-s)
