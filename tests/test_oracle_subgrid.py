"""Pins of the oracle's sub-grid PSF (``np64.subgrid_kernel``, the restatement of
lenstronomy 1.9.3 ``kernel_util.subgrid_kernel`` that src/gigalens/tf/simulator.py:47-50
calls; lenstronomy is not installed here).

Each building block is checked against an independent construction:
  * bilinear re-sampling with nearest extrapolation == scipy RegularGridInterpolator
    on clipped coordinates, with pixel centres at (i + 1/2)/n of the unit square;
  * even-factor averaging == the area overlap of each sub-pixel with the coarse
    pixels (sub-pixels on a coarse border are shared in halves, corners in quarters);
  * for even ss the correction loop converges to the unique fixed point
    ks = R (A R)^-1 k of "re-averaging reproduces the input", solved directly
    with numpy from the two independent matrices above;
  * for odd ss the final step makes the block means reproduce the input exactly.
Known answers: ss = 1 is the identity, a constant kernel gives a uniform one (odd ss), a
centred symmetric PSF stays centrosymmetric and transpose-symmetric, the
centroid of an asymmetric PSF is kept.  ``test_planted_mistakes_are_caught``
plants plausible mistakes in the oracle (on the CPU) and shows that each one
fails at least one of these checks; it also records that the reference's demo
image (a statistical check only) leaves most of them inside its chi^2 band.
"""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import np64
from tests import common as C

PSF = os.path.join(C.GOLDEN, "psf.npy")


def _centres(n):
    return (np.arange(n) + 0.5) / n


def _resize_matrix(n_in, n_out):
    """(n_out^2, n_in^2) matrix of bilinear re-sampling with nearest extrapolation,
    built column by column with scipy (independent of np64._interp_matrix)."""
    c_in, c_out = _centres(n_in), _centres(n_out)
    yy, xx = np.meshgrid(np.clip(c_out, c_in[0], c_in[-1]), np.clip(c_out, c_in[0], c_in[-1]), indexing="ij")
    pts = np.stack([yy.ravel(), xx.ravel()], axis=1)
    R = np.zeros((n_out * n_out, n_in * n_in))
    for j in range(n_in * n_in):
        e = np.zeros(n_in * n_in)
        e[j] = 1.0
        R[:, j] = RegularGridInterpolator((c_in, c_in), e.reshape(n_in, n_in), method="linear")(pts)
    return R


def _area_overlap_matrix(n_high, ss):
    """Averaging of an odd, centred sub-grid kernel (n_high) onto the coarse grid
    by an even factor ss: each sub-pixel contributes in proportion to its overlap
    with each coarse pixel, both grids centred on the same point."""
    n_low = int(round(n_high / ss + 0.5))
    if n_low % 2 == 0:
        n_low += 1
    # positions in sub-pixel units, origin at the common centre
    sub_c = np.arange(n_high) - (n_high - 1) / 2.0
    low_lo = (np.arange(n_low) - (n_low - 1) / 2.0) * ss - ss / 2.0
    w = np.zeros((n_low, n_high))
    for I in range(n_low):
        for i in range(n_high):
            a = max(sub_c[i] - 0.5, low_lo[I])
            b = min(sub_c[i] + 0.5, low_lo[I] + ss)
            w[I, i] = max(0.0, b - a)
    return np.kron(w, w), n_low


# ---- building blocks ---------------------------------------------------------
@pytest.mark.parametrize("n_in,n_out", [(13, 25), (11, 21), (21, 83), (11, 33)])
def test_resize_matches_scipy(n_in, n_out):
    rng = np.random.default_rng(n_in * 100 + n_out)
    v = rng.random((n_in, n_in))
    got = np64._re_size_array(_centres(n_in), _centres(n_in), v, _centres(n_out), _centres(n_out))
    ref = (_resize_matrix(n_in, n_out) @ v.ravel()).reshape(n_out, n_out)
    assert np.abs(got - ref).max() < 1e-14


@pytest.mark.parametrize("n_high,ss", [(25, 2), (21, 2), (83, 4), (51, 4)])
def test_even_averaging_is_area_overlap(n_high, ss):
    rng = np.random.default_rng(n_high)
    kh = rng.random((n_high, n_high))
    A, n_low = _area_overlap_matrix(n_high, ss)
    got = np64._averaging_even_kernel(kh, ss)
    assert got.shape == (n_low, n_low)
    assert np.abs(got.ravel() - A @ kh.ravel()).max() < 1e-13


# ---- the whole algorithm -----------------------------------------------------
def _fixed_point(k, ss):
    """Even ss: the converged sub-grid kernel solves A R t = k_n, ks = R t."""
    n = k.shape[0]
    n_out = n * ss - 1 if (n * ss) % 2 == 0 else n * ss
    R = _resize_matrix(n, n_out)
    A, n_low = _area_overlap_matrix(n_out, ss)
    assert n_low == n
    kn = k / k.sum()
    t = np.linalg.solve(A @ R, kn.ravel())
    return (R @ t).reshape(n_out, n_out)


def _check_even(ks, k, ss, tol=1e-6):
    ref = _fixed_point(k, ss)
    return ks.shape == ref.shape and np.abs(ks - ref).max() <= tol * ref.max()


def _check_odd(ks, k, ss, tol=1e-12):
    n = k.shape[0]
    back = ks.reshape(n, ss, n, ss).sum(axis=(1, 3))
    return abs(ks.sum() - 1.0) < tol and np.abs(back - k / k.sum()).max() <= tol


def _check_symmetric(subgrid, ss):
    from gigalens_amd import psf as gpsf
    g = gpsf.gaussian(0.1, 0.065, 11).astype(np.float64)
    gs = subgrid(g, ss)
    return np.allclose(gs, gs[::-1, ::-1], atol=1e-12 * gs.max()) and np.allclose(gs, gs.T, atol=1e-12 * gs.max())


def _centroid(a, pitch):
    c = (np.arange(a.shape[0]) - (a.shape[0] - 1) / 2.0) * pitch
    return np.array([(a.sum(1) * c).sum(), (a.sum(0) * c).sum()]) / a.sum()


def _check_centroid(ks, k, tol=0.02):
    # pitch of the sub-grid in coarse pixels: the unit square holds k.shape[0] coarse pixels
    return np.abs(_centroid(ks, k.shape[0] / ks.shape[0]) - _centroid(k, 1.0)).max() < tol


@pytest.mark.parametrize("ss", [2, 4])
def test_even_subgrid_is_the_fixed_point(ss):
    for k in (np.load(PSF).astype(np.float64), np64.gaussian_psf(0.1, 0.065, 11),
              np64.moffat_psf(0.08, 3.0, 0.04, 21)):
        ks = np64.subgrid_kernel(k, ss)
        assert _check_even(ks, k, ss), np.abs(ks - _fixed_point(k, ss)).max() / ks.max()


@pytest.mark.parametrize("ss", [3, 5])
def test_odd_subgrid_reaverages_exactly(ss):
    k = np.load(PSF).astype(np.float64)
    assert _check_odd(np64.subgrid_kernel(k, ss), k, ss)


def test_known_answers():
    k = np.load(PSF).astype(np.float64)
    assert np.array_equal(np64.subgrid_kernel(k, 1), k)
    # odd ss: a constant kernel stays uniform (even ss: the border coarse pixels lack
    # their outer half-shared sub-pixels, so the fixed point is not uniform; pinned above)
    for ss in (3, 5):
        u = np64.subgrid_kernel(np.ones((11, 11)), ss)
        assert np.allclose(u, 1.0 / u.size, rtol=1e-12)
    for ss in (2, 3, 4):
        assert _check_symmetric(np64.subgrid_kernel, ss)
        assert _check_centroid(np64.subgrid_kernel(k, ss), k)


# ---- planted mistakes --------------------------------------------------------
def _demo_chi2(ks):
    from gigalens_amd import workloads as wl
    k = np.load(PSF).astype(np.float32)
    cam = np64.Camera(0.065, 60, 2, k, None, subgrid=ks)
    lenses, ll, src = C.oracle_model_lists(wl.epl_shear_model(), wl.DEMO_TRUTH)
    img = np64.simulate(cam, lenses, ll, src)[0]
    obs = np.load(os.path.join(C.GOLDEN, "demo.npy"))
    return float(np64.forward_loglike(img[None], obs, 0.2, 100.0)[1][0])


def _nearest(x_in, y_in, v, x_out, y_out):
    ix = np.abs(x_out[:, None] - x_in[None]).argmin(1)
    iy = np.abs(y_out[:, None] - y_in[None]).argmin(1)
    return v[np.ix_(iy, ix)]


def _shifted(orig):
    def f(x_in, y_in, v, x_out, y_out):
        d = 0.5 / len(x_out)  # half a sub-pixel
        return orig(x_in, y_in, v, x_out + d, y_out + d)
    return f


def _no_correction(kernel, ss, odd=True, num_iter=100):
    # bilinear up-sampling and normalisation only
    n = kernel.shape[0]
    n_out = n * ss - 1 if (odd and (n * ss) % 2 == 0) else n * ss
    ks = np64._re_size_array(_centres(n), _centres(n), np.asarray(kernel, np.float64), _centres(n_out), _centres(n_out))
    return ks / ks.sum()


def _whole_pixels(kh, ss):
    # straddling sub-pixels given wholly to the lower coarse pixel
    n_low = int(round(len(kh) / ss + 0.5))
    n_low += n_low % 2 == 0
    n_high = n_low * ss - 1
    s = (n_high - len(kh)) // 2
    ke = np.zeros((n_high + 1, n_high + 1))
    ke[s:s + len(kh), s:s + len(kh)] = kh
    return ke.reshape(n_low, ss, n_low, ss).sum(axis=(1, 3))


@pytest.mark.parametrize("mistake", ["nearest", "shift", "no_correction", "whole_pixels"])
def test_planted_mistakes_are_caught(monkeypatch, mistake):
    k = np.load(PSF).astype(np.float64)
    good_chi2 = _demo_chi2(np64.subgrid_kernel(k, 2))
    if mistake == "nearest":
        monkeypatch.setattr(np64, "_re_size_array", _nearest)
    elif mistake == "shift":
        monkeypatch.setattr(np64, "_re_size_array", _shifted(np64._re_size_array))
    elif mistake == "no_correction":
        monkeypatch.setattr(np64, "subgrid_kernel", _no_correction)
    else:
        monkeypatch.setattr(np64, "_averaging_even_kernel", _whole_pixels)
    caught = []
    for ss in (2, 4):
        ks = np64.subgrid_kernel(k, ss)
        caught.append(not _check_even(ks, k, ss))
        caught.append(not _check_centroid(ks, k))
    caught.append(not _check_symmetric(np64.subgrid_kernel, 2))
    assert any(caught), mistake
    # the demo image alone: only gross mistakes leave its band (recorded, not required)
    chi2 = _demo_chi2(np64.subgrid_kernel(k, 2))
    print(f"{mistake}: demo chi2 {chi2:.4f} (correct {good_chi2:.4f}, band +-{3 * math.sqrt(2 / 3600) + 0.02:.4f})")
