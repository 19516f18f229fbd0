"""PSF pre-processing on the host (runs once per simulator).

``subgrid_kernel`` restates the behaviour of lenstronomy 1.9.3
``Util.kernel_util.subgrid_kernel`` as called at
``src/gigalens/tf/simulator.py:47-50`` (lenstronomy is a third-party
dependency of the reference and is not installed here): bilinear
up-sampling of the pixel PSF to the supersampled grid (odd size), then
``num_iter`` corrections so that averaging the result back onto the pixel
grid reproduces the input, normalised to unit sum.

``gaussian`` / ``moffat`` generate pixel-sampled analytic PSFs.
"""
from __future__ import annotations

import math

import numpy as np


def _lin_interp_matrix(x_in: np.ndarray, x_out: np.ndarray) -> np.ndarray:
    idx = np.clip(np.searchsorted(x_in, x_out, side="right") - 1, 0, len(x_in) - 2)
    a = (x_out - x_in[idx]) / (x_in[idx + 1] - x_in[idx])
    a = np.clip(a, 0.0, 1.0)  # constant extrapolation outside the input range
    L = np.zeros((len(x_out), len(x_in)))
    rows = np.arange(len(x_out))
    L[rows, idx] += 1.0 - a
    L[rows, idx + 1] += a
    return L


def _resize(values, x_in, y_in, x_out, y_out):
    return _lin_interp_matrix(y_in, y_out) @ values @ _lin_interp_matrix(x_in, x_out).T


def _average_even(kh: np.ndarray, ss: int) -> np.ndarray:
    """Coarse-grid average of an odd, centred kernel for an even factor: fine
    pixels on a coarse border contribute 1/2 (edge) or 1/4 (corner) to each
    neighbouring coarse pixel."""
    n_h = kh.shape[0]
    n_low = int(round(n_h / ss + 0.5))
    n_low += 1 - n_low % 2
    n_high = n_low * ss - 1
    if n_high != n_h:
        pad = (n_high - n_h) // 2
        kh = np.pad(kh, pad)
    # weight of fine index i in coarse index I along one axis
    w1 = np.zeros((n_low, n_high))
    for i in range(n_high):
        I, r = divmod(i, ss)
        if r < ss - 1:
            w1[I, i] = 1.0
        else:
            w1[I, i] += 0.5
            if I + 1 < n_low:
                w1[I + 1, i] += 0.5
    return w1 @ kh @ w1.T


def subgrid_kernel(kernel, subgrid_res: int, odd: bool = True, num_iter: int = 100) -> np.ndarray:
    kernel = np.asarray(kernel, dtype=np.float64)
    ss = int(subgrid_res)
    if ss == 1:
        return kernel.copy()
    nx, ny = kernel.shape
    x_in = (np.arange(nx) + 0.5) / nx
    y_in = (np.arange(ny) + 0.5) / nx
    nxn, nyn = nx * ss, ny * ss
    if odd:
        nxn -= 1 - nxn % 2
        nyn -= 1 - nyn % 2
    x_out = (np.arange(nxn) + 0.5) / nxn
    y_out = (np.arange(nyn) + 0.5) / nyn
    cur = kernel.copy()
    sub = _resize(cur, x_in, y_in, x_out, y_out)
    sub /= sub.sum()
    for _ in range(max(num_iter, 1)):
        if ss % 2 == 0:
            back = _average_even(sub, ss)
        else:
            back = sub.reshape(nx, ss, ny, ss).mean(axis=(1, 3))
        cur = cur + (kernel - back)
        sub = _resize(cur, x_in, y_in, x_out, y_out)
        sub /= sub.sum()
    if ss % 2 == 0:
        return sub
    back = sub.reshape(nx, ss, ny, ss).mean(axis=(1, 3))
    back /= back.sum()
    delta = np.kron(back - kernel / kernel.sum(), np.ones((ss, ss))) / ss ** 2
    out = sub - delta
    return out / out.sum()


def gaussian(fwhm: float, delta_pix: float, size: int) -> np.ndarray:
    """Unit-sum circular Gaussian sampled at pixel centres."""
    sigma = fwhm / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    r = (np.arange(size) - (size - 1) / 2.0) * delta_pix
    g = np.exp(-0.5 * (r[:, None] ** 2 + r[None, :] ** 2) / sigma ** 2)
    return (g / g.sum()).astype(np.float32)


def moffat(fwhm: float, beta: float, delta_pix: float, size: int) -> np.ndarray:
    """Unit-sum circular Moffat profile (1 + r^2/alpha^2)^-beta sampled at pixel centres."""
    alpha = fwhm / (2.0 * math.sqrt(2.0 ** (1.0 / beta) - 1.0))
    r = (np.arange(size) - (size - 1) / 2.0) * delta_pix
    m = (1.0 + (r[:, None] ** 2 + r[None, :] ** 2) / alpha ** 2) ** (-beta)
    return (m / m.sum()).astype(np.float32)
