"""Algorithmic operation counts of the fused lens-model kernels.

Counting rule: add/sub/mul = 1, FMA = 2, each transcendental (exp2, log2,
rsqrt, rcp, sqrt, log) = 1; comparisons, selects, LDS/global moves are not
counted.  Transcendentals are also reported separately.  The counts are the
operation lists of the implemented formulation (gigalens_amd/csrc/lens_math.h
and lens_kernels.hip), per sub-pixel ``S = (ss N)^2`` and per output pixel
``P = N^2``; the EPL series costs 3 flops per term (Clenshaw) and its length
``n_p`` is per particle.

Forward (simulate + log-likelihood), per particle:
  S * (8 coords + sum_c fwd_c + 1 per light (accumulate) + 2 per lens (beta))
  + P * (2 kf^2 conv + 1 scale + 12 likelihood)
Gradient launch (forward + adjoint), additionally:
  S * (sum_c adj_c + 2 kf^2/ss^2 adjoint conv + 2) + P * 15 likelihood adjoint
with kf = k + ss - 1 the PSF size after folding in the ss x ss pooling box.
"""
from __future__ import annotations

from typing import Dict, Iterable, Sequence

import numpy as np

from . import _native as nat

# (flops, transcendentals) per sub-pixel
_FWD = {
    nat.GL_SERSIC: (19, 3),
    nat.GL_SERSIC_ELLIPSE: (19, 3),
    nat.GL_SHEAR: (6, 0),
    nat.GL_SIS: (12, 1),
    nat.GL_SIE: (62, 2),
    nat.GL_TNFW: (55, 10),
    nat.GL_EPL: (44, 3),  # + 3 per series term
}
# adjoint recompute + reverse statistics (the forward is recomputed)
_ADJ = {
    nat.GL_SERSIC: (19 + 24, 3 + 1),
    nat.GL_SERSIC_ELLIPSE: (19 + 24, 3 + 1),
    nat.GL_SHEAR: (6 + 8, 0),
    nat.GL_SIS: (12 + 12, 1),
    nat.GL_SIE: (62 + 38, 2 + 3),
    nat.GL_TNFW: (55 + 40, 10 + 4),
    nat.GL_EPL: (44 + 12 + 59, 3 + 1),  # + 9 per series term (three series)
}


def series_length(e1: np.ndarray, e2: np.ndarray, niter: int) -> np.ndarray:
    """Per-particle EPL series length (epl.py:41 rule on the particle's own f)."""
    c = np.clip(np.hypot(np.asarray(e1, np.float64), np.asarray(e2, np.float64)), 0, 1)
    q = (1 - c) / (1 + c)
    f = (1 - q) / (1 + q)
    with np.errstate(divide="ignore", invalid="ignore"):
        lim = np.log(1e-12) / np.log(f) + 2.0
    n = np.where(f == 0, 1, np.where(lim > 1, np.ceil(lim) - 1, 0))
    return np.minimum(n, niter).astype(np.int64)


def conv_flops(num_pix: int, ss: int, kernel_size: int, sep_rank: int = 0, grad: bool = False) -> float:
    """Flops of the PSF+pool convolution of one particle (forward, plus the
    adjoint when ``grad``).  Direct path: 2 kf^2 per output pixel each way.
    Separable rank-R path: horizontal pass over the (ss N + k - 1) x N
    intermediate and vertical pass over N x N, 2 kf flops per tap each; the
    adjoint passes use ceil(kf/ss) taps over N x ssN and ssN x ssN."""
    N = num_pix
    k = kernel_size if kernel_size > 0 else 1
    kf = k + ss - 1
    if sep_rank <= 0:
        fwd = 2.0 * kf * kf * N * N
        adj = 2.0 * kf * kf * N * N
    else:
        m = -(-kf // ss)
        fwd = sep_rank * (2.0 * kf * (ss * N + k - 1) * N + 2.0 * kf * N * N)
        adj = sep_rank * (2.0 * m * N * ss * N + 2.0 * m * (ss * N) ** 2)
    return fwd + (adj if grad else 0.0)


def per_particle(kinds: Sequence[int], roles: Sequence[int], num_pix: int, ss: int, kernel_size: int,
                 n_terms: np.ndarray, grad: bool, sep_rank: int = 0) -> Dict[str, np.ndarray]:
    """Flops and transcendentals per particle (arrays over particles)."""
    S = (ss * num_pix) ** 2
    P = num_pix ** 2
    k = kernel_size if kernel_size > 0 else 1
    kf = k + ss - 1
    n_terms = np.asarray(n_terms, np.float64)
    f = np.full_like(n_terms, 8.0)
    tr = np.zeros_like(n_terms)
    for kind, role in zip(kinds, roles):
        a, t = _FWD[kind]
        f += a + (2 if role == nat.GL_ROLE_LENS else 1)
        tr += t
        if kind == nat.GL_EPL:
            f += 3 * n_terms
    flops = S * f + P * (1 + 12) + conv_flops(num_pix, ss, kernel_size, sep_rank, grad)
    trans = S * tr + P * 1
    if grad:
        g = np.full_like(n_terms, 8.0 + 2.0)
        gt = np.zeros_like(n_terms)
        for kind, role in zip(kinds, roles):
            a, t = _ADJ[kind]
            g += a
            gt += t
            if kind == nat.GL_EPL:
                g += 9 * n_terms
        flops = flops + S * g + P * 15
        trans = trans + S * gt
    return {"flops": flops, "transcendentals": trans}


def model_kinds(phys_model):
    kinds, roles = [], []
    for role, lst in ((nat.GL_ROLE_LENS, phys_model.lenses), (nat.GL_ROLE_LENS_LIGHT, phys_model.lens_light),
                      (nat.GL_ROLE_SOURCE, phys_model.source_light)):
        for prof in lst:
            kinds.append(prof._kind)
            roles.append(role)
    return kinds, roles


def survey_formula_per_particle(num_pix: int, ss: int, kernel_size: int, n_terms: np.ndarray) -> np.ndarray:
    """The reference-formulation estimate (naive EPL recurrence at 10 flops per
    term, unfused elementwise ops): forward S (71 + 10 n) + P (2 kf^2 + 10),
    sim+grad = forward + conv adjoint + 3 x ray/light part."""
    S = (ss * num_pix) ** 2
    P = num_pix ** 2
    kf = (kernel_size if kernel_size > 0 else 1) + ss - 1
    ray = S * (71 + 10 * np.asarray(n_terms, np.float64))
    conv = P * 2 * kf * kf
    return ray + conv + P * 10 + conv + 3 * ray


def shapelet_1d(n_max: int, interpolate: bool, deriv: bool = False):
    """(flops, transcendentals) of one 1-D shapelet evaluation up to n_max:
    table mode = index + n_max+1 lerps (3 flops; a slope when ``deriv``),
    direct mode = Hermite recurrence (4 flops per order) + Gaussian + scaling
    (+ 3 flops per order for the derivative)."""
    m = n_max + 1
    if interpolate:
        return 6 + 3 * m + (2 * m if deriv else 0), 0
    return 6 + 5 * m + (3 * m if deriv else 0), 1


def lstsq_per_particle(phys_model, num_pix: int, ss: int, kernel_size: int, n_terms: np.ndarray,
                       grad: bool, sep_rank: int = 0, adj_rank: int = 0) -> Dict[str, np.ndarray]:
    """Operation counts of the least-squares pipeline for one particle.
    ``flops`` are FP32 vector flops (basis, D-channel conv, image and
    likelihood, adjoint conv, gradient recompute), ``mfma`` the Gram product
    [X | obs]^T [X | obs] (2 flops per pair per pixel, upper triangle), ``fp64``
    the eigen-solve estimate (8 Jacobi sweeps of 6 (D+1)^3 flops).  The channel
    conv is counted in the form the plan runs: direct (2 kf^2 per output) or, for
    ``sep_rank`` R > 0, a horizontal pass over every sub-row and a vertical pass
    (2 kf R (ss N N + N N)); the adjoint conv likewise direct or, for ``adj_rank``
    > 0, its two polyphase passes (2 ceil(kf/ss) R ((ss N) N + (ss N)^2))."""
    S = (ss * num_pix) ** 2
    P = num_pix ** 2
    k = kernel_size if kernel_size > 0 else 1
    kf = k + ss - 1
    kinds, roles = model_kinds(phys_model)
    n_terms = np.asarray(n_terms, np.float64)
    D = 0
    f = np.full_like(n_terms, 8.0)
    tr = np.zeros_like(n_terms)
    g = np.full_like(n_terms, 8.0 + 2.0)
    gt = np.zeros_like(n_terms)
    for prof, kind, role in zip([*phys_model.lenses, *phys_model.lens_light, *phys_model.source_light], kinds, roles):
        if kind == nat.GL_SHAPELETS:
            ds = (prof.n_max + 1) * (prof.n_max + 2) // 2
            D += ds
            a, t = shapelet_1d(prof.n_max, prof.interpolate)
            f += 4 + 2 * a + ds
            tr += 2 * t
            ad, atr = shapelet_1d(prof.n_max, prof.interpolate, deriv=True)
            g += 4 + 2 * ad + 6 * ds + 8
            gt += 2 * atr
            continue
        if role != nat.GL_ROLE_LENS:
            D += 1
        a, t = _FWD[kind]
        f += a + (2 if role == nat.GL_ROLE_LENS else 0)
        tr += t
        a2, t2 = _ADJ[kind]
        g += a2
        gt += t2
        if kind == nat.GL_EPL:
            f += 3 * n_terms
            g += 9 * n_terms
    nsub = ss * num_pix
    if sep_rank > 0:
        conv = D * sep_rank * 2.0 * kf * (nsub * num_pix + P)
    else:
        conv = D * P * 2.0 * kf * kf
    if adj_rank > 0:
        mt = -(-kf // ss)
        adj = adj_rank * 2.0 * mt * (nsub * num_pix + S)
    else:
        adj = P * 2.0 * kf * kf
    flops = S * f + conv + P * (2.0 * D + 12)
    trans = S * tr + P
    mfma = P * (D + 1.0) * (D + 2.0) + P * (D + 1.0)
    fp64 = 8 * 6.0 * (D + 1) ** 3
    if grad:
        flops = flops + adj + S * g + P * 15
        trans = trans + S * gt
    return {"flops": flops, "transcendentals": trans, "mfma": np.full_like(n_terms, mfma),
            "fp64": np.full_like(n_terms, fp64), "channels": D, "conv_flops": float(conv)}
