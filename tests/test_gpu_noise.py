"""The device observation-noise sampler (gl_noise_poisson_gauss, csrc/noise.hip)
against the distributions it draws from: for each mean below and above the
switch between CDF inversion and PTRS at 10, 2^20 draws are compared with the
Poisson law (sample mean and variance within 5 standard errors, a chi-square
goodness-of-fit over bins with >= 50 expected counts, p > 1e-4); the Gaussian
background against N(0, sigma^2) (moments and a Kolmogorov-Smirnov test); and
the draws are a pure function of (seed, pixel)."""
import numpy as np
import pytest
import torch
from scipy import stats

from gigalens_amd.learned import poisson_gauss_noise

pytestmark = pytest.mark.gpu

N = 1 << 20
TEXP = 100.0


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("lam", [0.3, 4.0, 9.9, 10.0, 37.5, 1000.0, 1e5])
def test_poisson_law(lam, cuda):
    img = torch.full((N,), lam / TEXP, device=cuda)
    k = poisson_gauss_noise(img, TEXP, 0.0, seed=12345).double().cpu().numpy() * TEXP
    kr = np.rint(k)
    assert np.abs(k - kr).max() < 1e-3 * max(1.0, lam / 1e4)  # integer counts (float32 k / t * t)
    mean, var = kr.mean(), kr.var()
    se_mean = np.sqrt(lam / N)
    se_var = np.sqrt((lam + 2 * lam * lam) / N)
    _print(f"lam {lam}: mean {mean:.5g} ({(mean - lam) / se_mean:+.2f} se), var {var:.5g} "
           f"({(var - lam) / se_var:+.2f} se)")
    assert abs(mean - lam) < 5 * se_mean
    assert abs(var - lam) < 5 * se_var
    # chi-square over the support, tails pooled so that every bin expects >= 50
    lo, hi = int(stats.poisson.ppf(1e-6, lam)), int(stats.poisson.ppf(1 - 1e-6, lam))
    edges = [lo]
    acc = 0.0
    for v in range(lo, hi + 1):
        acc += stats.poisson.pmf(v, lam) * N
        if acc >= 50:
            edges.append(v + 1)
            acc = 0.0
    edges[-1] = hi + 1
    cdf = stats.poisson.cdf(np.array(edges) - 1, lam)
    exp = np.diff(cdf) * N
    exp[0] += stats.poisson.cdf(edges[0] - 1, lam) * N
    exp[-1] += stats.poisson.sf(edges[-1] - 1, lam) * N
    obs = np.histogram(np.clip(kr, edges[0], edges[-1] - 1), bins=np.array(edges) - 0.5)[0]
    chi2 = ((obs - exp) ** 2 / exp).sum()
    p = stats.chi2.sf(chi2, len(exp) - 1)
    _print(f"  chi2 {chi2:.1f} over {len(exp)} bins, p = {p:.3g}")
    assert p > 1e-4


def test_gaussian_background_and_determinism(cuda):
    sigma = 0.2
    img = torch.zeros(N, device=cuda)
    g = poisson_gauss_noise(img, TEXP, sigma, seed=7).double().cpu().numpy()
    se = sigma / np.sqrt(N)
    assert abs(g.mean()) < 5 * se
    assert abs(g.var() / sigma ** 2 - 1.0) < 5 * np.sqrt(2.0 / N)
    ks = stats.kstest(g / sigma, "norm")
    _print(f"gaussian: mean {g.mean():.2e} var {g.var():.5f} KS p = {ks.pvalue:.3g}")
    assert ks.pvalue > 1e-4
    # same seed -> same draws (also in place); another seed -> different draws
    img2 = torch.rand(4096, device=cuda) * 3.0
    a = poisson_gauss_noise(img2, TEXP, sigma, seed=99)
    b = img2.clone()
    poisson_gauss_noise(b, TEXP, sigma, seed=99, out=b)
    c = poisson_gauss_noise(img2, TEXP, sigma, seed=100)
    assert torch.equal(a, b)
    assert not torch.equal(a, c)
    # a prefix of the pixels gets the same draws as in the longer call
    d = poisson_gauss_noise(img2[:1000].contiguous(), TEXP, sigma, seed=99)
    assert torch.equal(d, a[:1000])


def test_nonpositive_and_nan_means_give_background_only(cuda):
    img = torch.tensor([0.0, -1.0, float("nan"), -0.0] * 256, device=cuda)
    out = poisson_gauss_noise(img, TEXP, 0.0, seed=3)
    assert torch.all(out == 0)
