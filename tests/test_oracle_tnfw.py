"""Pins of the oracle's truncated-NFW deflection (``oracle.np64.tnfw_deriv``,
restating src/gigalens/tf/profiles/mass/tnfw.py:14-58) by mathematics alone:

* r_trunc -> infinity gives the NFW deflection, whose closed form is
  alpha(R) = alpha_Rs h(x) / (x h(1)), h(x) = ln(x/2) + F(x), x = R/Rs
  (alpha_Rs is the deflection at R = Rs);
* at finite truncation, alpha(R) R is proportional to the mass inside the
  cylinder of radius R of the 3-D density
  rho(r) = rho_s / (x (1+x)^2) * tau^2 / (tau^2 + x^2)  (x = r/Rs, tau = r_t/Rs),
  with the same constant as the untruncated profile.  The cylinder mass is
  M_tot - 4 pi int_R^inf rho(r) r sqrt(r^2 - R^2) dr (scipy quad).
A dropped term, a wrong sign or a swapped tau/x in the restatement fails one
of the two.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import np64


def _alpha_radial(R, Rs=1.0, aR=1.0, rt=5.0):
    x = np.asarray(R, np.float64)
    ax, ay = np64.tnfw_deriv(x, np.zeros_like(x), Rs, aR, rt, 0.0, 0.0)
    return ax.reshape(-1)


def _h(x):
    return np.log(x / 2) + np64.tnfw_F(x)


@pytest.mark.parametrize("Rs,aR", [(1.0, 1.0), (0.7, 0.3)])
def test_tnfw_untruncated_limit_is_nfw(Rs, aR):
    R = np.array([0.01, 0.2, 0.9, 1.0, 1.3, 4.0, 12.0]) * Rs
    a = _alpha_radial(R, Rs, aR, rt=1e4 * Rs)
    x = R / Rs
    nfw = aR * _h(x) / (x * _h(np.array([1.0]))[0])
    # truncation at tau = 1e4 changes alpha by O(x^2/tau^2) (1.4e-6 at x = 12)
    np.testing.assert_allclose(a, nfw, rtol=1e-5)


def _cyl_mass(R, tau):
    rho = lambda r: tau ** 2 / (r * (1 + r) ** 2 * (tau ** 2 + r ** 2))
    tot = 4 * np.pi * integrate.quad(lambda r: rho(r) * r * r, 0, np.inf, limit=400)[0]
    out = integrate.quad(lambda r: rho(r) * r * np.sqrt(r * r - R * R), R, np.inf, limit=400)[0]
    return tot - 4 * np.pi * out


def _nfw_cyl_mass(R):
    return 4 * np.pi * _h(np.array([R]))[0]


@pytest.mark.parametrize("tau", [2.0, 5.0, 20.0])
def test_tnfw_matches_projected_mass_of_truncated_density(tau):
    # Rs = 1: alpha(R) R / M_cyl(R) must be the NFW constant alpha(1) / M_cyl,NFW(1)
    const = 1.0 / _nfw_cyl_mass(1.0)
    for R in (0.05, 0.5, 1.0, 1.7, 3.0, 10.0):
        a = _alpha_radial(np.array([R]), 1.0, 1.0, rt=tau)[0]
        m = _cyl_mass(R, tau)
        assert a * R == pytest.approx(const * m, rel=2e-7), (tau, R)


def test_tnfw_clamp_inside_001_rs():
    # R < 0.001 Rs uses x = 0.001 with the unclamped offset (tnfw.py:20-21)
    Rs = 0.8
    r_in = np.array([1e-5, 3e-4])
    a = _alpha_radial(r_in, Rs, 1.0, 4.0)
    a0 = _alpha_radial(np.array([0.001 * Rs]), Rs, 1.0, 4.0)[0]
    np.testing.assert_allclose(a / r_in, a0 / (0.001 * Rs), rtol=1e-12)
