"""Named workloads: model, prior and camera of the benchmark configurations.

Plain forward model ("60x60 EPL+shear"): 60x60 pixels at 0.065"/px,
supersample 2, Gaussian PSF FWHM 0.1" sampled as 11x11 (sub-grid 21x21);
EPL + external shear lens, elliptical Sersic lens light and source.
Prior (22 parameters, tf.nest order) -- where the benchmark text is silent
the reference's default prior (tests/conftest.py:20-73) is used:
lens-light centre ~ N(0, 0.05), source e1,e2 ~ TruncNormal(0, 0.15, -0.5, 0.5).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import distributions as tfd
from .model import PhysicalModel
from .profiles.light.sersic import SersicEllipse
from .profiles.light.shapelets import Shapelets
from .profiles.mass.epl import EPL
from .profiles.mass.shear import Shear
from .simulator import SimulatorConfig
from . import psf


def epl_shear_prior():
    """Prior of the 60x60 EPL+shear workload."""
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            theta_E=tfd.LogNormal(math.log(1.25), 0.25),
            gamma=tfd.TruncatedNormal(2, 0.25, 1.2, 2.8),
            e1=tfd.Normal(0, 0.1),
            e2=tfd.Normal(0, 0.1),
            center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05),
        )),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    lens_light = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(1.0), 0.15),
            n_sersic=tfd.Uniform(0.5, 8),
            e1=tfd.Normal(0, 0.1),
            e2=tfd.Normal(0, 0.1),
            center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05),
            Ie=tfd.LogNormal(math.log(500.0), 0.3),
        ))
    ])
    source = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(0.25), 0.15),
            n_sersic=tfd.Uniform(0.5, 4),
            e1=tfd.TruncatedNormal(0, 0.15, -0.5, 0.5),
            e2=tfd.TruncatedNormal(0, 0.15, -0.5, 0.5),
            center_x=tfd.Normal(0, 0.25),
            center_y=tfd.Normal(0, 0.25),
            Ie=tfd.LogNormal(math.log(150.0), 0.5),
        ))
    ])
    return tfd.JointDistributionSequential([lens, lens_light, source])


def reference_default_prior():
    """The reference's default prior, tests/conftest.py:20-73 / tf-demo.ipynb cell 2."""
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            theta_E=tfd.LogNormal(math.log(1.25), 0.25),
            gamma=tfd.TruncatedNormal(2, 0.25, 1, 3),
            e1=tfd.Normal(0, 0.1),
            e2=tfd.Normal(0, 0.1),
            center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05),
        )),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    lens_light = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(1.0), 0.15),
            n_sersic=tfd.Uniform(2, 6),
            e1=tfd.TruncatedNormal(0, 0.1, -0.3, 0.3),
            e2=tfd.TruncatedNormal(0, 0.1, -0.3, 0.3),
            center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05),
            Ie=tfd.LogNormal(math.log(500.0), 0.3),
        ))
    ])
    source = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(0.25), 0.15),
            n_sersic=tfd.Uniform(0.5, 4),
            e1=tfd.TruncatedNormal(0, 0.15, -0.5, 0.5),
            e2=tfd.TruncatedNormal(0, 0.15, -0.5, 0.5),
            center_x=tfd.Normal(0, 0.25),
            center_y=tfd.Normal(0, 0.25),
            Ie=tfd.LogNormal(math.log(150.0), 0.5),
        ))
    ])
    return tfd.JointDistributionSequential([lens, lens_light, source])


def epl_shear_model(niter: int = 50):
    return PhysicalModel([EPL(niter), Shear()], [SersicEllipse()], [SersicEllipse()])


def epl_shear_config():
    kernel = psf.gaussian(0.1, 0.065, 11)
    return SimulatorConfig(delta_pix=0.065, num_pix=60, supersample=2, kernel=kernel)


# demo truth of the reference notebook (tf-demo.ipynb cell 5)
DEMO_TRUTH = [[
    {"theta_E": 1.1, "gamma": 2.0, "e1": 0.1, "e2": 0.1, "center_x": 0.1, "center_y": 0.0},
    {"gamma1": -0.01, "gamma2": 0.03},
], [
    {"R_sersic": 0.8, "n_sersic": 2.5, "e1": 0.09534746574143645, "e2": 0.14849487967198177, "center_x": 0.1,
     "center_y": 0.0, "Ie": 499.3695906504067},
], [
    {"R_sersic": 0.25, "n_sersic": 1.5, "e1": 0., "e2": 0., "center_x": 0.09566681002252231,
     "center_y": -0.0639623054267272, "Ie": 149.58828877085668},
]]


def noisy_observation(clean: np.ndarray, background_rms: float, exp_time: float, seed: int) -> np.ndarray:
    """Observed image = Poisson(clean * t_exp) / t_exp + N(0, rms^2), seeded."""
    rng = np.random.default_rng(seed)
    lam = np.clip(np.asarray(clean, np.float64), 0, None) * exp_time
    img = rng.poisson(lam) / exp_time + rng.normal(0.0, background_rms, size=clean.shape)
    return img.astype(np.float32)


def shapelet_model(n_max: int = 10, interpolate: bool = True, niter: int = 50):
    """EPL + shear, Sersic lens light and shapelet source, all lights linear (least squares)."""
    return PhysicalModel([EPL(niter), Shear()], [SersicEllipse(use_lstsq=True)],
                         [Shapelets(n_max, use_lstsq=True, interpolate=interpolate)])


def shapelet_config():
    """100x100 px at 0.04"/px, supersample 4, 21x21 Moffat PSF (FWHM 0.08", beta 3)."""
    kernel = psf.moffat(0.08, 3.0, 0.04, 21)
    return SimulatorConfig(delta_pix=0.04, num_pix=100, supersample=4, kernel=kernel)


def shapelet_prior():
    """Prior of the shapelet workload: the EPL+shear lens of the 60x60 workload,
    a Sersic lens light and a shapelet source with linear amplitudes (so no Ie
    and no shapelet amplitudes among the sampled parameters; 17 in all)."""
    lens = epl_shear_prior().model[0]
    lens_light = tfd.JointDistributionSequential([tfd.JointDistributionNamed(dict(
        R_sersic=tfd.LogNormal(math.log(1.0), 0.15),
        n_sersic=tfd.Uniform(0.5, 8),
        e1=tfd.Normal(0, 0.1),
        e2=tfd.Normal(0, 0.1),
        center_x=tfd.Normal(0, 0.05),
        center_y=tfd.Normal(0, 0.05),
    ))])
    source = tfd.JointDistributionSequential([tfd.JointDistributionNamed(dict(
        beta=tfd.LogNormal(math.log(0.15), 0.2),
        center_x=tfd.Normal(0, 0.25),
        center_y=tfd.Normal(0, 0.25),
    ))])
    return tfd.JointDistributionSequential([lens, lens_light, source])


def shapelet_amplitudes(n_max: int, seed: int, lens_light_amp: float = 500.0, source_amp: float = 150.0):
    """Linear amplitudes of a synthetic truth: the lens light, then shapelet
    coefficients N(0, 1) * source_amp * exp(-order / 3) in the basis order."""
    rng = np.random.default_rng(seed)
    amps = [lens_light_amp]
    for order in range(n_max + 1):
        amps.extend(rng.normal(0.0, 1.0, size=order + 1) * source_amp * math.exp(-order / 3.0))
    return np.asarray(amps, np.float64)
