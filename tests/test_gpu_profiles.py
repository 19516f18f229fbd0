"""GPU parity of the SIS, SIE and truncated-NFW mass profiles
(src/gigalens/tf/profiles/mass/sis.py:12-18, sie.py:13-43, tnfw.py:14-58), of the
circular Sersic light profile (src/gigalens/tf/profiles/light/sersic.py:21-31) and
of a non-diagonal pixel-to-angle transform (src/gigalens/tf/simulator.py:19-29,
src/gigalens/simulator.py:91-125), through the LDS-resident fused kernel and
through the HBM-staged pipeline, and of the core-Sersic light profile
(src/gigalens/tf/profiles/light/sersic.py:75-121) through both, against the FP64
oracle.

Bars: images within 1e-4 of the image peak; log_prob within 1e-4 relative;
gradients against FP64 Richardson-extrapolated central differences of the
oracle (h = 1e-3 and 5e-4 relative) within 2e-3 of each component plus 2e-4 of
the largest.
"""
import math

import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, PhysicalModel
from gigalens_amd import distributions as tfd
from gigalens_amd import workloads as wl
from gigalens_amd.plan import Plan, layout_from_paths
from gigalens_amd import _native as nat
from gigalens_amd.profiles.light.sersic import CoreSersic, Sersic, SersicEllipse
from gigalens_amd.profiles.mass.epl import EPL
from gigalens_amd.profiles.mass.shear import Shear
from gigalens_amd.profiles.mass.sie import SIE
from gigalens_amd.profiles.mass.sis import SIS
from gigalens_amd.profiles.mass.tnfw import TNFW
from gigalens_amd.simulator import Camera

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _lights():
    prior = wl.epl_shear_prior()
    return prior.model[1], prior.model[2]


def _sie_problem():
    ll, src = _lights()
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            theta_E=tfd.LogNormal(math.log(1.2), 0.2), e1=tfd.Normal(0, 0.1), e2=tfd.Normal(0, 0.1),
            center_x=tfd.Normal(0, 0.05), center_y=tfd.Normal(0, 0.05))),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    prior = tfd.JointDistributionSequential([lens, ll, src])
    phys = PhysicalModel([SIE(), Shear()], [SersicEllipse()], [SersicEllipse()])
    return prior, phys


def _tnfw_problem():
    ll, src = _lights()
    base = wl.epl_shear_prior().model[0].model
    lens = tfd.JointDistributionSequential([
        base[0], base[1],
        tfd.JointDistributionNamed(dict(
            Rs=tfd.LogNormal(math.log(0.3), 0.2), alpha_Rs=tfd.LogNormal(math.log(0.15), 0.3),
            r_trunc=tfd.LogNormal(math.log(1.5), 0.2), center_x=tfd.Normal(0.6, 0.1),
            center_y=tfd.Normal(-0.4, 0.1))),
    ])
    prior = tfd.JointDistributionSequential([lens, ll, src])
    phys = PhysicalModel([EPL(), Shear(), TNFW()], [SersicEllipse()], [SersicEllipse()])
    return prior, phys


def _sis_problem():
    ll, src = _lights()
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            theta_E=tfd.LogNormal(math.log(1.2), 0.2), center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05))),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    prior = tfd.JointDistributionSequential([lens, ll, src])
    phys = PhysicalModel([SIS(), Shear()], [SersicEllipse()], [SersicEllipse()])
    return prior, phys


def _round_source_problem():
    """EPL + shear, Sersic-ellipse lens light, circular Sersic source (sersic.py:21-31)."""
    prior0 = wl.epl_shear_prior()
    src = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(0.25), 0.15), n_sersic=tfd.Uniform(0.5, 4),
            center_x=tfd.Normal(0, 0.25), center_y=tfd.Normal(0, 0.25), Ie=tfd.LogNormal(math.log(150.0), 0.5))),
    ])
    prior = tfd.JointDistributionSequential([prior0.model[0], prior0.model[1], src])
    phys = PhysicalModel([EPL(), Shear()], [SersicEllipse()], [Sersic()])
    return prior, phys


def _core_problem():
    """EPL + shear, core-Sersic lens light (sersic.py:75-121), Sersic source."""
    _, src = _lights()
    lens = wl.epl_shear_prior().model[0]
    ll = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            R_sersic=tfd.LogNormal(math.log(1.0), 0.15), n_sersic=tfd.Uniform(1.0, 4.0),
            Rb=tfd.LogNormal(math.log(0.1), 0.2), alpha=tfd.Uniform(1.0, 3.0), gamma=tfd.Uniform(0.1, 0.8),
            e1=tfd.Normal(0, 0.1), e2=tfd.Normal(0, 0.1), center_x=tfd.Normal(0, 0.05),
            center_y=tfd.Normal(0, 0.05), Ie=tfd.LogNormal(math.log(500.0), 0.3))),
    ])
    prior = tfd.JointDistributionSequential([lens, ll, src])
    phys = PhysicalModel([EPL(), Shear()], [CoreSersic()], [SersicEllipse()])
    return prior, phys


# a rotated and slightly sheared pixel grid (columns 0.0655"/px at -7 deg, rows 0.0645"/px)
ROTATED_T = np.array([[0.0655 * math.cos(-0.122), -0.0645 * math.sin(-0.122) + 0.004],
                      [0.0655 * math.sin(-0.122), 0.0645 * math.cos(-0.122)]], np.float32)


def _problem(name, bs=8):
    if name == "rotated":
        prior, phys = wl.epl_shear_prior(), wl.epl_shear_model()
    else:
        prior, phys = {"sie": _sie_problem, "tnfw": _tnfw_problem, "core": _core_problem, "sis": _sis_problem,
                       "round": _round_source_problem}[name]()
    cfg = wl.epl_shear_config()
    if name == "rotated":
        cfg.transform_pix2angle = ROTATED_T
    cam = Camera(cfg)
    ocam = C.oracle_camera(cfg, cam)
    truth_x = C.draw_prior(prior, 1, seed=0).numpy().astype(np.float64)
    clean = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, truth_x))[0]
    obs = wl.noisy_observation(clean, 0.2, 100.0, seed=1)
    x = C.draw_prior(prior, bs, seed=3)
    return dict(prior=prior, phys=phys, cfg=cfg, cam=cam, ocam=ocam, obs=obs, x=x, truth_x=truth_x)


def _plan(prob, staged):
    paths = [p for p, _ in tfd.leaves(prob["prior"])]
    return Plan(prob["phys"], prob["cam"], layout_from_paths(prob["phys"], paths), prior=prob["prior"],
                observed=prob["obs"], background_rms=0.2, exp_time=100.0, staged=staged)


CASES = [("sie", False), ("sie", True), ("tnfw", False), ("tnfw", True), ("core", False), ("core", True), ("sis", False),
         ("sis", True), ("round", False), ("round", True), ("rotated", False), ("rotated", True)]


@pytest.fixture(scope="module", params=CASES, ids=[f"{n}-{'staged' if s else 'fused'}" for n, s in CASES])
def case(request, cuda):
    name, staged = request.param
    prob = _problem(name)
    plan = _plan(prob, staged)
    fused = plan.info()["lds_bytes"] > 0
    assert fused == (not staged), "the fused kernel is instantiated for this model"
    return name, prob, plan


def test_profile_simulate_parity(case):
    name, prob, plan = case
    x = prob["x"]
    img = plan.simulate(x.cuda().contiguous()).cpu().numpy().astype(np.float64)
    ref = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prob["prior"], x.numpy().astype(np.float64)))
    err = np.abs(img - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    _print(f"{name}: simulate max err / peak {err.max():.2e}")
    assert np.all(err < 1e-4)


def test_profile_log_prob_parity(case):
    name, prob, plan = case
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).contiguous()
    lp, chi2 = plan.log_prob(z.cuda())[:2]
    lp_ref, chi_ref, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], z.numpy().astype(np.float64),
                                           prob["obs"], 0.2, 100.0)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    relc = np.abs(chi2.cpu().numpy() - chi_ref) / np.abs(chi_ref)
    _print(f"{name}: log_prob rel err {rel.max():.2e}, chi2 {relc.max():.2e}")
    assert rel.max() < 1e-4 and relc.max() < 1e-4


def test_profile_gradient_vs_fd(case):
    name, prob, plan = case
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    # particles near the truth, perturbed in the unconstrained space (stays inside every support)
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32)
    zt = pm.bij.inverse(C.nested_from_x(prob["prior"], xt))
    z = (zt.repeat(2, 1) + 0.02 * torch.randn(2, zt.shape[1], generator=torch.Generator().manual_seed(5))).contiguous()
    g = plan.log_prob(z.cuda(), want_grad=True)[2].cpu().numpy()
    zz = z.numpy().astype(np.float64)
    for i in range(2):
        gfd = np.zeros(zz.shape[1])
        for j in range(zz.shape[1]):
            # Richardson-extrapolated central differences (steps h and h/2): at these
            # particles the O(h^2) error of one central difference reaches 5e-2 of a
            # component at h = 1e-3 relative
            cd = []
            for hr in (1e-3, 5e-4):
                h = hr * max(abs(zz[i, j]), 1e-1)
                zp, zm = zz[i:i + 1].copy(), zz[i:i + 1].copy()
                zp[:, j] += h
                zm[:, j] -= h
                fp = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zp, prob["obs"], 0.2, 100.0)[0]
                fm = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zm, prob["obs"], 0.2, 100.0)[0]
                cd.append((fp - fm)[0] / (2 * h))
            gfd[j] = (4.0 * cd[1] - cd[0]) / 3.0
        scale = np.abs(gfd).max()
        err = np.abs(g[i] - gfd)
        _print(f"{name} particle {i}: max |g - fd| / max|fd| = {err.max() / scale:.2e}")
        assert np.all(err <= 2e-3 * np.abs(gfd) + 2e-4 * scale), (g[i], gfd)
