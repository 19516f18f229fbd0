"""Shapelet light with sampled amplitudes, ``Shapelets(n_max, use_lstsq=False)`` -- the
reference's default mode (src/gigalens/tf/profiles/light/shapelets.py:53-75), on the
GPU against the FP64 oracle.

* The reference's own test (tests/test_profiles.py:35-47): n_max = 5, table and direct
  modes, N(0, 1) amplitudes, centre 0, beta 1, at 5 x 5 points, rtol 1e-5 / atol 1e-4.
  The kernels evaluate light on an image grid, so the points are a 5 x 5 grid (no lens,
  no PSF, supersample 1; a sheared pixel-to-angle transform spreads them over
  [-1.6, 1.6]) and the light is the image divided by the pixel area.
* A model like the reference's shapelet demo (shapelets-demo.ipynb cells 4-7: EPL +
  shear, no lens light, direct mode, amplitudes ~ N(0, 500 / sqrt(i + 1)), 60 x 60 at
  0.065"/px) with n_max = 10 (66 amplitudes among 77 parameters), supersample 2 and a
  Gaussian PSF: images within 1e-4 of the peak, ForwardProbModel log_prob within 1e-4
  relative, gradient within 2e-3 of each component plus 2e-4 of the largest against
  Richardson-extrapolated FP64 central differences of the oracle.
"""
import math

import numpy as np
import pytest
import torch

from oracle import np64
from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator, PhysicalModel, SimulatorConfig
from gigalens_amd import distributions as tfd
from gigalens_amd import psf
from gigalens_amd.plan import Plan, layout_from_model
from gigalens_amd.profiles.light.shapelets import Shapelets
from gigalens_amd.profiles.mass.epl import EPL
from gigalens_amd.profiles.mass.shear import Shear
from gigalens_amd.simulator import Camera

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("interpolate", [True, False], ids=["table", "direct"])
def test_reference_shapelet_points(cuda, interpolate):
    T = np.array([[0.8, 0.12], [-0.1, 0.75]], np.float32)
    cfg = SimulatorConfig(delta_pix=0.8, num_pix=5, supersample=1, transform_pix2angle=T)
    shp = Shapelets(n_max=5, use_lstsq=False, interpolate=interpolate)
    phys = PhysicalModel([], [], [shp])
    cam = Camera(cfg)
    lay = layout_from_model(phys)
    plan = Plan(phys, cam, lay)
    rng = np.random.default_rng(11)
    bs = 8
    amps = rng.normal(size=(bs, shp.n_layers)).astype(np.float32)
    names = [k for _, _, k in lay]
    x = np.zeros((bs, len(lay)), np.float32)
    for j, n in enumerate(names):
        x[:, j] = amps[:, shp._amp_names.index(n)] if n.startswith("amp") else (1.0 if n == "beta" else 0.0)
    img = plan.simulate(torch.as_tensor(x).to(cuda).contiguous()).cpu().numpy().astype(np.float64)
    light = img / cam.conversion_factor
    gx, gy = cam.coords()
    ref = np64.shapelets_light(gx.astype(np.float64), gy.astype(np.float64), 0.0, 0.0, np.ones(bs),
                               amps.T.astype(np.float64), 5, interpolate)
    _print(f"shapelets n_max=5 ({'table' if interpolate else 'direct'}): max |light - ref| "
           f"{np.abs(light - ref).max():.2e}, points in [{gx.min():.2f}, {gx.max():.2f}]")
    assert np.allclose(light, ref, rtol=1e-5, atol=1e-4)


def _demo_problem(bs):
    n_max = 10
    shp = Shapelets(n_max=n_max, use_lstsq=False, interpolate=False)
    phys = PhysicalModel([EPL(), Shear()], [], [shp])
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(
            theta_E=tfd.LogNormal(math.log(1.0), 0.25), gamma=tfd.TruncatedNormal(2, 0.25, 1, 3),
            e1=tfd.Normal(0, 0.1), e2=tfd.Normal(0, 0.1), center_x=tfd.Normal(0, 0.025),
            center_y=tfd.Normal(0, 0.025))),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    src = dict(beta=tfd.LogNormal(math.log(0.1), 0.15), center_x=tfd.Normal(0, 0.01),
               center_y=tfd.Normal(0, 0.01))
    src.update({k: tfd.Normal(0, 500 / math.sqrt(i + 1)) for i, k in enumerate(shp._amp_names)})
    prior = tfd.JointDistributionSequential([lens, tfd.JointDistributionSequential([tfd.JointDistributionNamed(src)])])
    cfg = SimulatorConfig(delta_pix=0.065, num_pix=60, supersample=2, kernel=psf.gaussian(0.1, 0.065, 11))
    cam = Camera(cfg)
    ocam = C.oracle_camera(cfg, cam)
    truth_x = C.draw_prior(prior, 1, seed=121).numpy().astype(np.float64)
    clean = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, truth_x))[0]
    obs = (clean + np.random.default_rng(1).normal(size=clean.shape) *
           np.sqrt(0.1 ** 2 + np.clip(clean, 0, None) / 200)).astype(np.float32)
    x = C.draw_prior(prior, bs, seed=5)
    return dict(prior=prior, phys=phys, cfg=cfg, cam=cam, ocam=ocam, obs=obs, x=x, truth_x=truth_x)


def test_shapelet_demo_model(cuda):
    prob = _demo_problem(bs=4)
    prior = prob["prior"]
    assert len(tfd.leaves(prior)) == 77
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=4)
    assert sim.plan.info()["lds_bytes"] == 0  # the staged pipeline
    x = prob["x"]
    img = sim.simulate(x.to(cuda).contiguous()).cpu().numpy().astype(np.float64)
    ref = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prior, x.numpy().astype(np.float64)))
    err = np.abs(img - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    _print(f"demo shapelet model: image err / peak {err.max():.2e}")
    assert np.all(err < 1e-4)
    pm = ForwardProbModel(prior, prob["obs"], 0.1, 200.0)
    # particles near the truth (likelihood-dominated), perturbed in the unconstrained space
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32)
    zt = pm.bij.inverse(C.nested_from_x(prior, xt))
    z = (zt.repeat(2, 1) * (1 + 0.01 * torch.randn(2, zt.shape[1], generator=torch.Generator().manual_seed(3)))
         ).contiguous()
    lp, chi2, g = pm.log_prob_and_grad(LensSimulator(prob["phys"], prob["cfg"], bs=2), z.to(cuda))
    zz = z.numpy().astype(np.float64)
    lp_ref, chi_ref, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prior, zz, prob["obs"], 0.1, 200.0)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    _print(f"  log_prob rel err {rel.max():.2e}, chi2 {chi2.cpu().numpy()} vs {chi_ref}")
    assert rel.max() < 1e-4
    g = g.cpu().numpy().astype(np.float64)
    g_fd = C.oracle_grad_richardson(prob["ocam"], prob["phys"], prior, zz, prob["obs"], 0.1, 200.0)
    scale = np.abs(g_fd).max(axis=1, keepdims=True)
    e = np.abs(g - g_fd)
    _print(f"  gradient max |g - fd| / max|fd| = {(e / scale).max():.2e}")
    assert np.all(e <= 2e-3 * np.abs(g_fd) + 2e-4 * scale)
