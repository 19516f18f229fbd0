"""GPU parity of the least-squares (linear amplitude) path: lstsq_simulate
(tf/simulator.py:99-165) and BackwardProbModel.log_prob (tf/model.py:134-163)
against the FP64 oracle.  Tolerances: components and images within 1e-4 of
the peak, coefficients within 1e-3 relative (FP32 components, FP64 solve),
log_prob within 1e-4 relative, gradients vs FP64 central differences."""
import math

import numpy as np
import pytest
import torch

from oracle import np64
from tests import common as C
from gigalens_amd import BackwardProbModel, LensSimulator, PhysicalModel, SimulatorConfig
from gigalens_amd import distributions as tfd
from gigalens_amd import psf
from gigalens_amd.profiles.light.sersic import SersicEllipse
from gigalens_amd.profiles.light.shapelets import Shapelets
from gigalens_amd.profiles.mass.epl import EPL
from gigalens_amd.profiles.mass.shear import Shear

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def shapelet_prior(n_max):
    lens = tfd.JointDistributionSequential([
        tfd.JointDistributionNamed(dict(theta_E=tfd.LogNormal(math.log(1.25), 0.1), gamma=tfd.TruncatedNormal(2, 0.1, 1.2, 2.8),
                                        e1=tfd.Normal(0, 0.1), e2=tfd.Normal(0, 0.1),
                                        center_x=tfd.Normal(0, 0.05), center_y=tfd.Normal(0, 0.05))),
        tfd.JointDistributionNamed(dict(gamma1=tfd.Normal(0, 0.05), gamma2=tfd.Normal(0, 0.05))),
    ])
    ll = tfd.JointDistributionSequential([tfd.JointDistributionNamed(dict(
        R_sersic=tfd.LogNormal(math.log(1.0), 0.15), n_sersic=tfd.Uniform(2, 6),
        e1=tfd.TruncatedNormal(0, 0.1, -0.3, 0.3), e2=tfd.TruncatedNormal(0, 0.1, -0.3, 0.3),
        center_x=tfd.Normal(0, 0.05), center_y=tfd.Normal(0, 0.05)))])
    src = tfd.JointDistributionSequential([tfd.JointDistributionNamed(dict(
        beta=tfd.LogNormal(math.log(0.15), 0.2), center_x=tfd.Normal(0, 0.1), center_y=tfd.Normal(0, 0.1)))])
    return tfd.JointDistributionSequential([lens, ll, src])


def setup(num_pix, ss, delta, psf_size, n_max, interpolate, bs, seed=1):  # noqa: C901
    prior = shapelet_prior(n_max)
    phys = PhysicalModel([EPL(50), Shear()], [SersicEllipse(use_lstsq=True)],
                         [Shapelets(n_max, use_lstsq=True, interpolate=interpolate)])
    cfg = SimulatorConfig(delta_pix=delta, num_pix=num_pix, supersample=ss,
                          kernel=psf.moffat(2 * delta, 3.0, delta, psf_size))
    sim = LensSimulator(phys, cfg, bs=bs)
    ocam = C.oracle_camera(cfg, sim.camera)
    x = C.draw_prior(prior, bs, seed=seed)
    # observed image: a noisy lstsq-free model image built from prior draws with random amplitudes
    rng = np.random.default_rng(seed)
    truth = C.draw_prior(prior, 1, seed=seed + 100).numpy().astype(np.float64)
    stacked = C.oracle_lstsq(ocam, phys, C.nested_from_x(prior, truth), np.zeros((num_pix, num_pix)),
                             np.ones((num_pix, num_pix)))[0]
    amps = np.concatenate([[50.0], rng.normal(0, 30.0, size=stacked.shape[-1] - 1)])
    clean = np.einsum("hwd,d->hw", stacked[0], amps)
    obs = clean + rng.normal(0, 0.2, size=clean.shape)
    err = np.sqrt(0.2 ** 2 + np.clip(obs, 0, None) / 100.0)
    return dict(prior=prior, phys=phys, cfg=cfg, sim=sim, ocam=ocam, x=x, obs=obs.astype(np.float32), err=err)


@pytest.mark.parametrize("interpolate", [True, False])
def test_lstsq_simulate_small(interpolate, cuda):
    P = setup(30, 2, 0.08, 7, 4, interpolate, bs=4)
    sim, x = P["sim"], P["x"]
    nested = C.nested_from_x(P["prior"], x)
    st_ref, co_ref, im_ref = C.oracle_lstsq(P["ocam"], P["phys"], C.nested_from_x(P["prior"], x.numpy().astype(np.float64)),
                                            P["obs"], P["err"])
    stacked = sim.lstsq_simulate(nested, P["obs"], P["err"], return_stacked=True).cpu().numpy()
    coeffs = sim.lstsq_simulate(nested, P["obs"], P["err"], return_coeffs=True).cpu().numpy()
    img = sim.lstsq_simulate(nested, P["obs"], P["err"]).cpu().numpy()
    e_st = np.abs(stacked - st_ref).max() / np.abs(st_ref).max()
    e_im = (np.abs(img - im_ref).max(axis=(1, 2)) / np.abs(im_ref).max(axis=(1, 2))).max()
    e_co = np.abs(coeffs - co_ref).max() / np.abs(co_ref).max()
    _print(f"interp={interpolate}: stacked {e_st:.2e} image {e_im:.2e} coeffs {e_co:.2e}")
    assert stacked.shape == st_ref.shape
    assert e_st < 1e-4
    assert e_im < 1e-4
    assert e_co < 1e-3


def test_backward_log_prob_and_gradient(cuda):
    P = setup(30, 2, 0.08, 7, 4, True, bs=3)
    pm = BackwardProbModel(P["prior"], P["obs"], 0.2, 100.0)
    sim = P["sim"]
    z = pm.bij.inverse(C.nested_from_x(P["prior"], P["x"])).contiguous()
    lp, chi2, g = pm.log_prob_and_grad(sim, z.to(cuda))
    lp_ref, chi_ref, _ = C.oracle_backward_log_prob(P["ocam"], P["phys"], P["prior"], z.double().numpy(), P["obs"],
                                                    0.2, 100.0)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    _print("backward log_prob rel err:", rel.max(), "chi2", np.abs(chi2.cpu().numpy() - chi_ref).max() / chi_ref.max())
    assert rel.max() < 1e-4
    g = g.cpu().numpy()
    zz = z.double().numpy()
    for i in range(2):
        fd = np.zeros(zz.shape[1])
        for j in range(zz.shape[1]):
            h = 1e-4 * max(abs(zz[i, j]), 1e-1)
            zp, zm = zz[i:i + 1].copy(), zz[i:i + 1].copy()
            zp[0, j] += h
            zm[0, j] -= h
            fp = C.oracle_backward_log_prob(P["ocam"], P["phys"], P["prior"], zp, P["obs"], 0.2, 100.0)[0][0]
            fm = C.oracle_backward_log_prob(P["ocam"], P["phys"], P["prior"], zm, P["obs"], 0.2, 100.0)[0][0]
            fd[j] = (fp - fm) / (2 * h)
        scale = np.abs(fd).max()
        err = np.abs(g[i] - fd)
        _print(f"particle {i}: max |g - fd| / max|fd| = {err.max() / scale:.2e}")
        assert np.all(err <= 5e-3 * np.abs(fd) + 2e-3 * scale), (g[i], fd)


def test_lstsq_full_size_config(cuda):
    """100x100, ss 4, 21x21 Moffat (83x83 sub-grid), shapelets n_max=10 (67 channels)."""
    P = setup(100, 4, 0.04, 21, 10, True, bs=2, seed=3)
    sim, x = P["sim"], P["x"]
    nested = C.nested_from_x(P["prior"], x)
    st_ref, co_ref, im_ref = C.oracle_lstsq(P["ocam"], P["phys"], C.nested_from_x(P["prior"], x.numpy().astype(np.float64)),
                                            P["obs"], P["err"])
    img = sim.lstsq_simulate(nested, P["obs"], P["err"]).cpu().numpy()
    stacked = sim.lstsq_simulate(nested, P["obs"], P["err"], return_stacked=True).cpu().numpy()
    e_st = np.abs(stacked - st_ref).max() / np.abs(st_ref).max()
    e_im = (np.abs(img - im_ref).max(axis=(1, 2)) / np.abs(im_ref).max(axis=(1, 2))).max()
    _print(f"full size: stacked {e_st:.2e} image {e_im:.2e}")
    assert stacked.shape == (2, 100, 100, 67)
    assert e_st < 1e-4
    assert e_im < 1e-4


def test_backward_map_fused_matches_autograd(cuda):
    """MAP on the least-squares model: the fused step (pipeline + Adam in one
    pass) against autograd through log_prob with the same Keras Adam."""
    from gigalens_amd.inference import Adam, ModellingSequence
    P = setup(30, 2, 0.08, 7, 4, True, bs=16, seed=5)
    pm = BackwardProbModel(P["prior"], P["obs"], 0.2, 100.0)
    seq = ModellingSequence(P["phys"], pm, P["cfg"], device=cuda)
    z0 = pm.bij.inverse(C.nested_from_x(P["prior"], P["x"])).to(cuda).contiguous()
    z1 = seq.MAP(Adam(5e-3), start=z0.clone(), num_steps=5)
    sim = LensSimulator(P["phys"], P["cfg"], bs=16, device=cuda)
    z2 = seq._map_autograd(Adam(5e-3), z0.clone(), sim, 5)
    err = float((z1 - z2).abs().max())
    _print("least-squares MAP fused vs autograd max |z diff|:", err)
    assert err < 0.1 * 5 * 5e-3


def test_separable_channel_conv_matches_direct(cuda):
    """The staged plan's separable channel convolution (the least-squares default)
    against its direct one (psf_rank_tol = -1) at the full shapelet size (83x83
    Moffat, rank 7 at 1e-6)."""
    from gigalens_amd.plan import Plan
    P = setup(100, 4, 0.04, 21, 10, True, bs=2, seed=3)
    sim = P["sim"]
    x = P["x"].to(cuda).contiguous()
    direct = Plan(P["phys"], sim.camera, sim.layout, observed=P["obs"], noise_map=P["err"], psf_rank_tol=-1.0)
    sep = Plan(P["phys"], sim.camera, sim.layout, observed=P["obs"], noise_map=P["err"])
    assert direct.info()["sep_rank"] == 0 and sep.info()["sep_rank"] == 7
    _, _, st_d = direct.lstsq_simulate(x, want_image=False, want_stacked=True)
    _, _, st_s = sep.lstsq_simulate(x, want_image=False, want_stacked=True)
    e = float((st_d - st_s).abs().max() / st_d.abs().max())
    _print(f"separable vs direct channel conv: {e:.2e}")
    # the direct path sums 86x86 taps in FP32 (2e-5 of the peak from the FP64 oracle);
    # the separable one sums 7 x (86 + 86)
    assert e < 5e-5
