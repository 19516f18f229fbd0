"""Inference drivers on the GPU: the reference's invariants
(tests/tf/test_model.py:10-72) plus end-to-end checks."""
import math

import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator, PhysicalModel, SimulatorConfig
from gigalens_amd import distributions as tfd
from gigalens_amd import workloads as wl
from gigalens_amd.inference import Adam, ModellingSequence, PolynomialDecay
from gigalens_amd.profiles.light.sersic import SersicEllipse
from gigalens_amd.profiles.mass.epl import EPL
from gigalens_amd.profiles.mass.shear import Shear

pytestmark = pytest.mark.gpu


def _default():
    prior = wl.reference_default_prior()
    phys = PhysicalModel([EPL(), Shear()], [SersicEllipse()], [SersicEllipse()])
    data = (np.zeros((20, 20)), 0.1, 100)  # tests/conftest.py:83-85
    return prior, phys, data


def test_bij_roundtrip():
    """tests/tf/test_model.py:10-16"""
    prior, phys, _ = _default()
    model = ForwardProbModel(prior, np.ones((20, 20)), 1, 1)
    sample = prior.sample(5, seed=0)
    back = model.bij.forward(model.bij.inverse(sample))
    assert np.allclose(tfd.flatten(back), tfd.flatten(sample), rtol=1e-5, atol=1e-6)


def test_prior_jacobian_shape():
    """tests/tf/test_model.py:19-26"""
    prior, phys, _ = _default()
    model = ForwardProbModel(prior, np.ones((20, 20)), 1, 1)
    z = model.bij.inverse(prior.sample(5, seed=0))
    det = model.unconstraining_bij.forward_log_det_jacobian(model.pack_bij.forward(z))
    assert det.numel() == 5


def test_map_lr0_unchanged_lr_positive_changes(cuda):
    """tests/tf/test_model.py:29-42"""
    prior, phys, data = _default()
    pm = ForwardProbModel(prior, *data)
    cfg = SimulatorConfig(delta_pix=0.05, num_pix=20)
    seq = ModellingSequence(phys, pm, cfg)
    start = prior.sample(2, seed=1)
    ret = seq.MAP(Adam(0), start, n_samples=2, num_steps=5, seed=0)
    end = pm.bij.forward(ret.cpu())
    assert np.allclose(tfd.flatten(start), tfd.flatten(end), rtol=1e-5, atol=1e-6)
    ret = seq.MAP(Adam(1e-3), start, n_samples=2, num_steps=5, seed=0)
    end = pm.bij.forward(ret.cpu())
    assert not np.allclose(tfd.flatten(start), tfd.flatten(end))


def test_vi_lr0_unchanged_lr_positive_changes(cuda):
    """tests/tf/test_model.py:45-57"""
    prior, phys, data = _default()
    pm = ForwardProbModel(prior, *data)
    cfg = SimulatorConfig(delta_pix=0.05, num_pix=20)
    seq = ModellingSequence(phys, pm, cfg)
    start = pm.bij.inverse(prior.sample(2, seed=1))[0]
    q, losses = seq.SVI(Adam(0), start, n_vi=5, num_steps=5)
    assert np.allclose(q.mean().cpu(), start, atol=1e-6)
    q, losses = seq.SVI(Adam(1e-3), start, n_vi=5, num_steps=5)
    assert not np.allclose(q.mean().cpu(), start)
    assert losses.shape == (5,)


def test_vi_diagonal_lr0_unchanged_lr_positive_changes(cuda):
    """tests/tf/test_model.py:45-57 for the diagonal surrogate (full_rank=False): the
    scale parameters are a vector, lr 0 leaves mean and scale unchanged."""
    prior, phys, data = _default()
    pm = ForwardProbModel(prior, *data)
    cfg = SimulatorConfig(delta_pix=0.05, num_pix=20)
    seq = ModellingSequence(phys, pm, cfg)
    start = pm.bij.inverse(prior.sample(2, seed=1))[0]
    q, losses = seq.SVI(Adam(0), start, n_vi=5, num_steps=5, full_rank=False)
    assert q.raw.shape == (22,)
    assert np.allclose(q.mean().cpu(), start, atol=1e-6)
    assert np.allclose(q.stddev().cpu(), 1e-3, rtol=1e-5)
    q, losses = seq.SVI(Adam(1e-3), start, n_vi=5, num_steps=5, full_rank=False)
    assert not np.allclose(q.mean().cpu(), start)
    assert not np.allclose(q.stddev().cpu(), 1e-3, rtol=1e-5)
    assert losses.shape == (5,) and torch.isfinite(losses).all()


def test_hmc_sample_count(cuda):
    """tests/tf/test_model.py:60-72"""
    prior, phys, data = _default()
    pm = ForwardProbModel(prior, *data)
    cfg = SimulatorConfig(delta_pix=0.05, num_pix=20)
    seq = ModellingSequence(phys, pm, cfg)
    start = pm.bij.inverse(prior.sample(2, seed=1))[0]
    q, _ = seq.SVI(Adam(0), start, n_vi=2, num_steps=2)
    samples, stats = seq.HMC(q, n_hmc=3, init_eps=0.3, init_l=3, max_leapfrog_steps=5, num_burnin_steps=3,
                             num_results=5)
    assert len(samples) == 5
    assert samples.shape == (5, 3, 22)


def test_fused_map_equals_autograd_map(cuda):
    prob = C.epl_shear_problem(bs=8)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"])
    start = C.nested_from_x(prob["prior"], prob["x"][:8])
    z1 = seq.MAP(Adam(5e-3), start, n_samples=8, num_steps=4)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    z2 = seq._map_autograd(Adam(5e-3), pm.bij.inverse(start).to(cuda), sim, 4)
    assert torch.allclose(z1, z2, rtol=1e-4, atol=1e-5)


def test_map_reaches_noise_level(cuda):
    """Multi-start MAP from the prior on the 60x60 workload reaches mean r^2 <= 1.1."""
    prob = C.epl_shear_problem(bs=4)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"])
    z = seq.MAP(Adam(PolynomialDecay(1e-2, 300, 2e-3)), n_samples=1000, num_steps=300, seed=0)
    best = float(torch.nan_to_num(seq.chi2_history[-1], nan=1e9).min())
    print("best chi2 after MAP:", best, flush=True)
    assert best <= 1.1


def test_svi_hmc_recover_truth(cuda):
    """MAP -> SVI -> HMC on the 60x60 workload; the truth lies inside the HMC posterior."""
    prob = C.epl_shear_problem(bs=4)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"])
    z = seq.MAP(Adam(PolynomialDecay(1e-2, 300, 2e-3)), n_samples=1000, num_steps=300, seed=0)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=1000)
    lp, _ = pm.log_prob(sim, z)
    best = z[torch.argmax(torch.nan_to_num(lp, nan=-1e30))]
    q, losses = seq.SVI(Adam(PolynomialDecay(0.0, 200, 4e-3, 2)), best.cpu(), n_vi=200, num_steps=300)
    assert torch.isfinite(losses[-20:]).all()
    samples, stats = seq.HMC(q, n_hmc=32, num_burnin_steps=100, num_results=100, max_leapfrog_steps=20)
    assert samples.shape == (100, 32, 22)
    acc = stats["accept_rate"][-100:].mean()
    print("HMC acceptance:", acc, flush=True)
    assert 0.3 < acc <= 1.0
    x = pm.bij.forward(samples.reshape(-1, 22).cpu())
    flat = torch.stack([torch.as_tensor(v) for v in tfd.flatten(x)], 1)
    truth = torch.as_tensor(prob["truth_x"][0], dtype=torch.float32)
    mu, sd = flat.mean(0), flat.std(0)
    zscore = ((truth - mu) / sd).abs()
    print("max |z-score| of truth:", float(zscore.max()), flush=True)
    assert float(zscore[:6].max()) < 5.0  # EPL parameters
