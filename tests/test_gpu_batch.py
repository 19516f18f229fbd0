"""Several lens systems of one shape in one launch (gl_log_prob_lenses,
gl_map_adam_step_lenses, MultiLensMAP) against one plan per lens.
Bar: log_prob identical to the single-lens plans (same kernel, same data);
MAP trajectories within 1e-5 relative (the per-particle loss weight enters
as (-1/N^2) * w instead of -1/(n N^2))."""
import numpy as np
import pytest
import torch

from gigalens_amd import ForwardProbModel, LensSimulator
from gigalens_amd import workloads as wl
from gigalens_amd.batch import MultiLensMAP
from gigalens_amd.inference import Adam, ModellingSequence

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _lenses(cuda, n, seed=0):
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    cfg = wl.epl_shear_config()
    sim = LensSimulator(phys, cfg, bs=n, device=cuda)
    s = prior.sample(n, seed=seed)
    from gigalens_amd import distributions as tfd
    x = torch.stack([torch.as_tensor(v, dtype=torch.float32) for v in tfd.flatten(s)], 1).to(cuda)
    clean = sim.simulate(x).reshape(n, cfg.num_pix, cfg.num_pix).cpu().numpy()
    obs = np.stack([wl.noisy_observation(clean[k], 0.2, 100.0, seed=seed + 100 + k) for k in range(n)])
    return prior, phys, cfg, obs


def test_log_prob_lenses_matches_single_plans(cuda):
    L = 5
    prior, phys, cfg, obs = _lenses(cuda, L, seed=1)
    mlm = MultiLensMAP(phys, cfg, prior, obs, 0.2, 100.0, device=cuda)
    counts = [3, 7, 1, 4, 9]
    lens = torch.as_tensor(np.random.default_rng(0).permutation(np.repeat(np.arange(L), counts)), dtype=torch.int32,
                           device=cuda)
    z = mlm.bij.inverse(prior.sample(len(lens), seed=5)).to(cuda).float().contiguous()
    lp, chi2, g = mlm.plan.log_prob(z, want_grad=True, lens=lens)
    for k in range(L):
        pm = ForwardProbModel(prior, obs[k], 0.2, 100.0)
        sim = LensSimulator(phys, cfg, bs=1, device=cuda)
        sel = (lens == k).nonzero()[:, 0]
        lp1, chi1, g1 = pm.log_prob_and_grad(sim, z[sel].contiguous())
        assert torch.equal(lp[sel], lp1) and torch.equal(chi2[sel], chi1), k
        assert torch.equal(g[sel], g1), k
    bad = lens.clone()
    bad[0] = L
    lpb, chib, _ = mlm.plan.log_prob(z, lens=bad)
    assert torch.isnan(lpb[0]) and torch.isnan(chib[0])
    assert torch.equal(lpb[1:], lp[1:])


def test_multilens_map_matches_per_lens_map(cuda):
    L, n, steps = 3, 64, 20
    prior, phys, cfg, obs = _lenses(cuda, L, seed=2)
    mlm = MultiLensMAP(phys, cfg, prior, obs, 0.2, 100.0, device=cuda)
    lens = torch.arange(L, device=cuda, dtype=torch.int32).repeat_interleave(n)
    z0 = mlm.bij.inverse(prior.sample(L * n, seed=9)).to(cuda).float().contiguous()
    zb = mlm.run(z0, lens, steps, Adam(5e-3))
    for k in range(L):
        pm = ForwardProbModel(prior, obs[k], 0.2, 100.0)
        fitter = ModellingSequence(phys, pm, cfg, device=cuda)
        zk = fitter.MAP(optimizer=Adam(5e-3), start=z0[k * n:(k + 1) * n].clone(), num_steps=steps)
        err = float(((zb[k * n:(k + 1) * n] - zk).abs() / (zk.abs() + 1e-3)).max())
        lp_err = float((mlm.lp_history[-1, k * n:(k + 1) * n] - fitter.lp_history[-1]).abs().max()
                       / fitter.lp_history[-1].abs().max())
        _print(f"lens {k}: max rel z diff {err:.2e}, lp diff {lp_err:.2e}")
        assert err < 1e-5 and lp_err < 1e-5
