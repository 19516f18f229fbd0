"""HIP-graph capture of MAP steps (gigalens_amd.graphs): replays of a recorded
step sequence equal the eager launches bit for bit, on every replay, with eager
work and synchronisation between replays."""
import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator
from gigalens_amd.graphs import capture
from gigalens_amd.inference import Adam, ModellingSequence

pytestmark = pytest.mark.gpu


def test_map_graph_equals_eager(cuda):
    prob = C.epl_shear_problem(bs=64, seed=6)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"], device=cuda)
    z0 = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    z_e = seq.MAP(Adam(5e-3), start=z0.clone(), num_steps=20)
    lp_e = seq.lp_history.clone()
    z_g = seq.MAP(Adam(5e-3), start=z0.clone(), num_steps=20, graph=True)
    assert torch.equal(z_e, z_g) and torch.equal(lp_e, seq.lp_history)


def test_graph_replays_repeat(cuda):
    prob = C.epl_shear_problem(bs=32, seed=7)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=32, device=cuda)
    plan = pm.plan_for(sim)
    z0 = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    z, m, v = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lp = torch.empty(32, device=cuda)

    def steps():
        for s in range(1, 6):
            plan.map_adam_step(z, m, v, s, 5e-3, lp=lp)
    g = capture(steps)
    # eager reference
    ze, me, ve = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lpe = torch.empty(32, device=cuda)
    for s in range(1, 6):
        plan.map_adam_step(ze, me, ve, s, 5e-3, lp=lpe)
    for rep in range(3):
        z.copy_(z0)
        m.zero_()
        v.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = plan.log_prob(z0)  # eager work between replays
        assert torch.equal(z, ze) and torch.equal(lp, lpe), rep


def test_hmc_graph_leapfrog_equals_eager(cuda):
    """HMC with graph-captured leapfrog steps (one recorded step replayed L times,
    the step size in a device scalar) against the eager loop: identical samples and
    acceptance, with dual-averaging step-size adaptation changing the step size
    between trajectories."""
    prob = C.epl_shear_problem(bs=32, seed=8)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"], device=cuda)
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32)
    zt = pm.bij.inverse(C.nested_from_x(prob["prior"], xt))
    start = (zt.repeat(32, 1) + 1e-3 * torch.randn(32, zt.shape[1], generator=torch.Generator().manual_seed(2)))
    kw = dict(init_eps=1e-3, init_l=3, num_burnin_steps=10, num_results=6, max_leapfrog_steps=5, seed=4,
              start=start.to(cuda))
    s_e, st_e = seq.HMC(None, **kw)
    s_g, st_g = seq.HMC(None, graph=True, **kw)
    assert torch.equal(s_e, s_g)
    assert np.array_equal(st_e["accept_rate"], st_g["accept_rate"])
    assert np.array_equal(st_e["step_size"], st_g["step_size"])


def test_lens_batch_graph_equals_eager(cuda):
    """LensBatch.run(graph=True): the batched staged pipeline (lenses of ss 2, 3 and 4,
    separable convs, per-group launches) recorded into one HIP graph per call.  Two
    graph calls, with eager work and synchronisation between them, each against the
    eager run.  The statistics meet in atomic adds (order varies), so the bar is
    the batch-vs-per-lens one: z within 0.1 x steps x lr, lp within 1e-5 relative."""
    from gigalens_amd import SimulatorConfig
    from gigalens_amd import psf
    from gigalens_amd import workloads as wl
    from gigalens_amd.batch import LensBatch
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    kernel = psf.gaussian(0.1, 0.065, 11)
    lenses = []
    for k, (n_pix, ss, n) in enumerate([(48, 2, 9), (64, 3, 5), (100, 4, 3)]):
        cfg = SimulatorConfig(delta_pix=0.065, num_pix=n_pix, supersample=ss, kernel=kernel)
        obs = wl.noisy_observation(np.zeros((n_pix, n_pix)), 0.2, 100.0, seed=60 + k)
        lenses.append((cfg, obs, n))
    batch = LensBatch(phys, prior, lenses, 0.2, 100.0, device=cuda)
    z0 = batch.bij.inverse(prior.sample(batch.total, seed=5)).to(cuda).float().contiguous()
    steps, lr = 6, 5e-3
    z_e = batch.run(z0, steps, Adam(lr))
    lp_e = batch.lp_history.clone()
    for rep in range(2):
        torch.cuda.synchronize()
        _ = batch.log_prob(z0, want_grad=True)  # eager work between the graph calls
        torch.cuda.synchronize()
        z_g = batch.run(z0, steps, Adam(lr), graph=True)
        e_z = float((z_g - z_e).abs().max())
        e_lp = float(((batch.lp_history - lp_e).abs() / lp_e.abs()).max())
        print(f"graph call {rep}: max |z diff| {e_z:.2e}, lp history {e_lp:.2e}", flush=True)
        assert e_z < 0.1 * steps * lr and e_lp < 1e-5
