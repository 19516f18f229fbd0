"""The uneven-batch fitting pipeline end to end, at a small scale: the learned
initialiser trained on the device, its proposals screened by log-posterior, the
batched MAP on the banded persistent kernel recorded into HIP graphs, then the
device diagonal SVI -- the same calls bench.py's uneven workload makes, on four
lenses of different sizes and supersampling.  The reference fits one lens per
simulator (src/gigalens/tf/inference.py:17-73); the batching has no counterpart.

Checks: every step ran on the banded kernel (one launch); each lens's best
log-posterior over the MAP steps ends above its best start; each lens's best MAP
particle has the FP64 oracle's log_prob and chi2 (oracle/bridge.py) within 1e-4
relative; the SVI losses stay finite and the last steps' mean loss is below the
first steps' mean for every lens."""
import numpy as np
import pytest
import torch

from oracle import bridge as ob
from tests import common as C
from gigalens_amd import SimulatorConfig
from gigalens_amd import psf
from gigalens_amd import workloads as wl
from gigalens_amd.batch import LensBatch
from gigalens_amd.inference import Adam
from gigalens_amd.learned import LearnedInitializer, propose_batch
from gigalens_amd.simulator import Camera

pytestmark = pytest.mark.gpu

SPECS = [(48, 2, 96), (64, 3, 64), (80, 2, 128), (100, 4, 32)]  # (N, ss, particles)


def _print(*a):
    print(*a, flush=True)


def test_learned_map_svi_pipeline(cuda):
    prior, phys = wl.epl_shear_prior(), wl.epl_shear_model()
    kernel = psf.gaussian(0.1, 0.065, 11)
    lenses = []
    for k, (n_pix, ss, n) in enumerate(SPECS):
        cfg = SimulatorConfig(delta_pix=0.065, num_pix=n_pix, supersample=ss, kernel=kernel)
        ocam = C.oracle_camera(cfg, Camera(cfg))
        truth = C.draw_prior(prior, 1, seed=2000 + k).numpy().astype(np.float64)
        clean = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, truth))[0]
        lenses.append((cfg, wl.noisy_observation(clean, 0.2, 100.0, seed=3000 + k), n))
    images = [o for _, o, _ in lenses]

    init = LearnedInitializer(phys, wl.epl_shear_config(), prior, 0.2, 100.0, max_batch=256, seed=0, device=cuda)
    losses = init.train(n_images=64 * 256, batch=256, lr=1e-3, seed=10)
    assert all(np.isfinite(losses))

    batch = LensBatch(phys, prior, lenses, 0.2, 100.0, device=cuda, band=1)
    assert batch.info["lenses_band"] == len(SPECS) and batch.info["chunks"] == 0
    z0 = propose_batch(init, batch, images, prior_fraction=0.25, oversample=2, seed=5)
    lp0, _, _ = batch.log_prob(z0)
    lp0 = torch.nan_to_num(lp0, nan=-float("inf")).cpu()
    start_best = [float(lp0[int(batch.first[k]):int(batch.first[k]) + n].max()) for k, (*_, n) in enumerate(lenses)]

    z = batch.run(z0, 150, Adam(5e-3), graph=True, graph_steps=25)
    best_lp, best_chi = batch.per_lens_best()  # each lens's best particle at every step
    best_lp = np.maximum.accumulate(np.asarray(best_lp), axis=0)
    assert np.isfinite(best_lp[-1]).all()
    for k in range(len(SPECS)):
        _print(f"lens {k} {SPECS[k]}: best start lp {start_best[k]:.1f}, MAP best lp {best_lp[-1, k]:.1f}, "
               f"chi2 at the last step {float(best_chi[-1][k]):.3f}")
        assert best_lp[-1, k] > start_best[k]

    # each lens's best MAP particle against the FP64 oracle
    loc = batch.best_particles(z)
    for k, (cfg, obs, _) in enumerate(lenses):
        zk = loc[k:k + 1].contiguous()
        gen = LensBatch(phys, prior, [(cfg, obs, 1)], 0.2, 100.0, device=cuda, band=1)
        lp_k, chi_k, _ = gen.log_prob(zk)
        ocam = ob.oracle_camera(cfg)
        lp_ref, chi_ref, _ = ob.oracle_log_prob(ocam, phys, prior, zk.double().cpu().numpy(), obs, 0.2, 100.0)
        e_lp = abs(float(lp_k[0]) - float(lp_ref[0])) / abs(float(lp_ref[0]))
        e_chi = abs(float(chi_k[0]) - float(chi_ref[0])) / abs(float(chi_ref[0]))
        _print(f"lens {k}: MAP point lp {float(lp_k[0]):.3f} (oracle {float(lp_ref[0]):.3f}), rel {e_lp:.1e}; "
               f"chi2 rel {e_chi:.1e}")
        assert e_lp < 1e-4 and e_chi < 1e-4
        del gen

    # diagonal SVI from each lens's best MAP particle, every step a graph replay
    svi = LensBatch(phys, prior, [(c, o, 32) for c, o, _ in lenses], 0.2, 100.0, device=cuda, band=1)
    loc_f, log_scale, svi_losses = svi.svi_diag(loc, 100, Adam(5e-3), seed=2, graph=True)
    svi_losses = torch.as_tensor(svi_losses).cpu()
    assert torch.isfinite(svi_losses).all()
    first, last = svi_losses[:10].mean(0), svi_losses[-10:].mean(0)
    _print("SVI loss first10", first.tolist(), "last10", last.tolist())
    assert (last < first).all()
    assert torch.isfinite(loc_f).all() and torch.isfinite(log_scale).all()
