"""GPU parity of the learned initialiser's kernels (gl_cnn_*) against the CPU
oracle (oracle/cnn_ref.py, same bfloat16 rounding points), and what the
trained network gains over prior-sampled MAP starts.

Tolerances: network outputs within 1e-2 of their largest magnitude, the loss
within 1e-3 relative, each parameter block's gradient within 3e-2 of its
largest magnitude (bfloat16 rounding decisions can differ where the FP32
accumulation order differs), AdamW within 1e-6 relative."""
import math

import numpy as np
import pytest
import torch

from oracle import cnn_ref
from gigalens_amd import ForwardProbModel
from gigalens_amd import workloads as wl
from gigalens_amd.learned import ConvNet, LearnedInitializer

pytestmark = pytest.mark.gpu

CH = (32, 32, 64, 64)


def _print(*a):
    print(*a, flush=True)


def _net_and_batch(cuda, bs=4, d=22, seed=0):
    net = ConvNet(60, 60, d, CH, 256, max_batch=64, in_scale=0.2, seed=seed, device=cuda)
    rng = np.random.default_rng(seed + 1)
    img = torch.as_tensor(rng.gamma(1.0, 2.0, size=(bs, 60, 60)) + rng.normal(0, 0.2, size=(bs, 60, 60)),
                          dtype=torch.float32)
    tgt = torch.as_tensor(rng.normal(size=(bs, d)), dtype=torch.float32)
    return net, img, tgt


def test_forward_matches_oracle(cuda):
    net, img, _ = _net_and_batch(cuda)
    out = net.forward(img.to(cuda)).cpu()
    p = net.get_params().cpu()
    ref = cnn_ref.forward(cnn_ref.unpack(p, net.layout), img, 0.2, CH).detach()
    err = float((out - ref).abs().max() / ref.abs().max())
    _print("cnn forward max err / max:", err)
    assert err < 1e-2


def test_loss_and_gradient_match_oracle(cuda):
    net, img, tgt = _net_and_batch(cuda, bs=8, seed=3)
    loss, g = net.loss_grad(img.to(cuda), tgt.to(cuda), want_grad=True)
    p = net.get_params().cpu()
    lref, gref = cnn_ref.loss_and_grad(p, net.layout, img, tgt, 0.2, CH)
    rel = abs(float(loss) - float(lref)) / abs(float(lref))
    _print("cnn loss", float(loss), float(lref), rel)
    assert rel < 1e-3
    g = g.cpu()
    for name, shape, off in net.layout:
        n = int(np.prod(shape))
        a, b = g[off:off + n], gref[off:off + n]
        e = float((a - b).abs().max() / b.abs().max().clamp_min(1e-20))
        _print(f"  grad {name:8s} rel err {e:.2e}")
        assert e < 3e-2, name


# Channel counts and odd image sizes that take the convolution passes' other tile
# shapes: 32- and 64-channel output tiles, 32- and 64-channel K steps, several output
# chunks (weights reloaded per chunk) and several K steps, images whose rows and
# columns do not fill 16-pixel groups, and batches smaller than the resident grid.
_SHAPES = [((64, 96, 32, 128), 37, 45, 3), ((32, 64, 32, 64), 33, 29, 5), ((32, 64, 64, 32), 20, 24, 2)]


@pytest.mark.parametrize("channels,height,width,bs", _SHAPES)
def test_other_shapes_match_oracle(cuda, channels, height, width, bs):
    d = 7
    net = ConvNet(height, width, d, channels, 64, max_batch=8, in_scale=0.2, seed=11, device=cuda)
    rng = np.random.default_rng(12)
    img = torch.as_tensor(rng.gamma(1.0, 2.0, size=(bs, height, width)) + rng.normal(0, 0.2, size=(bs, height, width)),
                          dtype=torch.float32)
    tgt = torch.as_tensor(rng.normal(size=(bs, d)), dtype=torch.float32)
    p = net.get_params().cpu()
    out = net.forward(img.to(cuda)).cpu()
    ref = cnn_ref.forward(cnn_ref.unpack(p, net.layout), img, 0.2, channels).detach()
    err = float((out - ref).abs().max() / ref.abs().max())
    loss, g = net.loss_grad(img.to(cuda), tgt.to(cuda), want_grad=True)
    lref, gref = cnn_ref.loss_and_grad(p, net.layout, img, tgt, 0.2, channels)
    rel = abs(float(loss) - float(lref)) / abs(float(lref))
    _print(f"{channels} {height}x{width} bs {bs}: forward err {err:.2e}, loss rel {rel:.2e}")
    assert err < 1e-2
    assert rel < 1e-3
    g = g.cpu()
    for name, shape, off in net.layout:
        n = int(np.prod(shape))
        a, b = g[off:off + n], gref[off:off + n]
        e = float((a - b).abs().max() / b.abs().max().clamp_min(1e-20))
        _print(f"  grad {name:8s} rel err {e:.2e}")
        assert e < 3e-2, name


def test_adamw_matches_torch(cuda):
    net, img, tgt = _net_and_batch(cuda, bs=4, seed=5)
    p0 = net.get_params().cpu().clone()
    _, g = net.loss_grad(img.to(cuda), tgt.to(cuda), want_grad=True)
    net.adamw_step(1, 3e-4, weight_decay=0.01)
    net.loss_grad(img.to(cuda), tgt.to(cuda))
    _, g2 = net.loss_grad(img.to(cuda), tgt.to(cuda), want_grad=True)
    net.adamw_step(2, 3e-4, weight_decay=0.01)
    p2 = net.get_params().cpu()
    w = p0.clone().requires_grad_(True)
    opt = torch.optim.AdamW([w], lr=3e-4, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.01)
    for gg in (g, g2):
        opt.zero_grad()
        w.grad = gg.cpu().clone()
        opt.step()
    err = float((p2 - w.detach()).abs().max() / w.detach().abs().max())
    _print("adamw rel err", err)
    assert err < 1e-6


def test_training_beats_prior_starts(cuda):
    """A short training run (1200 x 256 simulated images) lowers the NLL, and on
    held-out lenses the network's proposals start MAP at higher log-posterior
    than prior draws and reach chi2 <= 1.1 in fewer steps."""
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    cfg = wl.epl_shear_config()
    init = LearnedInitializer(phys, cfg, prior, 0.2, 100.0, max_batch=256, seed=0, device=cuda)
    losses = init.train(n_images=1200 * 256, batch=256, lr=1e-3, seed=10)
    first, last = float(np.mean(losses[:10])), float(np.mean(losses[-10:]))
    _print(f"training NLL: first10 {first:.3f} last10 {last:.3f}")
    # an uninformed Gaussian (mean 0, log-std 0 on standardised z) scores 11 nats
    assert last < first - 3.0 and last < 9.0
    from gigalens_amd import LensSimulator
    from gigalens_amd.inference import ModellingSequence
    from gigalens_amd.learned import map_trace, first_step
    wins_lp, steps_l, steps_p = 0, [], []
    n_lens = 6
    for k in range(n_lens):
        images, z_true = init.simulate_batch(1, seed=1000 + k)
        obs = images[0].cpu().numpy()
        pm = ForwardProbModel(prior, obs, 0.2, 100.0)
        fitter = ModellingSequence(phys, pm, cfg, device=cuda)
        z_l = init.propose(images[0], 256, prior_fraction=0.25, seed=k)
        z_p = init.bij.inverse(prior.sample(256, seed=5000 + k)).to(cuda).float().contiguous()
        lp_l, chi_l, _ = map_trace(fitter, z_l, 200, 5e-3)
        lp_p, chi_p, _ = map_trace(fitter, z_p, 200, 5e-3)
        wins_lp += int(lp_l[0] > lp_p[0])
        sl = first_step(chi_l, lambda c: c <= 1.1)
        sp = first_step(chi_p, lambda c: c <= 1.1)
        steps_l.append(sl if sl is not None else 10 ** 6)
        steps_p.append(sp if sp is not None else 10 ** 6)
        _print(f"lens {k}: start lp learned {lp_l[0]:.1f} prior {lp_p[0]:.1f}; steps to chi2<=1.1 learned {sl} prior {sp}")
    assert wins_lp >= n_lens - 1
    assert np.median(steps_l) <= np.median(steps_p)


def test_compare_initialisation_small(cuda):
    """The held-out learned-vs-prior harness on 8 lenses x 256 particles."""
    from gigalens_amd.learned import compare_initialisation
    prior = wl.epl_shear_prior()
    init = LearnedInitializer(wl.epl_shear_model(), wl.epl_shear_config(), prior, 0.2, 100.0, max_batch=256,
                              seed=0, device=cuda)
    init.train(n_images=1200 * 256, batch=256, lr=1e-3, seed=10)
    r = compare_initialisation(init, n_lenses=8, n_particles=256, num_steps=200, lr=5e-3)
    _print(r)
    tl, tp = r["time_to_chi2"]["learned"], r["time_to_chi2"]["prior"]
    assert tl["reached"] >= tp["reached"]
    assert r["final_best_lp_gain_vs_prior_p50"]["learned"] > -1.0
