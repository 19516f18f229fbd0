"""The HBM-staged pipeline with fixed amplitudes, and batches of lens systems
with different cameras and particle counts (gl_batch_*).

Bars: staged vs the LDS-resident fused kernel on the 60x60 workload -- images
within 1e-5 of the peak, log_prob within 1e-5 relative, gradients within 1e-4
of the largest component (the two paths sum the same terms in a different
order); an image too large for the fused kernel (128x128, ss 4, 43x43 PSF)
against the FP64 oracle -- images within 1e-4 of the peak, log_prob within
1e-4 relative, gradient vs FP64 central differences; a batch of uneven lenses
against one staged plan per lens -- identical log_prob (same kernels and data), gradients
within 1e-5 (the statistics are summed with atomic adds, whose order varies)."""
import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator, SimulatorConfig
from gigalens_amd import distributions as tfd
from gigalens_amd import psf
from gigalens_amd import workloads as wl
from gigalens_amd.batch import LensBatch
from gigalens_amd.inference import Adam
from gigalens_amd.plan import Plan, layout_from_paths
from gigalens_amd.simulator import Camera

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _plans(prob, staged):
    prior, phys = prob["prior"], prob["phys"]
    paths = [p for p, _ in tfd.leaves(prior)]
    return Plan(phys, prob["cam"], layout_from_paths(phys, paths), prior=prior, observed=prob["obs"],
                background_rms=0.2, exp_time=100.0, staged=staged)


def test_staged_matches_fused(cuda):
    prob = C.epl_shear_problem(bs=12, seed=4)
    fused, staged = _plans(prob, False), _plans(prob, True)
    assert fused.info()["lds_bytes"] > 0 and staged.info()["lds_bytes"] == 0
    x = prob["x"].to(cuda).contiguous()
    im_f, im_s = fused.simulate(x), staged.simulate(x)
    e_im = float((im_f - im_s).abs().max() / im_f.abs().max())
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    lp_f, chi_f, g_f = fused.log_prob(z, want_grad=True)
    lp_s, chi_s, g_s = staged.log_prob(z, want_grad=True)
    e_lp = float(((lp_f - lp_s).abs() / lp_f.abs()).max())
    e_g = float(((g_f - g_s).abs().max(1).values / g_f.abs().max(1).values).max())
    gimg = torch.randn_like(im_f)
    v_f, v_s = fused.simulate_vjp(x, gimg), staged.simulate_vjp(x, gimg)
    e_v = float(((v_f - v_s).abs().max(1).values / v_f.abs().max(1).values).max())
    _print(f"staged vs fused: image {e_im:.2e} lp {e_lp:.2e} grad {e_g:.2e} vjp {e_v:.2e}")
    assert e_im < 1e-5 and e_lp < 1e-5 and e_g < 1e-4 and e_v < 1e-4
    # MAP steps: Adam moves each coordinate by about lr * sign(g) early on, so
    # gradient components near zero can step differently under FP32 summation
    # noise; compare the log posterior and bound z by a fraction of the steps taken
    zs, zf = z.clone(), z.clone()
    ms, vs, mf, vf = (torch.zeros_like(z) for _ in range(4))
    lpf, lps = torch.empty(z.shape[0], device=cuda), torch.empty(z.shape[0], device=cuda)
    for s in range(1, 11):
        fused.map_adam_step(zf, mf, vf, s, 5e-3, lp=lpf)
        staged.map_adam_step(zs, ms, vs, s, 5e-3, lp=lps)
    e_z = float((zf - zs).abs().max())
    e_lp10 = float(((lpf - lps).abs() / lpf.abs()).median())
    _print(f"10 MAP steps: max |z diff| {e_z:.2e} (10 steps of lr 5e-3), median lp rel diff {e_lp10:.2e}")
    assert e_z < 0.1 * 10 * 5e-3 and e_lp10 < 1e-4


def _big_problem(n_pix=128, ss=4, seed=7):
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    kernel = psf.gaussian(0.1, 0.065, 11)
    cfg = SimulatorConfig(delta_pix=0.065 * 60 / n_pix * 1.2, num_pix=n_pix, supersample=ss, kernel=kernel)
    cam = Camera(cfg)
    ocam = C.oracle_camera(cfg, cam)
    truth = C.draw_prior(prior, 1, seed=seed).numpy().astype(np.float64)
    clean = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, truth))[0]
    obs = wl.noisy_observation(clean, 0.2, 100.0, seed=seed + 1)
    return prior, phys, cfg, cam, ocam, obs


def test_large_image_staged_vs_oracle(cuda):
    prior, phys, cfg, cam, ocam, obs = _big_problem()
    pm = ForwardProbModel(prior, obs, 0.2, 100.0)
    sim = LensSimulator(phys, cfg, bs=2, device=cuda)
    plan = pm.plan_for(sim)
    assert plan.info()["lds_bytes"] == 0, "expected the staged path for 128x128 ss 4"
    x = C.draw_prior(prior, 2, seed=11)
    img = sim.plan.simulate(x.to(cuda).contiguous()).cpu().numpy()
    ref = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, x.numpy().astype(np.float64)))
    e_im = (np.abs(img - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))).max()
    z = pm.bij.inverse(C.nested_from_x(prior, x)).contiguous()
    lp, chi2, g = pm.log_prob_and_grad(sim, z.to(cuda))
    zz = z.double().numpy()
    lp_ref, chi_ref, _ = C.oracle_log_prob(ocam, phys, prior, zz, obs, 0.2, 100.0)
    e_lp = float(np.max(np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)))
    _print(f"128x128 ss4: image {e_im:.2e} lp {e_lp:.2e}")
    assert e_im < 1e-4 and e_lp < 1e-4
    g = g.cpu().numpy()[0]
    fd = np.zeros(zz.shape[1])
    for j in range(zz.shape[1]):
        h = 1e-4 * max(abs(zz[0, j]), 0.1)
        zp, zm = zz[:1].copy(), zz[:1].copy()
        zp[0, j] += h
        zm[0, j] -= h
        fd[j] = (C.oracle_log_prob(ocam, phys, prior, zp, obs, 0.2, 100.0)[0][0]
                 - C.oracle_log_prob(ocam, phys, prior, zm, obs, 0.2, 100.0)[0][0]) / (2 * h)
    err = np.abs(g - fd)
    _print(f"128x128 ss4 gradient: max |g - fd| / max |fd| = {err.max() / np.abs(fd).max():.2e}")
    assert np.all(err <= 5e-3 * np.abs(fd) + 2e-3 * np.abs(fd).max())


def test_uneven_batch_matches_per_lens_plans(cuda):
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    kernel = psf.gaussian(0.1, 0.065, 11)
    specs = [(48, 2, 17), (100, 4, 5), (64, 3, 9), (80, 2, 3)]
    lenses, plans = [], []
    paths = [p for p, _ in tfd.leaves(prior)]
    for k, (n_pix, ss, n) in enumerate(specs):
        cfg = SimulatorConfig(delta_pix=0.065, num_pix=n_pix, supersample=ss, kernel=kernel)
        sim = LensSimulator(phys, cfg, bs=1, device=cuda)
        truth = C.draw_prior(prior, 1, seed=20 + k).to(cuda)
        clean = sim.plan.simulate(truth).cpu().numpy()[0] if n_pix * ss <= 160 else None
        if clean is None:  # staged simulate for the large lenses
            sp = Plan(phys, Camera(cfg), layout_from_paths(phys, paths), staged=True)
            clean = sp.simulate(truth.contiguous()).cpu().numpy()[0]
        obs = wl.noisy_observation(clean, 0.2, 100.0, seed=40 + k)
        lenses.append((cfg, obs, n))
        # the same execution path as inside the batch: fused where the image fits LDS
        plans.append(Plan(phys, Camera(cfg), layout_from_paths(phys, paths), prior=prior, observed=obs,
                          background_rms=0.2, exp_time=100.0))
    batch = LensBatch(phys, prior, lenses, 0.2, 100.0, device=cuda)
    assert batch.total == sum(n for *_, n in specs)
    z = batch.bij.inverse(prior.sample(batch.total, seed=3)).to(cuda).float().contiguous()
    lp, chi2, g = batch.log_prob(z, want_grad=True)
    off = 0
    for k, (n_pix, ss, n) in enumerate(specs):
        lp1, chi1, g1 = plans[k].log_prob(z[off:off + n].contiguous(), want_grad=True)
        e = float((lp[off:off + n] - lp1).abs().max() / lp1.abs().max())
        eg = float((g[off:off + n] - g1).abs().max() / g1.abs().max())
        _print(f"lens {k} ({n_pix}px ss{ss}, {n} particles): lp diff {e:.2e}, grad diff {eg:.2e}")
        assert e < 1e-6 and eg < 1e-5
        off += n
    # a few MAP steps of the whole batch vs each lens alone
    zb = batch.run(z, 5, Adam(5e-3))
    off = 0
    for k, (_, _, n) in enumerate(specs):
        zk, mk, vk = z[off:off + n].clone(), torch.zeros((n, z.shape[1]), device=cuda), torch.zeros((n, z.shape[1]), device=cuda)
        for s in range(1, 6):
            plans[k].map_adam_step(zk, mk, vk, s, 5e-3)
        e = float((zb[off:off + n] - zk).abs().max())
        _print(f"lens {k}: MAP max |z diff| {e:.2e} after 5 steps of lr 5e-3")
        assert e < 0.1 * 5 * 5e-3  # gradients agree to ~1e-6 (atomic-add order); Adam steps ~lr sign(g)
        off += n


def test_conv_rows_variants_agree(cuda):
    """The staged conv's rows-per-lane tuning choice (2, 4, 8) does not change results."""
    prob = C.epl_shear_problem(bs=8, seed=12)
    prior, phys = prob["prior"], prob["phys"]
    paths = [p for p, _ in tfd.leaves(prior)]
    pm = ForwardProbModel(prior, prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prior, prob["x"])).to(cuda).contiguous()
    out = {}
    for r in (2, 4, 8):
        plan = Plan(phys, prob["cam"], layout_from_paths(phys, paths), prior=prior, observed=prob["obs"],
                    background_rms=0.2, exp_time=100.0, staged=True, conv_rows=r)
        out[r] = plan.log_prob(z, want_grad=True)
    for r in (2, 8):
        assert torch.equal(out[r][0], out[4][0]), r
        assert float((out[r][2] - out[4][2]).abs().max() / out[4][2].abs().max()) < 1e-5


def test_separable_adjoint_matches_direct(cuda):
    """The staged pipeline's separable channel and adjoint convs (default for a
    low-rank folded PSF) against their direct forms (psf_rank_tol = -1): the two sum
    the same terms of the rank-1 Gaussian PSF in a different order.  Bars: gradients
    within 1e-5 of the largest component, log_prob within 1e-6 relative."""
    prob = C.epl_shear_problem(bs=8, seed=6)
    prior, phys = prob["prior"], prob["phys"]
    paths = [p for p, _ in tfd.leaves(prior)]
    mk = lambda tol: Plan(phys, prob["cam"], layout_from_paths(phys, paths), prior=prior, observed=prob["obs"],
                          background_rms=0.2, exp_time=100.0, staged=True, psf_rank_tol=tol)
    sep, direct = mk(0.0), mk(-1.0)
    pm = ForwardProbModel(prior, prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prior, prob["x"])).to(cuda).contiguous()
    lp_s, _, g_s = sep.log_prob(z, want_grad=True)
    lp_d, _, g_d = direct.log_prob(z, want_grad=True)
    e_g = float(((g_s - g_d).abs().max(1).values / g_d.abs().max(1).values).max())
    _print(f"separable vs direct staged adjoint: grad {e_g:.2e}")
    e_lp = float(((lp_s - lp_d).abs() / lp_d.abs()).max())
    _print(f"separable vs direct staged log_prob: {e_lp:.2e}")
    assert sep.info()["sep_rank"] == 1 and direct.info()["sep_rank"] == 0
    assert e_lp < 1e-6
    assert e_g < 1e-5


def test_separable_rank5_ss3_matches_direct_and_oracle(cuda):
    """ss 3 with an 11x11 Gaussian PSF: the folded sub-grid PSF has rank 5 at 1e-6, so
    a staged plan runs the separable channel conv and the separable adjoint (both
    launched at compile-time ss 3).  Bars: image and log_prob within 1e-4 of the FP64
    oracle; log_prob within 1e-6 relative and gradients within 1e-5 of the largest
    component of the direct staged plan (psf_rank_tol = -1)."""
    prior, phys, cfg, cam, ocam, obs = _big_problem(n_pix=64, ss=3, seed=13)
    paths = [p for p, _ in tfd.leaves(prior)]
    mk = lambda tol: Plan(phys, cam, layout_from_paths(phys, paths), prior=prior, observed=obs,
                          background_rms=0.2, exp_time=100.0, staged=True, psf_rank_tol=tol)
    sep, direct = mk(0.0), mk(-1.0)
    info = sep.info()
    assert info["sep_rank"] == 5 and info["adj_rank"] == 5, info
    assert direct.info()["sep_rank"] == 0 and direct.info()["adj_rank"] == 0
    x = C.draw_prior(prior, 4, seed=14)
    img = sep.simulate(x.to(cuda).contiguous()).cpu().numpy()
    ref = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, x.numpy().astype(np.float64)))
    e_im = float((np.abs(img - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))).max())
    pm = ForwardProbModel(prior, obs, 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prior, x)).contiguous()
    lp_s, _, g_s = sep.log_prob(z.to(cuda), want_grad=True)
    lp_d, _, g_d = direct.log_prob(z.to(cuda), want_grad=True)
    lp_ref, _, _ = C.oracle_log_prob(ocam, phys, prior, z.double().numpy(), obs, 0.2, 100.0)
    e_lp = float(np.max(np.abs(lp_s.cpu().numpy() - lp_ref) / np.abs(lp_ref)))
    e_sd = float(((lp_s - lp_d).abs() / lp_d.abs()).max())
    e_g = float(((g_s - g_d).abs().max(1).values / g_d.abs().max(1).values).max())
    _print(f"64x64 ss3 rank 5: image vs oracle {e_im:.2e}, lp vs oracle {e_lp:.2e}, "
           f"separable vs direct lp {e_sd:.2e} grad {e_g:.2e}")
    assert e_im < 1e-4 and e_lp < 1e-4
    assert e_sd < 1e-6 and e_g < 1e-5
