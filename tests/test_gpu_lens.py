"""GPU parity of the fused lens-model kernels against the FP64 oracle.

Tolerances (stated per test): images within 1e-4 of the image peak,
log-probabilities within 1e-4 relative, gradients against FP64 central
finite differences of the oracle (h = 1e-3 relative).
"""
import math

import numpy as np
import pytest
import torch

from oracle import np64, torch_ref
from tests import common as C
from gigalens_amd import LensSimulator, ForwardProbModel
from gigalens_amd import workloads as wl
from gigalens_amd.simulator import Camera, SimulatorConfig
from gigalens_amd import psf

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def prob(cuda):
    return C.epl_shear_problem(bs=16)


def test_simulate_parity(prob, cuda):
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=16)
    x = prob["x"]
    img = sim.simulate(x.to(cuda)).cpu().numpy().astype(np.float64)
    ref = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prob["prior"], x.numpy().astype(np.float64)))
    err = np.abs(img - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    _print("simulate max err / peak per particle:", err.max())
    assert np.all(err < 1e-4)


def test_simulate_nested_params_truth(cuda):
    """Reference notebook truth (tf-demo.ipynb cell 5) through the nested-params API."""
    cfg = SimulatorConfig(delta_pix=0.065, num_pix=60, supersample=2,
                          kernel=np.load(C.GOLDEN + "/psf.npy").astype(np.float32))
    phys = wl.epl_shear_model()
    sim = LensSimulator(phys, cfg, bs=1)
    img = sim.simulate(wl.DEMO_TRUTH).cpu().numpy().astype(np.float64)
    ocam = C.oracle_camera(cfg, sim.camera)
    lenses, ll, src = C.oracle_model_lists(phys, wl.DEMO_TRUTH)
    ref = np64.simulate(ocam, lenses, ll, src)[0]
    err = np.abs(img - ref).max() / np.abs(ref).max()
    _print("truth image err / peak:", err)
    assert img.shape == (60, 60)
    assert err < 1e-4
    # statistical pin of the grid orientation / PSF / subgrid / pixel-area chain:
    # the reference's demo image is a noisy draw of this truth (bkg 0.2, t_exp 100)
    obs = np.load(C.GOLDEN + "/demo.npy")
    sig = np.sqrt(0.2 ** 2 + np.clip(img, 0, None) / 100)
    chi2 = np.mean(((img - obs) / sig) ** 2)
    _print("demo chi2 at truth:", chi2)
    assert abs(chi2 - 1.0) < 3 * math.sqrt(2.0 / 3600) + 0.02


def test_log_prob_parity(prob, cuda):
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=16)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    x = prob["x"]
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], x))
    lp, chi2 = pm.log_prob(sim, z.to(cuda))
    lp_ref, chi_ref, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], z.numpy().astype(np.float64),
                                           prob["obs"], 0.2, 100.0)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    relc = np.abs(chi2.cpu().numpy() - chi_ref) / np.abs(chi_ref)
    _print("log_prob rel err:", rel.max(), "chi2 rel err:", relc.max())
    assert rel.max() < 1e-4
    assert relc.max() < 1e-4


def _fd_grad(prob, z, h_rel=1e-3):
    zz = z.astype(np.float64)
    g = np.zeros_like(zz)
    for j in range(zz.shape[1]):
        h = h_rel * max(abs(zz[0, j]), 1e-1)
        zp, zm = zz.copy(), zz.copy()
        zp[:, j] += h
        zm[:, j] -= h
        fp = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zp, prob["obs"], 0.2, 100.0)[0]
        fm = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zm, prob["obs"], 0.2, 100.0)[0]
        g[:, j] = (fp - fm) / (2 * h)
    return g


def test_gradient_vs_fd(prob, cuda):
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=4)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    # particles near the truth so the likelihood terms dominate realistically
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32)
    x = xt.repeat(4, 1) * (1 + 0.02 * torch.randn(4, xt.shape[1], generator=torch.Generator().manual_seed(5)))
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], x)).contiguous()
    lp, chi2, g = pm.log_prob_and_grad(sim, z.to(cuda))
    g = g.cpu().numpy()
    for i in range(4):
        gfd = _fd_grad(prob, z[i:i + 1].numpy())[0]
        scale = np.abs(gfd).max()
        err = np.abs(g[i] - gfd)
        _print(f"particle {i}: max |g - fd| / max|fd| = {err.max() / scale:.2e}")
        assert np.all(err <= 2e-3 * np.abs(gfd) + 2e-4 * scale), (g[i], gfd)


def test_autograd_matches_fused(prob, cuda):
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    x = prob["x"][:8]
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], x)).to(cuda).requires_grad_(True)
    lp, chi2 = pm.log_prob(sim, z)
    w = torch.linspace(0.5, 1.5, 8, device=cuda)
    (lp * w).sum().backward()
    _, _, g = pm.log_prob_and_grad(sim, z.detach(), g_lp=w.contiguous())
    assert torch.allclose(z.grad, g, rtol=1e-5, atol=1e-5 * g.abs().max().item())


def test_chi2_gradient_vs_fd(prob, cuda):
    """Gradient of the chi^2 output (second log_prob output) by autograd."""
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=2)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32).repeat(2, 1)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], xt)).contiguous()
    zc = z.to(cuda).requires_grad_(True)
    lp, chi2 = pm.log_prob(sim, zc)
    chi2.sum().backward()
    g = zc.grad.cpu().numpy()[0]
    zz = z[:1].numpy().astype(np.float64)
    for j in (0, 5, 10, 15, 21):
        h = 1e-3 * max(abs(zz[0, j]), 1e-1)
        zp, zm = zz.copy(), zz.copy()
        zp[0, j] += h
        zm[0, j] -= h
        cp = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zp, prob["obs"], 0.2, 100.0)[1]
        cm = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], zm, prob["obs"], 0.2, 100.0)[1]
        fd = (cp - cm)[0] / (2 * h)
        assert abs(g[j] - fd) <= 5e-3 * abs(fd) + 1e-4, (j, g[j], fd)


def test_simulate_vjp_vs_torch_ref(prob, cuda):
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=3)
    x = torch.as_tensor(prob["truth_x"], dtype=torch.float32).repeat(3, 1)
    x = x * (1 + 0.01 * torch.randn(x.shape, generator=torch.Generator().manual_seed(7)))
    w = torch.randn(3, 60, 60, generator=torch.Generator().manual_seed(8))
    xc = x.to(cuda).requires_grad_(True)
    img = sim.simulate(xc)
    (img * w.to(cuda)).sum().backward()
    # FP64 autodiff of the op-for-op torch restatement
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SERSIC_ELLIPSE", {})]}
    ts = torch_ref.Sim(model, prob["ocam"], dtype=torch.float64)
    xr = x.double().requires_grad_(True)
    nested = C.nested_from_x(prob["prior"], xr)
    ref = ts.simulate(nested)
    (ref * w.double()).sum().backward()
    g, gr = xc.grad.cpu().double(), xr.grad
    rel = ((g - gr).abs() / (gr.abs() + 1e-3 * gr.abs().max())).max().item()
    _print("simulate vjp rel err:", rel)
    assert rel < 2e-3


@pytest.mark.parametrize("variant", ["psf_npy", "no_psf", "ss1", "ss3"])
def test_conv_variants(variant, cuda):
    """Static (ss,kf) specialisations and the generic path against the oracle."""
    kern = psf.gaussian(0.1, 0.065, 11)
    if variant == "psf_npy":
        cfg = SimulatorConfig(0.065, 60, 2, np.load(C.GOLDEN + "/psf.npy").astype(np.float32))
    elif variant == "no_psf":
        cfg = SimulatorConfig(0.065, 60, 2, None)
    elif variant == "ss1":
        cfg = SimulatorConfig(0.065, 40, 1, psf.gaussian(0.1, 0.065, 21))
    else:
        cfg = SimulatorConfig(0.08, 30, 3, kern)
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    x = C.draw_prior(prior, 4, seed=11)
    sim = LensSimulator(phys, cfg, bs=4)
    img = sim.simulate(x.to(cuda)).cpu().numpy().astype(np.float64)
    ocam = C.oracle_camera(cfg, sim.camera)
    ref = C.oracle_simulate(ocam, phys, C.nested_from_x(prior, x.numpy().astype(np.float64)))
    err = np.abs(img - ref).max() / np.abs(ref).max()
    _print(variant, "err/peak", err)
    assert err < 1e-4
    # adjoint of the same conv path: VJP vs FP64 torch autodiff
    w = torch.randn(img.shape, generator=torch.Generator().manual_seed(3))
    xc = x.to(cuda).requires_grad_(True)
    (sim.simulate(xc) * w.to(cuda)).sum().backward()
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SERSIC_ELLIPSE", {})]}
    xr = x.double().requires_grad_(True)
    ref_t = torch_ref.Sim(model, ocam, dtype=torch.float64).simulate(C.nested_from_x(prior, xr))
    (ref_t * w.double()).sum().backward()
    g, gr = xc.grad.cpu().double(), xr.grad
    rel = ((g - gr).abs() / (gr.abs() + 1e-3 * gr.abs().max())).max().item()
    _print(variant, "vjp rel", rel)
    assert rel < 2e-3


def test_map_adam_step(prob, cuda):
    """Fused MAP step == fused gradient + Keras Adam on the host."""
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    plan = pm.plan_for(sim)
    z0 = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"][:8])).to(cuda).contiguous()
    z = z0.clone()
    m = torch.zeros_like(z)
    v = torch.zeros_like(z)
    lp = torch.empty(8, device=cuda)
    chi2 = torch.empty(8, device=cuda)
    zr, mr, vr = z0.clone(), torch.zeros_like(z), torch.zeros_like(z)
    for step in (1, 2, 3):
        plan.map_adam_step(z, m, v, step, 5e-3, lp=lp, chi2=chi2)
        g_lp = torch.full((8,), -1.0 / (8 * 3600), device=cuda)
        _, _, g = plan.log_prob(zr, g_lp=g_lp, want_grad=True)
        mr = 0.9 * mr + 0.1 * g
        vr = 0.999 * vr + 0.001 * g * g
        lr_t = 5e-3 * math.sqrt(1 - 0.999 ** step) / (1 - 0.9 ** step)
        zr = zr - lr_t * mr / (torch.sqrt(vr) + 1e-7)
    torch.cuda.synchronize()
    assert torch.allclose(z, zr, rtol=1e-5, atol=1e-6)


def test_separable_psf_matches_direct(prob, cuda):
    """Rank-1 (separable) PSF path == direct 2-D path: log_prob and its gradient."""
    from gigalens_amd.plan import Plan, layout_from_paths
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=16)
    lay = layout_from_paths(prob["phys"], pm._paths)
    kw = dict(prior=prob["prior"], observed=prob["obs"], background_rms=0.2, exp_time=100.0)
    sep = Plan(prob["phys"], sim.camera, lay, psf_rank_tol=0.0, **kw)
    direct = Plan(prob["phys"], sim.camera, lay, psf_rank_tol=-1.0, **kw)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    w = torch.linspace(0.5, 1.5, 16, device=cuda)
    lp1, c1, g1 = sep.log_prob(z, g_lp=w, want_grad=True)
    lp2, c2, g2 = direct.log_prob(z, g_lp=w, want_grad=True)
    assert torch.allclose(lp1, lp2, rtol=1e-5)
    rel = ((g1 - g2).abs().max(1).values / g2.abs().max(1).values).max().item()
    _print("separable vs direct grad rel:", rel)
    assert rel < 1e-4
    x = prob["x"].to(cuda).contiguous()
    simp = Plan(prob["phys"], sim.camera, sim.layout, psf_rank_tol=0.0)
    simd = Plan(prob["phys"], sim.camera, sim.layout, psf_rank_tol=-1.0)
    i1, i2 = simp.simulate(x), simd.simulate(x)
    assert ((i1 - i2).abs().amax((1, 2)) / i2.abs().amax((1, 2))).max().item() < 1e-5


def test_gradient_64_particles_vs_fd_and_autodiff(cuda):
    """The batched-MAP configuration's gradient check: 64 prior particles (seed 3),
    dlp/dz of ForwardProbModel.log_prob from the hand-written adjoint against
    (a) FP64 autodiff of the reference's execution model (oracle/torch_ref.py:
    batch-coupled series, conv2d + avg_pool, torch autograd of log_prob) and
    (b) FP64 Richardson-extrapolated central differences of the oracle
    (h = 1e-3 and 5e-4 relative).  Bar: 2e-3 of each component plus 2e-4 of the
    particle's largest component."""
    prob = C.epl_shear_problem(bs=64, seed=3)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=64)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).contiguous()
    _, _, g = pm.log_prob_and_grad(sim, z.to(cuda))
    g = g.cpu().numpy().astype(np.float64)
    zz = z.numpy().astype(np.float64)
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SERSIC_ELLIPSE", {})]}
    tsim = torch_ref.Sim(model, prob["ocam"], dtype=torch.float64)
    zt = torch.tensor(zz, requires_grad=True)
    lp_t, _ = torch_ref.log_prob(tsim, C.oracle_nested_prior(prob["prior"]), zt, prob["obs"], 0.2, 100.0)
    lp_t.sum().backward()
    g_ad = zt.grad.numpy()
    g_fd = C.oracle_grad_richardson(prob["ocam"], prob["phys"], prob["prior"], zz, prob["obs"], 0.2, 100.0)
    for name, ref in (("autodiff", g_ad), ("richardson fd", g_fd)):
        scale = np.abs(ref).max(axis=1, keepdims=True)
        err = np.abs(g - ref)
        _print(f"64 particles vs {name}: max |g - ref| / max|ref| = {(err / scale).max():.2e}")
        assert np.all(err <= 2e-3 * np.abs(ref) + 2e-4 * scale)
