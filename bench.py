"""Benchmark: sim+grad particle-evals/s of multi-start MAP on one MI355X.

Workload (the batched-MAP configuration): 60x60 pixels at 0.065"/px,
supersample 2, Gaussian PSF FWHM 0.1" (11x11, sub-grid 21x21), EPL + shear
lens, elliptical Sersic lens light and source; 4,000 particles drawn from the
prior (seed 3); Keras-Adam lr 5e-3.  One step = one fused launch that
simulates every particle, evaluates log_prob, back-propagates with the
hand-written adjoint and applies the Adam update.  Inputs are resident in
HBM before the timed region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
Prints one JSON line.

``--workload cnn`` measures the learned initialiser's training step (1,024
simulated 60x60 images, BF16 MFMA network, AdamW).

``--workload uneven`` measures batched fitting of many lens systems with
different image sizes, supersampling and particle counts (lenses/hour).

``--workload shapelets`` measures the least-squares configuration instead
(100x100 px, supersample 4, 21x21 Moffat PSF, EPL+shear, Sersic lens light and
an n_max=10 shapelet source with linear amplitudes; 512 prior particles,
BackwardProbModel log_prob + gradient per step).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from gigalens_amd import BackwardProbModel, ForwardProbModel, LensSimulator  # noqa: E402
from gigalens_amd import distributions as tfd  # noqa: E402
from gigalens_amd import flops as fl  # noqa: E402
from gigalens_amd import workloads as wl  # noqa: E402

PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 (vector / matrix)
BASELINE_EVALS = 2200.0   # reference notebook MAP throughput (tf-demo.ipynb:808)


def measured_traffic():
    """HBM bytes per launch of the particle kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated)."""
    import re
    f = os.path.join(ROOT, "profiles", "particle_kernel_map4000.md")
    try:
        m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(f).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def flat_sample(prior, n, seed):
    s = prior.sample(n, seed=seed)
    return torch.stack([torch.as_tensor(v, dtype=torch.float32) for v in tfd.flatten(s)], dim=1)


def setup(bs, device):
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    cfg = wl.epl_shear_config()
    sim = LensSimulator(phys, cfg, bs=bs, device=device)
    truth = flat_sample(prior, 1, seed=0).to(device)
    clean = sim.simulate(truth).cpu().numpy()
    obs = wl.noisy_observation(clean, 0.2, 100.0, seed=1)
    pm = ForwardProbModel(prior, obs, 0.2, 100.0)
    x0 = flat_sample(prior, bs, seed=3)
    z0 = pm.bij.inverse(tfd.pack_as(prior, [x0[:, j] for j in range(x0.shape[1])])).to(device).contiguous()
    return prior, phys, cfg, sim, pm, obs, z0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(obs, n=1000, chunk=250, runs=3):
    """Oracle port (torch-CPU FP32, op-for-op restatement of the reference's TF
    path, autograd) timed on this host: log_prob + gradient of ``n`` prior
    particles (in batches of ``chunk`` to bound memory), one small warm-up, median
    of ``runs`` runs."""
    from oracle import torch_ref
    from oracle import np64
    prior = wl.epl_shear_prior()
    cfg = wl.epl_shear_config()
    ocam = np64.Camera(cfg.delta_pix, cfg.num_pix, cfg.supersample, cfg.kernel, None)
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SERSIC_ELLIPSE", {})]}
    kind = {1: "Normal", 2: "LogNormal", 3: "Uniform", 4: "TruncatedNormal"}

    def rec(node):
        if isinstance(node, tfd.JointDistributionSequential):
            return [rec(m) for m in node.model]
        if isinstance(node, tfd.JointDistributionNamed):
            return {k: rec(node.model[k]) for k in sorted(node.model)}
        return (kind[node.kind], tuple(node.args()))
    nested = rec(prior)
    sim = torch_ref.Sim(model, ocam, dtype=torch.float32)
    x = flat_sample(prior, 16 + n * runs, seed=3)
    z_all = torch.as_tensor(torch_ref.x_to_z(nested, x.double().numpy()), dtype=torch.float32)

    def run(bs, off):
        t = time.perf_counter()
        for a in range(off, off + bs, chunk):
            z = z_all[a:min(a + chunk, off + bs)].clone().requires_grad_(True)
            lp, _ = torch_ref.log_prob(sim, nested, z, obs, 0.2, 100.0)
            (-lp.mean()).backward()
        return time.perf_counter() - t
    run(16, 0)
    times = [run(n, 16 + i * n) for i in range(runs)]
    med = float(np.median(times))
    return {"value": n / med, "unit": "particle-evals/s", "cores": torch.get_num_threads(), "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{n} prior particles per run (batches of {chunk}), sim+log_prob+autograd grad in FP32 (torch "
                      f"CPU, op-for-op restatement of tf/simulator.py + tf/model.py), 1 warm-up of 16, median of "
                      f"{runs} runs"}


def oracle_error_map(prior, phys, cfg, pm, obs, z0, n=8):
    """The GPU path against the FP64 oracle on the workload's first ``n`` particles:
    image (from the same physical parameters), log_prob, chi2 and the gradient
    (Richardson-extrapolated central differences of the oracle)."""
    from oracle import bridge as ob
    z = z0[:n].contiguous()
    z64 = z.double().cpu().numpy()
    x64 = ob.z_to_x(prior, z64)
    sim = LensSimulator(phys, cfg, bs=n, device=z.device)
    img = sim.simulate(torch.as_tensor(x64, dtype=torch.float32, device=z.device)).double().cpu().numpy()
    lp, chi2, g = pm.log_prob_and_grad(sim, z)
    ocam = ob.oracle_camera(cfg)
    img_ref = ob.oracle_simulate(ocam, phys, ob.nested_from_x(prior, x64.astype(np.float32).astype(np.float64)))
    lp_ref, chi_ref, _ = ob.oracle_log_prob(ocam, phys, prior, z64, obs, 0.2, 100.0)
    g_ref = ob.oracle_grad_richardson(ocam, phys, prior, z64, obs, 0.2, 100.0)
    lp, chi2, g = (t.double().cpu().numpy() for t in (lp, chi2, g))
    peak = np.abs(img_ref).max(axis=(1, 2))
    return {"particles": n, "oracle": "FP64 (oracle/np64.py)",
            "image_max_abs": float(np.abs(img - img_ref).max()),
            "image_max_abs_over_peak": float((np.abs(img - img_ref).max(axis=(1, 2)) / peak).max()),
            "log_prob_max_abs": float(np.abs(lp - lp_ref).max()),
            "log_prob_max_rel": float((np.abs(lp - lp_ref) / np.abs(lp_ref)).max()),
            "chi2_max_rel": float((np.abs(chi2 - chi_ref) / np.abs(chi_ref)).max()),
            "grad_max_abs_over_particle_max": float((np.abs(g - g_ref).max(1) / np.abs(g_ref).max(1)).max())}


def shapelet_setup(bs, device, interpolate=True):
    """Shapelet workload: truth = prior draw (seed 0) with the lens light scaled
    to a peak of 100 and shapelet amplitudes from ``shapelet_amplitudes`` scaled
    to a source peak of 50; observed = Poisson + Gaussian noise (seed 1)."""
    prior = wl.shapelet_prior()
    phys = wl.shapelet_model(10, interpolate=interpolate)
    cfg = wl.shapelet_config()
    sim = LensSimulator(phys, cfg, bs=bs, device=device)
    truth = flat_sample(prior, 1, seed=0).to(device)
    N = cfg.num_pix
    st = sim.lstsq_simulate(truth, np.zeros((N, N), np.float32), np.ones((N, N), np.float32),
                            return_stacked=True)[0].double().cpu().numpy()
    amps = wl.shapelet_amplitudes(10, seed=0)
    amps[0] = 100.0 / max(st[..., 0].max(), 1e-30)
    src = st[..., 1:] @ amps[1:]
    amps[1:] *= 50.0 / max(np.abs(src).max(), 1e-30)
    clean = st @ amps
    obs = wl.noisy_observation(clean, 0.2, 100.0, seed=1)
    pm = BackwardProbModel(prior, obs, 0.2, 100.0)
    x0 = flat_sample(prior, bs, seed=3)
    z0 = pm.bij.inverse(tfd.pack_as(prior, [x0[:, j] for j in range(x0.shape[1])])).to(device).contiguous()
    return prior, phys, cfg, sim, pm, obs, z0


def shapelet_cpu_baseline(prior, phys, cfg, sim, obs, z0, n=1):
    """Oracle port timed on the host: BackwardProbModel log_prob (lstsq with
    torch.linalg.pinv) + autograd gradient for ``n`` particles, FP32, one timed
    run (about 30 s per particle on 16 cores, so no warm-up run)."""
    from oracle import torch_ref
    from oracle import np64
    cam = sim.camera
    ocam = np64.Camera(cfg.delta_pix, cfg.num_pix, cfg.supersample, cfg.kernel, None)
    model = {"lenses": [("EPL", {"niter": 50}), ("SHEAR", {})], "lens_light": [("SERSIC_ELLIPSE", {})],
             "source_light": [("SHAPELETS", {"n_max": 10, "interpolate": phys.source_light[0].interpolate})]}
    kind = {1: "Normal", 2: "LogNormal", 3: "Uniform", 4: "TruncatedNormal"}

    def rec(node):
        if isinstance(node, tfd.JointDistributionSequential):
            return [rec(m) for m in node.model]
        if isinstance(node, tfd.JointDistributionNamed):
            return {k: rec(node.model[k]) for k in sorted(node.model)}
        return (kind[node.kind], tuple(node.args()))
    nested = rec(prior)
    osim = torch_ref.Sim(model, ocam, dtype=torch.float32)
    zc = z0[:n].float().cpu()

    def run():
        z = zc.clone().requires_grad_(True)
        t = time.perf_counter()
        lp, _ = torch_ref.log_prob(osim, nested, z, obs, 0.2, 100.0, backward=True)
        (-lp.sum()).backward()
        return time.perf_counter() - t
    dt = run()
    return {"value": n / dt, "unit": "particle-evals/s", "cores": torch.get_num_threads(), "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{n} prior particle(s), BackwardProbModel log_prob + autograd gradient in FP32 (torch CPU "
                      f"restatement of tf/simulator.py lstsq_simulate + tf/model.py), 1 timed run"}


def oracle_error_shapelets(prior, phys, cfg, pm, obs, z0, n=2):
    """The least-squares path against the FP64 oracle (pinv solve) on the workload's
    first ``n`` particles: image and BackwardProbModel log_prob / chi2."""
    from oracle import bridge as ob
    from oracle import np64
    z = z0[:n].contiguous()
    z64 = z.double().cpu().numpy()
    x64 = ob.z_to_x(prior, z64)
    sim = LensSimulator(phys, cfg, bs=n, device=z.device)
    lp, chi2 = pm.log_prob(sim, z)
    err = np64.backward_err_map(np.asarray(obs, np.float64), 0.2, 100.0)
    img = sim.lstsq_simulate(torch.as_tensor(x64, dtype=torch.float32, device=z.device), obs,
                             err.astype(np.float32)).double().cpu().numpy()
    ocam = ob.oracle_camera(cfg)
    lp_ref, chi_ref, img_ref = ob.oracle_backward_log_prob(ocam, phys, prior, z64, obs, 0.2, 100.0)
    lp, chi2 = lp.double().cpu().numpy(), chi2.double().cpu().numpy()
    peak = np.abs(img_ref).max(axis=(1, 2))
    return {"particles": n, "oracle": "FP64 (oracle/np64.py, pinv rcond 1e-6)",
            "image_max_abs": float(np.abs(img - img_ref).max()),
            "image_max_abs_over_peak": float((np.abs(img - img_ref).max(axis=(1, 2)) / peak).max()),
            "log_prob_max_abs": float(np.abs(lp - lp_ref).max()),
            "log_prob_max_rel": float((np.abs(lp - lp_ref) / np.abs(lp_ref)).max()),
            "chi2_max_rel": float((np.abs(chi2 - chi_ref) / np.abs(chi_ref)).max())}


def run_shapelets(args, dev):
    bs = args.particles if args.particles != 4000 else 512
    prior, phys, cfg, sim, pm, obs, z0 = shapelet_setup(bs, dev)
    plan = pm.plan_for(sim)
    for _ in range(args.warmup):
        pm.log_prob_and_grad(sim, z0)
    torch.cuda.synchronize()
    K = args.steps
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(K):
        ev[s][0].record(stream)
        lp, chi2, g = pm.log_prob_and_grad(sim, z0)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    x0 = pm.bij.forward(z0.cpu())
    n = fl.series_length(x0[0][0]["e1"].numpy(), x0[0][0]["e2"].numpy(), 50)
    ksz = sim.camera.subgrid_kernel.shape[0]
    pinfo = plan.info()
    c = fl.lstsq_per_particle(phys, cfg.num_pix, cfg.supersample, ksz, n, grad=True, sep_rank=pinfo["sep_rank"],
                              adj_rank=pinfo["adj_rank"])
    fp32 = float(c["flops"].sum())
    achieved = fp32 / (step_ms * 1e-3) / 1e12
    value = bs * K / wall
    out = {
        "metric": "sim+grad particle-evals/s",
        "value": value,
        "unit": "particle-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (prior draws; observed = seeded truth with fixed amplitudes + Poisson + Gaussian noise)",
        "config": {"workload": "least-squares shapelets, 100x100 px, ss4, 83x83 sub-grid Moffat PSF, EPL+shear, "
                               f"Sersic lens light + n_max=10 shapelets (D={c['channels']}), {bs} particles, "
                               "BackwardProbModel log_prob + gradient",
                   "particles": bs, "num_pix": cfg.num_pix, "supersample": cfg.supersample,
                   "psf_subgrid": ksz, "params": z0.shape[1], "channels": c["channels"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "whole pipeline (mc_prologue .. mc_finish); the channel conv dominates ("
                               + (f"separable rank {pinfo['sep_rank']}: mc_sep_h + mc_sep_v)" if pinfo["sep_rank"]
                                  else "direct: mc_conv)"),
                     "step_ms": step_ms, "fp32_flop_per_step": fp32,
                     "conv_flop_per_step": c["conv_flops"] * bs,
                     "mfma_flop_per_step": float(c["mfma"].sum()), "fp64_flop_per_step": float(c["fp64"].sum()),
                     "transcendentals_per_step": float(c["transcendentals"].sum())},
        "chi2_median": float(chi2.median()),
    }
    if args.hmc:
        out["hmc"] = shapelet_hmc(prior, phys, cfg, pm, z0, dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = shapelet_cpu_baseline(prior, phys, cfg, sim, obs, z0)
        out["oracle_error"] = oracle_error_shapelets(prior, phys, cfg, pm, obs, z0)
    print(json.dumps(out), flush=True)


def shapelet_hmc(prior, phys, cfg, pm, z0, dev):
    """MAP of the 512 particles (fused steps, Keras Adam lr 5e-3, 200 steps), a
    diagonal SVI from the best particle (300 steps, 64 samples) for the mass
    matrix, then HMC with 256 chains x 1,000 leapfrog steps at step size 0.01
    (100 draws of 10 leapfrog steps) from the surrogate's draws."""
    from gigalens_amd.inference import Adam, ModellingSequence, PolynomialDecay
    seq = ModellingSequence(phys, pm, cfg, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    z = seq.MAP(Adam(5e-3), start=z0, num_steps=200)
    sim = LensSimulator(phys, cfg, bs=z.shape[0], device=dev)
    lp, chi2 = pm.log_prob(sim, z)
    torch.cuda.synchronize()
    t_map = time.perf_counter() - t0
    best = torch.argmax(torch.nan_to_num(lp.detach(), nan=-1e30))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q, losses = seq.SVI(Adam(PolynomialDecay(0.0, 100, 2e-3, 2)), z[best].detach().cpu(), n_vi=64, num_steps=300,
                        full_rank=False)
    torch.cuda.synchronize()
    t_svi = time.perf_counter() - t0
    t0 = time.perf_counter()
    samples, stats = seq.HMC(q, init_eps=0.01, init_l=10, n_hmc=256, num_burnin_steps=0, num_results=100,
                             adapt=False, graph=True)
    torch.cuda.synchronize()
    t_hmc = time.perf_counter() - t0
    return {"map_steps": 200, "map_seconds": t_map, "best_chi2": float(chi2.detach()[best]),
            "svi_steps": 300, "svi_samples": 64, "svi_seconds": t_svi,
            "chains": 256, "leapfrog_steps": 1000, "step_size": 0.01, "hmc_seconds": t_hmc, "hmc_graph": True,
            "leapfrog_particle_evals_per_s": 256 * 1000 / t_hmc,
            "accept_rate_mean": float(np.mean(stats["accept_rate"]))}


def run_cnn(args, dev):
    """Learned-initialiser training throughput: one step = simulate 1,024 prior
    images on the device (particle kernel + Poisson/Gaussian noise) and one
    AdamW step of the BF16 MFMA network on them."""
    from gigalens_amd.learned import LearnedInitializer
    from oracle import cnn_ref
    prior, phys, cfg = wl.epl_shear_prior(), wl.epl_shear_model(), wl.epl_shear_config()
    bs = 1024
    init = LearnedInitializer(phys, cfg, prior, 0.2, 100.0, max_batch=bs, seed=0, device=dev)
    K = args.steps
    init.train(n_images=bs * args.warmup, batch=bs, lr=1e-3, seed=3)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    imgs, z = init.simulate_batch(bs, seed=1)
    t = init.standardise(z)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(K):
        ev[s][0].record(stream)
        init.net.loss_grad(imgs, t)
        init.net.adamw_step(args.warmup + s + 1, 1e-3)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall_net = (time.perf_counter() - t0) / K
    net_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    t0 = time.perf_counter()
    init.train(n_images=bs * K, batch=bs, lr=1e-3, seed=4)
    torch.cuda.synchronize()
    wall_step = (time.perf_counter() - t0) / K
    if args.copy_probe:
        a = torch.empty(64 * 1024 * 1024, device=dev)
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
    N = cfg.num_pix
    ch = init.net.channels
    fwd = 0.0
    H, cin = N, 1
    for i, c in enumerate(ch):
        Ho = (H - 1) // (1, 2, 1, 2)[i] + 1
        fwd += 2.0 * Ho * Ho * c * 9 * cin
        H, cin = Ho, c
    first = 2.0 * N * N * ch[0] * 9
    flop_img = fwd + 2 * fwd - first
    achieved = flop_img * bs / (net_ms * 1e-3) / 1e12
    # algorithmic HBM bytes per image (bf16 activations): forward reads each layer's
    # input and writes Z and Y; the head reads Y4, Z4 and writes dZ4; each data
    # gradient reads dZ and Zprev and writes dZprev; each weight gradient reads dZ and
    # its input
    act, H, cin, x_in = [], N, 1, N * N * 2
    for i, c in enumerate(ch):
        Ho = (H - 1) // (1, 2, 1, 2)[i] + 1
        act.append(Ho * Ho * c * 2)
        H, cin = Ho, c
    ins = [x_in] + act[:-1]
    fwd_b = sum(ins) + 2 * sum(act)
    head_b = 3 * act[-1]
    dgrad_b = sum(act[l] + 2 * act[l - 1] for l in range(1, 4))
    wgrad_b = sum(act[l] + ins[l] for l in range(4))
    bytes_img = fwd_b + head_b + dgrad_b + wgrad_b
    hbm = bytes_img * bs / (net_ms * 1e-3) / 1e9
    out = {"metric": "learned-initialiser training images/s", "value": bs / wall_step, "unit": "images/s",
           "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": wall_step * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
           "data": "synthetic (prior draws simulated on the device, Poisson + Gaussian noise)",
           "config": {"workload": "BF16 CNN 4x conv3x3 (32/32/64/64, strides 1,2,1,2) + GAP + MLP 256 -> 2x22 on "
                                  "60x60 images, batch 1024, AdamW", "batch": bs, "num_pix": N},
           "roofline": {"bound": "hbm", "achieved": hbm, "peak": 8000.0, "unit": "GB/s", "frac": hbm / 8000.0,
                        "traffic": None, "algorithmic_bytes_per_step": bytes_img * bs,
                        "kernel": "network step (BF16 MFMA convolutions: wave-streaming forward and data gradient, LDS-transposed weight gradient; head; AdamW)",
                        "network_ms": net_ms, "network_wall_ms": wall_net * 1e3, "flop_per_image": flop_img,
                        "mfma_achieved_tflops": achieved, "mfma_peak_tflops": 2500.0, "mfma_frac": achieved / 2500.0,
                        "note": "traffic per step from rocprofv3 FETCH_SIZE/WRITE_SIZE is in the committed profile"}}
    if not args.no_cpu_baseline:
        n = 64
        p = init.net.get_params().cpu()
        x = imgs[:n].cpu()
        tt = t[:n].cpu()
        cnn_ref.loss_and_grad(p, init.net.layout, x, tt, 0.2, ch)
        t0 = time.perf_counter()
        cnn_ref.loss_and_grad(p, init.net.layout, x, tt, 0.2, ch)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "images/s", "cores": torch.get_num_threads(), "kind": "port",
                               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
                               "sample": f"{n} images, forward + NLL + autograd backward of the torch-CPU restatement "
                                         "(FP32 with the same bf16 rounding points), 1 warm-up, 1 timed run"}
        lref, gref = cnn_ref.loss_and_grad(p, init.net.layout, x, tt, 0.2, ch)
        loss, g = init.net.loss_grad(imgs[:n], t[:n], want_grad=True)
        gref = gref.double()
        out["oracle_error"] = {"images": n, "oracle": "torch-CPU FP32 restatement with the kernels' bf16 rounding "
                                                      "points (oracle/cnn_ref.py)",
                               "loss_rel": abs(float(loss) - float(lref)) / abs(float(lref)),
                               "grad_max_abs_over_max": float((g.double().cpu() - gref).abs().max()
                                                              / gref.abs().max())}
    print(json.dumps(out), flush=True)


def oracle_error_uneven(prior, phys, lens_set, batch, z, per_lens=4):
    """The batched log_prob and gradient against the FP64 oracle for ``per_lens``
    prior draws (seed 4242, in place of the first particles of z) of four lenses:
    the smallest and the largest image, the first lens on the resident-image (fused)
    path and the first on the staged path.  Prior draws rather than fitted particles:
    at a converged MAP the gradient is near zero and the posterior curvature large,
    where finite differences of the oracle lose their accuracy."""
    from oracle import bridge as ob
    z = z.clone()
    npix = [c.num_pix * c.supersample for c, _, _ in lens_set]
    sel = [int(np.argmin(npix)), int(np.argmax(npix))]
    for want in (True, False):
        k = next((i for i, f in enumerate(batch.fused) if f == want), None)
        if k is not None and k not in sel:
            sel.append(k)
    for k in sel:
        a, m = int(batch.first[k]), min(per_lens, lens_set[k][2])
        xk = flat_sample(prior, m, seed=4242 + k)
        zk = batch.bij.inverse(tfd.pack_as(prior, [xk[:, j] for j in range(xk.shape[1])]))
        z[a:a + m] = torch.as_tensor(zk, dtype=torch.float32).to(z.device)
    lp, chi2, g = batch.log_prob(z.contiguous(), want_grad=True)
    rows = []
    for k in sel:
        cfg, obs, cnt = lens_set[k]
        a, m = int(batch.first[k]), min(per_lens, cnt)
        z64 = z[a:a + m].double().cpu().numpy()
        ocam = ob.oracle_camera(cfg)
        lp_ref, chi_ref, _ = ob.oracle_log_prob(ocam, phys, prior, z64, obs, 0.2, 100.0)
        g_ref = ob.oracle_grad_richardson(ocam, phys, prior, z64, obs, 0.2, 100.0)
        lpk, chk, gk = (t[a:a + m].double().cpu().numpy() for t in (lp, chi2, g))
        rows.append({"lens": k, "num_pix": cfg.num_pix, "supersample": cfg.supersample, "fused": bool(batch.fused[k]),
                     "particles": m, "log_prob_max_rel": float((np.abs(lpk - lp_ref) / np.abs(lp_ref)).max()),
                     "chi2_max_rel": float((np.abs(chk - chi_ref) / np.abs(chi_ref)).max()),
                     "grad_max_abs_over_particle_max": float((np.abs(gk - g_ref).max(1)
                                                              / np.abs(g_ref).max(1)).max())})
    return {"oracle": "FP64 (oracle/np64.py)", "lenses": rows}


def uneven_lenses(args, dev):
    """The uneven-batch workload's lenses (seeds 2000..): (sim_config, observed image,
    particles) with N in {48,64,80,100,128}, ss in {2,3,4}, U{256..4096} particles
    (times ``--uneven-scale``), 0.065"/px, 11x11 Gaussian PSF FWHM 0.1"; observed =
    a prior draw + Poisson + Gaussian noise (seeds 3000..)."""
    from gigalens_amd import SimulatorConfig, psf
    from gigalens_amd.plan import Plan, layout_from_paths
    from gigalens_amd.simulator import Camera
    prior, phys = wl.epl_shear_prior(), wl.epl_shear_model()
    kernel = psf.gaussian(0.1, 0.065, 11)
    rng = np.random.default_rng(2000)
    specs = []
    for k in range(args.uneven_lenses):
        n_pix = int(rng.choice([48, 64, 80, 100, 128]))
        ss = int(rng.choice([2, 3, 4]))
        cnt = max(1, int(round(int(rng.integers(256, 4097)) * args.uneven_scale)))
        specs.append((n_pix, ss, cnt))
    paths = [pth for pth, _ in tfd.leaves(prior)]
    lenses = []
    for k, (n_pix, ss, cnt) in enumerate(specs):
        cfg = SimulatorConfig(delta_pix=0.065, num_pix=n_pix, supersample=ss, kernel=kernel)
        gen = Plan(phys, Camera(cfg), layout_from_paths(phys, paths), max_batch=1)
        truth = flat_sample(prior, 1, seed=2000 + k).to(dev).contiguous()
        clean = gen.simulate(truth).cpu().numpy()[0]
        del gen
        lenses.append((cfg, wl.noisy_observation(clean, 0.2, 100.0, seed=3000 + k), cnt))
    return prior, phys, lenses


def run_uneven(args, dev):
    """Uneven batched fitting: every lens of ``uneven_lenses`` fitted at once by the
    learned-init MAP -> diagonal SVI pipeline, on the path the batch autotuner picks
    (``--uneven-band auto``; by default every lens in one launch of the banded
    persistent kernel per step), tuned before the timed runs.  Starts: the learned initialiser
    (trained on the device on the 60x60 workload: ``--learned-images`` draws,
    ``--learned-epochs`` passes, AdamW ``--learned-lr``) predicts each lens's
    Gaussian from the central 60x60 pixels of its image; 3/4 of the candidates come
    from it and 1/4 from the prior, two candidate sets screened by log-posterior.
    MAP: 350 Keras-Adam steps (lr 5e-3), recorded into HIP graphs of 25 steps.
    SVI: each lens's best MAP particle starts a diagonal Gaussian, 1,500 steps of
    64 samples (lr 5e-3), every step (device sampling, batched log_prob + gradient,
    update) one replay of a HIP graph.  The wall time of the whole pipeline is
    measured (network training reported apart); per-lens latency: the pipeline
    completes every lens at once, and each lens's time to its MAP chi2 <= 1.1
    follows from its per-step chi2 history and the measured step times.  The
    sequential baseline fits ``--uneven-seq-lenses`` of the lenses one at a time
    through the same code (its own plan, learned starts, graph-captured MAP and
    SVI), each measured."""
    from gigalens_amd.batch import LensBatch
    from gigalens_amd.inference import Adam
    from gigalens_amd.learned import LearnedInitializer, propose_batch
    prior, phys, lenses = uneven_lenses(args, dev)
    L = len(lenses)
    total = sum(c for *_, c in lenses)
    n_map, n_svi, n_vi = args.uneven_map_steps, args.uneven_svi_steps, 64
    # learned initialiser (trained once; its time is reported apart)
    init = LearnedInitializer(phys, wl.epl_shear_config(), prior, 0.2, 100.0, max_batch=1024, seed=0, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    init.train(n_images=args.learned_images, batch=1024, lr=args.learned_lr, weight_decay=0.01, seed=10,
               epochs=args.learned_epochs)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"[uneven] {L} lenses, {total} particles; network trained in {t_train:.1f} s", flush=True)

    band = args.uneven_band if args.uneven_band == "auto" else int(args.uneven_band)

    def tune(lens_set):
        """The autotuner's choice for the MAP and SVI batches of these lenses (untimed; the
        pipeline's own batches then find it in the cache)."""
        out = {}
        for name, ls in (("map", lens_set), ("svi", [(c, o, n_vi) for c, o, _ in lens_set])):
            b = LensBatch(phys, prior, ls, 0.2, 100.0, device=dev, workspace_gib=args.uneven_workspace_gib,
                          band=band)
            out[name] = {"config": b.config, "info": b.info,
                         "timings_ms": b.tuning[1] if b.tuning is not None else None}
            del b
        return out

    def pipeline(lens_set, tag):
        images = [o for _, o, _ in lens_set]
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        batch = LensBatch(phys, prior, lens_set, 0.2, 100.0, device=dev, workspace_gib=args.uneven_workspace_gib,
                          band=band)
        svi = LensBatch(phys, prior, [(c, o, n_vi) for c, o, _ in lens_set], 0.2, 100.0, device=dev,
                        workspace_gib=args.uneven_workspace_gib, band=band)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t_start
        z0 = propose_batch(init, batch, images, prior_fraction=0.25, oversample=2, seed=5)
        torch.cuda.synchronize()
        t_prop = time.perf_counter() - t_start - t_setup
        print(f"[uneven] {tag}: setup {t_setup:.1f} s, proposals {t_prop:.1f} s", flush=True)
        z = batch.run(z0, n_map, Adam(5e-3), graph=True, graph_steps=25,
                      progress=(lambda a, n: print(f"[uneven] {tag}: MAP step {a}/{n}", flush=True))
                      if len(lens_set) > 1 else None)
        torch.cuda.synchronize()
        t_map = time.perf_counter() - t_start - t_setup - t_prop
        best_lp, best_chi = batch.per_lens_best()
        loc = batch.best_particles(z)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loc_f, log_scale, losses = svi.svi_diag(loc, n_svi, Adam(5e-3), seed=2, graph=True)
        torch.cuda.synchronize()
        t_svi = time.perf_counter() - t1
        wall = time.perf_counter() - t_start
        step_s = np.array([s for _, _, s in batch.step_seconds])
        # per-lens time to MAP chi2 <= 1.1: its first such step at the measured step times
        cum = np.concatenate([[0.0], np.cumsum(np.repeat(step_s, [b - a for a, b, _ in batch.step_seconds]))])
        hit = best_chi <= 1.1
        first = np.where(hit.any(0), np.argmax(hit, 0) + 1, -1)
        t_chi = np.array([t_setup + t_prop + cum[f] if f > 0 else np.nan for f in first])
        res = {"wall_s": wall, "setup_s": t_setup, "proposals_s": t_prop, "map_s": t_map, "svi_s": t_svi,
               "map_batch": {"config": batch.config, "lenses_band": batch.info["lenses_band"],
                             "lenses_fused": batch.info["lenses_fused"], "lenses_staged": batch.info["lenses_staged"],
                             "launches_per_step": int(batch.info["lenses_band"] > 0) + batch.info["lenses_fused"] +
                             batch.info["chunks"]},
               "map_step_s_mean": float(step_s.mean()), "svi_step_s": t_svi / n_svi,
               "fused_lenses": int(sum(batch.fused)), "staged_chunks": batch.num_chunks,
               "workspace_gib": batch.workspace_bytes / 2 ** 30,
               "map_chi2_reached": int((first > 0).sum()),
               "time_to_map_chi2_p50_s": float(np.nanmedian(t_chi)) if (first > 0).any() else None,
               "time_to_map_chi2_p95_s": float(np.nanpercentile(t_chi, 95)) if (first > 0).any() else None,
               "final_best_chi2_p50": float(np.median(best_chi[-1])),
               "svi_final_loss_finite": int(torch.isfinite(losses[-1]).sum()),
               "particle_evals_per_s": sum(c for *_, c in lens_set) / float(step_s.mean())}
        if tag == "batched" and not args.no_cpu_baseline:
            res["oracle_error"] = oracle_error_uneven(prior, phys, lens_set, batch, z)
        print(f"[uneven] {tag}: {json.dumps({k: v for k, v in res.items()})}", flush=True)
        del batch, svi
        return res

    t0 = time.perf_counter()
    tuned = tune(lenses)
    print(f"[uneven] autotuned in {time.perf_counter() - t0:.1f} s: map {tuned['map']['config']}, "
          f"svi {tuned['svi']['config']}", flush=True)
    for k in range(min(args.uneven_seq_lenses, L)):
        t0 = time.perf_counter()
        tk = tune([lenses[k]])
        print(f"[uneven] lens {k} autotuned in {time.perf_counter() - t0:.1f} s: map {tk['map']['config']}, "
              f"svi {tk['svi']['config']}", flush=True)
    b = pipeline(lenses, "batched")
    seq_rows = []
    for k in range(min(args.uneven_seq_lenses, L)):
        r = pipeline([lenses[k]], f"sequential lens {k}")
        seq_rows.append(r["wall_s"])
    out = {"metric": f"lenses/hour (uneven batch, learned-init MAP {n_map} + diagonal SVI {n_svi} steps)",
           "value": L / b["wall_s"] * 3600.0, "unit": "lenses/hour", "n_gpus": 1, "steps": n_map + n_svi,
           "warmup": 0, "ms_per_step": b["map_step_s_mean"] * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (prior draws, Poisson + Gaussian noise; lens seeds 2000.., noise seeds 3000..)",
           "config": {"workload": f"{L} lenses, N in {{48,64,80,100,128}}, ss 2-4, particles U{{256..4096}} x "
                                  f"{args.uneven_scale}, 11x11 Gaussian PSF; learned-init MAP {n_map} steps + diagonal "
                                  f"SVI {n_svi} steps x 64 samples, every step measured", "lenses": L,
                      "particles": total, "scale": args.uneven_scale},
           "batched": dict(b, latency_note="every lens's fit completes with the batch (wall_s); "
                                            "time_to_map_chi2_p50/p95 are per-lens"),
           "autotune": tuned,
           "learned_training_s": t_train,
           "lenses_per_hour_incl_training": L / (b["wall_s"] + t_train) * 3600.0}
    if seq_rows:
        seq = np.array(seq_rows)
        out["sequential"] = {"lenses": len(seq), "mean_s_per_lens": float(seq.mean()),
                             "lenses_per_hour": 3600.0 / float(seq.mean()),
                             "latency_p50_s": float(np.median(seq)), "latency_p95_s": float(np.percentile(seq, 95)),
                             "note": f"the first {len(seq)} lenses fitted one at a time, each measured"}
        out["speedup_vs_sequential"] = out["value"] / out["sequential"]["lenses_per_hour"]
    print(json.dumps(out), flush=True)


def run_uneven_step(args, dev):
    """One MAP step of every lens of the uneven configuration (``uneven_lenses``: 128
    lenses, full particle counts) on the path the batch autotuner picks (or
    ``--uneven-band``): W untimed steps, then K steps timed with HIP events on the
    launch stream.  Counted work: the flop model of the implemented formulation
    (``gigalens_amd/flops.py``) per particle with its own EPL series length, the
    separable PSF passes at each lens's rank, summed over the batch."""
    from gigalens_amd.batch import LensBatch
    from gigalens_amd.simulator import Camera
    prior, phys, lenses = uneven_lenses(args, dev)
    band = args.uneven_band if args.uneven_band == "auto" else int(args.uneven_band)
    batch = LensBatch(phys, prior, lenses, 0.2, 100.0, device=dev, workspace_gib=args.uneven_workspace_gib,
                      band=band)
    lib = batch._lib
    z0 = batch.bij.inverse(prior.sample(batch.total, seed=7)).to(dev).float().contiguous()
    z, m, v = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lp = torch.empty(batch.total, device=dev)
    chi2 = torch.empty(batch.total, device=dev)
    from gigalens_amd.plan import _ptr, _stream
    from gigalens_amd import _native as nat

    def step(s):
        nat.check(lib.gl_batch_map_adam_step(batch._h, _ptr(z), _ptr(m), _ptr(v), s, args.lr, 0.9, 0.999, 1e-7,
                                             _ptr(lp), _ptr(chi2), _stream()), "gl_batch_map_adam_step")
    for s in range(1, args.warmup + 1):
        step(s)
    torch.cuda.synchronize()
    K = args.steps
    stream = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record(stream)
    for s in range(K):
        step(args.warmup + s + 1)
    b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = a.elapsed_time(b) / K
    if args.copy_probe:
        c0 = torch.empty(64 * 1024 * 1024, device=dev)
        torch.empty_like(c0).copy_(c0)
        torch.cuda.synchronize()
    # counted work of one step
    kinds, roles = fl.model_kinds(phys)
    x0 = batch.bij.forward(z0.cpu())
    nt = fl.series_length(x0[0][0]["e1"].numpy(), x0[0][0]["e2"].numpy(), 50)
    flops, subpix = 0.0, 0.0
    for (cfg, _, n), first in zip(lenses, batch.first):
        cam = Camera(cfg)
        k = cam.subgrid_kernel.shape[0]
        ss = cfg.supersample
        kf = k + ss - 1
        W = np.zeros((kf, kf))
        for aa in range(ss):
            for bb in range(ss):
                W[aa:aa + k, bb:bb + k] += cam.subgrid_kernel[::-1, ::-1]
        sv = np.linalg.svd(W, compute_uv=False)
        rank = int((sv >= 1e-6 * sv[0]).sum())
        c = fl.per_particle(kinds, roles, cfg.num_pix, ss, k, nt[first:first + n], grad=True,
                            sep_rank=rank if rank <= 8 else 0)
        flops += float(c["flops"].sum())
        subpix += float(n) * (ss * cfg.num_pix) ** 2
    achieved = flops / (step_ms * 1e-3) / 1e12
    out = {"metric": "sim+grad particle-evals/s (uneven lens batch, one MAP step)",
           "value": batch.total / (step_ms * 1e-3), "unit": "particle-evals/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": step_ms, "wall_ms_per_step": wall / K * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (prior draws, Poisson + Gaussian noise; lens seeds 2000.., noise seeds 3000..)",
           "config": {"workload": f"{len(lenses)} lenses, N in {{48,64,80,100,128}}, ss 2-4, particles U{{256..4096}} "
                                  f"x {args.uneven_scale}, 11x11 Gaussian PSF; one Keras-Adam MAP step of every "
                                  "particle", "lenses": len(lenses), "particles": batch.total,
                      "path": batch.config, "batch": batch.info},
           "sub_pixels_per_step": subpix, "sub_pixels_per_s": subpix / (step_ms * 1e-3),
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                        "kernel": "gl::band_kernel (one launch per step)" if batch.info["lenses_fused"] +
                        batch.info["chunks"] == 0 else "all launches of the step",
                        "fp32_flop_per_step": flops}}
    if batch.tuning is not None:
        out["autotune_ms"] = batch.tuning[1]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=350)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=5e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="also launch a 256 MiB device copy (PMC calibration)")
    ap.add_argument("--workload", choices=["map", "shapelets", "cnn", "uneven", "uneven-step"], default="map")
    ap.add_argument("--uneven-lenses", type=int, default=128)
    ap.add_argument("--uneven-scale", type=float, default=1.0)
    ap.add_argument("--uneven-workspace-gib", type=float, default=48.0)
    ap.add_argument("--uneven-seq-lenses", type=int, default=16)
    ap.add_argument("--uneven-map-steps", type=int, default=350, help="MAP steps (the configuration: 350)")
    ap.add_argument("--uneven-svi-steps", type=int, default=1500, help="SVI steps (the configuration: 1500)")
    ap.add_argument("--uneven-band", default="auto",
                    help="lens-batch path: auto (autotuned), 1 (banded persistent kernel), 2, 0 (resident + staged)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-prior time-to-MAP evaluation")
    ap.add_argument("--learned-images", type=int, default=2 ** 20)
    ap.add_argument("--learned-lenses", type=int, default=256)
    ap.add_argument("--learned-lr", type=float, default=2e-3)
    ap.add_argument("--learned-epochs", type=int, default=4)
    ap.add_argument("--hmc", action="store_true", help="shapelets: also MAP + 256-chain HMC (1,000 leapfrog steps)")
    ap.add_argument("--learned-schedule", choices=["constant", "cosine"], default="constant")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    dev = torch.device("cuda:0")
    if args.workload == "shapelets":
        if args.steps == 350:
            args.steps = 10
        return run_shapelets(args, dev)
    if args.workload == "cnn":
        if args.steps == 350:
            args.steps = 100
        return run_cnn(args, dev)
    if args.workload == "uneven":
        return run_uneven(args, dev)
    if args.workload == "uneven-step":
        if args.steps == 350:
            args.steps = 10
        return run_uneven_step(args, dev)
    bs = args.particles
    prior, phys, cfg, sim, pm, obs, z0 = setup(bs, dev)
    plan = pm.plan_for(sim)

    # warm-up on a copy (also compiles/loads everything)
    zw, mw, vw = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lpw = torch.empty(bs, device=dev)
    chw = torch.empty(bs, device=dev)
    for s in range(1, args.warmup + 1):
        plan.map_adam_step(zw, mw, vw, s, args.lr, lp=lpw, chi2=chw)
    torch.cuda.synchronize()

    K = args.steps
    z, m, v = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lp_hist = torch.empty((K, bs), device=dev)
    chi_hist = torch.empty((K, bs), device=dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(K):
        ev[s][0].record()
        plan.map_adam_step(z, m, v, s + 1, args.lr, lp=lp_hist[s], chi2=chi_hist[s])
        ev[s][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))

    # the same K steps recorded into one HIP graph and replayed (per-step host work removed)
    from gigalens_amd.graphs import capture
    zg, mg, vg = z0.clone(), torch.zeros_like(z0), torch.zeros_like(z0)
    lpg = torch.empty((K, bs), device=dev)
    chg = torch.empty((K, bs), device=dev)

    def graph_steps():
        for s_ in range(K):
            plan.map_adam_step(zg, mg, vg, s_ + 1, args.lr, lp=lpg[s_], chi2=chg[s_])
    graph = capture(graph_steps)
    ga, gb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ga.record()
    graph.replay()
    gb.record()
    torch.cuda.synchronize()
    graph_ms = ga.elapsed_time(gb) / K
    graph_same = bool(torch.allclose(zg, z, rtol=0.0, atol=0.0, equal_nan=True))

    if args.copy_probe:
        a = torch.empty(64 * 1024 * 1024, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    # algorithmic work of one launch: per-particle series length at the start and end
    x_start = pm.bij.forward(z0.cpu())
    x_end = pm.bij.forward(z.cpu())
    kinds, roles = fl.model_kinds(phys)
    ksz = sim.camera.subgrid_kernel.shape[0]
    tot = 0.0
    tot_tr = 0.0
    tot_survey = 0.0
    for xs in (x_start, x_end):
        e1 = xs[0][0]["e1"].numpy()
        e2 = xs[0][0]["e2"].numpy()
        n = fl.series_length(e1, e2, 50)
        c = fl.per_particle(kinds, roles, cfg.num_pix, cfg.supersample, ksz, n, grad=True,
                            sep_rank=plan.info()["sep_rank"])
        tot += 0.5 * float(c["flops"].sum())
        tot_tr += 0.5 * float(c["transcendentals"].sum())
        tot_survey += 0.5 * float(fl.survey_formula_per_particle(cfg.num_pix, cfg.supersample, ksz, n).sum())
    achieved = tot / (kern_ms * 1e-3) / 1e12

    chi = chi_hist.cpu().numpy()
    best = np.nanmin(chi, axis=1)
    hit = np.nonzero(best <= 1.1)[0]
    ms_per_step = wall / K * 1e3
    value = bs * K / wall
    out = {
        "metric": "sim+grad particle-evals/s",
        "value": value,
        "unit": "particle-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / BASELINE_EVALS,
        "dtype": "f32",
        "data": "synthetic (prior draws; observed image = seeded prior draw + Poisson + Gaussian noise)",
        "config": {"workload": "multi-start MAP, 60x60 px, ss2, 21x21 PSF, EPL+shear + 2 Sersic, "
                               f"{bs} particles, Keras Adam lr {args.lr}",
                   "particles": bs, "num_pix": cfg.num_pix, "supersample": cfg.supersample,
                   "psf_subgrid": ksz, "params": z0.shape[1]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": measured_traffic() if bs == 4000 else None, "traffic_unit": "bytes",
                     "kernel": "gl::particle_kernel<512, EPL+SHEAR | SERSIC_ELLIPSE | SERSIC_ELLIPSE, grad>",
                     "kernel_ms": kern_ms, "flop_per_launch": tot, "transcendentals_per_launch": tot_tr,
                     "psf_path": ("separable rank %d" % plan.info()["sep_rank"]) if plan.info()["sep_rank"]
                     else "direct 2-D",
                     "reference_formulation_flop_per_launch": tot_survey,
                     "reference_formulation_achieved": tot_survey / (kern_ms * 1e-3) / 1e12,
                     "reference_formulation_frac": tot_survey / (kern_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                     "formulation_note": "achieved/frac count the flops the kernel executes (Clenshaw series, "
                                         "separable PSF, adjoint); reference_formulation_* count the reference's "
                                         "per-particle-eval operations (series, dense PSF, autodiff backward)"},
        "hip_graph": {"ms_per_step": graph_ms, "eager_ms_per_step": ms_per_step,
                      "particle_evals_per_s": bs / (graph_ms * 1e-3), "same_result_as_eager": graph_same},
        "time_to_map": {"chi2_threshold": 1.1,
                        "steps": int(hit[0]) + 1 if len(hit) else None,
                        "seconds": (int(hit[0]) + 1) * ms_per_step / 1e3 if len(hit) else None,
                        "best_chi2_final": float(best[-1]), "init": "prior"},
    }
    if not args.no_learned:
        out["time_to_map"]["learned_vs_prior"] = learned_time_to_map(prior, phys, cfg, dev, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(obs)
        out["oracle_error"] = oracle_error_map(prior, phys, cfg, pm, obs, z0)
    print(json.dumps(out), flush=True)


def learned_time_to_map(prior, phys, cfg, dev, args):
    """Learned initialiser, two training recipes on prior draws simulated on the
    device (batch 1024, AdamW, seed 10): the configured one (lr 3e-4, one pass over
    ``--learned-images``) and a tuned one (lr ``--learned-lr``, ``--learned-epochs``
    passes with fresh noise each).  Held-out lenses (seeds 1000..) are fitted by
    1,000-particle MAP from each network's screened proposals, from prior draws and
    from screened prior draws (one multi-lens launch per step for each); the first
    16 lenses are also fitted one at a time per strategy, each run timed."""
    from gigalens_amd.learned import LearnedInitializer, compare_initialisation
    recipes = {"configured": {"lr": 3e-4, "epochs": 1, "schedule": "constant"},
               "tuned": {"lr": args.learned_lr, "epochs": args.learned_epochs, "schedule": args.learned_schedule}}
    inits, training = {}, {}
    for name, rc in recipes.items():
        init = LearnedInitializer(phys, cfg, prior, 0.2, 100.0, max_batch=1024, seed=0, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = init.train(n_images=args.learned_images, batch=1024, lr=rc["lr"], weight_decay=0.01, seed=10,
                            schedule=rc["schedule"], epochs=rc["epochs"])
        torch.cuda.synchronize()
        inits[name] = init
        training[name] = {"images": args.learned_images, "epochs": rc["epochs"], "batch": 1024, "lr": rc["lr"],
                          "schedule": rc["schedule"], "seconds": time.perf_counter() - t0,
                          "nll_first10": float(np.mean(losses[:10])), "nll_last10": float(np.mean(losses[-10:]))}
    r = compare_initialisation(inits, n_lenses=args.learned_lenses, n_particles=1000, num_steps=350, lr=args.lr)
    r["training"] = training
    r["training_note"] = "each pass re-simulates the same prior draws with fresh noise; the target is a p50 time " \
                         "ratio vs prior starts <= 1/3"
    return r


if __name__ == "__main__":
    main()
