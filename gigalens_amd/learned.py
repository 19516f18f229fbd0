"""Learned initialiser for the multi-start MAP (no counterpart in the reference).

A small convolutional network -- 4 x (3x3 conv + GELU) with strides 1,2,1,2,
global average pool, a 256-wide hidden layer and a (mean, log-std) head --
maps an observed image to a diagonal Gaussian over the standardised
unconstrained parameters z of a prior.  Its forward and backward passes are
the BF16 MFMA kernels behind ``gl_cnn_*`` (FP32 accumulation); training data
are simulated on the device from the prior with fresh Poisson + Gaussian
noise every batch.  ``LearnedInitializer.propose`` draws MAP starting points
from the network and mixes in prior draws, which stay the fallback: every
start is judged by the simulator's log-posterior through the MAP itself
(``ModellingSequence.MAP``, src/gigalens/tf/inference.py:17-41).
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from . import distributions as dist
from .plan import _ptr, _stream
from .simulator import LensSimulator

STRIDES = (1, 2, 1, 2)


class ConvNet:
    """The network behind ``gl_cnn_create`` (weights, AdamW state and
    activation workspace live on the device)."""

    def __init__(self, height: int, width: int, n_out: int, channels: Sequence[int] = (32, 32, 64, 64),
                 hidden: int = 256, max_batch: int = 1024, in_scale: float = 0.2, seed: int = 0, device="cuda"):
        self.height, self.width, self.n_out = int(height), int(width), int(n_out)
        self.channels = tuple(int(c) for c in channels)
        self.hidden = int(hidden)
        self.max_batch = int(max_batch)
        self.in_scale = float(in_scale)
        self.device = torch.device(device)
        desc = nat.CnnDesc()
        desc.height, desc.width = self.height, self.width
        for i, c in enumerate(self.channels):
            desc.channels[i] = c
        desc.hidden, desc.n_out, desc.max_batch, desc.in_scale = self.hidden, self.n_out, self.max_batch, self.in_scale
        h = ctypes.c_void_p()
        nat.check(nat.lib().gl_cnn_create(ctypes.byref(desc), ctypes.byref(h)), "gl_cnn_create")
        self._h = h
        self.n_params = int(nat.lib().gl_cnn_num_params(h))
        self.layout = self.param_layout(self.channels, self.hidden, self.n_out)
        assert self.layout[-1][2] + int(np.prod(self.layout[-1][1])) == self.n_params
        self.set_params(self.init_params(seed))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                nat.lib().gl_cnn_destroy(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def param_layout(channels, hidden, n_out):
        """[(name, shape, offset)] of the parameter vector, in the order of the
        C ABI: per conv layer W[co][(kh*3+kw)*cin + ci] and b[co], then the head."""
        out = []
        off = 0
        cin = 1
        for i, c in enumerate(channels):
            out.append((f"conv{i}.w", (c, 9 * cin), off))
            off += c * 9 * cin
            out.append((f"conv{i}.b", (c,), off))
            off += c
            cin = c
        for name, shape in (("fc1.w", (hidden, cin)), ("fc1.b", (hidden,)), ("fc2.w", (2 * n_out, hidden)),
                            ("fc2.b", (2 * n_out,))):
            out.append((name, shape, off))
            off += int(np.prod(shape))
        return out

    def init_params(self, seed: int) -> torch.Tensor:
        """He-normal convolutions and hidden layer, a small output layer, zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n_params, np.float32)
        for name, shape, off in self.layout:
            n = int(np.prod(shape))
            if name.endswith(".w"):
                fan_in = shape[1]
                std = math.sqrt(2.0 / fan_in) if not name.startswith("fc2") else 0.1 / math.sqrt(fan_in)
                p[off:off + n] = rng.normal(0.0, std, size=n)
        return torch.as_tensor(p, device=self.device)

    def unpack(self, p: torch.Tensor) -> dict:
        return {name: p[off:off + int(np.prod(shape))].reshape(shape) for name, shape, off in self.layout}

    def set_params(self, p: torch.Tensor) -> None:
        p = p.to(self.device, torch.float32).contiguous()
        assert p.numel() == self.n_params
        nat.check(nat.lib().gl_cnn_set_params(self._h, _ptr(p), _stream()), "gl_cnn_set_params")
        torch.cuda.current_stream().synchronize()

    def get_params(self) -> torch.Tensor:
        p = torch.empty(self.n_params, device=self.device)
        nat.check(nat.lib().gl_cnn_get_params(self._h, _ptr(p), _stream()), "gl_cnn_get_params")
        return p

    def _images(self, images: torch.Tensor) -> torch.Tensor:
        images = images.to(self.device, torch.float32).contiguous()
        if images.dim() == 2:
            images = images[None]
        if images.shape[-2:] != (self.height, self.width):
            raise ValueError(f"images must be (bs, {self.height}, {self.width})")
        return images

    def forward(self, images: torch.Tensor) -> torch.Tensor:
        """(bs, 2 n_out): mean and log-std outputs."""
        images = self._images(images)
        bs = images.shape[0]
        out = torch.empty((bs, 2 * self.n_out), device=self.device)
        for o in range(0, bs, self.max_batch):
            n = min(self.max_batch, bs - o)
            nat.check(nat.lib().gl_cnn_forward(self._h, _ptr(images[o:o + n]), n, _ptr(out[o:o + n]), _stream()),
                      "gl_cnn_forward")
        return out

    def loss_grad(self, images: torch.Tensor, target: torch.Tensor, want_grad: bool = False):
        """Mean Gaussian NLL (device scalar) and, when asked, its gradient (P,)."""
        images = self._images(images)
        target = target.to(self.device, torch.float32).contiguous()
        bs = images.shape[0]
        if bs > self.max_batch or target.shape != (bs, self.n_out):
            raise ValueError("batch larger than max_batch or target shape mismatch")
        loss = torch.empty(1, device=self.device)
        grad = torch.empty(self.n_params, device=self.device) if want_grad else None
        nat.check(nat.lib().gl_cnn_loss_grad(self._h, _ptr(images), _ptr(target), bs, _ptr(loss), _ptr(grad),
                                             _stream()), "gl_cnn_loss_grad")
        return loss, grad

    def adamw_step(self, step: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.01) -> None:
        nat.check(nat.lib().gl_cnn_adamw_step(self._h, int(step), float(lr), float(beta1), float(beta2), float(eps),
                                              float(weight_decay), _stream()), "gl_cnn_adamw_step")


def poisson_gauss_noise(img: torch.Tensor, exp_time: float, background_rms: float, seed: int,
                        out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Poisson(max(img, 0) exp_time) / exp_time + background_rms N(0, 1) per pixel on the
    device (csrc/noise.hip); ``out`` may be ``img``."""
    img = img.contiguous()
    if img.device.type != "cuda" or img.dtype != torch.float32:
        raise ValueError("poisson_gauss_noise: img must be a float32 CUDA tensor")
    if out is None:
        out = torch.empty_like(img)
    elif out.shape != img.shape or out.dtype != torch.float32 or out.device != img.device or not out.is_contiguous():
        raise ValueError("poisson_gauss_noise: out must match img")
    nat.check(nat.lib().gl_noise_poisson_gauss(_ptr(img), img.numel(), float(exp_time), float(background_rms),
                                               ctypes.c_ulonglong(int(seed) % (2 ** 64)), _ptr(out), _stream()),
              "gl_noise_poisson_gauss")
    return out


def _flat(prior, sample) -> torch.Tensor:
    return torch.stack([torch.as_tensor(v, dtype=torch.float32) for v in dist.flatten(sample)], dim=1)


class LearnedInitializer:
    """q(z | image): a diagonal Gaussian over the unconstrained parameters,
    predicted by ``ConvNet`` and standardised with prior moments of z."""

    def __init__(self, phys_model, sim_config, prior, background_rms: float, exp_time: float,
                 channels=(32, 32, 64, 64), hidden: int = 256, max_batch: int = 1024, seed: int = 0,
                 n_stats: int = 65536, device="cuda"):
        self.phys_model = phys_model
        self.sim_config = sim_config
        self.prior = prior
        self.background_rms = float(background_rms)
        self.exp_time = float(exp_time)
        self.device = torch.device(device)
        self.bij = dist.ChainBijector(prior.experimental_default_event_space_bijector(), dist.PackBijector(prior))
        self.d = len(dist.leaves(prior))
        self.max_batch = int(max_batch)
        self.sim = LensSimulator(phys_model, sim_config, bs=self.max_batch, device=self.device)
        z = self.bij.inverse(prior.sample(n_stats, seed=seed + 7919)).double()
        self.z_mean = z.mean(0).float().to(self.device)
        self.z_std = z.std(0).clamp_min(1e-6).float().to(self.device)
        N = sim_config.num_pix
        self.net = ConvNet(N, N, self.d, channels, hidden, self.max_batch, self.background_rms, seed, self.device)
        self.losses: List[float] = []

    # ---- data ------------------------------------------------------------------
    def simulate_batch(self, n: int, seed: int, noise_seed: Optional[int] = None):
        """n prior draws (seed): noisy images (n, N, N) on the device and their z
        (n, d); the noise realisation follows ``noise_seed`` (default: seed)."""
        x_nested = self.prior.sample(n, seed=seed)
        z = self.bij.inverse(x_nested).to(self.device, torch.float32)
        x = _flat(self.prior, x_nested).to(self.device).contiguous()
        img = self.sim.plan.simulate(x)
        ns = seed if noise_seed is None else noise_seed
        noisy = poisson_gauss_noise(img, self.exp_time, self.background_rms, int(ns) * 2654435761 % (2 ** 64),
                                    out=img)
        return noisy, z

    def standardise(self, z: torch.Tensor) -> torch.Tensor:
        return (z - self.z_mean) / self.z_std

    # ---- training --------------------------------------------------------------
    def train(self, n_images: int = 2 ** 20, batch: int = 1024, lr: float = 3e-4, weight_decay: float = 0.01,
              seed: int = 10, log_every: int = 0, schedule: str = "constant", warmup: int = 50,
              epochs: int = 1) -> List[float]:
        """AdamW on the Gaussian NLL over ``n_images`` prior draws (batch k uses
        prior seed seed*1000003 + k), simulated on the device; each of the
        ``epochs`` passes draws fresh noise for the same lens configurations.
        ``schedule`` "cosine": linear warm-up over ``warmup`` steps to ``lr``,
        then cosine decay to 0 at the last step."""
        batch = min(int(batch), self.max_batch)
        per_epoch = max(1, int(n_images) // batch)
        steps = per_epoch * max(1, int(epochs))
        loss_acc = []
        for s in range(1, steps + 1):
            k, e = (s - 1) % per_epoch + 1, (s - 1) // per_epoch
            images, z = self.simulate_batch(batch, seed=seed * 1_000_003 + k,
                                            noise_seed=(seed * 1_000_003 + k) * 7919 + e)
            loss, _ = self.net.loss_grad(images, self.standardise(z))
            lr_s = lr
            if schedule == "cosine":
                lr_s = lr * min(1.0, s / max(1, warmup)) * 0.5 * (1.0 + math.cos(math.pi * (s - 1) / steps))
            self.net.adamw_step(len(self.losses) + len(loss_acc) + 1, lr_s, weight_decay=weight_decay)
            loss_acc.append(loss)
            if log_every and s % log_every == 0:
                print(f"step {s}/{steps}: loss {float(loss):.4f}", flush=True)
        vals = torch.cat(loss_acc).cpu().tolist()
        self.losses.extend(vals)
        return vals

    # ---- inference -------------------------------------------------------------
    def predict(self, images: torch.Tensor):
        """(mean, std) of z for each image (bs, d)."""
        out = self.net.forward(images)
        d = self.d
        mean = self.z_mean + self.z_std * out[:, :d]
        std = self.z_std * torch.exp(torch.clamp(out[:, d:], -7.0, 5.0))
        return mean, std

    def propose(self, image: torch.Tensor, n: int, prior_fraction: float = 0.25, widen: float = 1.0,
                seed: int = 0, prob_model=None, oversample: int = 8) -> torch.Tensor:
        """n starting points in z: (1 - prior_fraction) from the network's
        Gaussian (std scaled by ``widen``), the rest prior draws.  With a
        ``prob_model`` (whose observed image is ``image``) ``oversample * n``
        candidates are drawn and the n with the highest log-posterior kept."""
        if prob_model is not None and oversample > 1:
            cand = self.propose(image, oversample * n, prior_fraction, widen, seed)
            sim = LensSimulator(self.phys_model, self.sim_config, bs=cand.shape[0], device=self.device)
            lp, _, _ = prob_model.plan_for(sim).log_prob(cand)
            keep = torch.topk(torch.nan_to_num(lp, nan=-float("inf")), n).indices
            return cand[keep].contiguous()
        mean, std = self.predict(image)
        n_prior = int(round(prior_fraction * n))
        n_net = n - n_prior
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed) + 12345)
        z_net = mean[0] + widen * std[0] * torch.randn((n_net, self.d), device=self.device, generator=g)
        parts = [z_net]
        if n_prior > 0:
            parts.append(self.bij.inverse(self.prior.sample(n_prior, seed=seed + 1)).to(self.device, torch.float32))
        return torch.cat(parts, 0).contiguous()


def map_trace(fitter, start: torch.Tensor, num_steps: int, lr: float):
    """Run ``fitter.MAP`` from ``start`` and return (best lp per step, best chi2
    per step, seconds per step), all on the host."""
    from .inference import Adam
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fitter.MAP(optimizer=Adam(lr), start=start, num_steps=num_steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / num_steps
    lp = torch.nan_to_num(fitter.lp_history, nan=-float("inf")).max(1).values.cpu().numpy()
    chi = torch.nan_to_num(fitter.chi2_history, nan=float("inf")).min(1).values.cpu().numpy()
    return lp, chi, dt


def first_step(values: np.ndarray, ok) -> Optional[int]:
    idx = np.nonzero(ok(values))[0]
    return int(idx[0]) + 1 if len(idx) else None


def steps_to(trace: np.ndarray, ok) -> np.ndarray:
    """Per column of a (steps, lenses) trace: first 1-based step where ``ok``
    holds, or -1."""
    hit = ok(trace)
    first = np.argmax(hit, axis=0) + 1
    return np.where(hit.any(axis=0), first, -1)


def screen(log_prob_fn, z: torch.Tensor, lens: torch.Tensor, n_lenses: int, per_lens: int, keep: int):
    """Keep the ``keep`` highest-log-posterior candidates of each lens.  The
    candidates are grouped by lens (``per_lens`` each, lens-major order)."""
    lp, _ = log_prob_fn(z, lens)
    lp = torch.nan_to_num(lp, nan=-float("inf")).reshape(n_lenses, per_lens)
    idx = torch.topk(lp, keep, dim=1).indices + per_lens * torch.arange(n_lenses, device=z.device)[:, None]
    return z[idx.reshape(-1)].contiguous()


def compare_initialisation(init, n_lenses: int = 256, n_particles: int = 1000, num_steps: int = 350,
                           lr: float = 5e-3, prior_fraction: float = 0.25, oversample: int = 8, seed0: int = 1000,
                           chi2_threshold: float = 1.1, measure_lenses: int = 16):
    """Held-out evaluation of learned vs prior-sampled MAP starts.

    ``init`` is a ``LearnedInitializer`` or a dict name -> initializer (several
    trained networks against the same prior runs).  Lens k is a prior draw
    simulated with fresh noise (seed seed0 + k).  Start strategies, each fitting all
    lenses in one multi-lens MAP (``n_particles`` starts per lens, Keras Adam ``lr``):

      * ``prior``          -- prior draws (the reference's MAP, tf/inference.py:21-24)
      * ``learned``        -- ``oversample * n_particles`` candidates, (1 - prior_fraction)
                              from the network's Gaussian and the rest prior draws,
                              screened by the log-posterior (keep the best n_particles);
                              ``learned:<name>`` for each network of a dict
      * ``prior_screened`` -- the same screening applied to prior draws only
                              (what the screening alone buys)

    Per lens: steps until the best particle reaches chi2 <= ``chi2_threshold``
    and until it comes within 1 nat of the ``prior`` run's final best
    log-posterior; seconds = steps times the measured time of one single-lens MAP
    step of ``n_particles``, plus the measured per-lens start-up cost (network
    inference, screening evaluation).  ``measured_single_lens``: the first
    ``measure_lenses`` lenses fitted one at a time per strategy -- start-up, then
    ``num_steps`` fused MAP steps -- each run timed, and each lens's time to the
    chi2 threshold read from its own run (start-up + its steps at that run's
    measured step rate)."""
    from .batch import MultiLensMAP
    from .inference import Adam, ModellingSequence
    from .model import ForwardProbModel
    inits = dict(init) if isinstance(init, dict) else {"": init}
    first = next(iter(inits.values()))
    dev = first.device
    d = first.d
    imgs = []
    for k in range(n_lenses):
        im, _ = first.simulate_batch(1, seed=seed0 + k)
        imgs.append(im[0])
    images = torch.stack(imgs).contiguous()
    mlm = MultiLensMAP(first.phys_model, first.sim_config, first.prior, images.cpu().numpy(), first.background_rms,
                       first.exp_time, device=dev)
    lens = torch.arange(n_lenses, device=dev, dtype=torch.int32).repeat_interleave(n_particles)
    K = oversample * n_particles
    lens_c = torch.arange(n_lenses, device=dev, dtype=torch.int32).repeat_interleave(K)
    n_prior = int(round(prior_fraction * K))
    n_net = K - n_prior

    def learned_candidates(net, imgs_, n_l, gen, fb_seed):
        mean, std = net.predict(imgs_)
        z_net = (mean[:, None, :] + std[:, None, :] * torch.randn((n_l, n_net, d), device=dev, generator=gen))
        z_net[:, 0] = mean  # the network's point estimate is one of the candidates
        z_fb = net.bij.inverse(net.prior.sample(n_l * n_prior, seed=fb_seed)).to(dev, torch.float32)
        return torch.cat([z_net, z_fb.reshape(n_l, n_prior, d)], 1).reshape(-1, d).contiguous()

    starts, setup = {}, {}
    for name, net in inits.items():
        key = "learned" if name == "" else f"learned:{name}"
        g = torch.Generator(device=dev)
        g.manual_seed(seed0 + 7)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        z_c = learned_candidates(net, images, n_lenses, g, seed0 + 11)
        starts[key] = screen(mlm.log_prob, z_c, lens_c, n_lenses, K, n_particles)
        torch.cuda.synchronize()
        setup[key] = (time.perf_counter() - t0) / n_lenses
    z_pc = first.bij.inverse(first.prior.sample(n_lenses * K, seed=seed0 + 17)).to(dev, torch.float32).contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    starts["prior_screened"] = screen(mlm.log_prob, z_pc, lens_c, n_lenses, K, n_particles)
    torch.cuda.synchronize()
    setup["prior_screened"] = (time.perf_counter() - t0) / n_lenses
    starts["prior"] = first.bij.inverse(first.prior.sample(n_lenses * n_particles, seed=seed0 + 13)).to(
        dev, torch.float32)
    setup["prior"] = 0.0
    runs = {}
    for name, z0 in starts.items():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mlm.run(z0, lens, num_steps, Adam(lr))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        best_lp, best_chi = mlm.per_lens_best(lens)
        runs[name] = {"best_lp": best_lp, "best_chi": best_chi, "batched_seconds": wall}
    # one single-lens MAP step of n_particles (the per-lens cost of each step)
    pm = ForwardProbModel(first.prior, images[0].cpu().numpy(), first.background_rms, first.exp_time)
    fitter = ModellingSequence(first.phys_model, pm, first.sim_config, device=dev)
    fitter.MAP(optimizer=Adam(lr), start=starts["prior"][:n_particles], num_steps=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fitter.MAP(optimizer=Adam(lr), start=starts["prior"][:n_particles], num_steps=50)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / 50
    target = runs["prior"]["best_lp"][-1] - 1.0

    def stats(s, extra):
        ok = s > 0
        if not ok.any():
            return {"reached": 0}
        sec = s[ok] * t_step + extra
        return {"reached": int(ok.sum()), "steps_p50": float(np.median(s[ok])),
                "steps_p95": float(np.percentile(s[ok], 95)), "seconds_p50": float(np.median(sec)),
                "seconds_p95": float(np.percentile(sec, 95))}
    res = {"lenses": n_lenses, "particles": n_particles, "steps": num_steps, "lr": lr,
           "prior_fraction": prior_fraction, "oversample": oversample, "chi2_threshold": chi2_threshold,
           "single_lens_step_seconds": t_step, "setup_seconds_per_lens": setup,
           "time_to_chi2": {}, "time_to_prior_best_lp_within_1nat": {}, "final_best_lp_gain_vs_prior_p50": {},
           "batched_seconds": {}}
    steps_lp = {}
    for name, r in runs.items():
        s_chi = steps_to(r["best_chi"], lambda c: c <= chi2_threshold)
        s_lp = steps_to(r["best_lp"], lambda v: v >= target[None, :])
        steps_lp[name] = s_lp
        res["time_to_chi2"][name] = stats(s_chi, setup[name])
        res["time_to_prior_best_lp_within_1nat"][name] = stats(s_lp, setup[name])
        res["final_best_lp_gain_vs_prior_p50"][name] = float(np.median(r["best_lp"][-1] - runs["prior"]["best_lp"][-1]))
        res["batched_seconds"][name] = r["batched_seconds"]
    for name in runs:
        if name == "prior":
            continue
        both = (steps_lp[name] > 0) & (steps_lp["prior"] > 0)
        ratio = (steps_lp[name][both] * t_step + setup[name]) / (steps_lp["prior"][both] * t_step)
        res["time_to_prior_best_lp_within_1nat"][name]["time_ratio_vs_prior_p50"] = \
            float(np.median(ratio)) if both.any() else None
    # single-lens runs, each timed
    m = min(measure_lenses, n_lenses)
    if m > 0:
        meas = {name: [] for name in runs}
        for k in range(m):
            pm_k = ForwardProbModel(first.prior, images[k].cpu().numpy(), first.background_rms, first.exp_time)
            fit_k = ModellingSequence(first.phys_model, pm_k, first.sim_config, device=dev)
            sim_k = fit_k._sim(K)
            for name in runs:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if name.startswith("learned"):
                    net = inits["" if name == "learned" else name.split(":", 1)[1]]
                    g = torch.Generator(device=dev)
                    g.manual_seed(seed0 + 7 + 101 * k)
                    z_c = learned_candidates(net, images[k:k + 1], 1, g, seed0 + 11 + 101 * k)
                elif name == "prior_screened":
                    z_c = first.bij.inverse(first.prior.sample(K, seed=seed0 + 17 + 101 * k)).to(
                        dev, torch.float32).contiguous()
                else:
                    z_c = None
                if z_c is not None:
                    lp_c, _ = pm_k.log_prob(sim_k, z_c)
                    z0 = z_c[torch.topk(torch.nan_to_num(lp_c, nan=-float("inf")), n_particles).indices]
                else:
                    z0 = first.bij.inverse(first.prior.sample(n_particles, seed=seed0 + 13 + 101 * k)).to(
                        dev, torch.float32)
                torch.cuda.synchronize()
                t_setup = time.perf_counter() - t0
                t0 = time.perf_counter()
                fit_k.MAP(optimizer=Adam(lr), start=z0.contiguous(), num_steps=num_steps)
                torch.cuda.synchronize()
                t_map = time.perf_counter() - t0
                chi = torch.nan_to_num(fit_k.chi2_history, nan=float("inf")).min(1).values.cpu().numpy()
                hit = np.nonzero(chi <= chi2_threshold)[0]
                meas[name].append(t_setup + (hit[0] + 1) * t_map / num_steps if len(hit) else np.nan)
        res["measured_single_lens"] = {"lenses": m, "chi2_threshold": chi2_threshold,
                                       "note": "each lens fitted alone per strategy; seconds to its best chi2 <= "
                                               "threshold, start-up and MAP timed on that lens's run"}
        for name, v in meas.items():
            v = np.array(v)
            ok = np.isfinite(v)
            row = {"reached": int(ok.sum())}
            if ok.any():
                row.update(seconds_p50=float(np.median(v[ok])), seconds_p95=float(np.percentile(v[ok], 95)))
            if name != "prior":
                p = np.array(meas["prior"])
                both = ok & np.isfinite(p)
                row["time_ratio_vs_prior_p50"] = float(np.median(v[both] / p[both])) if both.any() else None
            res["measured_single_lens"][name] = row
    return res


def central_view(image, size: int, background_rms: float, seed: int) -> torch.Tensor:
    """The central size x size pixels of a centred image on the same pixel grid:
    a crop when the image is larger, else the image inside a border of Gaussian
    background noise (rms ``background_rms``, seeded).  Both grids are centred on the
    optical axis, so pixel centres coincide when N - size is even."""
    img = torch.as_tensor(image, dtype=torch.float32)
    N = img.shape[-1]
    if (N - size) % 2:
        raise ValueError("the image and the view must differ by an even number of pixels")
    if N >= size:
        o = (N - size) // 2
        return img[o:o + size, o:o + size].contiguous()
    g = torch.Generator().manual_seed(int(seed))
    out = background_rms * torch.randn((size, size), generator=g)
    o = (size - N) // 2
    out[o:o + N, o:o + N] = img
    return out


def propose_batch(init: "LearnedInitializer", batch, images: Sequence, prior_fraction: float = 0.25,
                  oversample: int = 2, seed: int = 0) -> torch.Tensor:
    """Starting points for every particle of a ``LensBatch`` (lens-major): each
    lens's network Gaussian (predicted from the central view of its observed image)
    with ``prior_fraction`` of prior draws as the fallback; ``oversample`` candidate
    sets are evaluated through the batch's log_prob and each lens keeps its best
    ``count`` candidates."""
    dev = batch.device
    views = torch.stack([central_view(im, init.sim_config.num_pix, init.background_rms, seed + k)
                         for k, im in enumerate(images)]).to(dev)
    mean, std = init.predict(views)
    lens = batch.lens.long()
    counts = torch.as_tensor(batch.counts, device=dev)
    first = torch.as_tensor(batch.first, device=dev)
    pos = torch.arange(batch.total, device=dev) - first[lens]
    n_prior = torch.round(prior_fraction * counts.float()).long()
    use_prior = pos >= (counts - n_prior)[lens]
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed) + 4242)
    cands, lps = [], []
    for r in range(max(1, int(oversample))):
        z = mean[lens] + std[lens] * torch.randn((batch.total, init.d), device=dev, generator=g)
        if r == 0:
            z[pos == 0] = mean  # the network's point estimate is one of the candidates
        zp = init.bij.inverse(init.prior.sample(batch.total, seed=seed + 100 + r)).to(dev, torch.float32)
        z = torch.where(use_prior[:, None], zp, z).contiguous()
        lp, _, _ = batch.log_prob(z)
        cands.append(z)
        lps.append(torch.nan_to_num(lp, nan=-1e30).clamp(-1e11, 1e11))
    R = len(cands)
    zc = torch.cat(cands)
    lpc = torch.cat(lps).double()
    lens_c = lens.repeat(R)
    order = torch.argsort(lens_c.double() * 1e12 - lpc)  # by lens, then descending log-posterior
    pick = order[R * first[lens] + pos]
    return zc[pick].contiguous()
