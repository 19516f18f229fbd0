"""Fitting many lens systems at once (no counterpart in the reference, which
builds one simulator and one probabilistic model per lens).

``LensBatch`` takes lens systems with their own image sizes, supersampling,
PSFs, observed images and particle counts: each lens is an HBM-staged plan and
every stage of a log_prob / MAP step is one launch over all lenses
(``gl_batch_*``).

``MultiLensMAP`` runs the multi-start MAP of ``ModellingSequence.MAP``
(src/gigalens/tf/inference.py:17-41) for several lens systems that share a
physical model, camera and prior but have their own observed images: one plan
holds all observed images, every particle carries the index of its lens, and
each Adam step is one fused launch over all particles of all lenses.  With
per-particle loss weights 1 / (particles of its lens) every lens follows
exactly the optimisation it would follow alone.
"""
from __future__ import annotations

from typing import Optional

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _native as nat
from . import distributions as dist
from .inference import Adam
from .plan import Plan, _ptr, _stream, layout_from_paths
from .simulator import Camera


class MultiLensMAP:
    def __init__(self, phys_model, sim_config, prior, observed_images, background_rms: float, exp_time: float,
                 device="cuda"):
        self.device = torch.device(device)
        self.prior = prior
        obs = np.ascontiguousarray(np.asarray(observed_images, np.float32))
        if obs.ndim == 2:
            obs = obs[None]
        self.n_lenses = obs.shape[0]
        self.bij = dist.ChainBijector(prior.experimental_default_event_space_bijector(), dist.PackBijector(prior))
        paths = [p for p, _ in dist.leaves(prior)]
        self.plan = Plan(phys_model, Camera(sim_config), layout_from_paths(phys_model, paths), prior=prior,
                         observed=obs, background_rms=background_rms, exp_time=exp_time,
                         likelihood=nat.GL_LIK_FORWARD)
        self.lp_history: Optional[torch.Tensor] = None
        self.chi2_history: Optional[torch.Tensor] = None

    def log_prob(self, z: torch.Tensor, lens: torch.Tensor):
        lp, chi2, _ = self.plan.log_prob(z.contiguous(), lens=lens)
        return lp, chi2

    def run(self, z0: torch.Tensor, lens: torch.Tensor, num_steps: int = 350, optimizer: Optional[Adam] = None,
            record_every: int = 1, graph: bool = False) -> torch.Tensor:
        """Adam (Keras semantics) from z0 (n, d); lens (n,) int32.  Returns z;
        ``lp_history`` / ``chi2_history`` hold the per-particle values of every
        recorded step (before that step's update)."""
        optimizer = optimizer if optimizer is not None else Adam(1e-2)
        z = z0.to(self.device, torch.float32).contiguous().clone()
        lens = lens.to(self.device, torch.int32).contiguous()
        n = z.shape[0]
        counts = torch.bincount(lens.long(), minlength=self.n_lenses).float()
        if int(lens.min()) < 0 or int(lens.max()) >= self.n_lenses:
            raise ValueError("lens index out of range")
        weight = (1.0 / counts[lens.long()]).contiguous()
        m, v = torch.zeros_like(z), torch.zeros_like(z)
        nrec = (num_steps - 1) // record_every + 1
        lp_h = torch.empty((nrec, n), device=self.device)
        chi_h = torch.empty((nrec, n), device=self.device)
        scratch = torch.empty((2, n), device=self.device)

        def steps():
            for s in range(num_steps):
                rec = s % record_every == 0
                lp = lp_h[s // record_every] if rec else scratch[0]
                chi = chi_h[s // record_every] if rec else scratch[1]
                self.plan.map_adam_step(z, m, v, optimizer.iterations + 1, optimizer.lr(optimizer.iterations),
                                        optimizer.beta_1, optimizer.beta_2, optimizer.epsilon, lp=lp, chi2=chi,
                                        lens=lens, weight=weight)
                optimizer.iterations += 1
        if graph:
            from .graphs import capture
            capture(steps).replay()
        else:
            steps()
        self.lp_history, self.chi2_history = lp_h, chi_h
        return z

    def per_lens_best(self, lens: torch.Tensor):
        """(best lp per recorded step, best chi2 per recorded step), each (steps, lenses), on the host."""
        lens = lens.to(self.device).long()
        L = self.n_lenses
        lp = torch.nan_to_num(self.lp_history, nan=-float("inf"))
        chi = torch.nan_to_num(self.chi2_history, nan=float("inf"))
        idx = lens[None].expand_as(lp)
        best_lp = torch.full((lp.shape[0], L), -float("inf"), device=self.device).scatter_reduce(
            1, idx, lp, reduce="amax")
        best_chi = torch.full((chi.shape[0], L), float("inf"), device=self.device).scatter_reduce(
            1, idx, chi, reduce="amin")
        return best_lp.cpu().numpy(), best_chi.cpu().numpy()


class LensBatch:
    """Lens systems with different cameras and particle counts in one launch per
    stage.  ``lenses`` is a sequence of (sim_config, observed_image, n_particles);
    all share the physical model, the prior and the noise model
    (``ForwardProbModel``, tf/model.py:53-86).  Particles are lens-major."""

    def __init__(self, phys_model, prior, lenses: Sequence, background_rms: float, exp_time: float, device="cuda"):
        self.device = torch.device(device)
        self.prior = prior
        self.bij = dist.ChainBijector(prior.experimental_default_event_space_bijector(), dist.PackBijector(prior))
        paths = [p for p, _ in dist.leaves(prior)]
        self.d = len(paths)
        self.plans = []
        self.counts = []
        for cfg, obs, n in lenses:
            # the batch owns the workspace of its particles: the plan itself keeps one particle's
            self.plans.append(Plan(phys_model, Camera(cfg), layout_from_paths(phys_model, paths), prior=prior,
                                   observed=np.asarray(obs, np.float32), background_rms=background_rms,
                                   exp_time=exp_time, likelihood=nat.GL_LIK_FORWARD, staged=True, max_batch=1))
            self.counts.append(int(n))
        self.n_lenses = len(self.plans)
        handles = (ctypes.c_void_p * self.n_lenses)(*[p._h.value for p in self.plans])
        counts = (ctypes.c_int * self.n_lenses)(*self.counts)
        h = ctypes.c_void_p()
        lib = nat.lib()
        nat.check(lib.gl_batch_create(handles, self.n_lenses, counts, ctypes.byref(h)), "gl_batch_create")
        self._h = h
        self._lib = lib
        self.total = int(lib.gl_batch_size(h))
        self.lens = torch.arange(self.n_lenses, device=self.device, dtype=torch.int32).repeat_interleave(
            torch.as_tensor(self.counts, device=self.device))
        self.lp_history = None
        self.chi2_history = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.gl_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def log_prob(self, z: torch.Tensor, want_grad: bool = False):
        z = z.to(self.device, torch.float32).contiguous()
        if z.shape != (self.total, self.d):
            raise ValueError(f"z must be ({self.total}, {self.d}): every particle of every lens, lens-major")
        lp = torch.empty(self.total, device=self.device)
        chi2 = torch.empty(self.total, device=self.device)
        grad = torch.empty_like(z) if want_grad else None
        nat.check(self._lib.gl_batch_log_prob(self._h, _ptr(z), _ptr(lp), _ptr(chi2), _ptr(grad), _stream()),
                  "gl_batch_log_prob")
        return lp, chi2, grad

    def run(self, z0: torch.Tensor, num_steps: int = 350, optimizer: Optional[Adam] = None,
            record_every: int = 1, graph: bool = False) -> torch.Tensor:
        """Keras-Adam MAP of every lens at once from z0 (total, d); histories as in
        MultiLensMAP; ``graph=True`` records the steps into one HIP graph first."""
        optimizer = optimizer if optimizer is not None else Adam(1e-2)
        z = z0.to(self.device, torch.float32).contiguous().clone()
        if z.shape != (self.total, self.d):
            raise ValueError(f"z0 must be ({self.total}, {self.d}): every particle of every lens, lens-major")
        m, v = torch.zeros_like(z), torch.zeros_like(z)
        nrec = (num_steps - 1) // record_every + 1
        lp_h = torch.empty((nrec, self.total), device=self.device)
        chi_h = torch.empty((nrec, self.total), device=self.device)
        scratch = torch.empty((2, self.total), device=self.device)

        def steps():
            for s in range(num_steps):
                rec = s % record_every == 0
                lp = lp_h[s // record_every] if rec else scratch[0]
                chi = chi_h[s // record_every] if rec else scratch[1]
                nat.check(self._lib.gl_batch_map_adam_step(
                    self._h, _ptr(z), _ptr(m), _ptr(v), optimizer.iterations + 1,
                    float(optimizer.lr(optimizer.iterations)), float(optimizer.beta_1), float(optimizer.beta_2),
                    float(optimizer.epsilon), _ptr(lp), _ptr(chi), _stream()), "gl_batch_map_adam_step")
                optimizer.iterations += 1
        if graph:
            from .graphs import capture
            capture(steps).replay()
        else:
            steps()
        self.lp_history, self.chi2_history = lp_h, chi_h
        return z

    def svi_diag(self, loc: torch.Tensor, num_steps: int = 1500, optimizer: Optional[Adam] = None,
                 init_scale: float = 1e-3, seed: int = 2):
        """Diagonal-Gaussian SVI of every lens at once (the surrogate of
        tf/inference.py:43-73 restricted to a diagonal scale).  The batch's
        particle count per lens is the number of Monte-Carlo samples per step.
        loc: (lenses, d) starting means.  Reparameterised ELBO gradient from one
        batched log_prob + gradient launch per step:
          d/d loc = mean_s grad_z lp,  d/d log sigma = mean_s grad_z lp * eps * sigma + 1.
        Returns (loc, log_sigma, losses (steps, lenses))."""
        optimizer = optimizer if optimizer is not None else Adam(1e-3)
        L, d = loc.shape
        counts = torch.as_tensor(self.counts, device=self.device)
        loc = loc.to(self.device, torch.float32).clone()
        logs = torch.full_like(loc, float(np.log(init_scale)))
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        idx = self.lens.long()
        inv_n = (1.0 / counts.float())[:, None]
        m = [torch.zeros_like(loc), torch.zeros_like(logs)]
        v = [torch.zeros_like(loc), torch.zeros_like(logs)]
        losses = torch.empty((num_steps, L), device=self.device)
        for s in range(num_steps):
            eps = torch.randn((self.total, d), device=self.device, generator=g)
            sig = torch.exp(logs)
            z = (loc[idx] + sig[idx] * eps).contiguous()
            lp, _, gz = self.log_prob(z, want_grad=True)
            # loss = -ELBO = -(mean lp + sum log sigma) + const
            lp_mean = torch.zeros(L, device=self.device).index_add_(0, idx, lp) * inv_n[:, 0]
            losses[s] = -(lp_mean + logs.sum(1))
            g_loc = -torch.zeros_like(loc).index_add_(0, idx, gz) * inv_n
            g_logs = -(torch.zeros_like(logs).index_add_(0, idx, gz * eps * sig[idx]) * inv_n + 1.0)
            lr_t = optimizer.lr(optimizer.iterations) * np.sqrt(1 - optimizer.beta_2 ** (s + 1)) / \
                (1 - optimizer.beta_1 ** (s + 1))
            for k, (par, gr) in enumerate(((loc, g_loc), (logs, g_logs))):
                m[k].mul_(optimizer.beta_1).add_((1 - optimizer.beta_1) * gr)
                v[k].mul_(optimizer.beta_2).add_((1 - optimizer.beta_2) * gr * gr)
                par.sub_(lr_t * m[k] / (torch.sqrt(v[k]) + optimizer.epsilon))
            optimizer.iterations += 1
        return loc, logs, losses
