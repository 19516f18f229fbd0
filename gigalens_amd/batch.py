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
    """Lens systems with different cameras and particle counts, fitted together.
    ``lenses`` is a sequence of (sim_config, observed_image, n_particles); all
    share the physical model, the prior and the noise model (``ForwardProbModel``,
    tf/model.py:53-86).  Particles are lens-major.  ``band`` 1 puts every lens the
    banded persistent kernel takes into one launch per call (2: only the lenses the
    resident-image kernel cannot hold; 0: none), with at most ``band_rows`` sub-rows
    per band and ``band_lds_bytes`` of LDS per workgroup (0: defaults); ``"auto"``
    times the candidate configurations on this batch and keeps the fastest
    (``autotune.tune_batch``; ``tuning`` holds the timings).  Other lenses
    whose image fits the resident-image kernel run there (one fused launch each); the
    rest run the HBM-staged pipeline in chunks whose workspaces share an arena of at
    most ``workspace_gib``.  ``info`` reports where the lenses went."""

    def __init__(self, phys_model, prior, lenses: Sequence, background_rms: float, exp_time: float, device="cuda",
                 workspace_gib: float = 32.0, fused: bool = True, band=0, band_rows: int = 0,
                 band_lds_bytes: int = 0):
        self.device = torch.device(device)
        self.prior = prior
        self.bij = dist.ChainBijector(prior.experimental_default_event_space_bijector(), dist.PackBijector(prior))
        paths = [p for p, _ in dist.leaves(prior)]
        self.d = len(paths)
        self.plans = []
        self.counts = []
        for cfg, obs, n in lenses:
            # staged lenses keep their workspace in the batch's arena: the plan itself holds one particle's
            self.plans.append(Plan(phys_model, Camera(cfg), layout_from_paths(phys_model, paths), prior=prior,
                                   observed=np.asarray(obs, np.float32), background_rms=background_rms,
                                   exp_time=exp_time, likelihood=nat.GL_LIK_FORWARD, staged=not fused, max_batch=1))
            self.counts.append(int(n))
        self.n_lenses = len(self.plans)
        self.fused = [p.info()["lds_bytes"] > 0 for p in self.plans]
        self._lib = nat.lib()
        self._h = None
        self.workspace_gib = float(workspace_gib)
        self.total = int(sum(self.counts))
        self.first = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)
        self.lens = torch.arange(self.n_lenses, device=self.device, dtype=torch.int32).repeat_interleave(
            torch.as_tensor(self.counts, device=self.device))
        self.lp_history = None
        self.chi2_history = None
        self.tuning = None
        if band == "auto":
            from .autotune import tune_batch
            z = self.bij.inverse(prior.sample(self.total, seed=12345)).to(self.device, torch.float32).contiguous()
            self.tuning = tune_batch(self, z)
        else:
            self.configure(int(band), int(band_rows), int(band_lds_bytes))

    def configure(self, band: int = 0, band_rows: int = 0, band_lds_bytes: int = 0) -> None:
        """(Re)build the device batch over the same plans with these ``gl_batch_opts``."""
        lib = self._lib
        if self._h is not None and self._h.value:
            lib.gl_batch_destroy(self._h)
            self._h = None
        handles = (ctypes.c_void_p * self.n_lenses)(*[p._h.value for p in self.plans])
        counts = (ctypes.c_int * self.n_lenses)(*self.counts)
        h = ctypes.c_void_p()
        opts = nat.BatchOpts(workspace_bytes=int(self.workspace_gib * (1 << 30)), band=int(band),
                             band_rows=int(band_rows), band_lds_bytes=int(band_lds_bytes))
        nat.check(lib.gl_batch_create_ex(handles, self.n_lenses, counts, ctypes.byref(opts), ctypes.byref(h)),
                  "gl_batch_create_ex")
        self._h = h
        assert int(lib.gl_batch_size(h)) == self.total
        info = nat.BatchInfo()
        nat.check(lib.gl_batch_query(h, ctypes.byref(info)), "gl_batch_query")
        self.info = {name: getattr(info, name) for name, _ in nat.BatchInfo._fields_}
        self.config = {"band": int(band), "band_rows": int(band_rows), "band_lds_bytes": int(band_lds_bytes)}
        self.workspace_bytes = int(lib.gl_batch_workspace_bytes(h))
        self.num_chunks = int(lib.gl_batch_num_chunks(h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.gl_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def log_prob(self, z: torch.Tensor, want_grad: bool = False, out=None):
        """(lp, chi2, dlp/dz or None) of every particle; ``out`` = (lp, chi2, grad)
        buffers to write instead of new tensors."""
        z = z.to(self.device, torch.float32).contiguous()
        if z.shape != (self.total, self.d):
            raise ValueError(f"z must be ({self.total}, {self.d}): every particle of every lens, lens-major")
        if out is None:
            lp = torch.empty(self.total, device=self.device)
            chi2 = torch.empty(self.total, device=self.device)
            grad = torch.empty_like(z) if want_grad else None
        else:
            lp, chi2, grad = out
        nat.check(self._lib.gl_batch_log_prob(self._h, _ptr(z), _ptr(lp), _ptr(chi2), _ptr(grad), _stream()),
                  "gl_batch_log_prob")
        return lp, chi2, grad

    def run(self, z0: torch.Tensor, num_steps: int = 350, optimizer: Optional[Adam] = None,
            record_every: int = 1, graph: bool = False, graph_steps: int = 0, progress=None) -> torch.Tensor:
        """Keras-Adam MAP of every lens at once from z0 (total, d); histories as in
        MultiLensMAP.  ``graph=True`` records the steps into HIP graphs first (of
        ``graph_steps`` steps each, default all) and replays them; ``step_seconds``
        then holds the measured time of each graph (per step).  ``progress(done,
        total)`` is called after each graph."""
        optimizer = optimizer if optimizer is not None else Adam(1e-2)
        z = z0.to(self.device, torch.float32).contiguous().clone()
        if z.shape != (self.total, self.d):
            raise ValueError(f"z0 must be ({self.total}, {self.d}): every particle of every lens, lens-major")
        m, v = torch.zeros_like(z), torch.zeros_like(z)
        nrec = (num_steps - 1) // record_every + 1
        lp_h = torch.empty((nrec, self.total), device=self.device)
        chi_h = torch.empty((nrec, self.total), device=self.device)
        scratch = torch.empty((2, self.total), device=self.device)

        def steps(s0, s1):
            for s in range(s0, s1):
                rec = s % record_every == 0
                lp = lp_h[s // record_every] if rec else scratch[0]
                chi = chi_h[s // record_every] if rec else scratch[1]
                nat.check(self._lib.gl_batch_map_adam_step(
                    self._h, _ptr(z), _ptr(m), _ptr(v), s + 1,
                    float(optimizer.lr(optimizer.iterations + s)), float(optimizer.beta_1), float(optimizer.beta_2),
                    float(optimizer.epsilon), _ptr(lp), _ptr(chi), _stream()), "gl_batch_map_adam_step")
        self.step_seconds = []
        if graph:
            from .graphs import capture
            gs = graph_steps if graph_steps > 0 else num_steps
            for s0 in range(0, num_steps, gs):
                s1 = min(num_steps, s0 + gs)
                g = capture(lambda: steps(s0, s1))
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record()
                g.replay()
                ev1.record()
                self.step_seconds.append((s0, s1, ev0, ev1))
                del g
                if progress is not None:
                    progress(s1, num_steps)
            torch.cuda.synchronize()
            self.step_seconds = [(s0, s1, a.elapsed_time(b) * 1e-3 / (s1 - s0)) for s0, s1, a, b in self.step_seconds]
        else:
            steps(0, num_steps)
        optimizer.iterations += num_steps
        self.lp_history, self.chi2_history = lp_h, chi_h
        return z

    def per_lens_best(self):
        """(best lp, best chi2) of every lens at every recorded step, each (steps, lenses), on the host."""
        idx = self.lens.long()[None].expand(self.lp_history.shape[0], -1)
        lp = torch.nan_to_num(self.lp_history, nan=-float("inf"))
        chi = torch.nan_to_num(self.chi2_history, nan=float("inf"))
        L = self.n_lenses
        best_lp = torch.full((lp.shape[0], L), -float("inf"), device=self.device).scatter_reduce(
            1, idx, lp, reduce="amax")
        best_chi = torch.full((chi.shape[0], L), float("inf"), device=self.device).scatter_reduce(
            1, idx, chi, reduce="amin")
        return best_lp.cpu().numpy(), best_chi.cpu().numpy()

    def best_particles(self, z: torch.Tensor) -> torch.Tensor:
        """Each lens's highest-log-posterior particle of z (total, d) -> (lenses, d)."""
        lp, _, _ = self.log_prob(z)
        lp = torch.nan_to_num(lp, nan=-float("inf"))
        best = torch.full((self.n_lenses,), -float("inf"), device=self.device).scatter_reduce(
            0, self.lens.long(), lp, reduce="amax")
        is_best = lp == best[self.lens.long()]
        pos = torch.arange(self.total, device=self.device)
        first = torch.full((self.n_lenses,), self.total, device=self.device, dtype=torch.long).scatter_reduce(
            0, self.lens.long(), torch.where(is_best, pos, torch.full_like(pos, self.total)), reduce="amin")
        return z[first.clamp(max=self.total - 1)].contiguous()

    def svi_diag(self, loc: torch.Tensor, num_steps: int = 1500, optimizer: Optional[Adam] = None,
                 init_scale: float = 1e-3, seed: int = 2, graph: bool = False):
        """Diagonal-Gaussian SVI of every lens at once (the surrogate of
        tf/inference.py:43-73 restricted to a diagonal scale): every lens of this
        batch holds the same number of particles, used as the Monte-Carlo samples of
        each step.  One step = ``gl_svi_diag_sample`` (on-device Philox normals),
        the batched log_prob + gradient, ``gl_svi_diag_update`` (per-lens ELBO
        gradient + Keras Adam); ``graph=True`` records one step into a HIP graph and
        replays it (a constant learning rate).  loc: (lenses, d) starting means.
        Returns (loc, log_scale, losses (steps, lenses))."""
        optimizer = optimizer if optimizer is not None else Adam(1e-3)
        n_vi = self.counts[0]
        if any(c != n_vi for c in self.counts):
            raise ValueError("svi_diag needs the same number of particles (samples) for every lens")
        if graph and callable(optimizer.learning_rate):
            raise ValueError("graph=True records a constant learning rate")
        L, d = self.n_lenses, self.d
        loc = loc.to(self.device, torch.float32).contiguous().clone()
        if loc.shape != (L, d):
            raise ValueError(f"loc must be ({L}, {d})")
        raw = torch.full_like(loc, float(np.log(init_scale - 1e-6)))
        m = torch.zeros((2, L, d), device=self.device)
        v = torch.zeros((2, L, d), device=self.device)
        step = torch.zeros(L, device=self.device, dtype=torch.int32)
        z = torch.empty((self.total, d), device=self.device)
        eps = torch.empty_like(z)
        lp = torch.empty(self.total, device=self.device)
        chi2 = torch.empty(self.total, device=self.device)
        gz = torch.empty_like(z)
        losses = torch.empty((num_steps, L), device=self.device)
        lib = self._lib

        def one(s):
            nat.check(lib.gl_svi_diag_sample(_ptr(loc), _ptr(raw), L, n_vi, d, ctypes.c_ulonglong(int(seed)),
                                             _ptr(step), _ptr(z), _ptr(eps), _stream()), "gl_svi_diag_sample")
            self.log_prob(z, out=(lp, chi2, gz))
            nat.check(lib.gl_svi_diag_update(_ptr(lp), _ptr(gz), _ptr(eps), L, n_vi, d, _ptr(loc), _ptr(raw), _ptr(m),
                                             _ptr(v), _ptr(step), float(optimizer.lr(optimizer.iterations + s)),
                                             float(optimizer.beta_1), float(optimizer.beta_2),
                                             float(optimizer.epsilon), _ptr(losses), num_steps, _stream()),
                      "gl_svi_diag_update")
        if graph:
            from .graphs import capture
            g = capture(lambda: one(0))
            for _ in range(num_steps):
                g.replay()
        else:
            for s in range(num_steps):
                one(s)
        optimizer.iterations += num_steps
        log_scale = torch.log(torch.exp(raw) + 1e-6)
        return loc, log_scale, losses
