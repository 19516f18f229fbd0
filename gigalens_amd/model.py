"""Physical and probabilistic models (mirrors ``gigalens.model`` and
``gigalens.tf.model``).

``ForwardProbModel.log_prob(simulator, z)`` returns ``(lp, chi2)`` like
src/gigalens/tf/model.py:53-86 through the ``gigalens::log_prob`` operator
(ops.py) and is differentiable: its backward is the hand-written adjoint
kernel, which recomputes the ray-trace instead of storing per-particle grids.
"""
from __future__ import annotations

import os
from abc import ABC, abstractmethod
from collections import OrderedDict
from typing import List

import torch

from . import _native as nat
from . import distributions as dist
from . import ops
from .plan import Plan, layout_from_paths


class PhysicalModel:
    """``gigalens.model.PhysicalModel`` (src/gigalens/model.py:7-29)."""

    def __init__(self, lenses: List, lens_light: List, source_light: List):
        self.lenses = lenses
        self.lens_light = lens_light
        self.source_light = source_light


class ProbabilisticModel(ABC):
    """``gigalens.model.ProbabilisticModel`` (src/gigalens/model.py:32-58)."""

    def __init__(self, prior, bij=None, *args):
        self.prior = prior
        self.bij = bij

    @abstractmethod
    def log_prob(self, simulator, z):
        pass


class _ProbModelBase(ProbabilisticModel):
    _likelihood = nat.GL_LIK_FORWARD
    _fldj_at_x = False

    def __init__(self, prior, observed_image=None, background_rms=None, exp_time=None):
        super().__init__(prior)
        self.observed_image = torch.as_tensor(observed_image, dtype=torch.float32)
        self.background_rms = float(background_rms)
        self.exp_time = float(exp_time)
        self.pack_bij = dist.PackBijector(prior)
        self.unconstraining_bij = prior.experimental_default_event_space_bijector()
        self.bij = dist.ChainBijector(self.unconstraining_bij, self.pack_bij)
        self._paths = [p for p, _ in dist.leaves(prior)]
        self._plans = OrderedDict()

    #: pick the fastest execution variant of each new plan by timing (autotune.py); also
    #: switched on by the environment variable GIGALENS_AMD_AUTOTUNE=1
    autotune = False

    #: plans kept per prob model (least recently used evicted first; a plan holds its
    #: device constants and, for staged plans, a workspace sized for its batch)
    plan_cache_size = 2

    @staticmethod
    def _camera_key(simulator):
        cam = simulator.camera
        k = None if cam.subgrid_kernel is None else cam.subgrid_kernel.tobytes()
        return (cam.num_pix, cam.supersample, cam.transform_sub.tobytes(), cam.origin, cam.conversion_factor, k)

    def plan_for(self, simulator, epl_batch_series=False) -> Plan:
        """The device plan of (physical model, camera, batch size, series rule): built on
        first use, then reused by every simulator of the same model, camera and batch."""
        sbs = int(getattr(simulator, "bs", 0) or 0)
        key = (id(simulator.phys_model), self._camera_key(simulator), sbs, bool(epl_batch_series))
        hit = self._plans.pop(key, None)
        if hit is None:
            lay = layout_from_paths(simulator.phys_model, self._paths)
            kw = dict(prior=self.prior, observed=self.observed_image.cpu().numpy(),
                      background_rms=self.background_rms, exp_time=self.exp_time, likelihood=self._likelihood,
                      fldj_at_x=self._fldj_at_x, epl_batch_series=epl_batch_series)
            # staged plans size their workspace for the simulator's batch (one chunk per call),
            # capped by the library at a 64 GiB workspace
            if 0 < sbs <= 1024:
                kw["max_batch"] = sbs
            variant = {}
            if self.autotune or os.environ.get("GIGALENS_AMD_AUTOTUNE") == "1":
                from .autotune import tune
                bs = max(1, min(int(getattr(simulator, "bs", 256)), 1024))
                z = self.bij.inverse(self.prior.sample(bs, seed=12345)).to(simulator.device, torch.float32)
                variant, _ = tune(simulator.phys_model, simulator.camera, lay, z.contiguous(), kw)
            while len(self._plans) >= max(1, self.plan_cache_size):
                self._plans.pop(next(iter(self._plans)))  # least recently used
            # the entry keeps the physical model alive, so its id stays unique while cached
            hit = (simulator.phys_model, Plan(simulator.phys_model, simulator.camera, lay, **kw, **variant))
        self._plans[key] = hit
        return hit[1]

    def log_prob(self, simulator, z):
        """Log posterior of unconstrained ``z`` (bs, d): ``(lp, chi2)``."""
        plan = self.plan_for(simulator)
        z = torch.as_tensor(z, dtype=torch.float32, device=simulator.device).contiguous()
        return ops.log_prob(plan.handle, z)

    def log_prob_and_grad(self, simulator, z, g_lp=None):
        """Fused value and gradient in one launch: ``(lp, chi2, g_lp * dlp/dz)``."""
        plan = self.plan_for(simulator)
        z = torch.as_tensor(z, dtype=torch.float32, device=simulator.device).contiguous()
        return plan.log_prob(z, g_lp=g_lp, want_grad=True)


class ForwardProbModel(_ProbModelBase):
    """Noise from the simulated image; no least squares (tf/model.py:9-86)."""


class BackwardProbModel(_ProbModelBase):
    """Noise from the observed image; linear light amplitudes by least squares
    (tf/model.py:89-163).  Its Jacobian term is evaluated at the constrained
    parameters, as in tf/model.py:160."""

    _likelihood = nat.GL_LIK_BACKWARD
    _fldj_at_x = True

    def __init__(self, prior, observed_image, background_rms, exp_time):
        super().__init__(prior, observed_image, background_rms, exp_time)
        obs = self.observed_image
        self.err_map = torch.sqrt(self.background_rms ** 2 + torch.clamp(obs, min=0) / self.exp_time)

