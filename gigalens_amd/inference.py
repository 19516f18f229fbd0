"""Modelling sequence: multi-start MAP -> SVI -> HMC (mirrors
``gigalens.inference`` and ``gigalens.tf.inference``).

All three drivers evaluate the log posterior and its gradient through the
fused HIP kernels.  MAP runs one fused launch per step (log_prob + adjoint +
Keras-Adam update, tf/inference.py:28-35).  SVI calls the fused
value-and-gradient entry point once per step with the reparameterisation as
torch ops on the same stream.  HMC keeps its whole state, adaptation included,
on the device (``DeviceHMC``, csrc/hmc.hip).
"""
from __future__ import annotations

import ctypes
import math
from typing import Callable, Optional, Union

import numpy as np
import torch

from . import distributions as tfd
from .simulator import LensSimulator


# ---------------------------------------------------------------------------
# Optimiser objects with Keras semantics
# ---------------------------------------------------------------------------
class PolynomialDecay:
    """``tf.keras.optimizers.schedules.PolynomialDecay`` (used in tf-demo.ipynb cells 12, 14)."""

    def __init__(self, initial_learning_rate, decay_steps, end_learning_rate=1e-4, power=1.0):
        self.initial_learning_rate = float(initial_learning_rate)
        self.decay_steps = float(decay_steps)
        self.end_learning_rate = float(end_learning_rate)
        self.power = float(power)

    def __call__(self, step):
        s = min(float(step), self.decay_steps)
        frac = 1.0 - s / self.decay_steps
        return (self.initial_learning_rate - self.end_learning_rate) * frac ** self.power + self.end_learning_rate


class Adam:
    """Keras Adam: ``lr_t = lr sqrt(1-b2^t)/(1-b1^t)``, ``x -= lr_t m / (sqrt(v) + eps)``
    with eps outside the square root (epsilon default 1e-7)."""

    def __init__(self, learning_rate: Union[float, Callable] = 1e-3, beta_1=0.9, beta_2=0.999, epsilon=1e-7):
        self.learning_rate = learning_rate
        self.beta_1, self.beta_2, self.epsilon = float(beta_1), float(beta_2), float(epsilon)
        self.iterations = 0
        self._state = {}

    def lr(self, step0: int) -> float:
        lr = self.learning_rate
        return float(lr(step0)) if callable(lr) else float(lr)

    def apply_gradients(self, grads_and_vars):
        """Update torch tensors in place (minimisation)."""
        t = self.iterations + 1
        lr = self.lr(self.iterations)
        lr_t = lr * math.sqrt(1 - self.beta_2 ** t) / (1 - self.beta_1 ** t)
        with torch.no_grad():
            for g, var in grads_and_vars:
                st = self._state.get(id(var))
                if st is None:
                    st = (torch.zeros_like(var), torch.zeros_like(var))
                    self._state[id(var)] = st
                m, v = st
                m.mul_(self.beta_1).add_(g, alpha=1 - self.beta_1)
                v.mul_(self.beta_2).addcmul_(g, g, value=1 - self.beta_2)
                var.sub_(lr_t * m / (v.sqrt() + self.epsilon))
        self.iterations += 1


# ---------------------------------------------------------------------------
# Surrogate posteriors
# ---------------------------------------------------------------------------
def _fill_scale_tril_inverse(L: torch.Tensor, diag_shift: float) -> torch.Tensor:
    d = L.shape[0]
    raw = L.clone()
    idx = torch.arange(d)
    raw[idx, idx] = torch.log(torch.clamp(L[idx, idx] - diag_shift, min=1e-30))
    return raw


class MultivariateNormalTriL:
    """Full-rank Gaussian surrogate ``N(loc, L L^T)`` with ``L = FillScaleTriL(diag
    Exp, diag_shift)`` (tf/inference.py:56-63).  ``loc`` and the unconstrained
    ``raw`` are the trainable tensors."""

    def __init__(self, loc: torch.Tensor, scale: torch.Tensor, diag_shift: float = 1e-6, full_rank: bool = True):
        self.d = loc.shape[-1]
        self.diag_shift = diag_shift
        self.full_rank = full_rank
        self.loc = loc.detach().clone().requires_grad_(True)
        if full_rank:
            L0 = torch.diag(scale.to(loc))
            self.raw = _fill_scale_tril_inverse(L0, diag_shift).requires_grad_(True)
        else:
            self.raw = torch.log(torch.clamp(scale.to(loc) - diag_shift, min=1e-30)).requires_grad_(True)

    def trainable_variables(self):
        return [self.loc, self.raw]

    def scale_tril(self) -> torch.Tensor:
        if self.full_rank:
            L = torch.tril(self.raw, diagonal=-1)
            return L + torch.diag(torch.exp(torch.diagonal(self.raw)) + self.diag_shift)
        return torch.diag(torch.exp(self.raw) + self.diag_shift)

    def sample(self, n, generator: Optional[torch.Generator] = None) -> torch.Tensor:
        eps = torch.randn((int(n), self.d), generator=generator, device="cpu").to(self.loc)
        return self.loc + eps @ self.scale_tril().T

    def sample_and_log_prob(self, n, generator: Optional[torch.Generator] = None):
        """Reparameterised samples z = loc + L eps and log q(z) evaluated from eps,
        -|eps|^2/2 - log det L - d/2 log 2 pi: the same value and total gradient
        as log_prob(z), without the triangular solve (which loses all precision in
        FP32 once L is ill-conditioned)."""
        eps = torch.randn((int(n), self.d), generator=generator, device="cpu").to(self.loc)
        L = self.scale_tril()
        z = self.loc + eps @ L.T
        lq = -0.5 * (eps ** 2).sum(1) - torch.log(torch.diagonal(L)).sum() - 0.5 * self.d * math.log(2 * math.pi)
        return z, lq

    def log_prob(self, z: torch.Tensor) -> torch.Tensor:
        L = self.scale_tril()
        diff = (z - self.loc).T
        sol = torch.linalg.solve_triangular(L, diff, upper=False)
        return (-0.5 * (sol ** 2).sum(0) - torch.log(torch.diagonal(L)).sum()
                - 0.5 * self.d * math.log(2 * math.pi))

    def mean(self) -> torch.Tensor:
        return self.loc.detach()

    def covariance(self) -> torch.Tensor:
        L = self.scale_tril().detach()
        return L @ L.T

    def stddev(self) -> torch.Tensor:
        return torch.sqrt(torch.diagonal(self.covariance()))


# ---------------------------------------------------------------------------
# The modelling sequence
# ---------------------------------------------------------------------------
class ModellingSequenceInterface:
    """``gigalens.inference.ModellingSequenceInterface`` (src/gigalens/inference.py:10-109)."""

    def __init__(self, phys_model, prob_model, sim_config):
        self.phys_model = phys_model
        self.prob_model = prob_model
        self.sim_config = sim_config


class ModellingSequence(ModellingSequenceInterface):
    """``gigalens.tf.inference.ModellingSequence`` (src/gigalens/tf/inference.py:16-141)."""

    def __init__(self, phys_model, prob_model, sim_config, device=None):
        super().__init__(phys_model, prob_model, sim_config)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.chi2_history = None
        self.lp_history = None

    def _sim(self, bs):
        return LensSimulator(self.phys_model, self.sim_config, bs=bs, device=self.device)

    def _to_z(self, start):
        if torch.is_tensor(start) and start.dim() == 2:
            return start.to(self.device, torch.float32).contiguous()
        return self.prob_model.bij.inverse(start).to(self.device, torch.float32).contiguous()

    # ---- MAP ----------------------------------------------------------------
    def MAP(self, optimizer=None, start=None, n_samples=500, num_steps=350, seed=0, progress=False,
            record_every=1, graph=False):
        """Multi-start MAP (tf/inference.py:17-41): returns unconstrained z (n, d).

        With a Keras-style ``Adam`` each step is one fused launch (log_prob,
        adjoint, Adam update).  ``chi2_history`` and ``lp_history`` hold the
        per-particle mean r^2 and log posterior of every recorded step (device
        tensors, values before that step's update).  ``graph=True`` records the
        ``num_steps`` launches (with their learning rates) into one HIP graph
        and replays it: the same kernels and results without per-step host work."""
        optimizer = optimizer if optimizer is not None else Adam(1e-2)
        if start is None:
            start = self.prob_model.prior.sample(n_samples, seed=seed)
        z = self._to_z(start).clone()
        n = z.shape[0]
        sim = self._sim(n)
        pm = self.prob_model
        if not isinstance(optimizer, Adam) or not hasattr(pm, "plan_for"):
            return self._map_autograd(optimizer, z, sim, num_steps)
        plan = pm.plan_for(sim)
        m, v = torch.zeros_like(z), torch.zeros_like(z)
        nrec = (num_steps - 1) // record_every + 1
        hist = torch.empty((nrec, n), device=self.device)
        lp_hist = torch.empty((nrec, n), device=self.device)
        scratch = torch.empty((2, n), device=self.device)
        def steps(it, show):
            for s in it:
                step = optimizer.iterations + 1
                rec = s % record_every == 0
                row = hist[s // record_every] if rec else scratch[0]
                lp = lp_hist[s // record_every] if rec else scratch[1]
                plan.map_adam_step(z, m, v, step, optimizer.lr(optimizer.iterations), optimizer.beta_1,
                                   optimizer.beta_2, optimizer.epsilon, lp=lp, chi2=row)
                optimizer.iterations += 1
                if show and s % 10 == 0:
                    it.set_description(f"Chi Squared: {float(torch.nan_to_num(row, nan=float('inf')).min()):.4f}")
        if graph:
            # the progress text needs a host read per step, which a graph cannot hold: the
            # steps are recorded without it and the final value is reported after the replay
            from .graphs import capture
            capture(lambda: steps(range(num_steps), False)).replay()
            if progress:
                print(f"Chi Squared: {float(torch.nan_to_num(hist[-1], nan=float('inf')).min()):.4f}", flush=True)
        else:
            it = range(num_steps)
            if progress:
                from tqdm.auto import trange
                it = trange(num_steps)
            steps(it, progress)
        self.chi2_history = hist
        self.lp_history = lp_hist
        return z

    def _map_autograd(self, optimizer, z, sim, num_steps):
        npix = float(self.sim_config.num_pix ** 2)
        z = z.clone().requires_grad_(True)
        for _ in range(num_steps):
            lp, chi2 = self.prob_model.log_prob(sim, z)
            loss = torch.mean(-lp / npix)
            g, = torch.autograd.grad(loss, [z])
            if isinstance(optimizer, Adam):
                optimizer.apply_gradients([(g, z)])
            else:
                optimizer.zero_grad()
                z.grad = g
                optimizer.step()
        return z.detach()

    # ---- SVI ----------------------------------------------------------------
    def SVI(self, optimizer=None, start=None, n_vi=250, init_scales=1e-3, num_steps=500, seed=2, full_rank=True):
        """Stochastic VI with a Gaussian surrogate (tf/inference.py:43-73):
        reverse-KL ELBO, ``n_vi`` reparameterised samples per step.  Returns
        ``(q_z, losses)``.  ``full_rank=False`` gives the diagonal surrogate."""
        optimizer = optimizer if optimizer is not None else Adam(1e-3)
        start = torch.as_tensor(start, dtype=torch.float32).reshape(-1).to(self.device)
        d = start.shape[0]
        scale = torch.full((d,), float(init_scales)) if np.size(init_scales) == 1 else \
            torch.as_tensor(init_scales, dtype=torch.float32)
        q = MultivariateNormalTriL(start, scale.to(self.device), full_rank=full_rank)
        sim = self._sim(n_vi)
        gen = torch.Generator().manual_seed(int(seed))
        losses = []
        for _ in range(num_steps):
            z, lq = q.sample_and_log_prob(n_vi, generator=gen)
            lp, _ = self.prob_model.log_prob(sim, z.contiguous())
            loss = torch.mean(lq - lp)
            grads = torch.autograd.grad(loss, q.trainable_variables())
            optimizer.apply_gradients(list(zip(grads, q.trainable_variables())))
            losses.append(loss.detach())
        return q, torch.stack(losses).cpu()

    # ---- HMC ----------------------------------------------------------------
    def HMC(self, q_z, init_eps=0.3, init_l=3, n_hmc=50, num_burnin_steps=250, num_results=750,
            max_leapfrog_steps=30, seed=3, adapt=True, target_accept=0.75, start=None, graph=False):
        """Preconditioned HMC (tf/inference.py:75-141) on the device: momentum ~ N(0,
        Sigma^-1) with Sigma the surrogate covariance (kinetic energy p^T Sigma p / 2),
        dual-averaging step-size adaptation and ChEES trajectory-length adaptation during
        80% of the burn-in, all state kept on the GPU (``DeviceHMC``).  ``adapt=False``
        runs fixed ``init_eps`` / ``init_l``.  ``q_z=None`` gives identity mass; ``start``
        (n_hmc, d) overrides the chains' initial states (default: draws from q_z).
        ``graph=True`` records one whole iteration (momentum draw, every leapfrog step,
        Metropolis correction, adaptation) into a HIP graph and replays it.  Returns
        ``(samples (num_results, n_hmc, d), stats)``."""
        gen = torch.Generator().manual_seed(int(seed))
        if q_z is None:
            d0 = start.shape[1]
            cov64 = torch.eye(d0, dtype=torch.float64, device=self.device)
        else:
            chol = q_z.scale_tril().detach().to(self.device, torch.float64)
            cov64 = chol @ chol.T
        if start is not None:
            z = torch.as_tensor(start, dtype=torch.float32).to(self.device).contiguous()
            n_hmc = z.shape[0]
        else:
            z = q_z.sample(n_hmc, generator=gen).detach().contiguous()
        sim = self._sim(n_hmc)
        plan = self.prob_model.plan_for(sim)
        hmc = DeviceHMC(plan_target(plan, n_hmc), n_hmc, z.shape[1], cov64, init_eps, init_l,
                        max_leapfrog_steps if adapt else int(init_l), num_burnin_steps, num_results, adapt,
                        target_accept, seed, self.device)
        return hmc.run(z, graph=graph)


def plan_target(plan, n: int):
    """Target log-density and gradient of a device plan for ``DeviceHMC``: resident-image
    plans skip leapfrog steps beyond the trajectory on the device (gl_log_prob_masked);
    staged plans evaluate every recorded step (unchanged positions give unchanged values)."""
    from . import _native as nat
    from .plan import _ptr, _stream
    lib = nat.lib()
    masked = plan.info()["lds_bytes"] > 0
    chi2 = torch.empty(n, device="cuda")

    def lp_grad(z, lp, g, k, limit):
        if limit is not None and masked:
            nat.check(lib.gl_log_prob_masked(plan._h, _ptr(z), n, _ptr(lp), _ptr(chi2), None, None, _ptr(g),
                                             _ptr(limit), int(k), _stream()), "gl_log_prob_masked")
        else:
            nat.check(lib.gl_log_prob(plan._h, _ptr(z), n, _ptr(lp), _ptr(chi2), None, None, _ptr(g), _stream()),
                      "gl_log_prob")
    return lp_grad


class DeviceHMC:
    """Preconditioned HMC with dual-averaging step size and ChEES trajectory-length
    adaptation, state on the device (the kernel stack of tf/inference.py:98-131; the
    arithmetic is in csrc/hmc.hip).  ``lp_grad(z, lp, g, k, limit)`` writes the target
    log-density (n,) and its gradient (n, d) at z on the current stream; k is the
    leapfrog step and ``limit`` the device scalar L (the call may skip k >= L; k = -1,
    limit None for the initial evaluation)."""

    def __init__(self, lp_grad, n: int, d: int, cov, init_eps=0.3, init_l=3, max_leapfrog_steps=30,
                 num_burnin_steps=250, num_results=750, adapt=True, target_accept=0.75, seed=3, device="cuda"):
        self.lp_grad = lp_grad
        self.n, self.d = int(n), int(d)
        dev = torch.device(device)
        self.device = dev
        cov64 = torch.as_tensor(cov, dtype=torch.float64, device=dev)
        self.cov = cov64.float().contiguous()
        L = torch.linalg.cholesky(cov64)
        self.chol_inv = torch.linalg.solve_triangular(L, torch.eye(self.d, dtype=torch.float64, device=dev),
                                                      upper=False).float().contiguous()
        self.init_eps, self.init_l = float(init_eps), int(init_l)
        self.max_l = max(1, int(max_leapfrog_steps))
        self.burnin, self.n_results = int(num_burnin_steps), int(num_results)
        self.adapt = bool(adapt)
        self.n_adapt = int(num_burnin_steps * 0.8)  # tf/inference.py:112
        self.target = float(target_accept)
        self.seed = int(seed) * 0x9E3779B97F4A7C15 % (1 << 64)

    def _launch_iteration(self, B):
        from . import _native as nat
        from .plan import _ptr, _stream
        lib = nat.lib()
        n, d = self.n, self.d
        nat.check(lib.gl_hmc_begin(n, d, _ptr(self.chol_inv), ctypes.c_ulonglong(self.seed), _ptr(B["si"]),
                                   _ptr(B["sf"]), _ptr(B["z"]), _ptr(B["lp"]), _ptr(B["g"]), _ptr(B["p"]),
                                   _ptr(B["z0"]), _ptr(B["lp0"]), _ptr(B["g0"]), _ptr(B["p0"]), _stream()),
                  "gl_hmc_begin")
        limit = B["si"][1:2]
        for k in range(self.max_l):
            nat.check(lib.gl_hmc_position(n, d, _ptr(self.cov), _ptr(B["si"]), _ptr(B["sf"]), k, _ptr(B["z"]),
                                          _ptr(B["p"]), _stream()), "gl_hmc_position")
            self.lp_grad(B["z"], B["lp"], B["g"], k, limit)
            nat.check(lib.gl_hmc_momentum(n, d, _ptr(B["si"]), _ptr(B["sf"]), k, _ptr(B["p"]), _ptr(B["g"]),
                                          _stream()), "gl_hmc_momentum")
        nat.check(lib.gl_hmc_end(n, d, _ptr(self.cov), ctypes.c_ulonglong(self.seed), _ptr(B["si"]), _ptr(B["z"]),
                                 _ptr(B["lp"]), _ptr(B["g"]), _ptr(B["p"]), _ptr(B["z0"]), _ptr(B["lp0"]),
                                 _ptr(B["g0"]), _ptr(B["p0"]), _ptr(B["zp"]), _ptr(B["vel"]), _ptr(B["acc"]),
                                 _ptr(B["samples"]), self.burnin, self.n_results, _stream()), "gl_hmc_end")
        nat.check(lib.gl_hmc_adapt(n, d, _ptr(B["si"]), _ptr(B["sf"]), _ptr(B["z0"]), _ptr(B["zp"]), _ptr(B["vel"]),
                                   _ptr(B["acc"]), int(self.adapt), self.n_adapt, self.target, self.max_l,
                                   _ptr(B["eps_h"]), _ptr(B["acc_h"]), _ptr(B["l_h"]), _ptr(B["t_h"]), B["total"],
                                   _stream()), "gl_hmc_adapt")

    def run(self, z0: torch.Tensor, graph: bool = False):
        n, d, dev = self.n, self.d, self.device
        total = self.burnin + self.n_results
        f = lambda *shape: torch.zeros(shape, device=dev, dtype=torch.float32)  # noqa: E731
        B = {"z": z0.to(dev, torch.float32).contiguous().clone(), "lp": f(n), "g": f(n, d), "p": f(n, d),
             "z0": f(n, d), "lp0": f(n), "g0": f(n, d), "p0": f(n, d), "zp": f(n, d), "vel": f(n, d), "acc": f(n),
             "samples": f(max(1, self.n_results), n, d), "eps_h": f(total), "acc_h": f(total),
             "l_h": torch.zeros(total, device=dev, dtype=torch.int32), "t_h": f(total), "total": total}
        eps, T = self.init_eps, self.init_eps * self.init_l
        u0 = 0.5  # base-2 Halton sequence, first element
        L0 = min(self.max_l, max(1, math.ceil(u0 * T / eps))) if self.adapt else self.init_l
        B["si"] = torch.tensor([0, L0, 0, 0], device=dev, dtype=torch.int32)
        B["sf"] = torch.tensor([eps, u0, math.log(T), 0.0, 0.0, math.log(10 * eps), 0.0, math.log(T)], device=dev,
                               dtype=torch.float32)
        self.lp_grad(B["z"], B["lp"], B["g"], -1, None)
        if graph:
            from .graphs import capture
            g = capture(lambda: self._launch_iteration(B))
            for _ in range(total):
                g.replay()
        else:
            for _ in range(total):
                self._launch_iteration(B)
        stats = {"accept_rate": B["acc_h"].cpu().numpy(), "step_size": B["eps_h"].cpu().numpy(),
                 "leapfrog_steps": B["l_h"].cpu().numpy(), "trajectory_length": B["t_h"].cpu().numpy()}
        self.buffers = B
        return B["samples"][:self.n_results], stats
