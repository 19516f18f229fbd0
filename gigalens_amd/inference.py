"""Modelling sequence: multi-start MAP -> SVI -> HMC (mirrors
``gigalens.inference`` and ``gigalens.tf.inference``).

All three drivers evaluate the log posterior and its gradient through the
fused HIP kernels.  MAP runs one fused launch per step (log_prob + adjoint +
Keras-Adam update, tf/inference.py:28-35).  SVI and HMC call the fused
value-and-gradient entry point once per log-density evaluation; the few
small vector operations around it (reparameterisation, leapfrog updates) run
as torch ops on the same stream.
"""
from __future__ import annotations

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
        """Preconditioned HMC (tf/inference.py:75-141): momentum ~ N(0, Sigma^-1)
        with Sigma the surrogate covariance (kinetic energy p^T Sigma p / 2),
        dual-averaging step-size adaptation and ChEES trajectory-length
        adaptation during 80% of the burn-in.  ``adapt=False`` runs fixed
        ``init_eps`` / ``init_l``.  ``q_z=None`` gives plain HMC (identity mass);
        ``start`` (n_hmc, d) overrides the chains' initial states (default: draws
        from q_z).  ``graph=True`` records one leapfrog step (position update,
        log_prob + gradient launch, momentum update) into a HIP graph on static
        buffers, with the step size in a device scalar, and replays it L times
        per trajectory: the same arithmetic in the same order as the eager loop.
        Returns ``(samples (num_results, n_hmc, d), stats)``."""
        gen = torch.Generator().manual_seed(int(seed))
        if q_z is None:
            d0 = start.shape[1]
            chol = torch.eye(d0, dtype=torch.float64, device=self.device)
        else:
            # Sigma = L L^T with L the surrogate's scale_tril (positive diagonal): L is
            # already the Cholesky factor; build Sigma in FP64 for the kinetic energy
            chol = q_z.scale_tril().detach().to(self.device, torch.float64)
        cov = (chol @ chol.T).to(torch.float32)
        chol_inv = torch.linalg.solve_triangular(chol, torch.eye(chol.shape[0], dtype=torch.float64,
                                                                  device=self.device), upper=False)
        chol_inv = chol_inv.to(torch.float32)
        if start is not None:
            z = torch.as_tensor(start, dtype=torch.float32).to(self.device).contiguous()
            n_hmc = z.shape[0]
        else:
            z = q_z.sample(n_hmc, generator=gen).detach().contiguous()
        sim = self._sim(n_hmc)
        pm = self.prob_model
        d = z.shape[1]

        def lp_grad(zz):
            lp, _, g = pm.log_prob_and_grad(sim, zz.contiguous())
            return lp, g

        lp, g = lp_grad(z)
        leap = None
        if graph:
            from .graphs import capture
            st_z, st_p, st_g, st_lp = z.clone(), torch.zeros_like(z), g.clone(), lp.clone()
            eps_t = torch.zeros((), dtype=torch.float32, device=self.device)

            def step():
                st_z.add_((eps_t * st_p) @ cov)
                lpn_, gn_ = lp_grad(st_z)
                st_lp.copy_(lpn_)
                st_g.copy_(gn_)
                st_p.add_(eps_t * st_g)
            leap = capture(step)
        n_adapt = int(num_burnin_steps * 0.8)
        eps = float(init_eps)
        traj = float(init_eps) * int(init_l)  # trajectory length (ChEES adapts this)
        # dual averaging (Hoffman & Gelman 2014, Alg. 5) state
        mu = math.log(10 * eps)
        h_bar, log_eps_bar = 0.0, 0.0
        gamma, t0, kappa = 0.05, 10.0, 0.75
        # ChEES (Hoffman, Radul & Sountsov 2021): Adam on log trajectory length
        log_traj = math.log(traj)
        ch_m, ch_v = 0.0, 0.0
        samples, accept_hist, eps_hist = [], [], []
        total = num_burnin_steps + num_results
        for it in range(total):
            jitter = (it * 0.6180339887498949) % 1.0 if adapt else 1.0
            L = int(min(max_leapfrog_steps, max(1, math.ceil(jitter * traj / eps)))) if adapt else int(init_l)
            p = torch.randn((n_hmc, d), generator=gen).to(self.device) @ chol_inv
            # p ~ N(0, Sigma^-1):  p = eps_std @ L^-1 with Sigma = L L^T
            z0, lp0, g0, p0 = z, lp, g, p
            if leap is not None:
                st_z.copy_(z)
                st_p.copy_(p + 0.5 * eps * g)
                eps_t.fill_(eps)
                for _ in range(L):
                    leap.replay()
                zn, pn, gn, lpn = st_z.clone(), st_p.clone(), st_g.clone(), st_lp.clone()
            else:
                zn, pn, gn = z.clone(), p + 0.5 * eps * g, g
                for _ in range(L):
                    zn = zn + eps * pn @ cov
                    lpn, gn = lp_grad(zn)
                    pn = pn + eps * gn
            pn = pn - 0.5 * eps * gn
            h0 = -lp0 + 0.5 * ((p0 @ cov) * p0).sum(1)
            h1 = -lpn + 0.5 * ((pn @ cov) * pn).sum(1)
            log_acc = torch.clamp(-(h1 - h0), max=0.0)
            log_acc = torch.nan_to_num(log_acc, nan=-float("inf"))
            u = torch.rand(n_hmc, generator=gen).to(self.device)
            acc = torch.log(u) < log_acc
            acc_prob = torch.exp(log_acc)
            z = torch.where(acc[:, None], zn, z0)
            lp = torch.where(acc, lpn, lp0)
            g = torch.where(acc[:, None], gn, g0)
            if adapt and it < n_adapt:
                a = float(acc_prob.mean())
                m_ = it + 1
                h_bar = (1 - 1 / (m_ + t0)) * h_bar + (target_accept - a) / (m_ + t0)
                log_eps = mu - math.sqrt(m_) / gamma * h_bar
                w = m_ ** (-kappa)
                log_eps_bar = w * log_eps + (1 - w) * log_eps_bar
                eps = math.exp(log_eps)
                # ChEES gradient of E[(|z'-mean|^2 - |z-mean|^2)^2] w.r.t. trajectory length
                zc = z0 - z0.mean(0)
                znc = zn - zn.mean(0)
                diff = (znc ** 2).sum(1) - (zc ** 2).sum(1)
                vel = (pn @ cov)
                term = acc_prob * diff * (znc * vel).sum(1) * jitter
                term = term[torch.isfinite(term)]
                gch = float(term.mean()) * traj if term.numel() else 0.0
                if math.isfinite(gch):
                    ch_m = 0.9 * ch_m + 0.1 * gch
                    ch_v = 0.999 * ch_v + 0.001 * gch * gch
                    log_traj += 0.025 * (ch_m / (1 - 0.9 ** m_)) / (math.sqrt(ch_v / (1 - 0.999 ** m_)) + 1e-8)
                log_traj = min(max(log_traj, math.log(eps)), math.log(max_leapfrog_steps * eps))
                traj = math.exp(log_traj)
            elif adapt and it == n_adapt:
                eps = math.exp(log_eps_bar)
            if it >= num_burnin_steps:
                samples.append(z.detach().clone())
            accept_hist.append(float(acc.float().mean()))
            eps_hist.append(eps)
        stats = {"accept_rate": np.array(accept_hist), "step_size": np.array(eps_hist)}
        return torch.stack(samples), stats
