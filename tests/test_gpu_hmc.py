"""The device HMC (DeviceHMC, csrc/hmc.hip: preconditioned HMC with dual-averaging
step size and ChEES trajectory length, the kernel stack of
src/gigalens/tf/inference.py:98-131) on Gaussian targets whose answers are known.

* A 10-d Gaussian with scales 0.5..2 (identity mass): after adaptation the mean
  acceptance is at the 0.75 target, and the samples' mean and variance match.
* ChEES on a Gaussian whitened by the mass matrix (mass matrix = target covariance,
  every direction at unit frequency): the adapted trajectory length sits where the
  expected ChEES gradient with respect to T changes sign from + to -.  That gradient
  is computed here independently, by Monte Carlo over leapfrog trajectories of the
  whitened oscillator (x, p) -> L steps of size eps, with the Metropolis acceptance
  weights and the Halton jitter u; with exact dynamics and uniform jitter the
  criterion is 4 d sin^2 averaged over the trajectory time, 1/2 - sin 2T / 4T, first
  maximal at T = 2.2467, and the leapfrog steps and ceil(u T / eps) move that root.
* Scale equivariance: a target scaled by s with the step size scaled by s gives the
  same run in scaled units (log step size and log T shift by log s in exact
  arithmetic): early iterations agree to 1e-3, the adapted T scales by s.
* One iteration recorded into a HIP graph and replayed equals the eager launches.
"""
import math

import numpy as np
import pytest
import torch

from gigalens_amd.inference import DeviceHMC

pytestmark = pytest.mark.gpu


def gaussian(scales):
    s2 = torch.as_tensor(np.asarray(scales, np.float32) ** 2, device="cuda")

    def lp_grad(z, lp, g, k, limit):
        lp.copy_(-0.5 * (z * z / s2).sum(1))
        g.copy_(-z / s2)
    return lp_grad


def run(scales, cov, n=256, burnin=600, results=400, init_eps=0.3, init_l=3, seed=7, graph=False, z0_scale=1.0):
    d = len(scales)
    hmc = DeviceHMC(gaussian(scales), n, d, cov, init_eps=init_eps, init_l=init_l, max_leapfrog_steps=30,
                    num_burnin_steps=burnin, num_results=results, adapt=True, target_accept=0.75, seed=seed)
    g = torch.Generator(device="cuda").manual_seed(1)
    z0 = z0_scale * torch.randn((n, d), device="cuda", generator=g) * torch.as_tensor(scales, device="cuda")
    samples, stats = hmc.run(z0.float(), graph=graph)
    return samples, stats


def test_gaussian_reaches_target_acceptance(cuda):
    scales = np.linspace(0.5, 2.0, 10)
    samples, stats = run(scales, np.eye(10))
    acc = float(stats["accept_rate"][-400:].mean())
    eps = stats["step_size"]
    x = samples.reshape(-1, 10).double().cpu().numpy()
    print(f"10-d Gaussian: acceptance {acc:.3f}, final step size {eps[-1]:.3f}, "
          f"T {stats['trajectory_length'][-1]:.2f}", flush=True)
    assert abs(acc - 0.75) < 0.07
    assert np.all(eps[-400:] == eps[-1])  # adaptation stopped: the averaged step size
    assert np.all(np.abs(x.mean(0)) < 0.15 * scales)
    assert np.all(np.abs(x.var(0) / scales ** 2 - 1.0) < 0.15)


def halton2(i):
    f, r = 1.0, 0.0
    while i > 0:
        f *= 0.5
        r += f * (i & 1)
        i >>= 1
    return r


def chees_stable_roots(sig, eps, max_l, n_mc=20000, seed=0):
    """T where E[ChEES gradient in T] crosses zero downwards, for N(0, diag(sig^2))
    sampled with mass matrix diag(sig^2) and leapfrog step eps."""
    rng = np.random.default_rng(seed)
    d = len(sig)
    x0, p0 = rng.standard_normal((n_mc, d)), rng.standard_normal((n_mc, d))
    x, p = x0.copy(), p0.copy()
    h0 = 0.5 * (x0 ** 2 + p0 ** 2).sum(1)
    s0 = (sig ** 2 * x0 ** 2).sum(1)
    per_l = []
    for _ in range(max_l):
        p = p - 0.5 * eps * x
        x = x + eps * p
        p = p - 0.5 * eps * x
        acc = np.minimum(1.0, np.exp(h0 - 0.5 * (x ** 2 + p ** 2).sum(1)))
        ds = (sig ** 2 * x ** 2).sum(1) - s0
        dot = (sig ** 2 * x * p).sum(1)  # z' . velocity, z = sig x, velocity = sig p
        per_l.append((acc * ds * dot).sum() / acc.sum())
    us = [halton2(i) for i in range(2, 258)]

    def grad(T):
        return sum(u * per_l[min(max_l, max(1, math.ceil(u * T / eps))) - 1] for u in us)
    Ts = np.linspace(eps, max_l * eps, 600)
    g = np.array([grad(T) for T in Ts])
    return [0.5 * (Ts[i] + Ts[i + 1]) for i in range(len(Ts) - 1) if g[i] > 0 >= g[i + 1]]


def test_chees_trajectory_length_and_scale(cuda):
    d = 20
    rng = np.random.default_rng(3)
    A = rng.normal(size=(d, d))
    evals = np.linalg.eigvalsh(A @ A.T / d + np.eye(d))
    scales = np.sqrt(evals)
    # start inside the first root's basin: T0 = 0.75 x 2
    _, st1 = run(scales, np.diag(evals), n=1024, burnin=800, results=10, init_eps=0.75, init_l=2)
    T1, eps1 = float(st1["trajectory_length"][-1]), float(st1["step_size"][-1])
    roots = chees_stable_roots(scales, eps1, 30)
    print(f"ChEES whitened Gaussian: averaged T {T1:.3f}, step size {eps1:.3f}, "
          f"expected-gradient stable roots {np.round(roots, 3)}", flush=True)
    assert 1.2 < roots[0] < 2.3
    assert abs(T1 / roots[0] - 1.0) < 0.1
    # scale equivariance: target x 4, step size x 4, same mass matrix.  Powers of two
    # scale the arithmetic exactly except the logarithms and exponentials of the
    # adaptation (1-ulp differences), so the first iterations agree to 1e-3 (a rare
    # flipped accept decision aside) and the adapted T scales by 4 statistically
    s = 4.0
    _, sa = run(scales, np.diag(evals), burnin=40, results=1, init_eps=0.75, init_l=2)
    _, sb = run(s * scales, np.diag(evals), burnin=40, results=1, init_eps=0.75 * s, init_l=2)
    ta, tb = sa["trajectory_length"][:30], sb["trajectory_length"][:30]
    close = np.abs(tb / (s * ta) - 1.0) < 1e-3
    print(f"scale x{s}: {int(close.sum())}/30 early trajectory lengths within 1e-3 of x{s}", flush=True)
    assert close.mean() >= 0.9
    _, st4 = run(s * scales, np.diag(evals), n=1024, burnin=800, results=10, init_eps=0.75 * s, init_l=2,
                 seed=11)
    T4 = float(st4["trajectory_length"][-1])
    print(f"scaled target (x{s}): averaged T {T4:.3f} vs {s * T1:.3f}", flush=True)
    assert abs(T4 / (s * T1) - 1.0) < 0.15


def test_graph_equals_eager(cuda):
    scales = np.linspace(0.5, 2.0, 6)
    se, ste = run(scales, np.eye(6), n=64, burnin=50, results=20, graph=False)
    sg, stg = run(scales, np.eye(6), n=64, burnin=50, results=20, graph=True)
    assert torch.equal(se, sg)
    for k in ("accept_rate", "step_size", "leapfrog_steps", "trajectory_length"):
        assert np.array_equal(ste[k], stg[k])


def test_leapfrog_step_composition_and_reversibility(cuda):
    """gl_leapfrog_step on the 60x60 EPL+shear posterior: equal to the composition of
    its parts through the plan's own log_prob (kick, drift with the mass matrix,
    gradient, kick) and time-reversible: ten steps, momentum negated, ten steps back
    return to the start."""
    from tests import common as C
    from gigalens_amd import ForwardProbModel, LensSimulator
    prob = C.epl_shear_problem(bs=8)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    plan = pm.plan_for(sim)
    xt = torch.as_tensor(prob["truth_x"], dtype=torch.float32)
    x = xt.repeat(8, 1) * (1 + 0.01 * torch.randn(8, xt.shape[1], generator=torch.Generator().manual_seed(2)))
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], x)).to(cuda).contiguous()
    d = z.shape[1]
    g = torch.Generator(device="cuda").manual_seed(4)
    # inverse mass ~ 1e-6 (the posterior's stiffest directions have curvature ~1e7, so
    # eps sqrt(curvature cov) stays well inside the leapfrog's stability range 2)
    A = torch.randn((d, d), device="cuda", generator=g) * 1e-4
    cov = (A @ A.T + 1e-6 * torch.eye(d, device="cuda")).contiguous()
    p = torch.randn((8, d), device="cuda", generator=g) * 1e3
    _, _, grad = plan.log_prob(z, want_grad=True)
    eps = 0.05
    # composition through the plan's log_prob
    p1 = p + 0.5 * eps * grad
    z1 = z + eps * (p1 @ cov)
    lp_ref, _, g1 = plan.log_prob(z1.contiguous(), want_grad=True)
    p2 = p1 + 0.5 * eps * g1
    zl, pl, gl = z.clone(), p.clone(), grad.clone()
    lp, _ = plan.leapfrog_step(zl, pl, gl, eps, cov=cov)
    assert torch.allclose(zl, z1, rtol=1e-6, atol=1e-6)
    assert torch.allclose(lp, lp_ref, rtol=1e-5)
    assert torch.allclose(pl, p2, rtol=1e-4, atol=1e-3)
    # reversibility
    zr, pr, gr = z.clone(), p.clone(), grad.clone()
    for _ in range(10):
        plan.leapfrog_step(zr, pr, gr, eps, cov=cov)
    pr.neg_()
    for _ in range(10):
        plan.leapfrog_step(zr, pr, gr, eps, cov=cov)
    err = float((zr - z).abs().max() / z.abs().max())
    print(f"leapfrog 10 + 10 reversed: max |z - z0| / max |z0| = {err:.2e}", flush=True)
    assert err < 1e-4
    assert torch.allclose(-pr, p, rtol=1e-3, atol=1e-2 * float(p.abs().max()))
