"""The hot path as PyTorch custom operators (namespace ``gigalens``).

Each operator takes a device plan by its handle (an int: ``Plan.handle``) and the
caller's tensors, and launches the HIP kernels through the C ABI on torch's current
stream.  Registered with ``torch.library`` so the dispatcher sees them: fake
(shape-only) implementations for tracing, and autograd formulas whose backward is
the hand-written adjoint kernel (it recomputes the ray-traced coordinates instead of
storing per-particle grids).

  gigalens::simulate(plan, x) -> img                       LensSimulator.simulate
                                                          (src/gigalens/tf/simulator.py:65-97)
  gigalens::simulate_vjp(plan, x, g_img) -> grad_x         its reverse mode
  gigalens::log_prob(plan, z) -> (lp, chi2)                ForwardProbModel / BackwardProbModel
                                                          .log_prob (src/gigalens/tf/model.py:53-86,
                                                          134-163)
  gigalens::log_prob_grad(plan, z, g_lp, g_chi2) -> (lp, chi2, grad_z)
  gigalens::map_adam_step(plan, z!, m!, v!, step, lr, beta1, beta2, epsilon) -> (lp, chi2)
                                                          one MAP train_step with Keras Adam
                                                          (src/gigalens/tf/inference.py:28-35)
"""
from __future__ import annotations

import weakref
from typing import Optional, Tuple

import torch
from torch import Tensor

_PLANS: "weakref.WeakValueDictionary[int, object]" = weakref.WeakValueDictionary()


def register(plan) -> int:
    """Make a plan reachable from the operators by its handle."""
    _PLANS[plan.handle] = plan
    return plan.handle


def _plan(handle: int):
    plan = _PLANS.get(int(handle))
    if plan is None:
        raise RuntimeError(f"gigalens: no live plan with handle {handle:#x}")
    return plan


@torch.library.custom_op("gigalens::simulate", mutates_args=())
def simulate(plan: int, x: Tensor) -> Tensor:
    return _plan(plan).simulate(x)


@simulate.register_fake
def _(plan: int, x: Tensor) -> Tensor:
    N = _plan(plan).N
    return x.new_empty((x.shape[0], N, N))


@torch.library.custom_op("gigalens::simulate_vjp", mutates_args=())
def simulate_vjp(plan: int, x: Tensor, g_img: Tensor) -> Tensor:
    return _plan(plan).simulate_vjp(x, g_img.contiguous())


@simulate_vjp.register_fake
def _(plan: int, x: Tensor, g_img: Tensor) -> Tensor:
    return torch.empty_like(x)


def _simulate_setup(ctx, inputs, output):
    plan, x = inputs
    ctx.plan = plan
    ctx.save_for_backward(x)


def _simulate_backward(ctx, g_img):
    (x,) = ctx.saved_tensors
    return None, simulate_vjp(ctx.plan, x, g_img)


simulate.register_autograd(_simulate_backward, setup_context=_simulate_setup)


@torch.library.custom_op("gigalens::log_prob", mutates_args=())
def log_prob(plan: int, z: Tensor) -> Tuple[Tensor, Tensor]:
    lp, chi2, _ = _plan(plan).log_prob(z)
    return lp, chi2


@log_prob.register_fake
def _(plan: int, z: Tensor) -> Tuple[Tensor, Tensor]:
    return z.new_empty(z.shape[0]), z.new_empty(z.shape[0])


@torch.library.custom_op("gigalens::log_prob_grad", mutates_args=())
def log_prob_grad(plan: int, z: Tensor, g_lp: Tensor, g_chi2: Optional[Tensor]) -> Tuple[Tensor, Tensor, Tensor]:
    return _plan(plan).log_prob(z, g_lp=g_lp.contiguous(), g_chi2=None if g_chi2 is None else g_chi2.contiguous(),
                                want_grad=True)


@log_prob_grad.register_fake
def _(plan: int, z: Tensor, g_lp: Tensor, g_chi2: Optional[Tensor]) -> Tuple[Tensor, Tensor, Tensor]:
    return z.new_empty(z.shape[0]), z.new_empty(z.shape[0]), torch.empty_like(z)


def _log_prob_setup(ctx, inputs, output):
    plan, z = inputs
    ctx.plan = plan
    ctx.save_for_backward(z)


def _log_prob_backward(ctx, g_lp, g_chi2):
    (z,) = ctx.saved_tensors
    if g_lp is None:
        g_lp = torch.zeros(z.shape[0], device=z.device, dtype=z.dtype)
    _, _, grad = log_prob_grad(ctx.plan, z, g_lp, g_chi2)
    return None, grad


log_prob.register_autograd(_log_prob_backward, setup_context=_log_prob_setup)


@torch.library.custom_op("gigalens::map_adam_step", mutates_args=("z", "m", "v"))
def map_adam_step(plan: int, z: Tensor, m: Tensor, v: Tensor, step: int, lr: float, beta1: float, beta2: float,
                  epsilon: float) -> Tuple[Tensor, Tensor]:
    lp = torch.empty(z.shape[0], device=z.device, dtype=torch.float32)
    chi2 = torch.empty_like(lp)
    _plan(plan).map_adam_step(z, m, v, step, lr, beta1, beta2, epsilon, lp=lp, chi2=chi2)
    return lp, chi2


@map_adam_step.register_fake
def _(plan: int, z: Tensor, m: Tensor, v: Tensor, step: int, lr: float, beta1: float, beta2: float,
      epsilon: float) -> Tuple[Tensor, Tensor]:
    return z.new_empty(z.shape[0]), z.new_empty(z.shape[0])
