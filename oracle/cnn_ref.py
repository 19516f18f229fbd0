"""Torch-CPU restatement of the learned initialiser's network (the oracle of
``gl_cnn_*``).

TEST INFRASTRUCTURE ONLY.  Nothing in ``gigalens_amd`` imports this module.

The network has no counterpart in the reference.  This restatement uses
torch's own ``conv2d`` (cross-correlation, padding 1), ``gelu(approximate=
'tanh')``, ``mean`` and ``linear``, and applies bfloat16 rounding at the same
points as the HIP kernels:

  * the input asinh(img / in_scale) and the convolution weights,
  * Z = conv + bias (its gradient dZ is rounded too) and Y = gelu(Z) (forward
    value only; the gradient passes unchanged),

with FP32 arithmetic everywhere else.  Gradients come from torch autograd.

Pinning: ``tests/test_cnn_oracle.py`` checks the unrounded restatement in
float64 against central finite differences, checks the two rounding
functions against a bit-level round-to-nearest-even on uint32 patterns, and
checks the convolution convention (weight layout [co][(kh*3+kw)*cin + ci],
output size (H-1)//s + 1) against an explicit loop on a small input.
"""
from __future__ import annotations

from typing import Dict, Sequence

import numpy as np
import torch
import torch.nn.functional as F

STRIDES = (1, 2, 1, 2)


def bf16(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(x.dtype)


class _RoundBoth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return bf16(g)


class _RoundForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return g


def unpack(p: torch.Tensor, layout) -> Dict[str, torch.Tensor]:
    return {name: p[off:off + int(np.prod(shape))].reshape(shape) for name, shape, off in layout}


def forward(params: Dict[str, torch.Tensor], images: torch.Tensor, in_scale: float, channels: Sequence[int],
            rounding: bool = True) -> torch.Tensor:
    """(bs, 2 d) network outputs for images (bs, H, W)."""
    rb = _RoundBoth.apply if rounding else (lambda t: t)
    rf = _RoundForward.apply if rounding else (lambda t: t)
    x = torch.asinh(images / in_scale)[:, None]
    if rounding:
        x = bf16(x)
    cin = 1
    for i, c in enumerate(channels):
        w = params[f"conv{i}.w"]
        if rounding:
            w = _RoundForward.apply(w)
        w4 = w.reshape(c, 3, 3, cin).permute(0, 3, 1, 2)
        z = F.conv2d(x, w4, params[f"conv{i}.b"], stride=STRIDES[i], padding=1)
        z = rb(z)
        x = rf(F.gelu(z, approximate="tanh"))
        cin = c
    h0 = x.mean(dim=(2, 3))
    h1 = F.gelu(F.linear(h0, params["fc1.w"], params["fc1.b"]), approximate="tanh")
    return F.linear(h1, params["fc2.w"], params["fc2.b"])


def nll(out: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean over the batch of sum_j [ls + 0.5 (t - mu)^2 exp(-2 ls)], ls = clamp(raw, -7, 5)."""
    d = target.shape[1]
    mu, ls = out[:, :d], torch.clamp(out[:, d:], -7.0, 5.0)
    return (ls + 0.5 * (target - mu) ** 2 * torch.exp(-2.0 * ls)).sum(1).mean()


def loss_and_grad(p: torch.Tensor, layout, images: torch.Tensor, target: torch.Tensor, in_scale: float,
                  channels: Sequence[int], rounding: bool = True):
    p = p.detach().clone().requires_grad_(True)
    loss = nll(forward(unpack(p, layout), images, in_scale, channels, rounding), target)
    (g,) = torch.autograd.grad(loss, [p])
    return loss.detach(), g


def conv_loop(x: np.ndarray, w: np.ndarray, b: np.ndarray, stride: int) -> np.ndarray:
    """Explicit-loop 3x3 cross-correlation with zero padding 1 on NHWC input
    (bs, H, W, cin) and weights (cout, 9 cin) in the [(kh*3+kw)*cin + ci]
    order: the convention the kernels implement."""
    bs, H, W, cin = x.shape
    cout = w.shape[0]
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    out = np.zeros((bs, Ho, Wo, cout))
    for ho in range(Ho):
        for wo in range(Wo):
            acc = np.tile(b, (bs, 1)).astype(np.float64)
            for kh in range(3):
                for kw in range(3):
                    hi, wi = ho * stride + kh - 1, wo * stride + kw - 1
                    if 0 <= hi < H and 0 <= wi < W:
                        acc += x[:, hi, wi, :] @ w[:, (kh * 3 + kw) * cin:(kh * 3 + kw + 1) * cin].T
            out[:, ho, wo, :] = acc
    return out
