"""Declarative prior distributions and default event-space bijectors.

The reference builds priors from TensorFlow-Probability objects
(tests/conftest.py:20-73, tf-demo.ipynb cell 2) and unconstrains them with
``prior.experimental_default_event_space_bijector()`` (tf/model.py:50).
This module offers the same names and semantics for the four families that
path uses, as small declarative objects the C ABI can evaluate inside the
kernels (each leaf carries a family code and its arguments):

    Normal(loc, scale)                     bijector: Identity
    LogNormal(loc, scale)                  bijector: Exp
    Uniform(low, high)                     bijector: Sigmoid(low, high)
    TruncatedNormal(loc, scale, low, high) bijector: Sigmoid(low, high)

plus ``JointDistributionNamed`` (dict) and ``JointDistributionSequential``
(list).  Flattening follows ``tf.nest``: lists in order, dict keys sorted.
Samples are torch tensors (float32).
"""
from __future__ import annotations

import math
from typing import Any, Dict, List, Optional, Sequence, Tuple

import torch

from . import _native as nat

Nest = Any


def _gen(seed) -> Optional[torch.Generator]:
    if seed is None:
        return None
    g = torch.Generator()
    g.manual_seed(int(seed))
    return g


class Distribution:
    kind: int = 0

    def sample(self, n=None, seed=None):  # pragma: no cover - abstract
        raise NotImplementedError

    def args(self) -> Tuple[float, ...]:  # pragma: no cover - abstract
        raise NotImplementedError


class Normal(Distribution):
    kind = nat.GL_PRIOR_NORMAL

    def __init__(self, loc, scale):
        self.loc, self.scale = float(loc), float(scale)

    def args(self):
        return (self.loc, self.scale)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        shape = () if n is None else (int(n),)
        return self.loc + self.scale * torch.randn(shape, generator=g)

    def log_prob(self, x):
        x = torch.as_tensor(x, dtype=torch.float32)
        r = (x - self.loc) / self.scale
        return -0.5 * r * r - math.log(self.scale) - 0.5 * math.log(2 * math.pi)


class LogNormal(Distribution):
    kind = nat.GL_PRIOR_LOGNORMAL

    def __init__(self, loc, scale):
        self.loc, self.scale = float(loc), float(scale)

    def args(self):
        return (self.loc, self.scale)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        shape = () if n is None else (int(n),)
        return torch.exp(self.loc + self.scale * torch.randn(shape, generator=g))

    def log_prob(self, x):
        x = torch.as_tensor(x, dtype=torch.float32)
        lx = torch.log(x)
        r = (lx - self.loc) / self.scale
        return -0.5 * r * r - math.log(self.scale) - 0.5 * math.log(2 * math.pi) - lx


class Uniform(Distribution):
    kind = nat.GL_PRIOR_UNIFORM

    def __init__(self, low, high):
        self.low, self.high = float(low), float(high)

    def args(self):
        return (self.low, self.high)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        shape = () if n is None else (int(n),)
        return self.low + (self.high - self.low) * torch.rand(shape, generator=g)

    def log_prob(self, x):
        x = torch.as_tensor(x, dtype=torch.float32)
        inside = (x >= self.low) & (x < self.high)
        return torch.where(inside, torch.full_like(x, -math.log(self.high - self.low)),
                           torch.full_like(x, -math.inf))


class TruncatedNormal(Distribution):
    kind = nat.GL_PRIOR_TRUNCNORMAL

    def __init__(self, loc, scale, low, high):
        self.loc, self.scale, self.low, self.high = float(loc), float(scale), float(low), float(high)
        a = (self.low - self.loc) / self.scale / math.sqrt(2)
        b = (self.high - self.loc) / self.scale / math.sqrt(2)
        self.logZ = math.log(0.5 * (math.erf(b) - math.erf(a)))

    def args(self):
        return (self.loc, self.scale, self.low, self.high)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        shape = () if n is None else (int(n),)
        ca = 0.5 * (1 + math.erf((self.low - self.loc) / self.scale / math.sqrt(2)))
        cb = 0.5 * (1 + math.erf((self.high - self.loc) / self.scale / math.sqrt(2)))
        u = ca + (cb - ca) * torch.rand(shape, generator=g, dtype=torch.float64)
        x = self.loc + self.scale * torch.special.ndtri(u)
        return x.clamp(self.low, self.high).to(torch.float32)

    def log_prob(self, x):
        x = torch.as_tensor(x, dtype=torch.float32)
        r = (x - self.loc) / self.scale
        lp = -0.5 * r * r - math.log(self.scale) - 0.5 * math.log(2 * math.pi) - self.logZ
        inside = (x >= self.low) & (x <= self.high)
        return torch.where(inside, lp, torch.full_like(x, -math.inf))


# ---------------------------------------------------------------------------
# Joint distributions and nesting
# ---------------------------------------------------------------------------
class JointDistributionNamed:
    def __init__(self, model: Dict[str, Distribution]):
        self.model = dict(model)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        return {k: self.model[k].sample(n, generator=g) for k in sorted(self.model)}

    def log_prob(self, value):
        return sum(self.model[k].log_prob(value[k]) for k in sorted(self.model))


class JointDistributionSequential:
    def __init__(self, model: Sequence[Any]):
        self.model = list(model)

    def sample(self, n=None, seed=None, generator=None):
        g = generator or _gen(seed)
        return [m.sample(n, generator=g) for m in self.model]

    def log_prob(self, value):
        return sum(m.log_prob(v) for m, v in zip(self.model, value))

    def experimental_default_event_space_bijector(self):
        return UnconstrainingBijector(self)


def leaves(prior) -> List[Tuple[Tuple, Distribution]]:
    """Leaves of a nested prior in tf.nest flatten order with their paths."""
    out: List[Tuple[Tuple, Distribution]] = []

    def rec(node, path):
        if isinstance(node, JointDistributionSequential):
            for i, m in enumerate(node.model):
                rec(m, path + (i,))
        elif isinstance(node, JointDistributionNamed):
            for k in sorted(node.model):
                rec(node.model[k], path + (k,))
        elif isinstance(node, Distribution):
            out.append((path, node))
        else:
            raise TypeError(f"unsupported prior node {type(node)}")
    rec(prior, ())
    return out


def flatten(nest) -> List[Any]:
    """tf.nest.flatten: lists/tuples in order, dict keys sorted."""
    out = []

    def rec(n):
        if isinstance(n, dict):
            for k in sorted(n):
                rec(n[k])
        elif isinstance(n, (list, tuple)):
            for v in n:
                rec(v)
        else:
            out.append(n)
    rec(nest)
    return out


def pack_as(prior, cols: Sequence[Any]):
    """Rebuild the nested structure of a prior sample from flat leaves."""
    it = iter(cols)

    def rec(node):
        if isinstance(node, JointDistributionSequential):
            return [rec(m) for m in node.model]
        if isinstance(node, JointDistributionNamed):
            return {k: rec(node.model[k]) for k in sorted(node.model)}
        return next(it)
    return rec(prior)


# ---------------------------------------------------------------------------
# Bijectors (TFP semantics, used on the host for start points and results;
# the kernels apply the same maps on the device)
# ---------------------------------------------------------------------------
def _leaf_forward(d: Distribution, z: torch.Tensor) -> torch.Tensor:
    if d.kind == nat.GL_PRIOR_NORMAL:
        return z
    if d.kind == nat.GL_PRIOR_LOGNORMAL:
        return torch.exp(z)
    lo, hi = d.args()[-2], d.args()[-1]
    return hi * torch.sigmoid(z) + lo * torch.sigmoid(-z)


def _leaf_inverse(d: Distribution, x: torch.Tensor) -> torch.Tensor:
    if d.kind == nat.GL_PRIOR_NORMAL:
        return x
    if d.kind == nat.GL_PRIOR_LOGNORMAL:
        return torch.log(x)
    lo, hi = d.args()[-2], d.args()[-1]
    s = (x - lo) / (hi - lo)
    return torch.log(s) - torch.log1p(-s)


def _leaf_fldj(d: Distribution, z: torch.Tensor) -> torch.Tensor:
    if d.kind == nat.GL_PRIOR_NORMAL:
        return torch.zeros_like(z)
    if d.kind == nat.GL_PRIOR_LOGNORMAL:
        return z
    lo, hi = d.args()[-2], d.args()[-1]
    return math.log(hi - lo) - torch.nn.functional.softplus(-z) - torch.nn.functional.softplus(z)


class UnconstrainingBijector:
    """Structured default event-space bijector of a joint prior: maps nested
    unconstrained values to nested physical values (tf/model.py:50)."""

    def __init__(self, prior):
        self.prior = prior
        self._leaves = leaves(prior)

    def forward(self, nest):
        flat = flatten(nest)
        return pack_as(self.prior, [_leaf_forward(d, torch.as_tensor(v, dtype=torch.float32))
                                    for (_, d), v in zip(self._leaves, flat)])

    def inverse(self, nest):
        flat = flatten(nest)
        return pack_as(self.prior, [_leaf_inverse(d, torch.as_tensor(v, dtype=torch.float32))
                                    for (_, d), v in zip(self._leaves, flat)])

    def forward_log_det_jacobian(self, nest):
        flat = flatten(nest)
        return sum(_leaf_fldj(d, torch.as_tensor(v, dtype=torch.float32)) for (_, d), v in zip(self._leaves, flat))


class PackBijector:
    """(bs, d) tensor <-> nested structure (tf/model.py:42-49)."""

    def __init__(self, prior):
        self.prior = prior
        self.size = len(leaves(prior))

    def forward(self, z):
        z = torch.as_tensor(z)
        return pack_as(self.prior, [z[..., j] for j in range(self.size)])

    def inverse(self, nest):
        flat = [torch.as_tensor(v, dtype=torch.float32) for v in flatten(nest)]
        return torch.stack(flat, dim=-1)


class ChainBijector:
    """unconstraining_bij o pack_bij: flat unconstrained z <-> nested physical x."""

    def __init__(self, unconstraining, pack):
        self.unconstraining = unconstraining
        self.pack = pack

    def forward(self, z):
        return self.unconstraining.forward(self.pack.forward(z))

    def inverse(self, x):
        return self.pack.inverse(self.unconstraining.inverse(x))


def prior_table(prior):
    """(kinds[d], args[d*4]) for the C ABI, in flatten order."""
    kinds, args = [], []
    for _, d in leaves(prior):
        kinds.append(d.kind)
        a = list(d.args()) + [0.0] * (4 - len(d.args()))
        args.extend(a)
    return kinds, args
