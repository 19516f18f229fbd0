"""Torch-CPU op-for-op restatement of the reference TensorFlow path.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Nothing in
``gigalens_amd`` imports this module.

It follows the reference's *execution model* -- unfused elementwise tensor
ops over the full (bs, ssN, ssN) grid, the batch-coupled EPL series length,
``conv2d`` at sub-pixel resolution followed by ``avg_pool2d``, the Gaussian
log-likelihood, prior + bijector terms -- and differentiates it with torch
autograd.  It therefore serves as (1) the stand-in for the reference's
autodiff gradient (TensorFlow is not installed here) and (2) the labelled
CPU baseline.  It runs in float32 (like the reference) or float64.

References (paths relative to the reference repository root):
  grid                 src/gigalens/simulator.py:91-125
  simulate             src/gigalens/tf/simulator.py:57-97
  lstsq_simulate       src/gigalens/tf/simulator.py:99-165
  ForwardProbModel     src/gigalens/tf/model.py:53-86
  BackwardProbModel    src/gigalens/tf/model.py:134-163
  EPL                  src/gigalens/tf/profiles/mass/epl.py:23-68
  Shear                src/gigalens/tf/profiles/mass/shear.py:12-14
  SIS / SIE            src/gigalens/tf/profiles/mass/sis.py:14-18, sie.py:13-43
  SersicEllipse        src/gigalens/tf/profiles/light/sersic.py:33-72
  Shapelets            src/gigalens/tf/profiles/light/shapelets.py:20-85
"""
from __future__ import annotations

import math
from typing import Any, Dict, List, Sequence, Tuple

import numpy as np
import torch
import torch.nn.functional as F

from . import np64


def _b(v, dt):
    return torch.as_tensor(v, dtype=dt).reshape(-1, 1, 1)


def _rotate(x, y, phi):
    c, s = torch.cos(phi), torch.sin(phi)
    return x * c + y * s, -x * s + y * c


def epl_deriv(x, y, theta_E, gamma, e1, e2, center_x, center_y, niter=50):
    phi = torch.atan2(e2, e1) / 2
    c = torch.clamp(torch.sqrt(e1 ** 2 + e2 ** 2), 0, 1)
    q = (1 - c) / (1 + c)
    theta_E_conv = theta_E / torch.sqrt((1.0 + q ** 2) / (2.0 * q))
    b = theta_E_conv * torch.sqrt((1 + q ** 2) / 2)
    t = gamma - 1
    x, y = x - center_x, y - center_y
    x, y = _rotate(x, y, phi)
    R = torch.clamp(torch.sqrt((q * x) ** 2 + y ** 2), 1e-10, 1e10)
    ang = torch.atan2(y, q * x)
    f = (1 - q) / (1 + q)
    Cs, Ss = torch.cos(ang), torch.sin(ang)
    Cs2, Ss2 = torch.cos(2 * ang), torch.sin(2 * ang)
    n_it = np64.epl_series_length(float(torch.max(f).detach()), niter)
    lx, ly, fx, fy = Cs, Ss, Cs, Ss
    for n in range(1, n_it + 1):
        pref = -f * (2 * n - (2 - t)) / (2 * n + (2 - t))
        lx, ly = pref * (Cs2 * lx - Ss2 * ly), pref * (Ss2 * lx + Cs2 * ly)
        fx, fy = fx + lx, fy + ly
    pref = (2 * b) / (1 + q) * torch.pow(b / R, t - 1)
    return _rotate(fx * pref, fy * pref, -phi)


def shear_deriv(x, y, gamma1, gamma2):
    return gamma1 * x + gamma2 * y, gamma2 * x - gamma1 * y


def sis_deriv(x, y, theta_E, center_x, center_y):
    dx, dy = x - center_x, y - center_y
    R = torch.sqrt(dx ** 2 + dy ** 2)
    a = torch.where(R == 0, torch.zeros_like(R), theta_E / R)
    return a * dx, a * dy


def sie_deriv(x, y, theta_E, e1, e2, center_x, center_y):
    phi = torch.atan2(e2, e1) / 2
    c = torch.clamp(torch.sqrt(e1 ** 2 + e2 ** 2), max=0.9999)
    q = (1 - c) / (1 + c)
    b = theta_E / torch.sqrt((1.0 + q ** 2) / (2.0 * q)) * torch.sqrt((1 + q ** 2) / 2)
    x, y = _rotate(x - center_x, y - center_y, phi)
    psi = torch.sqrt(q ** 2 * x ** 2 + y ** 2)
    sq = torch.sqrt(1.0 - q ** 2)
    fx = b / sq * torch.atan(sq * x / psi)
    fy = b / sq * torch.atanh(sq * y / psi)
    return _rotate(fx, fy, -phi)


MASS = {"EPL": epl_deriv, "SHEAR": shear_deriv, "SIS": sis_deriv, "SIE": sie_deriv}


def sersic_distance(x, y, cx, cy, e1=None, e2=None):
    if e1 is None:
        e1 = torch.zeros_like(cx)
    if e2 is None:
        e2 = torch.zeros_like(cx)
    phi = torch.atan2(e2, e1) / 2
    c = torch.sqrt(e1 ** 2 + e2 ** 2)
    q = (1 - c) / (1 + c)
    dx, dy = x - cx, y - cy
    cp, sp = torch.cos(phi), torch.sin(phi)
    xt1 = (cp * dx + sp * dy) * torch.sqrt(q)
    xt2 = (-sp * dx + cp * dy) / torch.sqrt(q)
    return torch.sqrt(xt1 ** 2 + xt2 ** 2)


def sersic_ellipse_light(x, y, R_sersic, n_sersic, e1, e2, center_x, center_y, Ie=None):
    R = sersic_distance(x, y, center_x, center_y, e1, e2)
    bn = 1.9992 * n_sersic - 0.3271
    ret = torch.exp(-bn * ((R / R_sersic) ** (1 / n_sersic) - 1.0))
    return ret if Ie is None else Ie * ret


def sersic_light(x, y, R_sersic, n_sersic, center_x, center_y, Ie=None):
    R = sersic_distance(x, y, center_x, center_y)
    bn = 1.9992 * n_sersic - 0.3271
    ret = torch.exp(-bn * ((R / R_sersic) ** (1 / n_sersic) - 1.0))
    return ret if Ie is None else Ie * ret


def _interp(x, table):
    ny = table.shape[-1]
    xi = (ny - 1) * (x - np64.TABLE_LO) / (np64.TABLE_HI - np64.TABLE_LO)
    xc = torch.clamp(xi, 0, ny - 1)
    lo = torch.floor(xc)
    hi = torch.clamp(lo + 1, max=ny - 1)
    lo = torch.clamp(hi - 1, min=0)
    a = xc - lo
    li, hi_ = lo.long(), hi.long()
    out = (1 - a)[None] * table[:, li] + a[None] * table[:, hi_]
    oob = (xi < 0) | (xi > ny - 1)
    return torch.where(oob[None], torch.zeros_like(out), out)


def shapelet_basis(x, y, center_x, center_y, beta, n_max, interpolate=True):
    u = (x - center_x) / beta
    v = (y - center_y) / beta
    orders = np64.shapelet_orders(n_max)
    if interpolate:
        tab = torch.as_tensor(np64.shapelet_table(n_max), dtype=x.dtype)
        hx, hy = _interp(u, tab), _interp(v, tab)
        return torch.stack([hx[a] * hy[b] for a, b in orders])
    N = torch.arange(0, n_max + 1, dtype=x.dtype)
    pref = 1.0 / torch.sqrt(2 ** N * math.sqrt(math.pi) * torch.exp(torch.lgamma(N + 1)))

    def herm(z):
        h = [torch.ones_like(z), 2 * z]
        for i in range(2, n_max + 1):
            h.append(2 * (z * h[-1] - (i - 1) * h[-2]))
        return [pref[i] * h[i] for i in range(n_max + 1)]
    XX, YY = herm(u), herm(v)
    fac = torch.exp(-(u ** 2 + v ** 2) / 2)
    return torch.stack([fac * XX[a] * YY[b] for a, b in orders])


class Sim:
    """``LensSimulator`` (src/gigalens/tf/simulator.py:11-55) for a model given as
    lists of (kind, options) descriptors."""

    def __init__(self, model: Dict[str, List[Tuple[str, Dict]]], cam: np64.Camera, dtype=torch.float32):
        self.model = model
        self.cam = cam
        self.dt = dtype
        self.X = torch.as_tensor(cam.x, dtype=dtype)[None]
        self.Y = torch.as_tensor(cam.y, dtype=dtype)[None]
        self.kernel = None
        if cam.kernel is not None:
            self.kernel = torch.as_tensor(np.ascontiguousarray(cam.kernel[::-1, ::-1]), dtype=dtype)[None, None]
        self.cf = cam.conversion_factor

    def _beta(self, lens_params):
        bx, by = self.X, self.Y
        for (kind, opt), p in zip(self.model["lenses"], lens_params):
            kw = {k: _b(v, self.dt) if not torch.is_tensor(v) else v.reshape(-1, 1, 1) for k, v in p.items()}
            if kind == "EPL":
                ax, ay = epl_deriv(self.X, self.Y, niter=opt.get("niter", 50), **kw)
            else:
                ax, ay = MASS[kind](self.X, self.Y, **kw)
            bx, by = bx - ax, by - ay
        return bx, by

    @staticmethod
    def _kw(p, drop=()):
        return {k: v.reshape(-1, 1, 1) for k, v in p.items() if k not in drop}

    def _light(self, kind, opt, x, y, p, unit=False):
        if kind in ("SERSIC_ELLIPSE", "SERSIC"):
            kw = self._kw(p, drop=("Ie",) if unit else ())
            fn = sersic_ellipse_light if kind == "SERSIC_ELLIPSE" else sersic_light
            out = fn(x, y, **kw)
            return out[None] if unit else out
        if kind == "SHAPELETS":
            B = shapelet_basis(x, y, p["center_x"].reshape(-1, 1, 1), p["center_y"].reshape(-1, 1, 1),
                               p["beta"].reshape(-1, 1, 1), opt["n_max"], opt.get("interpolate", True))
            if unit:
                return B
            amps = torch.stack([p[k] for k in sorted(p) if k.startswith("amp")]).reshape(len(B), -1, 1, 1)
            return (B * amps).sum(0)
        raise KeyError(kind)

    def _split(self, params):
        if len(self.model["lens_light"]) > 0:
            return params[0], params[1], params[2]
        return params[0], [], params[1]

    def _conv_pool(self, img):
        # img (bs*C, 1, H, W)
        if self.kernel is not None:
            k = self.kernel.shape[-1]
            pb = (k - 1) // 2
            img = F.conv2d(F.pad(img, (pb, k - 1 - pb, pb, k - 1 - pb)), self.kernel)
        if self.cam.ss != 1:
            img = F.avg_pool2d(img, self.cam.ss, self.cam.ss)
        return img

    def simulate(self, params, no_deflection=False):
        lp, llp, slp = self._split(params)
        bx, by = self._beta(lp)
        if no_deflection:
            bx, by = self.X, self.Y
        img = torch.zeros((), dtype=self.dt)
        for (kind, opt), p in zip(self.model["lens_light"], llp):
            img = img + self._light(kind, opt, self.X, self.Y, p)
        for (kind, opt), p in zip(self.model["source_light"], slp):
            img = img + self._light(kind, opt, bx, by, p)
        img = torch.where(torch.isnan(img), torch.zeros_like(img), img)
        out = self._conv_pool(img[:, None])[:, 0]
        return out * self.cf

    def lstsq_simulate(self, params, observed_image, err_map, return_stacked=False, return_coeffs=False,
                       no_deflection=False):
        lp, llp, slp = self._split(params)
        bx, by = self._beta(lp)
        if no_deflection:
            bx, by = self.X, self.Y
        comps = []
        for (kind, opt), p in zip(self.model["lens_light"], llp):
            comps.append(self._light(kind, opt, self.X, self.Y, p, unit=True))
        for (kind, opt), p in zip(self.model["source_light"], slp):
            comps.append(self._light(kind, opt, bx, by, p, unit=True))
        bs = bx.shape[0]
        comps = [c.expand(c.shape[0], bs, *c.shape[-2:]) for c in comps]
        img = torch.cat(comps, 0)  # (D, bs, H, W)
        img = torch.where(torch.isnan(img), torch.zeros_like(img), img)
        D = img.shape[0]
        ret = self._conv_pool(img.permute(1, 0, 2, 3).reshape(bs * D, 1, *img.shape[-2:]))
        ret = ret.reshape(bs, D, *ret.shape[-2:]).permute(0, 2, 3, 1)
        ret = torch.where(torch.isnan(ret), torch.zeros_like(ret), ret)
        if return_stacked:
            return ret
        W = (1 / torch.as_tensor(err_map, dtype=self.dt))[..., None]
        Y = (torch.as_tensor(observed_image, dtype=self.dt) * W[..., 0]).reshape(1, -1, 1)
        X = (ret * W).reshape(bs, -1, D)
        Xt = X.transpose(1, 2)
        coeffs = (torch.linalg.pinv(Xt @ X, rtol=1e-6) @ Xt @ Y)[..., 0]
        if return_coeffs:
            return coeffs
        return (ret * coeffs[:, None, None, :]).sum(-1)


# --------------------------------------------------------------------------
# Prior / bijector / log_prob (src/gigalens/tf/model.py)
# --------------------------------------------------------------------------
def flatten_prior(prior) -> List[Tuple[Tuple, str, Tuple[float, ...]]]:
    """tf.nest flatten order of a nested prior: lists in order, dict keys sorted.
    Leaves are (kind, args).  Returns [(path, kind, args)]."""
    out = []

    def rec(node, path):
        if isinstance(node, dict):
            for k in sorted(node):
                rec(node[k], path + (k,))
        elif isinstance(node, (list,)) and not (len(node) == 2 and isinstance(node[0], str)):
            for i, v in enumerate(node):
                rec(v, path + (i,))
        else:
            out.append((path, node[0], tuple(node[1])))
    rec(prior, ())
    return out


def _logpdf(kind, args, x):
    if kind == "Normal":
        loc, sc = args
        return -0.5 * ((x - loc) / sc) ** 2 - math.log(sc) - 0.5 * math.log(2 * math.pi)
    if kind == "LogNormal":
        loc, sc = args
        lx = torch.log(x)
        return -0.5 * ((lx - loc) / sc) ** 2 - math.log(sc) - 0.5 * math.log(2 * math.pi) - lx
    if kind == "Uniform":
        lo, hi = args
        inside = (x >= lo) & (x < hi)
        return torch.where(inside, torch.full_like(x, -math.log(hi - lo)), torch.full_like(x, -math.inf))
    if kind == "TruncatedNormal":
        loc, sc, lo, hi = args
        Z = 0.5 * (math.erf((hi - loc) / sc / math.sqrt(2)) - math.erf((lo - loc) / sc / math.sqrt(2)))
        lp = -0.5 * ((x - loc) / sc) ** 2 - math.log(sc) - 0.5 * math.log(2 * math.pi) - math.log(Z)
        inside = (x >= lo) & (x <= hi)
        return torch.where(inside, lp, torch.full_like(x, -math.inf))
    raise KeyError(kind)


def _bij(kind, args, z):
    if kind == "Normal":
        return z, torch.zeros_like(z)
    if kind == "LogNormal":
        return torch.exp(z), z
    lo, hi = args[-2], args[-1]
    x = lo + (hi - lo) * torch.sigmoid(z)
    fldj = math.log(hi - lo) - F.softplus(-z) - F.softplus(z)
    return x, fldj


def unflatten(prior, cols):
    """Rebuild the nested params structure from per-leaf column tensors."""
    it = iter(cols)

    def rec(node):
        if isinstance(node, dict):
            return {k: rec(node[k]) for k in sorted(node)}
        if isinstance(node, list) and not (len(node) == 2 and isinstance(node[0], str)):
            return [rec(v) for v in node]
        return next(it)
    return rec(prior)


def log_prob(sim: Sim, prior, z: torch.Tensor, observed_image, background_rms, exp_time, backward=False):
    """``ForwardProbModel.log_prob`` (src/gigalens/tf/model.py:53-86) or, with
    ``backward=True``, ``BackwardProbModel.log_prob`` (model.py:134-163,
    including its fldj evaluated at the constrained x, model.py:160)."""
    leaves = flatten_prior(prior)
    xs, fl, lpri = [], 0.0, 0.0
    for j, (path, kind, args) in enumerate(leaves):
        x, f = _bij(kind, args, z[:, j])
        xs.append(x)
        if backward:
            _, f = _bij(kind, args, x)
        fl = fl + f
        lpri = lpri + _logpdf(kind, args, x)
    params = unflatten(prior, xs)
    obs = torch.as_tensor(observed_image, dtype=sim.dt)
    if not backward:
        im = sim.simulate(params)
        err = torch.sqrt(background_rms ** 2 + torch.clamp(im, 0, math.inf) / exp_time)
    else:
        err = torch.sqrt(background_rms ** 2 + torch.clamp(obs, 0, math.inf) / exp_time)
        im = sim.lstsq_simulate(params, obs, err)
    if im.dim() == 2:
        im = im[None]
    r = (obs - im) / err
    ll = (-0.5 * r ** 2 - torch.log(err) - 0.5 * math.log(2 * math.pi)).sum((-2, -1))
    return ll + lpri + fl, (r ** 2).mean((-2, -1))


def params_to_x(prior, params) -> np.ndarray:
    """Nested physical params -> (bs, d) in flatten order."""
    leaves = flatten_prior(prior)
    cols = []
    for path, _, _ in leaves:
        node = params
        for k in path:
            node = node[k]
        cols.append(np.asarray(node, np.float64).reshape(-1))
    return np.stack(cols, 1)


def x_to_z(prior, x: np.ndarray) -> np.ndarray:
    leaves = flatten_prior(prior)
    return np.stack([np64.bijector_inverse(k, a, x[:, j]) for j, (_, k, a) in enumerate(leaves)], 1)
