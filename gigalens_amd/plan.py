"""Compile a PhysicalModel + camera (+ prior + data) into a device plan.

A plan is the immutable device state of ``LensSimulator.__init__``
(src/gigalens/tf/simulator.py:11-55) and of the probabilistic model's
constants (src/gigalens/tf/model.py:29-51, 109-132): PSF folded with the
pooling box, observed image, prior table, component table.  Calls on a plan
are stream-ordered launches of the HIP kernels; nothing here computes.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as nat

Layout = List[Tuple[int, int, str]]  # column -> (role, component index within role, param name)

_CANON = {
    nat.GL_EPL: ["theta_E", "gamma", "e1", "e2", "center_x", "center_y"],
    nat.GL_SHEAR: ["gamma1", "gamma2"],
    nat.GL_SIS: ["theta_E", "center_x", "center_y"],
    nat.GL_SIE: ["theta_E", "e1", "e2", "center_x", "center_y"],
    nat.GL_TNFW: ["Rs", "alpha_Rs", "r_trunc", "center_x", "center_y"],
    nat.GL_SERSIC: ["R_sersic", "n_sersic", "center_x", "center_y", "Ie"],
    nat.GL_SERSIC_ELLIPSE: ["R_sersic", "n_sersic", "e1", "e2", "center_x", "center_y", "Ie"],
    nat.GL_CORE_SERSIC: ["R_sersic", "n_sersic", "Rb", "alpha", "gamma", "e1", "e2", "center_x", "center_y", "Ie"],
    nat.GL_SHAPELETS: ["beta", "center_x", "center_y"],
}


def components(phys_model):
    """(role, index, profile) for lenses, lens light, source (model.py:21-29)."""
    out = []
    for role, lst in ((nat.GL_ROLE_LENS, phys_model.lenses), (nat.GL_ROLE_LENS_LIGHT, phys_model.lens_light),
                      (nat.GL_ROLE_SOURCE, phys_model.source_light)):
        for i, prof in enumerate(lst):
            out.append((role, i, prof))
    return out


def model_param_names(prof) -> List[str]:
    """Parameter keys a params dict of this profile carries.  Light profiles
    carry ``Ie`` when not solved by least squares (sersic.py:26,66)."""
    k = prof._kind
    names = list(_CANON[k])
    if k in (nat.GL_SERSIC, nat.GL_SERSIC_ELLIPSE, nat.GL_CORE_SERSIC) and getattr(prof, "use_lstsq", False):
        names.remove("Ie")
    if k == nat.GL_SHAPELETS and not prof.use_lstsq:
        names += list(prof._amp_names)
    return names


def layout_from_model(phys_model) -> Layout:
    """Flatten order of a params structure (lens list, lens light, source),
    dict keys sorted (tf.nest order)."""
    lay: Layout = []
    for role, i, prof in components(phys_model):
        for k in sorted(model_param_names(prof)):
            lay.append((role, i, k))
    return lay


def layout_from_paths(phys_model, paths: Sequence[Tuple]) -> Layout:
    """Map prior leaf paths (i, j, key) to (role, component, key).  With no
    lens light the structure is (lens, source) (tf/simulator.py:67-72)."""
    has_ll = len(phys_model.lens_light) > 0
    lay: Layout = []
    for p in paths:
        if len(p) != 3:
            raise ValueError(f"prior leaf path {p} is not (group, component, name)")
        g, j, key = p
        role = g if has_ll else (nat.GL_ROLE_LENS if g == 0 else nat.GL_ROLE_SOURCE)
        lay.append((role, int(j), str(key)))
    return lay


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev(t: torch.Tensor, name: str) -> torch.Tensor:
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    return t


def _devn(t, name: str, numel: int):
    """Optional float32 device tensor of exactly ``numel`` elements (None passes)."""
    if t is None:
        return None
    _dev(t, name)
    if t.numel() != numel:
        raise ValueError(f"{name} must hold {numel} elements, got {t.numel()}")
    return t


def _lens(t: torch.Tensor, bs: int) -> torch.Tensor:
    if not t.is_cuda or t.dtype != torch.int32 or not t.is_contiguous() or t.numel() != bs:
        raise TypeError("lens must be a contiguous int32 CUDA tensor with one entry per particle")
    return t


class Plan:
    def __init__(self, phys_model, camera, layout: Layout, prior=None, observed=None, background_rms=1.0,
                 exp_time=1.0, likelihood=nat.GL_LIK_FORWARD, fldj_at_x=False, epl_batch_series=False,
                 psf_rank_tol=0.0, max_batch=0, noise_map=None, staged=False, conv_rows=0):
        self.layout = layout
        self.d = len(layout)
        self.N = camera.num_pix
        self.camera = camera
        comps = components(phys_model)
        cd = (nat.ComponentDesc * len(comps))()
        col_of = {(r, i, k): c for c, (r, i, k) in enumerate(layout)}
        used = set()
        for n, (role, i, prof) in enumerate(comps):
            c = cd[n]
            c.kind = prof._kind
            c.role = role
            names = _CANON[prof._kind]
            for j in range(nat.GL_MAX_COMPONENT_PARAMS):
                c.param_col[j] = -1
                c.param_const[j] = 0.0
            for j, name in enumerate(names):
                key = (role, i, name)
                if key in col_of:
                    c.param_col[j] = col_of[key]
                    used.add(key)
                elif name == "Ie" and getattr(prof, "use_lstsq", False):
                    c.param_const[j] = 1.0
                else:
                    raise ValueError(f"parameter {name!r} of {prof} (role {role}, index {i}) is not in the layout")
            c.niter = getattr(prof, "niter", 0)
            c.n_max = getattr(prof, "n_max", 0)
            c.interpolate = int(getattr(prof, "interpolate", 0))
            c.use_lstsq = int(getattr(prof, "use_lstsq", False))
            c.amp_col0 = -1
            if prof._kind == nat.GL_SHAPELETS and not prof.use_lstsq:
                # sampled amplitudes amp00.. (shapelets.py:32-36): contiguous columns in basis order
                try:
                    acols = [col_of[(role, i, a)] for a in prof._amp_names]
                except KeyError as e:
                    raise ValueError(f"shapelet amplitude {e.args[0][2]!r} is not in the layout") from None
                if acols != list(range(acols[0], acols[0] + len(acols))):
                    raise ValueError("shapelet amplitudes must occupy contiguous columns in basis order")
                c.amp_col0 = acols[0]
                used.update((role, i, a) for a in prof._amp_names)
        extra = [k for k in col_of if k not in used]
        if extra:
            raise ValueError(f"layout columns not used by the physical model: {extra}")
        desc = nat.PlanDesc()
        desc.num_pix = camera.num_pix
        desc.supersample = camera.supersample
        for j in range(4):
            desc.transform_sub[j] = float(camera.transform_sub.reshape(-1)[j])
        desc.origin[0], desc.origin[1] = camera.origin
        desc.conversion_factor = camera.conversion_factor
        self._kernel = None
        if camera.subgrid_kernel is not None:
            k = np.ascontiguousarray(camera.subgrid_kernel, dtype=np.float32)
            self._kernel = k
            desc.kernel_size = k.shape[0]
            desc.kernel = k.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        else:
            desc.kernel_size = 0
        desc.n_components = len(comps)
        desc.components = cd
        desc.n_params = self.d
        if prior is not None:
            from .distributions import prior_table
            kinds, args = prior_table(prior)
            if len(kinds) != self.d:
                raise ValueError("prior size does not match the parameter layout")
        else:
            kinds, args = [0] * self.d, [0.0] * (4 * self.d)
        self._pk = (ctypes.c_int * self.d)(*kinds)
        self._pa = (ctypes.c_float * (4 * self.d))(*args)
        desc.prior_kind = self._pk
        desc.prior_args = self._pa
        self.n_observed = 1
        if observed is not None:
            obs = np.ascontiguousarray(np.asarray(observed, dtype=np.float32))
            if obs.ndim == 3 and obs.shape[1:] == (self.N, self.N):
                self.n_observed = obs.shape[0]
            elif obs.shape != (self.N, self.N):
                raise ValueError(f"observed image must be {self.N}x{self.N} or (lenses, {self.N}, {self.N})")
            self._obs = obs
            desc.observed = obs.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            desc.n_observed = self.n_observed
        desc.background_rms = float(background_rms)
        desc.exp_time = float(exp_time)
        desc.likelihood = int(likelihood)
        desc.fldj_at_x = int(bool(fldj_at_x))
        desc.epl_batch_series = int(bool(epl_batch_series))
        desc.psf_rank_tol = float(psf_rank_tol)
        desc.max_batch = int(max_batch)
        desc.staged = int(bool(staged))
        desc.conv_rows = int(conv_rows)
        if noise_map is not None:
            nm = np.ascontiguousarray(np.asarray(noise_map, dtype=np.float32))
            if nm.size != self.n_observed * self.N * self.N:
                raise ValueError(f"noise map must hold {self.n_observed} image(s) of {self.N}x{self.N}")
            self._nm = nm
            desc.noise_map = nm.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.depth = sum(getattr(prof, "depth", 1) if getattr(prof, "use_lstsq", False) or prof._kind == nat.GL_SHAPELETS
                         else 1 for role, i, prof in comps if role != nat.GL_ROLE_LENS)
        lib = nat.lib()
        h = ctypes.c_void_p()
        nat.check(lib.gl_plan_create(ctypes.byref(desc), ctypes.byref(h)), "gl_plan_create")
        self._h = h
        self._lib = lib
        from . import ops
        ops.register(self)

    @property
    def handle(self) -> int:
        """The plan's handle: how the ``gigalens::*`` operators (ops.py) refer to it."""
        return int(self._h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.gl_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def info(self) -> dict:
        pi = nat.PlanInfo()
        nat.check(self._lib.gl_plan_query(self._h, ctypes.byref(pi)), "gl_plan_query")
        return {n: getattr(pi, n) for n, _ in pi._fields_}

    # ---- launches --------------------------------------------------------
    def simulate(self, x: torch.Tensor) -> torch.Tensor:
        x = _dev(x, "x")
        bs = x.shape[0]
        img = torch.empty((bs, self.N, self.N), device=x.device, dtype=torch.float32)
        nat.check(self._lib.gl_simulate(self._h, _ptr(x), bs, _ptr(img), _stream()), "gl_simulate")
        return img

    def simulate_vjp(self, x: torch.Tensor, g_img: torch.Tensor) -> torch.Tensor:
        x = _dev(x, "x")
        g_img = _dev(g_img, "g_img")
        grad = torch.empty_like(x)
        nat.check(self._lib.gl_simulate_vjp(self._h, _ptr(x), x.shape[0], _ptr(g_img), _ptr(grad), _stream()),
                  "gl_simulate_vjp")
        return grad

    def log_prob(self, z: torch.Tensor, g_lp=None, g_chi2=None, want_grad=False, lens=None):
        """lp, chi2 (and dlp/dz when ``want_grad``); ``lens`` (bs int32) picks
        each particle's observed image of a multi-lens plan."""
        z = _dev(z, "z")
        bs = z.shape[0]
        if z.dim() != 2 or z.shape[1] != self.d:
            raise ValueError(f"z must be (bs, {self.d})")
        g_lp = _devn(g_lp, "g_lp", bs)
        g_chi2 = _devn(g_chi2, "g_chi2", bs)
        lp = torch.empty(bs, device=z.device, dtype=torch.float32)
        chi2 = torch.empty(bs, device=z.device, dtype=torch.float32)
        grad = torch.empty_like(z) if want_grad else None
        if lens is None:
            nat.check(self._lib.gl_log_prob(self._h, _ptr(z), bs, _ptr(lp), _ptr(chi2), _ptr(g_lp), _ptr(g_chi2),
                                            _ptr(grad), _stream()), "gl_log_prob")
        else:
            lens = _lens(lens, bs)
            nat.check(self._lib.gl_log_prob_lenses(self._h, _ptr(z), _ptr(lens), bs, _ptr(lp), _ptr(chi2), _ptr(g_lp),
                                                   _ptr(g_chi2), _ptr(grad), _stream()), "gl_log_prob_lenses")
        return lp, chi2, grad

    def leapfrog_step(self, z, p, grad, eps: float, cov=None, lp=None, chi2=None):
        """One leapfrog step of preconditioned HMC in place (gl_leapfrog_step): p +=
        eps/2 grad, z += eps cov p, grad = dlp/dz at the new z, p += eps/2 grad.
        ``grad`` must hold dlp/dz at z on entry; ``cov`` (d x d) the inverse mass
        matrix (None: identity).  Returns (lp, chi2) at the new z."""
        z = _dev(z, "z")
        bs = z.shape[0]
        if z.dim() != 2 or z.shape[1] != self.d:
            raise ValueError(f"z must be (bs, {self.d})")
        p = _devn(p, "p", z.numel())
        grad = _devn(grad, "grad", z.numel())
        if p is None or grad is None:
            raise ValueError("p and grad are required")
        cov = _devn(cov, "cov", self.d * self.d)
        lp = _devn(lp, "lp", bs) if lp is not None else torch.empty(bs, device=z.device, dtype=torch.float32)
        chi2 = _devn(chi2, "chi2", bs) if chi2 is not None else torch.empty(bs, device=z.device, dtype=torch.float32)
        nat.check(self._lib.gl_leapfrog_step(self._h, bs, _ptr(cov), float(eps), _ptr(z), _ptr(p), _ptr(lp),
                                             _ptr(chi2), _ptr(grad), _stream()), "gl_leapfrog_step")
        return lp, chi2

    def lstsq_simulate(self, x: torch.Tensor, want_image=True, want_coeffs=False, want_stacked=False):
        """Least-squares simulation (tf/simulator.py:99-165): (image, coeffs, stacked)."""
        x = _dev(x, "x")
        bs = x.shape[0]
        D = self.depth
        img = torch.empty((bs, self.N, self.N), device=x.device) if want_image else None
        coef = torch.empty((bs, D), device=x.device) if want_coeffs else None
        st = torch.empty((bs, self.N, self.N, D), device=x.device) if want_stacked else None
        nat.check(self._lib.gl_lstsq_simulate(self._h, _ptr(x), bs, _ptr(img), _ptr(coef), _ptr(st), _stream()),
                  "gl_lstsq_simulate")
        return img, coef, st

    def map_adam_step(self, z, m, v, step: int, lr: float, beta1=0.9, beta2=0.999, epsilon=1e-7, lp=None,
                      chi2=None, lens=None, weight=None):
        z = _dev(z, "z")
        bs = z.shape[0]
        if z.dim() != 2 or z.shape[1] != self.d:
            raise ValueError(f"z must be (bs, {self.d})")
        m = _devn(m, "m", z.numel())
        v = _devn(v, "v", z.numel())
        if m is None or v is None:
            raise ValueError("m and v (Adam moments) are required")
        lp = _devn(lp, "lp", bs)
        chi2 = _devn(chi2, "chi2", bs)
        if lens is None and weight is None:
            nat.check(self._lib.gl_map_adam_step(self._h, _ptr(z), _ptr(m), _ptr(v), bs, int(step), float(lr),
                                                 float(beta1), float(beta2), float(epsilon), _ptr(lp), _ptr(chi2),
                                                 _stream()), "gl_map_adam_step")
        else:
            lens = None if lens is None else _lens(lens, bs)
            weight = _devn(weight, "weight", bs)
            nat.check(self._lib.gl_map_adam_step_lenses(self._h, _ptr(z), _ptr(m), _ptr(v), _ptr(lens), _ptr(weight),
                                                        bs, int(step), float(lr), float(beta1), float(beta2),
                                                        float(epsilon), _ptr(lp), _ptr(chi2), _stream()),
                      "gl_map_adam_step_lenses")
