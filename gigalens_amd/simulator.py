"""Simulator API (mirrors ``gigalens.simulator`` and ``gigalens.tf.simulator``).

``LensSimulator(phys_model, sim_config, bs)`` keeps the reference's
constructor and methods (src/gigalens/tf/simulator.py:10-165); the work is
done by the HIP kernels behind ``gigalens_amd.plan.Plan``, reached through the
``gigalens::simulate`` operator (ops.py), whose backward is the adjoint kernel.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Optional

import numpy as np
import torch

from . import _native as nat
from . import ops
from . import psf as _psf
from .plan import Plan, components, layout_from_model, model_param_names


@dataclass
class SimulatorConfig:
    """``gigalens.simulator.SimulatorConfig`` (src/gigalens/simulator.py:11-28)."""

    delta_pix: float
    num_pix: int
    supersample: Optional[int] = 1
    kernel: Optional[Any] = None
    transform_pix2angle: Optional[np.ndarray] = None


class Camera:
    """Grid and PSF state derived from a SimulatorConfig
    (src/gigalens/tf/simulator.py:18-55, src/gigalens/simulator.py:91-125).

    The sub-pixel (row i, column j) sits at ``origin + T_sub @ (j, i)`` with
    ``T_sub = T / ss`` and ``origin = T_sub @ (lo, lo)``, ``lo = -(ss N - 1)/2``;
    the kernels evaluate these coordinates on the fly.
    """

    def __init__(self, cfg: SimulatorConfig):
        self.supersample = int(cfg.supersample or 1)
        self.num_pix = int(cfg.num_pix)
        T = (np.eye(2, dtype=np.float32) * np.float32(cfg.delta_pix) if cfg.transform_pix2angle is None
             else np.asarray(cfg.transform_pix2angle, dtype=np.float32))
        self.conversion_factor = float(np.float32(np.linalg.det(T.astype(np.float64))))
        self.transform_sub = (T / np.float32(self.supersample)).astype(np.float32)
        n = self.supersample * self.num_pix
        lo = np.arange(0, n, dtype=np.float32)
        lo = float(np.min(lo - np.mean(lo)))
        o = self.transform_sub @ np.array([lo, lo], dtype=np.float32)
        self.origin = (float(o[0]), float(o[1]))
        self.subgrid_kernel = None
        if cfg.kernel is not None:
            self.subgrid_kernel = np.asarray(
                _psf.subgrid_kernel(np.asarray(cfg.kernel, dtype=np.float64), self.supersample, odd=True),
                dtype=np.float32)

    def coords(self):
        """(x, y) sub-pixel grids, float32 (get_coords, simulator.py:91-125)."""
        n = self.supersample * self.num_pix
        jj, ii = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64))
        M = self.transform_sub.astype(np.float64)
        x = self.origin[0] + M[0, 0] * jj + M[0, 1] * ii
        y = self.origin[1] + M[1, 0] * jj + M[1, 1] * ii
        return x.astype(np.float32), y.astype(np.float32)


def _as_batch(v, bs, device):
    t = torch.as_tensor(v, dtype=torch.float32, device=device).reshape(-1)
    if t.numel() == 1 and bs > 1:
        t = t.expand(bs)
    return t


class LensSimulator:
    """``gigalens.tf.simulator.LensSimulator`` (src/gigalens/tf/simulator.py:10-165)."""

    def __init__(self, phys_model, sim_config: SimulatorConfig, bs: int, device=None):
        self.phys_model = phys_model
        self.sim_config = sim_config
        self.bs = int(bs)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.camera = Camera(sim_config)
        self.supersample = self.camera.supersample
        self.numPix = self.camera.num_pix
        self.conversion_factor = self.camera.conversion_factor
        self.depth = sum(getattr(x, "depth", 1) for x in phys_model.lens_light) + \
            sum(getattr(x, "depth", 1) for x in phys_model.source_light)
        self.layout = layout_from_model(phys_model)
        self._plan = None

    @staticmethod
    def get_coords(supersample, num_pix, transform_pix2angle):
        """``LensSimulatorInterface.get_coords`` (src/gigalens/simulator.py:91-125)."""
        n = supersample * num_pix
        lo = np.arange(0, n, dtype=np.float32)
        lo = float(np.min(lo - np.mean(lo)))
        M = np.asarray(transform_pix2angle, dtype=np.float32)
        ra0, dec0 = (M @ np.array([lo, lo], dtype=np.float32)).astype(np.float64)
        jj, ii = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64))
        M = M.astype(np.float64)
        x = jj * M[0, 0] + ii * M[0, 1] + ra0
        y = jj * M[1, 0] + ii * M[1, 1] + dec0
        return float(ra0), float(dec0), x.astype(np.float32), y.astype(np.float32)

    @property
    def plan(self) -> Plan:
        if self._plan is None:
            self._plan = Plan(self.phys_model, self.camera, self.layout)
        return self._plan

    def params_to_x(self, params) -> torch.Tensor:
        """Nested params (lens list, [lens light list,] source list) -> (bs, d)."""
        if len(self.phys_model.lens_light) > 0:
            groups = (params[0], params[1], params[2])
        else:
            groups = (params[0], [], params[1])
        by_role = {nat.GL_ROLE_LENS: groups[0], nat.GL_ROLE_LENS_LIGHT: groups[1], nat.GL_ROLE_SOURCE: groups[2]}
        bs = 1
        for role, i, key in self.layout:
            bs = max(bs, torch.as_tensor(by_role[role][i][key]).numel())
        cols = [_as_batch(by_role[role][i][key], bs, self.device) for role, i, key in self.layout]
        return torch.stack(cols, dim=1).contiguous()

    def undeflected(self):
        """(plan, light columns) of the lights alone: ``no_deflection`` evaluates the
        sources on the image-plane grid (tf/simulator.py:74-75, 116-117), which is the
        same model without its lenses; the lens parameters take no part."""
        if getattr(self, "_undeflected", None) is None:
            from .model import PhysicalModel
            phys = PhysicalModel([], self.phys_model.lens_light, self.phys_model.source_light)
            lay = layout_from_model(phys)
            col_of = {k: c for c, k in enumerate(self.layout)}
            self._undeflected = (phys, lay, [col_of[k] for k in lay])
        return self._undeflected

    def simulate(self, params, no_deflection=False):
        """Images (bs, N, N), squeezed like the reference (tf/simulator.py:65-97).
        ``no_deflection`` evaluates the sources on the undeflected grid (:74-75)."""
        x = params if torch.is_tensor(params) else self.params_to_x(params)
        if no_deflection:
            phys, lay, cols = self.undeflected()
            if getattr(self, "_plan_nd", None) is None:
                self._plan_nd = Plan(phys, self.camera, lay)
            img = ops.simulate(self._plan_nd.handle, x[:, cols].contiguous())
            return img.squeeze()
        img = ops.simulate(self.plan.handle, x.contiguous())
        return img.squeeze()

    def lstsq_simulate(self, params, observed_image, err_map, return_stacked=False, return_coeffs=False,
                       no_deflection=False):
        from .lstsq import lstsq_simulate
        return lstsq_simulate(self, params, observed_image, err_map, return_stacked=return_stacked,
                              return_coeffs=return_coeffs, no_deflection=no_deflection)
