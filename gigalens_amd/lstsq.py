"""Least-squares simulation, ``LensSimulator.lstsq_simulate``
(src/gigalens/tf/simulator.py:99-165), on the multi-channel HIP pipeline."""
from __future__ import annotations

import numpy as np
import torch

from .plan import Plan


def lstsq_simulate(sim, params, observed_image, err_map, return_stacked=False, return_coeffs=False,
                   no_deflection=False):
    """Unit-amplitude components, weighted least squares for the amplitudes
    (pseudo-inverse, rcond 1e-6), and the resulting image.  Returns the
    stacked components (bs, N, N, D) when ``return_stacked``, else the
    coefficients (bs, D) when ``return_coeffs``, else the image (squeezed)."""
    if no_deflection:
        raise NotImplementedError("no_deflection is not supported by the fused kernels")
    obs = np.asarray(torch.as_tensor(observed_image, dtype=torch.float32).cpu(), dtype=np.float32)
    err = np.asarray(torch.as_tensor(err_map, dtype=torch.float32).cpu(), dtype=np.float32)
    key = (obs.tobytes(), err.tobytes())
    cache = sim.__dict__.setdefault("_lstsq_plans", {})
    if key not in cache:
        cache.clear()
        cache[key] = Plan(sim.phys_model, sim.camera, sim.layout, observed=obs, noise_map=err,
                          max_batch=getattr(sim, "max_batch", 0))
    plan = cache[key]
    x = params if torch.is_tensor(params) else sim.params_to_x(params)
    img, coef, st = plan.lstsq_simulate(x.contiguous(), want_image=not (return_stacked or return_coeffs),
                                        want_coeffs=return_coeffs and not return_stacked,
                                        want_stacked=return_stacked)
    if return_stacked:
        return st
    if return_coeffs:
        return coef
    return img.squeeze()
