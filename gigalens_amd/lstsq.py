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
    coefficients (bs, D) when ``return_coeffs``, else the image (squeezed).
    ``no_deflection`` evaluates the sources on the undeflected grid
    (tf/simulator.py:116-117)."""
    obs_t = torch.as_tensor(observed_image, dtype=torch.float32)
    err_t = torch.as_tensor(err_map, dtype=torch.float32)
    # the plan holds the observed image and noise map: reuse it while the caller passes
    # the same tensors (compared on the device, no host copy), rebuild it otherwise
    cache = sim.__dict__.setdefault("_lstsq_plans", {})
    key = bool(no_deflection)
    hit = cache.get(key)
    if hit is not None and hit[0].shape == obs_t.shape and hit[1].shape == err_t.shape and \
            torch.equal(hit[0], obs_t.to(hit[0].device)) and torch.equal(hit[1], err_t.to(hit[1].device)):
        plan = hit[2]
    else:
        if no_deflection:
            phys, lay, _ = sim.undeflected()
        else:
            phys, lay = sim.phys_model, sim.layout
        plan = Plan(phys, sim.camera, lay, observed=obs_t.cpu().numpy(), noise_map=err_t.cpu().numpy(),
                    max_batch=getattr(sim, "max_batch", 0))
        cache[key] = (obs_t.detach().clone(), err_t.detach().clone(), plan)
    x = params if torch.is_tensor(params) else sim.params_to_x(params)
    if no_deflection:
        x = x[:, sim.undeflected()[2]]
    img, coef, st = plan.lstsq_simulate(x.contiguous(), want_image=not (return_stacked or return_coeffs),
                                        want_coeffs=return_coeffs and not return_stacked,
                                        want_stacked=return_stacked)
    if return_stacked:
        return st
    if return_coeffs:
        return coef
    return img.squeeze()
