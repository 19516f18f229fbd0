"""Autotuner: picks the fastest of the equivalent execution variants of a plan,
and the path and tile shape of a lens batch.

Every variant computes the same model (results agree to FP32 rounding):

* the LDS-resident fused kernel with the separable (low-rank) PSF path
  (``psf_rank_tol=0``) or the direct 2-D path (``psf_rank_tol=-1``), when the
  image fits in LDS;
* the HBM-staged pipeline with 2, 4 or 8 output rows per lane in its channel
  convolution (``staged=True, conv_rows=...``), or with its separable
  channel convolution (``staged=True, psf_rank_tol=1e-6``).

``tune`` builds each candidate plan, warms it up and times ``log_prob`` with its
gradient on a batch of the caller's z; the fastest configuration is kept per
key (image size, supersampling, PSF size, profile kinds, likelihood, batch size)
in memory and in ``~/.cache/gigalens_amd/autotune.json``.

``tune_batch`` does the same for a ``LensBatch``: every lens on the banded
persistent kernel with bands of 64, 32 or 16 sub-rows (the band is the tile of
the supersampled image a workgroup holds in LDS; taller bands mean fewer
barriers and passes, shorter ones a smaller LDS footprint), only the lenses the
resident-image kernel cannot hold on it, or none (resident-image and staged
launches); the key is the batch's lens shapes and particle counts.  No
counterpart in the reference (TensorFlow picks its own kernels).
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, List, Optional, Tuple

import torch

from . import _native as nat
from .plan import Plan

CACHE_PATH = os.path.join(os.path.expanduser("~"), ".cache", "gigalens_amd", "autotune.json")
_memory: Dict[str, dict] = {}


def _key(phys_model, camera, likelihood, bs) -> str:
    kinds = []
    for lst in (phys_model.lenses, phys_model.lens_light, phys_model.source_light):
        kinds.append(",".join(str(p._kind) + ("l" if getattr(p, "use_lstsq", False) else "") for p in lst))
    k = 0 if camera.subgrid_kernel is None else camera.subgrid_kernel.shape[0]
    return f"N{camera.num_pix}-ss{camera.supersample}-k{k}-{'|'.join(kinds)}-lik{int(likelihood)}-bs{int(bs)}"


def _load() -> Dict[str, dict]:
    try:
        with open(CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, rec: dict) -> None:
    data = _load()
    data[key] = rec
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        with open(CACHE_PATH, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    except OSError:
        pass


def candidates(fused_fits: bool, lstsq: bool) -> List[dict]:
    c: List[dict] = []
    if fused_fits and not lstsq:
        c += [{"psf_rank_tol": 0.0}, {"psf_rank_tol": -1.0}]
    c += [{"staged": True, "conv_rows": r} for r in (2, 4, 8)]
    c += [{"staged": True, "psf_rank_tol": 1e-6}]
    return c


def _time(plan: Plan, z: torch.Tensor, repeats: int) -> float:
    plan.log_prob(z, want_grad=True)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(repeats):
        plan.log_prob(z, want_grad=True)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / repeats


def tune(phys_model, camera, layout, z: torch.Tensor, plan_kwargs: dict, repeats: int = 5,
         use_cache: bool = True) -> Tuple[dict, Dict[str, float]]:
    """Fastest variant for ``Plan(phys_model, camera, layout, **plan_kwargs, **variant)``
    timed on log_prob + gradient of z (bs, d).  Returns (variant kwargs, timings ms)."""
    lik = plan_kwargs.get("likelihood", nat.GL_LIK_FORWARD)
    key = _key(phys_model, camera, lik, z.shape[0])
    if use_cache:
        rec = _memory.get(key) or _load().get(key)
        if rec is not None:
            _memory[key] = rec
            return dict(rec["best"]), dict(rec["timings"])
    base = Plan(phys_model, camera, layout, **plan_kwargs)
    fused_fits = base.info()["lds_bytes"] > 0
    lstsq = any(getattr(p, "use_lstsq", False) for p in (*phys_model.lens_light, *phys_model.source_light))
    del base
    timings: Dict[str, float] = {}
    best: Optional[dict] = None
    best_ms = float("inf")
    for cand in candidates(fused_fits, lstsq):
        name = json.dumps(cand, sort_keys=True)
        try:
            plan = Plan(phys_model, camera, layout, **plan_kwargs, **cand)
        except nat.GlError:
            continue
        ms = _time(plan, z, repeats)
        del plan
        timings[name] = ms
        if ms < best_ms:
            best, best_ms = cand, ms
    rec = {"best": best, "timings": timings}
    _memory[key] = rec
    if use_cache:
        _store(key, rec)
    return dict(best), timings


def batch_candidates() -> List[dict]:
    """Batch configurations ``tune_batch`` times (``gl_batch_opts`` fields)."""
    return [{"band": 1, "band_rows": 0}, {"band": 1, "band_rows": 32}, {"band": 1, "band_rows": 16},
            {"band": 2, "band_rows": 0}, {"band": 0, "band_rows": 0}]


def _batch_key(batch) -> str:
    import hashlib
    h = hashlib.sha1()
    for plan, n in zip(batch.plans, batch.counts):
        i = plan.info()
        h.update(f"{i['num_pix']},{i['supersample']},{i['kernel_size']},{n};".encode())
    return f"batch-{batch.n_lenses}-{batch.total}-{h.hexdigest()[:16]}"


def tune_batch(batch, z: torch.Tensor, repeats: int = 2, use_cache: bool = True) -> Tuple[dict, Dict[str, float]]:
    """Fastest configuration of ``batch`` (a ``LensBatch``) for log_prob + gradient of
    z (total, d); the batch is left in that configuration.  Returns (configuration,
    timings ms by configuration)."""
    key = _batch_key(batch)
    if use_cache:
        rec = _memory.get(key) or _load().get(key)
        if rec is not None:
            _memory[key] = rec
            batch.configure(**rec["best"])
            return dict(rec["best"]), dict(rec["timings"])
    timings: Dict[str, float] = {}
    best: Optional[dict] = None
    best_ms = float("inf")
    out = None
    for cand in batch_candidates():
        try:
            batch.configure(**cand)
        except nat.GlError:
            continue
        if out is None:
            out = batch.log_prob(z, want_grad=True)
        batch.log_prob(z, want_grad=True, out=out)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(repeats):
            batch.log_prob(z, want_grad=True, out=out)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / repeats
        timings[json.dumps(cand, sort_keys=True)] = ms
        if ms < best_ms:
            best, best_ms = cand, ms
    batch.configure(**best)
    rec = {"best": best, "timings": timings}
    _memory[key] = rec
    if use_cache:
        _store(key, rec)
    return dict(best), timings
