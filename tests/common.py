"""Shared helpers of the tests: workloads in both the product's and the
oracle's vocabulary."""
import math
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

import numpy as np
import torch

from oracle import np64  # noqa: F401
from oracle.bridge import (KIND_NAME, nested_from_x, oracle_backward_log_prob, oracle_camera,  # noqa: F401
                           oracle_grad_richardson, oracle_log_prob, oracle_lstsq, oracle_lstsq_lists,
                           oracle_model_lists, oracle_nested_prior, oracle_prior, oracle_simulate)
from gigalens_amd import distributions as tfd
from gigalens_amd import workloads as wl
from gigalens_amd.simulator import Camera


def draw_prior(prior, n, seed):
    s = prior.sample(n, seed=seed)
    return torch.stack([torch.as_tensor(v, dtype=torch.float32) for v in tfd.flatten(s)], dim=1)


def epl_shear_problem(bs=16, seed=2, obs_seed=1):
    """60x60 EPL+shear workload: prior, model, config, observed image."""
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    cfg = wl.epl_shear_config()
    cam = Camera(cfg)
    ocam = oracle_camera(cfg, cam)
    truth_x = draw_prior(prior, 1, seed=0).numpy().astype(np.float64)
    clean = oracle_simulate(ocam, phys, nested_from_x(prior, truth_x))[0]
    obs = wl.noisy_observation(clean, 0.2, 100.0, seed=obs_seed)
    x = draw_prior(prior, bs, seed=seed)
    return dict(prior=prior, phys=phys, cfg=cfg, cam=cam, ocam=ocam, obs=obs, x=x, truth_x=truth_x)


