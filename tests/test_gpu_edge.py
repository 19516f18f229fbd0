"""Edge cases of the fused path on the GPU: a single particle, a ragged batch
large enough for the longest-first launch order (bitwise equal to the same
particles in unordered chunks: every particle's arithmetic is independent of
its workgroup), a NaN particle that stays isolated, and error returns for
empty batches and null pointers."""
import ctypes

import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel
from gigalens_amd import _native as nat
from gigalens_amd import distributions as tfd
from gigalens_amd.plan import Plan, layout_from_paths

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(cuda):
    prob = C.epl_shear_problem(bs=2003, seed=11)
    prior, phys = prob["prior"], prob["phys"]
    plan = Plan(phys, prob["cam"], layout_from_paths(phys, [p for p, _ in tfd.leaves(prior)]), prior=prior,
                observed=prob["obs"], background_rms=0.2, exp_time=100.0)
    pm = ForwardProbModel(prior, prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prior, prob["x"])).to(cuda).contiguous()
    return prob, plan, z


def _chunked(plan, z, sizes):
    lps, gs = [], []
    o = 0
    for n in sizes:
        lp, _, g = plan.log_prob(z[o:o + n].contiguous(), want_grad=True)
        lps.append(lp)
        gs.append(g)
        o += n
    return torch.cat(lps), torch.cat(gs)


def test_ragged_batch_with_launch_order_matches_chunks(setup):
    prob, plan, z = setup
    assert plan.info()["lds_bytes"] > 0  # fused kernel
    lp, _, g = plan.log_prob(z, want_grad=True)  # 2003 > 1024: ordered launch
    lp_c, g_c = _chunked(plan, z, [1000, 1000, 3])  # <= 1024 each: unordered launches
    assert torch.equal(lp, lp_c)
    assert torch.equal(g, g_c)
    lp1, _, g1 = plan.log_prob(z[5:6].contiguous(), want_grad=True)  # bs = 1
    assert torch.equal(lp1, lp[5:6]) and torch.equal(g1, g[5:6])


def test_nan_particle_is_isolated(setup):
    prob, plan, z = setup
    zn = z.clone()
    zn[17, 3] = float("nan")
    lp, chi2, g = plan.log_prob(zn, want_grad=True)
    lp0, chi0, g0 = plan.log_prob(z, want_grad=True)
    assert torch.isnan(lp[17])
    keep = torch.ones(z.shape[0], dtype=torch.bool, device=z.device)
    keep[17] = False
    assert torch.equal(lp[keep], lp0[keep]) and torch.equal(chi2[keep], chi0[keep])
    assert torch.equal(g[keep], g0[keep])


def test_error_returns(setup):
    prob, plan, z = setup
    L = nat.lib()
    lp = torch.empty(4, device=z.device)
    chi = torch.empty(4, device=z.device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    zp = ctypes.c_void_p(z.data_ptr())
    rc = L.gl_log_prob(plan._h, zp, 0, ctypes.c_void_p(lp.data_ptr()), ctypes.c_void_p(chi.data_ptr()), None, None,
                       None, s)
    assert rc == nat.GL_ERR_ARG and "bs" in L.gl_last_error().decode()
    rc = L.gl_log_prob(plan._h, None, 4, ctypes.c_void_p(lp.data_ptr()), ctypes.c_void_p(chi.data_ptr()), None, None,
                       None, s)
    assert rc == nat.GL_ERR_ARG
    rc = L.gl_simulate(plan._h, zp, 4, None, s)
    assert rc == nat.GL_ERR_ARG
    torch.cuda.synchronize()
