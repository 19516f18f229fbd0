"""The drop-in boundary's contracts on the GPU: plan reuse and eviction across
simulators (the reference builds a new LensSimulator per driver call,
src/gigalens/tf/inference.py:21,45,90), and calls on one plan from several
streams, which the library orders so the plan's scratch is never shared by two
streams at once."""
import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator
from gigalens_amd import distributions as tfd
from gigalens_amd.inference import Adam, ModellingSequence
from gigalens_amd.plan import Plan, layout_from_paths

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(cuda):
    return C.epl_shear_problem(bs=64, seed=3)


def test_plan_cache_reuses_and_evicts(prob, cuda):
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    s1 = LensSimulator(prob["phys"], prob["cfg"], bs=16)
    s2 = LensSimulator(prob["phys"], prob["cfg"], bs=16)
    assert pm.plan_for(s1) is pm.plan_for(s2)
    for bs in (8, 4, 2):
        pm.plan_for(LensSimulator(prob["phys"], prob["cfg"], bs=bs))
    assert len(pm._plans) == pm.plan_cache_size
    seq = ModellingSequence(prob["phys"], pm, prob["cfg"])
    z0 = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"][:8]))
    for _ in range(3):
        seq.MAP(Adam(1e-3), start=z0, num_steps=2)
    assert len(pm._plans) <= pm.plan_cache_size


@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
def test_calls_on_two_streams_are_ordered(prob, cuda, staged):
    paths = [p for p, _ in tfd.leaves(prob["prior"])]
    plan = Plan(prob["phys"], prob["cam"], layout_from_paths(prob["phys"], paths), prior=prob["prior"],
                observed=prob["obs"], background_rms=0.2, exp_time=100.0, staged=staged, max_batch=64)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    za, zb = z[:32].contiguous(), z[32:].contiguous()
    ref_a = plan.log_prob(za, want_grad=True)
    ref_b = plan.log_prob(zb, want_grad=True)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(4):
        with torch.cuda.stream(sa):
            oa = plan.log_prob(za, want_grad=True)
        with torch.cuda.stream(sb):
            ob = plan.log_prob(zb, want_grad=True)
        outs.append((oa, ob))
    torch.cuda.synchronize()
    for oa, ob in outs:
        for got, ref in ((oa, ref_a), (ob, ref_b)):
            for g, r in zip(got, ref):
                # staged statistics meet in atomic adds: the last bits vary from run to run
                assert float((g - r).abs().max()) <= 2e-6 * float(r.abs().max())
