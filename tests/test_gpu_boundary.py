"""The drop-in boundary's contracts on the GPU: plan reuse and eviction across
simulators (the reference builds a new LensSimulator per driver call,
src/gigalens/tf/inference.py:21,45,90), and calls on one plan from several
streams, which the library orders so the plan's scratch is never shared by two
streams at once."""
import numpy as np
import pytest
import torch

from oracle import np64
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


def test_simulate_no_deflection(prob, cuda):
    """simulate(params, no_deflection=True) (src/gigalens/tf/simulator.py:74-75): the
    sources on the undeflected grid; image within 1e-4 of the peak of the FP64 oracle,
    and through autograd the lens parameters get exactly zero gradient."""
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    x = prob["x"][:8].to(cuda).contiguous().requires_grad_(True)
    img = sim.simulate(x, no_deflection=True)
    lenses, ll, src = C.oracle_model_lists(prob["phys"], C.nested_from_x(prob["prior"],
                                                                          prob["x"][:8].numpy().astype(np.float64)))
    ref = np64.simulate(prob["ocam"], lenses, ll, src, no_deflection=True)
    got = img.detach().cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    print(f"no_deflection simulate: err / peak {err.max():.2e}", flush=True)
    assert np.all(err < 1e-4)
    deflected = sim.simulate(x.detach()).cpu().numpy()
    assert np.abs(deflected - got).max() > 1e-2 * np.abs(got).max()
    img.sum().backward()
    lens_cols = [c for c, (role, _, _) in enumerate(sim.layout) if role == 0]
    assert torch.all(x.grad[:, lens_cols] == 0) and torch.all(torch.isfinite(x.grad))
    assert float(x.grad.abs().max()) > 0


def test_lstsq_simulate_no_deflection(cuda):
    """lstsq_simulate(..., no_deflection=True) (tf/simulator.py:116-117) against the
    oracle's least squares on the undeflected grid: components within 1e-4 of the peak,
    image within 1e-4 of the peak."""
    from tests.test_gpu_lstsq import setup
    S = setup(30, 2, 0.08, 11, 4, True, bs=3)
    sim = S["sim"]
    x = S["x"].to(cuda).contiguous()
    st = sim.lstsq_simulate(x, S["obs"], S["err"], return_stacked=True, no_deflection=True).cpu().numpy()
    img = sim.lstsq_simulate(x, S["obs"], S["err"], no_deflection=True).cpu().numpy()
    lenses, ll, src = C.oracle_lstsq_lists(S["phys"], C.nested_from_x(S["prior"], S["x"].numpy().astype(np.float64)))
    st_ref, img_ref = [], []
    for i in range(3):
        def pick(d):
            return {k: (v if k == "niter" else np.atleast_1d(np.asarray(v, np.float64))[i:i + 1]) for k, v in d.items()}
        L = [(k, pick(p)) for k, p in lenses]
        LL = [(it[0], pick(it[1])) + tuple(it[2:]) for it in ll]
        SS = [(it[0], pick(it[1])) + tuple(it[2:]) for it in src]
        stacked = np64.lstsq_components(S["ocam"], L, LL, SS, no_deflection=True)
        im, _ = np64.lstsq_solve(stacked, S["obs"], S["err"])
        st_ref.append(stacked[0])
        img_ref.append(im[0])
    st_ref, img_ref = np.stack(st_ref), np.stack(img_ref)
    e_st = np.abs(st - st_ref).max() / np.abs(st_ref).max()
    e_im = np.abs(img - img_ref).max() / np.abs(img_ref).max()
    print(f"no_deflection lstsq: components err {e_st:.2e}, image err {e_im:.2e}", flush=True)
    assert e_st < 1e-4 and e_im < 1e-4
