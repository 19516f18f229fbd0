"""The autotuner times the equivalent variants of a plan and picks one; the
tuned plan computes the same log posterior (within the FP32 agreement of the
variants: 1e-5 relative) as the default plan."""
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator
from gigalens_amd import autotune

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_variant(cuda, tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "autotune.json"))
    autotune._memory.clear()
    prob = C.epl_shear_problem(bs=256, seed=9)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=256, device=cuda)
    default_plan = pm.plan_for(sim)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    lp0, _, _ = default_plan.log_prob(z)
    pm2 = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    pm2.autotune = True
    plan = pm2.plan_for(LensSimulator(prob["phys"], prob["cfg"], bs=256, device=cuda))
    rec = list(autotune._memory.values())[0]
    print("autotune timings (ms):", rec["timings"], "best:", rec["best"], flush=True)
    assert len(rec["timings"]) == 6  # fused separable / direct, staged conv rows x 3, staged separable
    lp1, _, _ = plan.log_prob(z)
    assert float(((lp1 - lp0).abs() / lp0.abs()).max()) < 1e-5
    # cached: a second tuning of the same key does not time again
    best, timings = autotune.tune(prob["phys"], sim.camera, plan.layout, z, {"prior": prob["prior"]})
    assert best == rec["best"]
