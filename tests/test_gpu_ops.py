"""The ``gigalens::*`` PyTorch operators on the GPU: torch.library.opcheck (schema,
fake implementation, autograd registration), autograd through the operators equal to
the fused value-and-gradient launch, and the MAP-step operator equal to the plan's
MAP step."""
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(cuda):
    prob = C.epl_shear_problem(bs=6, seed=3)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=6)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    return prob, sim, pm, z


def test_opcheck(setup, cuda):
    prob, sim, pm, z = setup
    x = prob["x"].to(cuda).contiguous().requires_grad_(True)
    torch.library.opcheck(torch.ops.gigalens.simulate.default, (sim.plan.handle, x),
                          test_utils=("test_schema", "test_autograd_registration", "test_faketensor"))
    plan = pm.plan_for(sim)
    zz = z.clone().requires_grad_(True)
    torch.library.opcheck(torch.ops.gigalens.log_prob.default, (plan.handle, zz),
                          test_utils=("test_schema", "test_autograd_registration", "test_faketensor"))


def test_autograd_through_operators(setup, cuda):
    prob, sim, pm, z = setup
    zz = z.clone().requires_grad_(True)
    lp, chi2 = pm.log_prob(sim, zz)
    w = torch.linspace(0.5, 1.5, z.shape[0], device=cuda)
    (lp * w + 3.0 * chi2).sum().backward()
    plan = pm.plan_for(sim)
    lp2, chi22, g = plan.log_prob(z, g_lp=w.contiguous(), g_chi2=torch.full_like(w, 3.0), want_grad=True)
    assert torch.equal(lp.detach(), lp2) and torch.equal(chi2.detach(), chi22)
    assert torch.allclose(zz.grad, g, rtol=1e-6, atol=1e-6 * float(g.abs().max()))
    # simulate's backward is the adjoint kernel
    x = prob["x"].to(cuda).contiguous().requires_grad_(True)
    img = sim.simulate(x)
    gi = torch.randn_like(img)
    (img * gi).sum().backward()
    assert torch.allclose(x.grad, sim.plan.simulate_vjp(x.detach(), gi.contiguous()), rtol=1e-6, atol=1e-6)


def test_map_step_operator(setup, cuda):
    prob, sim, pm, z = setup
    plan = pm.plan_for(sim)
    za, ma, va = z.clone(), torch.zeros_like(z), torch.zeros_like(z)
    zb, mb, vb = z.clone(), torch.zeros_like(z), torch.zeros_like(z)
    for s in (1, 2, 3):
        lp_a, chi_a = torch.ops.gigalens.map_adam_step(plan.handle, za, ma, va, s, 5e-3, 0.9, 0.999, 1e-7)
        lp_b = torch.empty(z.shape[0], device=cuda)
        chi_b = torch.empty_like(lp_b)
        plan.map_adam_step(zb, mb, vb, s, 5e-3, lp=lp_b, chi2=chi_b)
    for a, b in ((za, zb), (ma, mb), (va, vb), (lp_a, lp_b), (chi_a, chi_b)):
        assert torch.equal(a, b)
