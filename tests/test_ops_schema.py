"""The hot path's PyTorch operators are registered with their schemas (no GPU
needed: nothing is launched)."""
import torch

import gigalens_amd  # noqa: F401  (registers the operators)
from gigalens_amd import ops  # noqa: F401


def test_operators_registered():
    want = {
        "simulate": "gigalens::simulate(SymInt plan, Tensor x) -> Tensor",
        "simulate_vjp": "gigalens::simulate_vjp(SymInt plan, Tensor x, Tensor g_img) -> Tensor",
        "log_prob": "gigalens::log_prob(SymInt plan, Tensor z) -> (Tensor, Tensor)",
        "log_prob_grad": "gigalens::log_prob_grad(SymInt plan, Tensor z, Tensor g_lp, Tensor? g_chi2) -> "
                         "(Tensor, Tensor, Tensor)",
    }
    for name, schema in want.items():
        assert str(getattr(torch.ops.gigalens, name).default._schema) == schema
    s = str(torch.ops.gigalens.map_adam_step.default._schema)
    assert "Tensor(a1!) z" in s and "Tensor(a2!) m" in s and "Tensor(a3!) v" in s
