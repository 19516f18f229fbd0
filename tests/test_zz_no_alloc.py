"""gl_map_adam_step_noalloc (the MAP step whose scratch is a static device array)
against gl_map_adam_step, through the C ABI: two calls in a row, the second with a
smaller batch, both above 1,024 particles so the launch-order scratch is used, and a
plan with the batch-coupled EPL series (its series counts also live in the scratch).
Bar: bit-identical z, m, v, lp and chi2 (the same kernels; the launch order does not
change any particle's arithmetic)."""
import ctypes

import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel
from gigalens_amd import _native as nat
from gigalens_amd import distributions as tfd
from gigalens_amd.plan import Plan, _ptr, _stream, layout_from_paths

pytestmark = pytest.mark.gpu


def _plan(prob, epl_batch_series):
    paths = [p for p, _ in tfd.leaves(prob["prior"])]
    return Plan(prob["phys"], prob["cam"], layout_from_paths(prob["phys"], paths), prior=prob["prior"],
                observed=prob["obs"], background_rms=0.2, exp_time=100.0, epl_batch_series=epl_batch_series)


def _step(lib, fn, plan, z, m, v, step, lp, chi2):
    rc = getattr(lib, fn)(plan._h, _ptr(z), _ptr(m), _ptr(v), z.shape[0], step, 5e-3, 0.9, 0.999, 1e-7, _ptr(lp),
                          _ptr(chi2), _stream())
    nat.check(rc, fn)


@pytest.mark.parametrize("epl_batch_series", [False, True])
def test_noalloc_map_step_matches(cuda, epl_batch_series):
    prob = C.epl_shear_problem(bs=2000, seed=3)
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z_all = pm.bij.inverse(C.nested_from_x(prob["prior"], prob["x"])).to(cuda).contiguous()
    plan = _plan(prob, epl_batch_series)
    assert plan.info()["lds_bytes"] > 0
    lib = nat.lib()
    for bs in (2000, 1200):
        print(f"noalloc MAP step, bs {bs}, epl_batch_series {epl_batch_series}", flush=True)
        outs = []
        for fn in ("gl_map_adam_step", "gl_map_adam_step_noalloc"):
            z = z_all[:bs].clone()
            m, v = torch.zeros_like(z), torch.zeros_like(z)
            lp = torch.empty(bs, device=cuda)
            chi2 = torch.empty(bs, device=cuda)
            for step in (1, 2):
                _step(lib, fn, plan, z, m, v, step, lp, chi2)
            torch.cuda.synchronize()
            outs.append([t.cpu() for t in (z, m, v, lp, chi2)])
        for a, b in zip(*outs):
            assert torch.equal(torch.nan_to_num(a, nan=123.0), torch.nan_to_num(b, nan=123.0))


def test_noalloc_refuses_what_it_cannot_hold(cuda):
    prob = C.epl_shear_problem(bs=2, seed=3)
    plan = _plan(prob, False)
    lib = nat.lib()
    z = torch.zeros((2, 22), device=cuda)
    with pytest.raises(nat.GlError, match="exceeds the static scratch"):
        nat.check(lib.gl_map_adam_step_noalloc(plan._h, _ptr(z), _ptr(z), _ptr(z), 16385, 1, 1e-3, 0.9, 0.999, 1e-7,
                                               None, None, _stream()), "gl_map_adam_step_noalloc")
    paths = [p for p, _ in tfd.leaves(prob["prior"])]
    staged = Plan(prob["phys"], prob["cam"], layout_from_paths(prob["phys"], paths), prior=prob["prior"],
                  observed=prob["obs"], background_rms=0.2, exp_time=100.0, staged=True)
    with pytest.raises(nat.GlError, match="resident-image plans"):
        nat.check(lib.gl_map_adam_step_noalloc(staged._h, _ptr(z), _ptr(z), _ptr(z), 2, 1, 1e-3, 0.9, 0.999, 1e-7,
                                               None, None, _stream()), "gl_map_adam_step_noalloc")
