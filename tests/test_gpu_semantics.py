"""GPU parity of reference semantics at their edges, against the FP64 oracle.

* The reference's batch-coupled EPL series length (src/gigalens/tf/profiles/mass/epl.py:41):
  one loop count for the whole batch from the largest f; one particle with |e| >= 1
  (f = 1, log f = 0) makes the count -inf, so every particle keeps only the first term.
* The NaN -> 0 pixel policy (src/gigalens/tf/simulator.py:81): a light profile with
  |e| > 1 has q < 0 and sqrt(q) = NaN everywhere; the summed sub-pixels (lens light +
  source) are zeroed before the convolution, so the whole image is zero.  A light
  centred exactly on a sub-pixel node (R = 0 there) stays finite.
  Each particle's result and gradient is its own: degenerate particles do not change
  the others'.

Bars: images within 1e-4 of the image peak, log_prob within 1e-4 relative, gradients
within 2e-3 of each component plus 2e-4 of the largest (Richardson-extrapolated FP64
central differences of the oracle).
"""
import numpy as np
import pytest
import torch

from tests import common as C
from gigalens_amd import ForwardProbModel, LensSimulator
from gigalens_amd import distributions as tfd
from gigalens_amd.plan import Plan, layout_from_paths

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _col(prior, path):
    paths = [p for p, _ in tfd.leaves(prior)]
    return paths.index(path)


@pytest.fixture(scope="module")
def prob(cuda):
    return C.epl_shear_problem(bs=8, seed=4)


def _batch_plan(prob, epl_batch_series):
    paths = [p for p, _ in tfd.leaves(prob["prior"])]
    return Plan(prob["phys"], prob["cam"], layout_from_paths(prob["phys"], paths), prior=prob["prior"],
                observed=prob["obs"], background_rms=0.2, exp_time=100.0, epl_batch_series=epl_batch_series)


@pytest.mark.parametrize("degenerate", [False, True], ids=["prior", "one_e_ge_1"])
def test_epl_batch_coupled_series(prob, cuda, degenerate):
    x = prob["x"].clone()
    if degenerate:
        x[2, _col(prob["prior"], (0, 0, "e1"))] = 1.2
        x[2, _col(prob["prior"], (0, 0, "e2"))] = 0.0
    xd = x.numpy().astype(np.float64)
    nested = C.nested_from_x(prob["prior"], xd)
    plan = _batch_plan(prob, True)
    img = plan.simulate(x.to(cuda).contiguous()).cpu().numpy().astype(np.float64)
    ref = C.oracle_simulate(prob["ocam"], prob["phys"], nested, per_particle_series=False)
    err = np.abs(img - ref).max(axis=(1, 2)) / np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-30)
    _print(f"batch-coupled series ({'degenerate' if degenerate else 'prior'}): image err / peak {err.max():.2e}")
    assert np.all(err < 1e-4)
    if degenerate:
        # b = 0 at |e| >= 1 and gamma < 2: 0 * 0^(gamma-2) is NaN, the whole image is zeroed
        if float(x[2, _col(prob["prior"], (0, 0, "gamma"))]) < 2.0:
            assert np.all(ref[2] == 0.0) and np.all(img[2] == 0.0)
        # the quirk is visible: the truncated series differs from the per-particle one
        per = C.oracle_simulate(prob["ocam"], prob["phys"], nested, per_particle_series=True)
        keep = [i for i in range(8) if i != 2]
        assert np.abs(per[keep] - ref[keep]).max() > 1e-2 * np.abs(ref).max()
    # log_prob through the prob model's batch-coupled plan
    pm = ForwardProbModel(prob["prior"], prob["obs"], 0.2, 100.0)
    z = pm.bij.inverse(C.nested_from_x(prob["prior"], x)).contiguous()
    lp, chi2, _ = plan.log_prob(z.to(cuda))
    lp_ref, chi_ref, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prob["prior"], z.numpy().astype(np.float64),
                                           prob["obs"], 0.2, 100.0, per_particle_series=False)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    _print(f"  log_prob rel err {rel.max():.2e}")
    assert rel.max() < 1e-4


def test_nan_pixels_and_isolation(prob, cuda):
    """Particle 1: source |e| > 1 (NaN source light -> zeroed).  Particle 3: lens light
    |e| > 1.  Particles 5 and 6: lens-light centre exactly on a sub-pixel node, with
    n = 0.55 (I is C^1 in the position there) and n = 4 (a cusp: the one-sided
    derivatives are infinite; the kernels take the symmetric derivative, 0, for that
    sub-pixel, and central differences of the likelihood around the cusp are not a
    gradient, so the two centre coordinates of particle 6 are not compared)."""
    prior = prob["prior"]
    x = prob["x"].clone()
    x[1, _col(prior, (2, 0, "e1"))] = 0.9
    x[1, _col(prior, (2, 0, "e2"))] = 0.8
    x[3, _col(prior, (1, 0, "e1"))] = -1.1
    gx, gy = prob["cam"].coords()
    for i, n, (r, c) in ((5, 0.55, (57, 63)), (6, 4.0, (61, 58))):
        x[i, _col(prior, (1, 0, "center_x"))] = float(gx[r, c])
        x[i, _col(prior, (1, 0, "center_y"))] = float(gy[r, c])
        x[i, _col(prior, (1, 0, "n_sersic"))] = n
    # the degenerate source (outside its truncated-normal prior's support) goes through
    # simulate only; log_prob below takes the particles inside the prior's support
    xd = x.numpy().astype(np.float64)
    sim = LensSimulator(prob["phys"], prob["cfg"], bs=8)
    img = sim.simulate(x.to(cuda).contiguous()).cpu().numpy().astype(np.float64)
    ref = C.oracle_simulate(prob["ocam"], prob["phys"], C.nested_from_x(prior, xd))
    assert np.isfinite(img).all()
    err = np.abs(img - ref).max(axis=(1, 2)) / np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-30)
    _print(f"NaN-pixel particles: image err / peak {err.max():.2e}")
    assert np.all(err < 1e-4)
    # NaN -> 0 acts on the summed sub-pixel (lens light + source, tf/simulator.py:77-81):
    # a NaN source or lens light zeroes the whole image
    assert np.all(img[1] == 0.0) and np.all(img[3] == 0.0) and np.all(ref[1] == 0.0)

    # log_prob and gradient of the particles inside the prior's support
    pm = ForwardProbModel(prior, prob["obs"], 0.2, 100.0)
    pick = [0, 3, 5, 6]
    z = pm.bij.inverse(C.nested_from_x(prior, x[pick])).contiguous()
    sim4 = LensSimulator(prob["phys"], prob["cfg"], bs=len(pick))
    lp, _, g = pm.log_prob_and_grad(sim4, z.to(cuda))
    zz = z.numpy().astype(np.float64)
    lp_ref, _, _ = C.oracle_log_prob(prob["ocam"], prob["phys"], prior, zz, prob["obs"], 0.2, 100.0)
    rel = np.abs(lp.cpu().numpy() - lp_ref) / np.abs(lp_ref)
    _print(f"  log_prob rel err {rel.max():.2e}")
    assert rel.max() < 1e-4
    g = g.cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all()
    g_fd = C.oracle_grad_richardson(prob["ocam"], prob["phys"], prior, zz, prob["obs"], 0.2, 100.0)
    cusp = [_col(prior, (1, 0, "center_x")), _col(prior, (1, 0, "center_y"))]
    for k, i in enumerate(pick):
        cols = [j for j in range(zz.shape[1]) if not (i == 6 and j in cusp)]
        scale = np.abs(g_fd[k, cols]).max()
        e = np.abs(g[k, cols] - g_fd[k, cols])
        _print(f"  particle {i}: gradient max |g - fd| / max|fd| = {(e / scale).max():.2e}")
        assert np.all(e <= 2e-3 * np.abs(g_fd[k, cols]) + 2e-4 * scale)
    # isolation: the healthy particle's result does not depend on its neighbours
    lp1, _, g1 = pm.log_prob_and_grad(LensSimulator(prob["phys"], prob["cfg"], bs=1), z[:1].to(cuda).contiguous())
    assert float(lp1[0]) == float(lp[0]) and np.array_equal(g1.cpu().numpy()[0], g[0].astype(np.float32))
