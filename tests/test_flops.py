"""The operation counts behind bench.py's roofline figures (gigalens_amd/flops.py):
the least-squares pipeline's channel and adjoint convolutions are counted in the
form the plan runs.  CPU only."""
import numpy as np

from gigalens_amd import flops as fl
from gigalens_amd import workloads as wl


def test_lstsq_conv_counted_in_the_form_the_plan_runs():
    phys = wl.shapelet_model(10, interpolate=True)
    N, ss, k = 100, 4, 83
    kf, nsub, P = k + ss - 1, ss * N, N * N
    n = np.full(3, 20.0)
    direct = fl.lstsq_per_particle(phys, N, ss, k, n, grad=True)
    sep = fl.lstsq_per_particle(phys, N, ss, k, n, grad=True, sep_rank=7, adj_rank=7)
    D = direct["channels"]
    assert D == 67
    # direct: 2 kf^2 per output and channel; separable: 2 kf R (ss N N + N N)
    assert direct["conv_flops"] == D * P * 2.0 * kf * kf
    assert sep["conv_flops"] == D * 7 * 2.0 * kf * (nsub * N + P)
    # the adjoint: direct 2 kf^2 per pixel, separable 2 ceil(kf/ss) R ((ss N) N + (ss N)^2)
    mt = -(-kf // ss)
    adj_direct, adj_sep = P * 2.0 * kf * kf, 7 * 2.0 * mt * (nsub * N + nsub * nsub)
    diff = (direct["flops"] - sep["flops"])[0]
    assert np.isclose(diff, (direct["conv_flops"] - sep["conv_flops"]) + (adj_direct - adj_sep), rtol=1e-12)
    # everything else is the same work
    assert np.array_equal(direct["mfma"], sep["mfma"]) and np.array_equal(direct["fp64"], sep["fp64"])
    assert np.array_equal(direct["transcendentals"], sep["transcendentals"])
    # forward-only counts leave the adjoint out
    fwd = fl.lstsq_per_particle(phys, N, ss, k, n, grad=False, sep_rank=7, adj_rank=7)
    assert fwd["flops"][0] < sep["flops"][0]
