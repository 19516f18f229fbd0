"""CPU pins of the learned initialiser's oracle (oracle/cnn_ref.py) and of the
host-side parameter layout."""
import numpy as np
import pytest
import torch

from oracle import cnn_ref
from gigalens_amd.learned import ConvNet


def _params(layout, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    n = layout[-1][2] + int(np.prod(layout[-1][1]))
    return torch.as_tensor(rng.normal(0, scale, size=n))


def test_layout_sizes():
    lay = ConvNet.param_layout((32, 32, 64, 64), 256, 22)
    n = lay[-1][2] + int(np.prod(lay[-1][1]))
    expect = (32 * 9 + 32) + (32 * 288 + 32) + (64 * 288 + 64) + (64 * 576 + 64) + (256 * 64 + 256) + (44 * 256 + 44)
    assert n == expect
    offs = [o for _, _, o in lay]
    assert offs == sorted(offs) and offs[0] == 0


@pytest.mark.parametrize("stride", [1, 2])
@pytest.mark.parametrize("H", [7, 8])
def test_conv_convention_matches_loop(stride, H):
    rng = np.random.default_rng(H * 10 + stride)
    cin, cout = 3, 5
    x = rng.normal(size=(2, H, H + 1, cin))
    w = rng.normal(size=(cout, 9 * cin))
    b = rng.normal(size=cout)
    ref = cnn_ref.conv_loop(x, w, b, stride)
    w4 = torch.as_tensor(w).reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    out = torch.nn.functional.conv2d(torch.as_tensor(x).permute(0, 3, 1, 2), w4, torch.as_tensor(b), stride=stride,
                                     padding=1).permute(0, 2, 3, 1).numpy()
    assert out.shape == ref.shape == (2, (H - 1) // stride + 1, (H + 1 - 1) // stride + 1, cout)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(0)
    f = (rng.normal(size=100000) * 10.0 ** rng.integers(-20, 20, size=100000)).astype(np.float32)
    # exact ties: low 16 bits 0x8000, both parities of bit 16
    ties = (rng.integers(0, 2 ** 31, size=1000, dtype=np.uint64).astype(np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x8000)
    f = np.concatenate([f, ties.view(np.float32)])
    f = f[np.isfinite(f)]
    u = f.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    got = cnn_ref.bf16(torch.as_tensor(f)).numpy()
    np.testing.assert_array_equal(got, r)


def test_oracle_gradient_matches_finite_differences():
    torch.manual_seed(0)
    ch = (4, 4, 8, 8)
    lay = ConvNet.param_layout(ch, 16, 3)
    p = _params(lay, 1)
    img = torch.as_tensor(np.random.default_rng(2).normal(0, 1.0, size=(2, 9, 9)))
    tgt = torch.as_tensor(np.random.default_rng(3).normal(size=(2, 3)))
    loss, g = cnn_ref.loss_and_grad(p, lay, img, tgt, 0.5, ch, rounding=False)

    def f(pp):
        return float(cnn_ref.nll(cnn_ref.forward(cnn_ref.unpack(pp, lay), img, 0.5, ch, rounding=False), tgt))
    rng = np.random.default_rng(4)
    idx = rng.choice(p.numel(), size=40, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.clone(), p.clone()
        pp[i] += h
        pm[i] -= h
        fd = (f(pp) - f(pm)) / (2 * h)
        assert abs(fd - float(g[i])) <= 1e-6 + 1e-5 * abs(fd), (i, fd, float(g[i]))


def test_rounded_oracle_close_to_unrounded():
    """The bf16 rounding points perturb the loss at the bf16 level only."""
    ch = (32, 32, 64, 64)
    lay = ConvNet.param_layout(ch, 64, 5)
    p = _params(lay, 5, scale=0.05).float()
    img = torch.as_tensor(np.random.default_rng(6).normal(0, 2.0, size=(2, 20, 20))).float()
    tgt = torch.as_tensor(np.random.default_rng(7).normal(size=(2, 5))).float()
    l0, g0 = cnn_ref.loss_and_grad(p, lay, img, tgt, 0.2, ch, rounding=False)
    l1, g1 = cnn_ref.loss_and_grad(p, lay, img, tgt, 0.2, ch, rounding=True)
    assert abs(float(l1 - l0)) <= 2e-2 * abs(float(l0))
    assert float((g1 - g0).abs().max()) <= 5e-2 * float(g0.abs().max())
