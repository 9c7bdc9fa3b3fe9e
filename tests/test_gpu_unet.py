"""U-Net training and inference kernels (BF16 MFMA) against the CPU oracle.

The oracle (oracle/unet_oracle.py, rounding=True) rounds to BF16 where the
device stores BF16 and accumulates in float64 where the device accumulates
in FP32, so the comparison tolerances below cover FP32 accumulation order
and the occasional one-ulp BF16 rounding flip that it causes.
"""
import numpy as np
import pytest

import unet_oracle as uo

pytestmark = pytest.mark.gpu


def _case(seed, n, batch):
    rng = np.random.default_rng(seed)
    params = uo.init_params(seed).astype(np.float32)
    inp = rng.standard_normal((batch, 2, n ** 3)).astype(np.float32)
    target = (0.1 * rng.standard_normal((batch, n ** 3))).astype(np.float32)
    return params, inp, target


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


# rows of up to 64 points take the x-accumulator kernel (16-point tiles; 12, 20 and 40 leave
# the last tile part empty), 68 the per-tap kernel on its first level
@pytest.mark.parametrize("n,batch", [(16, 2), (12, 3), (4, 1), (20, 2), (40, 1), (64, 1), (68, 1)])
def test_forward_matches_oracle(cfm, gpu, n, batch):
    params, inp, _ = _case(10 + n, n, batch)
    with cfm.UNet(n, batch, params=params) as net:
        out = net.forward(inp)
    ref, _ = uo.forward(params.astype(np.float64), inp, n, rounding=True)
    assert out.shape == (batch, n, n, n)
    assert np.all(np.isfinite(out))
    # both forward kernels measured 0.17 %, 0.59 % and 0.52 % at n = 20, 40 and 64 (the BF16
    # rounding flips grow with the grid), well under 1 % for either
    assert _rel(out.reshape(batch, -1), ref) < 1e-2


# 64 and 32: the training size's grids, where every level takes the all-taps weight-gradient
# kernel (whole x rows per 64-point block) and the taller forward tiles
@pytest.mark.parametrize("n,batch", [(16, 2), (12, 3), (8, 1), (32, 1), (64, 1)])
def test_gradient_matches_oracle(cfm, gpu, n, batch):
    params, inp, target = _case(20 + n, n, batch)
    with cfm.UNet(n, batch, params=params) as net:
        loss, g = net.gradient(inp, target)
    rl, rg, _ = uo.loss_and_grad(params.astype(np.float64), inp, target, n, rounding=True)
    assert loss == pytest.approx(rl, rel=5e-3)
    assert _rel(g, rg) < 1e-2
    # per layer: weights and biases separately
    off = 0
    for li, (ci, co) in enumerate(uo.CONVS):
        for name, cnt in (("W", co * ci * 27), ("b", co)):
            e = _rel(g[off: off + cnt], rg[off: off + cnt])
            assert e < 3e-2, f"conv {li} {name}: rel err {e:.3g}"
            off += cnt
    assert _rel(g[off:], rg[off:]) < 1e-2


def test_gradient_is_deterministic(cfm, gpu):
    params, inp, target = _case(31, 16, 2)
    with cfm.UNet(16, 2, params=params) as net:
        l1, g1 = net.gradient(inp, target)
        l2, g2 = net.gradient(inp, target)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_training_reduces_loss(cfm, gpu):
    n, batch = 16, 2
    rng = np.random.default_rng(40)
    inp = rng.standard_normal((batch, 2, n ** 3)).astype(np.float32)
    # a smooth target the network can fit: a blurred copy of channel 0
    f = inp[:, 0].reshape(batch, n, n, n).astype(np.float64)
    for ax in (1, 2, 3):
        f = (np.roll(f, 1, ax) + f + np.roll(f, -1, ax)) / 3.0
    target = (0.5 * f).reshape(batch, -1).astype(np.float32)
    with cfm.UNet(n, batch, seed=3) as net:
        first = net.train_step(inp, target, 2e-3)
        for _ in range(60):
            last = net.train_step(inp, target, 2e-3)
        p = net.params()
    assert np.all(np.isfinite(p))
    assert last < 0.5 * first


def test_params_roundtrip(cfm, gpu):
    p = cfm.unet_init_params(9)
    with cfm.UNet(8, 1, params=p) as net:
        assert np.array_equal(net.params(), p)
