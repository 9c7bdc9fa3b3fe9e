"""Pins of the U-Net oracle (oracle/unet_oracle.py) on the CPU.

The float64 oracle is checked against independent routes to the same
numbers: brute-force loops, torch's conv3d / pooling / interpolation, torch
autograd for the whole network, central differences, and the adjoint
identity of the convolution.  The BF16 rounding helper is checked against
hand-worked ties and torch's bfloat16 conversion.
"""
import numpy as np
import pytest

import unet_oracle as uo


@pytest.fixture(scope="module")
def tt():
    """torch and torch.nn.functional, imported only when these CPU tests run
    (so that a GPU test process never loads the HIP runtime torch bundles)."""
    import torch
    import torch.nn.functional as F
    return torch, F


def test_bf16_rounding_known_answers(tt):
    torch, F = tt
    one = 1.0
    vals = np.array([one, one + 2.0 ** -8, one + 3 * 2.0 ** -8, one + 2.0 ** -7, -2.5, 0.0, 3.0e38])
    exp = np.array([one, one, one + 2.0 ** -6, one + 2.0 ** -7, -2.5, 0.0,
                    float(torch.tensor(3.0e38).to(torch.bfloat16).float())])
    assert np.array_equal(uo.bf16(vals), exp)


def test_bf16_rounding_matches_torch(tt):
    torch, F = tt
    x = np.random.default_rng(0).standard_normal(100000) * 10.0 ** np.random.default_rng(1).uniform(-8, 8, 100000)
    t = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(uo.bf16(x), t)


def _brute_conv(X, W, b):
    B, n, _, _, Ci = X.shape
    Co = W.shape[0]
    Y = np.zeros((B, n, n, n, Co))
    for bi in range(B):
        for z in range(n):
            for y in range(n):
                for x in range(n):
                    for co in range(Co):
                        s = b[co]
                        for kz in range(3):
                            for ky in range(3):
                                for kx in range(3):
                                    zz, yy, xx = z + kz - 1, y + ky - 1, x + kx - 1
                                    if 0 <= zz < n and 0 <= yy < n and 0 <= xx < n:
                                        s += W[co, :, kz * 9 + ky * 3 + kx] @ X[bi, zz, yy, xx]
                        Y[bi, z, y, x, co] = s
    return Y


def test_conv_brute_force_and_torch(tt):
    torch, F = tt
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2, 4, 4, 4, 3))
    W = rng.standard_normal((5, 3, 27))
    b = rng.standard_normal(5)
    Y = uo.conv3(X, W, b)
    assert np.allclose(Y, _brute_conv(X, W, b), rtol=1e-12, atol=1e-12)
    Xt = torch.from_numpy(np.moveaxis(X, -1, 1).copy())
    Wt = torch.from_numpy(W.reshape(5, 3, 3, 3, 3).copy())
    Yt = F.conv3d(Xt, Wt, torch.from_numpy(b), padding=1).numpy()
    assert np.allclose(Y, np.moveaxis(Yt, 1, -1), rtol=1e-12, atol=1e-12)


def test_conv_adjoint_identity():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, 6, 6, 6, 4))
    Y = rng.standard_normal((1, 6, 6, 6, 7))
    W = rng.standard_normal((7, 4, 27))
    lhs = np.sum(uo.conv3(X, W) * Y)
    rhs = np.sum(X * uo.conv3_dgrad(Y, W))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # weight gradient is the adjoint in W
    V = rng.standard_normal(W.shape)
    assert np.sum(uo.conv3(X, V) * Y) == pytest.approx(np.sum(V * uo.conv3_wgrad(X, Y)), rel=1e-12)


def test_pool_and_upsample_match_torch(tt):
    torch, F = tt
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2, 8, 8, 8, 3))
    Xt = torch.from_numpy(np.moveaxis(X, -1, 1).copy())
    assert np.allclose(uo.pool(X), np.moveaxis(F.avg_pool3d(Xt, 2).numpy(), 1, -1), rtol=1e-14)
    assert np.array_equal(uo.upsample(X), np.moveaxis(F.interpolate(Xt, scale_factor=2, mode="nearest").numpy(),
                                                      1, -1))
    # backward operators are the adjoints
    Y = rng.standard_normal((2, 16, 16, 16, 3))
    assert np.sum(uo.upsample(X) * Y) == pytest.approx(np.sum(X * uo.upsample_bwd(Y)), rel=1e-12)
    Z = rng.standard_normal((2, 4, 4, 4, 3))
    assert np.sum(uo.pool(X) * Z) == pytest.approx(np.sum(X * uo.pool_bwd(Z)), rel=1e-12)


def _torch_net(tt, params, inp, target, n):
    torch, F = tt
    convs, (hw, hb) = uo.split_params(params)
    P = [torch.tensor(p, dtype=torch.float64, requires_grad=True)
         for w, b in convs for p in (w.reshape(w.shape[0], w.shape[1], 3, 3, 3), b)]
    hw_t = torch.tensor(hw, requires_grad=True)
    hb_t = torch.tensor(float(hb), dtype=torch.float64, requires_grad=True)
    x = torch.from_numpy(np.asarray(inp, dtype=np.float64).reshape(-1, 2, n, n, n))

    def c(i, t):
        return torch.relu(F.conv3d(t, P[2 * i], P[2 * i + 1], padding=1))

    a = c(1, c(0, x))
    s1 = a
    a = c(3, c(2, F.avg_pool3d(s1, 2)))
    s2 = a
    a = c(5, c(4, F.avg_pool3d(s2, 2)))
    a = torch.cat([F.interpolate(a, scale_factor=2, mode="nearest"), s2], dim=1)
    a = c(7, c(6, a))
    a = torch.cat([F.interpolate(a, scale_factor=2, mode="nearest"), s1], dim=1)
    a = c(9, c(8, a))
    pred = torch.einsum("bchwd,c->bhwd", a, hw_t) + hb_t
    t = torch.from_numpy(np.asarray(target, dtype=np.float64).reshape(pred.shape))
    loss = ((pred - t) ** 2).mean()
    loss.backward()
    g = [p.grad.numpy().reshape(-1) for p in P] + [hw_t.grad.numpy(), np.atleast_1d(hb_t.grad.numpy())]
    return loss.item(), np.concatenate(g), pred.detach().numpy().reshape(pred.shape[0], -1)


def _case(seed, n, batch):
    rng = np.random.default_rng(seed)
    params = uo.init_params(seed)
    inp = rng.standard_normal((batch, 2, n ** 3))
    target = rng.standard_normal((batch, n ** 3)) * 0.1
    return params, inp, target


def test_network_matches_torch_autograd(tt):
    torch, F = tt
    n = 8
    params, inp, target = _case(5, n, 2)
    loss, g, pred = uo.loss_and_grad(params, inp, target, n)
    tl, tg, tp = _torch_net(tt, params, inp, target, n)
    assert loss == pytest.approx(tl, rel=1e-12)
    assert np.allclose(pred, tp, rtol=1e-11, atol=1e-13)
    assert np.linalg.norm(g - tg) <= 1e-10 * np.linalg.norm(tg)
    assert g.size == uo.param_count()


def test_network_central_differences():
    n = 4
    params, inp, target = _case(6, n, 1)
    _, g, _ = uo.loss_and_grad(params, inp, target, n)
    rng = np.random.default_rng(7)
    eps = 1e-6
    for _ in range(4):
        d = rng.standard_normal(params.size)
        lp = uo.loss_and_grad(params + eps * d, inp, target, n)[0]
        lm = uo.loss_and_grad(params - eps * d, inp, target, n)[0]
        fd = (lp - lm) / (2 * eps)
        assert fd == pytest.approx(float(g @ d), rel=1e-6)


def test_rounding_mode_is_close_but_not_exact():
    n = 8
    params, inp, target = _case(8, n, 1)
    l0, g0, p0 = uo.loss_and_grad(params, inp, target, n, rounding=False)
    l1, g1, p1 = uo.loss_and_grad(params, inp, target, n, rounding=True)
    assert not np.array_equal(p0, p1)
    assert np.linalg.norm(p1 - p0) <= 2e-2 * np.linalg.norm(p0)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)
    assert l1 == pytest.approx(l0, rel=2e-2)
