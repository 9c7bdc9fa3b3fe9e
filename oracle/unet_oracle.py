"""CPU oracle of the warm-start U-Net (forward, loss and backward) in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
training and inference kernels; never used by the product path.

The network has no counterpart in the reference (the reference's CG always
starts from x0 = 0, CG/src/single-stream/standard.cu:182-199); it is the
learned warm start added beside the plain solver.  Layout and parameter
order follow the product's C ABI:

  * activations channels-last [B, n, n, n, C], spatial order (z, y, x);
  * conv weights W[Cout][Cin][27], tap = 9(dz+1) + 3(dy+1) + (dx+1), zero
    padding, output Y[p] = b + sum_tap W[:, :, tap] X[p + off(tap)];
  * encoder 2->16->16, pool, 16->32->32, pool, 32->64->64, nearest upsample,
    concat [up, skip] -> 32->32, upsample, concat -> 16->16, 1x1 head 16->1;
  * loss = mean((pred - target)^2).

``rounding=False`` is the exact float64 network (pinned in the tests against
torch autograd and central differences).  ``rounding=True`` rounds to BF16
at the points where the device stores BF16: input, weights, every activation
after bias+ReLU, pooled/upsampled tensors, and every activation gradient
(head output gradient, dgrad outputs, upsample/pool gradients, skip sums).
Products are accumulated in float64 here (FP32 on the device).
"""
import numpy as np

CONVS = ((2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16))
HEAD_IN = 16


def param_count():
    return sum(co * ci * 27 + co for ci, co in CONVS) + HEAD_IN + 1


def split_params(flat):
    """flat -> list of (W [Cout,Cin,27], b [Cout]) and (head_w [16], head_b)."""
    flat = np.asarray(flat, dtype=np.float64)
    out, off = [], 0
    for ci, co in CONVS:
        w = flat[off: off + co * ci * 27].reshape(co, ci, 27)
        off += co * ci * 27
        b = flat[off: off + co]
        off += co
        out.append((w, b))
    hw = flat[off: off + HEAD_IN]
    hb = flat[off + HEAD_IN]
    assert off + HEAD_IN + 1 == flat.size
    return out, (hw, hb)


def join_grads(conv_grads, head_grads):
    parts = []
    for gw, gb in conv_grads:
        parts += [gw.ravel(), gb.ravel()]
    parts += [np.asarray(head_grads[0]).ravel(), np.atleast_1d(head_grads[1])]
    return np.concatenate(parts)


def bf16(a):
    """Round to the nearest BF16 value (ties to even), via float32."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps():
    for tap in range(27):
        yield tap, tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1


def conv3(X, W, b=None):
    """Y[B,n,n,n,Cout] = b + sum_tap W[:,:,tap] X[shifted by off(tap)] (zero padded)."""
    B, n = X.shape[0], X.shape[1]
    Xp = np.zeros((B, n + 2, n + 2, n + 2, X.shape[4]))
    Xp[:, 1:-1, 1:-1, 1:-1] = X
    Y = np.zeros((B, n, n, n, W.shape[0]))
    for tap, dz, dy, dx in _taps():
        Xs = Xp[:, 1 + dz: 1 + dz + n, 1 + dy: 1 + dy + n, 1 + dx: 1 + dx + n]
        Y += Xs @ W[:, :, tap].T
    if b is not None:
        Y += b
    return Y


def conv3_wgrad(X, dZ):
    """dW[co, ci, tap] = sum_p dZ[p, co] X[p + off(tap), ci]."""
    B, n, C = X.shape[0], X.shape[1], X.shape[4]
    Xp = np.zeros((B, n + 2, n + 2, n + 2, C))
    Xp[:, 1:-1, 1:-1, 1:-1] = X
    dZf = dZ.reshape(-1, dZ.shape[4])
    dW = np.zeros((dZ.shape[4], C, 27))
    for tap, dz, dy, dx in _taps():
        Xs = Xp[:, 1 + dz: 1 + dz + n, 1 + dy: 1 + dy + n, 1 + dx: 1 + dx + n].reshape(-1, C)
        dW[:, :, tap] = dZf.T @ Xs
    return dW


def conv3_dgrad(dZ, W):
    """dX[q, ci] = sum_tap sum_co W[co, ci, tap] dZ[q - off(tap), co]."""
    B, n = dZ.shape[0], dZ.shape[1]
    dXp = np.zeros((B, n + 2, n + 2, n + 2, W.shape[1]))
    for tap, dz, dy, dx in _taps():
        dXp[:, 1 + dz: 1 + dz + n, 1 + dy: 1 + dy + n, 1 + dx: 1 + dx + n] += dZ @ W[:, :, tap]
    return dXp[:, 1:-1, 1:-1, 1:-1]


def pool(X):
    B, n, C = X.shape[0], X.shape[1], X.shape[4]
    m = n // 2
    return X.reshape(B, m, 2, m, 2, m, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dY):
    return upsample(dY) * 0.125


def upsample(X):
    return X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def upsample_bwd(dY):
    B, n, C = dY.shape[0], dY.shape[1], dY.shape[4]
    m = n // 2
    return dY.reshape(B, m, 2, m, 2, m, 2, C).sum(axis=(2, 4, 6))


def to_channels_last(inp, n):
    """[B, 2, n^3] planes -> [B, n, n, n, 2]."""
    a = np.asarray(inp, dtype=np.float64).reshape(-1, 2, n, n, n)
    return np.moveaxis(a, 1, -1)


def forward(params, inp, n, rounding=False):
    """Returns (pred [B, n^3], cache)."""
    R = bf16 if rounding else (lambda a: a)
    convs, (hw, hb) = split_params(params)
    Ws = [R(w) for w, _ in convs]
    bs = [b for _, b in convs]
    X0 = R(to_channels_last(inp, n))

    def cr(i, X):
        return R(np.maximum(conv3(X, Ws[i], bs[i]), 0.0))

    c = {"X0": X0}
    c["A0"] = cr(0, X0)
    c["S1"] = cr(1, c["A0"])
    c["Q1"] = R(pool(c["S1"]))
    c["A2"] = cr(2, c["Q1"])
    c["S2"] = cr(3, c["A2"])
    c["Q2"] = R(pool(c["S2"]))
    c["A4"] = cr(4, c["Q2"])
    c["A5"] = cr(5, c["A4"])
    c["U2"] = np.concatenate([upsample(c["A5"]), c["S2"]], axis=-1)
    c["A6"] = cr(6, c["U2"])
    c["A7"] = cr(7, c["A6"])
    c["U1"] = np.concatenate([upsample(c["A7"]), c["S1"]], axis=-1)
    c["A8"] = cr(8, c["U1"])
    c["A9"] = cr(9, c["A8"])
    pred = c["A9"] @ hw + hb
    c["Ws"] = Ws
    return pred.reshape(pred.shape[0], -1), c


def loss_and_grad(params, inp, target, n, rounding=False):
    """Returns (loss, flat gradient, pred)."""
    R = bf16 if rounding else (lambda a: a)
    convs, (hw, hb) = split_params(params)
    pred, c = forward(params, inp, n, rounding)
    Ws = c["Ws"]
    t = np.asarray(target, dtype=np.float64).reshape(pred.shape)
    N = pred.size
    d = pred - t
    loss = float(np.sum(d * d) / N)
    dpred = (2.0 * d / N).reshape(c["A9"].shape[:4])
    g_hw = np.tensordot(dpred, c["A9"], axes=([0, 1, 2, 3], [0, 1, 2, 3]))
    g_hb = float(dpred.sum())
    grads = [None] * len(CONVS)

    def back(i, Y, dY, X, want_dx=True):
        dZ = np.where(Y > 0.0, dY, 0.0)
        grads[i] = (conv3_wgrad(X, dZ), dZ.sum(axis=(0, 1, 2, 3)))
        return R(conv3_dgrad(dZ, Ws[i])) if want_dx else None

    dA9 = R(dpred[..., None] * hw)
    dA8 = back(9, c["A9"], dA9, c["A8"])
    dU1 = back(8, c["A8"], dA8, c["U1"])
    dA7 = R(upsample_bwd(dU1[..., :32]))
    dS1_skip = dU1[..., 32:]
    dA6 = back(7, c["A7"], dA7, c["A6"])
    dU2 = back(6, c["A6"], dA6, c["U2"])
    dA5 = R(upsample_bwd(dU2[..., :64]))
    dS2_skip = dU2[..., 64:]
    dA4 = back(5, c["A5"], dA5, c["A4"])
    dQ2 = back(4, c["A4"], dA4, c["Q2"])
    dS2 = R(R(pool_bwd(dQ2)) + dS2_skip)
    dA2 = back(3, c["S2"], dS2, c["A2"])
    dQ1 = back(2, c["A2"], dA2, c["Q1"])
    dS1 = R(R(pool_bwd(dQ1)) + dS1_skip)
    dA0 = back(1, c["S1"], dS1, c["A0"])
    back(0, c["A0"], dA0, c["X0"], want_dx=False)
    return loss, join_grads(grads, (g_hw, g_hb)), pred


def init_params(seed):
    """Random parameters for tests (not the product's initialiser)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in CONVS:
        a = np.sqrt(6.0 / (27 * ci))
        parts += [rng.uniform(-a, a, co * ci * 27), rng.uniform(-0.05, 0.05, co)]
    parts += [rng.uniform(-0.3, 0.3, HEAD_IN), rng.uniform(-0.1, 0.1, 1)]
    return np.concatenate(parts)
