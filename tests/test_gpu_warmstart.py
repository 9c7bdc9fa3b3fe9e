"""Learned warm start: what the trained network gains over plain CG.

Loads the committed weights (trained on the GPU by cpu-free-model_amd/
warmstart.py on 20000 generated systems of seed TRAIN_SEED) and solves
held-out systems (seed EVAL_SEED) twice in one batched plan: from x0 = 0 and
from the network's guess (inference, two damped Jacobi sweeps, guard).  Both
must reach 1e-8 in the true residual; the learned start must need fewer
iterations; a useless network must fall back to exactly the plain solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import warmstart as ws

pytestmark = pytest.mark.gpu

NSYS = 128  # held-out systems (the benchmark evaluates 256, the training run 1000)


@pytest.fixture(scope="module")
def held_out(gpu):
    systems = ws.generate(ws.EVAL_SEED, range(3000, 3000 + NSYS))
    params, meta = ws.load()
    return systems, params, meta


def _true_rel_residual(s, x):
    rp, col, val, f = s[:4]
    A = sp.csr_matrix((val, col, rp), shape=(f.size, f.size))
    return np.linalg.norm(f - A @ x) / np.linalg.norm(f)


def test_learned_start_needs_fewer_iterations(gpu, held_out):
    systems, params, meta = held_out
    bt = ws.Batch(systems, guard=True)
    try:
        st0, it0, s0, _, _ = bt.solve_plain()
        x_plain = bt.solutions().copy()
        with gpu.UNet(meta["n"], 8, params=params) as net:
            _, (st1, it1, s1, h1, _) = bt.solve_learned(net, meta["u_scale"], hist_cap=1)
        x_learned = bt.solutions()
        rejected = bt.plan.x0_rejected()
    finally:
        bt.close()
    assert st0 == gpu.OK and st1 == gpu.OK
    assert not rejected.any()
    bb = np.array([s[3] @ s[3] for s in systems])
    r0 = np.sqrt(h1[:, 0] / bb)  # ||b - A x0|| / ||b|| of the smoothed guesses
    assert np.all(r0 <= 1.0) and np.percentile(r0, 95) < 0.5, np.percentile(r0, [50, 95, 100])
    assert it1.mean() <= 0.93 * it0.mean(), (it0.mean(), it1.mean())
    assert (it1 < it0).mean() >= 0.9  # nearly every system gains, not only the mean
    for i, s in enumerate(systems):
        assert _true_rel_residual(s, x_plain[i]) <= 2e-8
        assert _true_rel_residual(s, x_learned[i]) <= 2e-8
        assert np.linalg.norm(x_learned[i] - x_plain[i]) <= 1e-6 * np.linalg.norm(x_plain[i])


def test_useless_network_falls_back_to_plain(gpu, held_out):
    systems, params, meta = held_out
    bad = params.copy()
    bad[-1] = 1e4  # head bias: a huge constant guess
    bt = ws.Batch(systems[:4], guard=True)
    try:
        st0, it0, _, h0, _ = bt.solve_plain(hist_cap=800)
        x_plain = bt.solutions().copy()
        with gpu.UNet(meta["n"], 2, params=bad) as net:
            _, (st1, it1, _, h1, _) = bt.solve_learned(net, meta["u_scale"], hist_cap=800)
        x_fb = bt.solutions()
        rejected = bt.plan.x0_rejected()
    finally:
        bt.close()
    assert st0 == gpu.OK and st1 == gpu.OK
    assert rejected.all()
    assert np.array_equal(it0, it1)
    assert np.array_equal(h0, h1)
    assert np.array_equal(x_plain, x_fb)
