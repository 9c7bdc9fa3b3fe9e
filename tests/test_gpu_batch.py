"""Batched CG on the GPU (one persistent launch for many systems).

Every batch member must equal, bit for bit, the oracle's Pipelined CG on that
system alone (same chunk-tree dot products, same fused updates), whatever the
workgroup shape, grid size or the other members of the batch; per-system
tolerance, iteration limits and breakdown must not leak between systems.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _systems(gpu, seed, nsys, lmin, lmax):
    out = []
    for sid in range(nsys):
        rp, col, val, f, dims = gpu.varcoef3d(seed, sid, lmin, lmax)
        out.append((rp, col, val, f))
    return out


def _batch(gpu, systems, **kw):
    srp, rp, col, val = gpu.block_diag([s[:3] for s in systems])
    b = np.concatenate([s[3] for s in systems])
    p = gpu.cg_params(method=gpu.CG_PIPELINED, **kw)
    return srp, rp, col, val, b, p


def _check_members(orc, systems, srp, x, iters, hist, tol, max_iter, quirk=True, x0=None):
    for i, (rp, col, val, f) in enumerate(systems):
        r0, r1 = srp[i], srp[i + 1]
        st, ox, oh, oit = orc.cg(orc.Csr(rp, col, val), f, x0=None if x0 is None else x0[r0:r1], tol=tol,
                                 max_iter=max_iter, variant="pipelined", quirk=quirk)
        assert iters[i] == oit, (i, iters[i], oit)
        assert np.array_equal(x[r0:r1], ox), i
        assert np.array_equal(hist[i, : oit + 1], oh), i


@pytest.mark.parametrize("quirk", [1, 0])
def test_batch_members_equal_oracle(gpu, orc, quirk):
    systems = _systems(gpu, 3, 12, 5, 14)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-10, max_iter=2000, restart_quirk=quirk)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=2001)
        x = plan.x()
    assert st == gpu.OK and np.all(stat == gpu.OK)
    _check_members(orc, systems, srp, x, iters, hist, 1e-10, 2000, bool(quirk))


@pytest.mark.parametrize("block,grid", [(256, 0), (512, 0), (1024, 0), (256, 1), (1024, 3), (512, 7)])
def test_batch_shapes_do_not_change_bits(gpu, orc, block, grid):
    systems = _systems(gpu, 5, 9, 9, 20)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-9, max_iter=3000, block_threads=block,
                                     grid_blocks=grid)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=3001)
        x = plan.x()
    assert st == gpu.OK
    _check_members(orc, systems, srp, x, iters, hist, 1e-9, 3000)


def test_batch_equals_solo_gpu(gpu):
    systems = _systems(gpu, 8, 6, 10, 24)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-8, max_iter=5000)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=0)
        x = plan.x()
    for i, (srp_i, s) in enumerate(zip(srp, systems)):
        sst, sx, sh, sit, st_ = gpu.cg(*s[:3], s[3], method=gpu.CG_PIPELINED, tol=1e-8, max_iter=5000)
        assert sit == iters[i] and np.array_equal(sx, x[srp[i]:srp[i + 1]])


def test_batch_fixed_iterations_and_limits(gpu, orc):
    systems = _systems(gpu, 9, 5, 6, 12)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=0.0, max_iter=25)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=26)
        x = plan.x()
    assert st == gpu.OK and np.all(iters == 25)
    _check_members(orc, systems, srp, x, iters, hist, 0.0, 25)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-12, max_iter=7)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve()
    assert st == gpu.ENOTCONV and np.all(stat == gpu.ENOTCONV) and np.all(iters == 7)


def test_batch_breakdown_is_contained(gpu, orc):
    systems = _systems(gpu, 4, 4, 6, 10)
    n = 700  # an indefinite diagonal system in the middle of the batch
    bad = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
           np.where(np.arange(n) % 2 == 0, 1.0, -1.0), np.ones(n))
    allsys = systems[:2] + [bad] + systems[2:]
    srp, rp, col, val, b, p = _batch(gpu, allsys, tol=1e-10, max_iter=2000)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=2001)
        x = plan.x()
    assert st == gpu.EBREAKDOWN
    assert stat[2] == gpu.EBREAKDOWN and np.all(np.delete(stat, 2) == gpu.OK)
    good = [0, 1, 3, 4]
    for i in good:
        rp_, col_, val_, f_ = allsys[i]
        st_o, ox, oh, oit = orc.cg(orc.Csr(rp_, col_, val_), f_, tol=1e-10, max_iter=2000, variant="pipelined")
        assert iters[i] == oit and np.array_equal(x[srp[i]:srp[i + 1]], ox)


def test_batch_nonzero_x0(gpu, orc):
    systems = _systems(gpu, 12, 5, 6, 12)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-10, max_iter=2000)
    x0 = gpu.uniform(1, 2, int(srp[-1]))
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(x0)
        st, iters, stat, hist, t = plan.solve(hist_cap=2001)
        x = plan.x()
    _check_members(orc, systems, srp, x, iters, hist, 1e-10, 2000, x0=x0)


def test_batch_autotune_and_cache(gpu, orc, tmp_path):
    systems = _systems(gpu, 6, 16, 8, 24)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-9, max_iter=3000)
    cache = str(tmp_path / "tune.txt")
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        blk, grid = plan.tune(trial_iters=10, cache_path=cache)
        assert blk in (256, 512, 1024) and grid >= 1
        fmt = plan.matrix_format()  # the 7-point systems are banded: SELL-64 and DIA both compete
        assert fmt in ("sell", "dia")
        st, iters, stat, hist, t = plan.solve(hist_cap=3001)
        x = plan.x()
        assert t["block_threads"] == blk and t["grid_blocks"] == grid
    _check_members(orc, systems, srp, x, iters, hist, 1e-9, 3000)
    with open(cache) as fh:
        assert len(fh.read().strip().splitlines()) == 1
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:  # cache hit, no new line
        assert plan.matrix_format() == "sell"  # solves start on SELL-64
        assert plan.tune(trial_iters=10, cache_path=cache) == (blk, grid)
        assert plan.matrix_format() == fmt
    with open(cache) as fh:
        assert len(fh.read().strip().splitlines()) == 1


def test_batch_converges_true_residual(gpu):
    systems = _systems(gpu, 0, 24, 16, 40)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-8, max_iter=10000)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve()
        x = plan.x()
    assert st == gpu.OK
    for i, (rp_, col_, val_, f_) in enumerate(systems):
        M = sp.csr_matrix((val_, col_, rp_))
        xi = x[srp[i]:srp[i + 1]]
        r = f_ - M @ xi
        assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(f_)


@pytest.mark.parametrize("block", [256, 1024])
def test_batch_x0_guard_rejects_bad_guesses(gpu, orc, block):
    """Guard on: a guess with ||b - A x0|| > ||b|| restarts from zero and then
    equals the oracle's run from zero bit for bit; an acceptable guess is
    kept and equals the oracle's run from that guess."""
    systems = _systems(gpu, 21, 8, 6, 18)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-9, max_iter=3000, x0_guard=1, block_threads=block)
    rng = np.random.default_rng(0)
    x0 = np.zeros(b.size)
    bad = np.array([i % 3 == 0 for i in range(len(systems))])
    for i, (rp_, col_, val_, f) in enumerate(systems):
        r0, r1 = srp[i], srp[i + 1]
        if bad[i]:  # noise only raises the residual: ||b - A x0||^2 ~ ||b||^2 + ||A x0||^2
            x0[r0:r1] = 1e-2 * rng.standard_normal(r1 - r0)
        else:  # half the solution: b - A x0 ~ b / 2
            x0[r0:r1] = 0.5 * orc.cg(orc.Csr(rp_, col_, val_), f, tol=1e-12, max_iter=5000)[1]
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(x0)
        st, iters, stat, hist, t = plan.solve(hist_cap=3001)
        x = plan.x()
        rej = plan.x0_rejected()
    assert st == gpu.OK
    assert np.array_equal(rej.astype(bool), bad)
    for i, (rp_, col_, val_, f) in enumerate(systems):
        r0, r1 = srp[i], srp[i + 1]
        ost, ox, oh, oit = orc.cg(orc.Csr(rp_, col_, val_), f, x0=None if bad[i] else x0[r0:r1], tol=1e-9,
                                  max_iter=3000, variant="pipelined")
        assert iters[i] == oit, i
        assert np.array_equal(x[r0:r1], ox), i
        assert np.array_equal(hist[i, : oit + 1], oh), i


def test_batch_x0_guard_off_keeps_bad_guess(gpu, orc):
    systems = _systems(gpu, 22, 3, 6, 10)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-9, max_iter=3000, x0_guard=0)
    x0 = 1e2 * np.random.default_rng(1).standard_normal(b.size)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(x0)
        st, iters, stat, hist, t = plan.solve(hist_cap=3001)
        x = plan.x()
        assert not plan.x0_rejected().any()
    _check_members(orc, systems, srp, x, iters, hist, 1e-9, 3000, x0=x0)


def test_batch_warm_start_writes_network_output(gpu):
    """x0 after warm_start == scale * net(features), for a system count that
    is not a multiple of the network's batch (short last group)."""
    n, nsys, scale = 8, 5, 0.37
    systems, feats = [], []
    rng = np.random.default_rng(3)
    for sid in range(nsys):
        rp, col, val, f, dims = gpu.varcoef3d(9, sid, n, n)
        systems.append((rp, col, val, f))
        feats.append(rng.standard_normal((2, n ** 3)).astype(np.float32))
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=0.0, max_iter=0)
    params = gpu.unet_init_params(4)
    params[-1] = 0.25  # head bias, so that no output is trivially zero
    with gpu.UNet(n, 2, params=params) as net, gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.load_features(np.concatenate([f_.reshape(-1) for f_ in feats]))
        tw = plan.warm_start(net, scale)
        plan.solve()  # zero iterations: x = x0
        x = plan.x()
        expect = []
        for g in range(0, nsys, 2):
            grp = feats[g: g + 2] + [np.zeros((2, n ** 3), np.float32)] * (2 - len(feats[g: g + 2]))
            out = net.forward(np.stack(grp))
            expect += [scale * out[j].reshape(-1).astype(np.float64) for j in range(min(2, nsys - g))]
    assert tw["iterations"] == 3
    assert np.array_equal(x, np.concatenate(expect))


def test_batch_warm_start_rejects_mismatched_sizes(gpu):
    systems = _systems(gpu, 23, 2, 6, 10)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-8, max_iter=100)
    with gpu.UNet(8, 1) as net, gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        with pytest.raises(gpu.CfmError) as e:
            plan.warm_start(net, 1.0)  # no features loaded
        assert e.value.status == gpu.EINVAL
        plan.load_features(np.zeros(2 * b.size, np.float32))
        with pytest.raises(gpu.CfmError) as e:
            plan.warm_start(net, 1.0)  # systems are not 8^3 cubes
        assert e.value.status == gpu.EINVAL


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_batch_smooth_x0_equals_oracle(gpu, orc, sweeps):
    systems = _systems(gpu, 24, 5, 5, 20)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=0.0, max_iter=0)
    x0 = np.random.default_rng(5).standard_normal(b.size)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(x0)
        t = plan.smooth_x0(sweeps, 0.8)
        plan.solve()  # zero iterations: x = x0
        x = plan.x()
    assert t["iterations"] == sweeps
    for i, (rp_, col_, val_, f) in enumerate(systems):
        r0, r1 = srp[i], srp[i + 1]
        ref = orc.jacobi_smooth(orc.Csr(rp_, col_, val_), f, x0[r0:r1], sweeps, 0.8)
        assert np.array_equal(x[r0:r1], ref), i


@pytest.mark.parametrize("fmt", [1, 2])
def test_batch_sell_and_dia_equal_oracle(gpu, orc, fmt):
    """Both storages of the batch (SELL-64 and per-system DIA) give the oracle's bits."""
    systems = _systems(gpu, 31, 10, 4, 17)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-10, max_iter=2000, matrix_format=fmt)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=2001)
        x = plan.x()
    assert st == gpu.OK
    _check_members(orc, systems, srp, x, iters, hist, 1e-10, 2000)


def test_batch_dia_rejects_unbanded(gpu):
    n = 2000
    M = sp.random(n, n, density=0.004, random_state=7, format="csr")
    M = (M + M.T + sp.eye(n) * 50).tocsr()
    M.sort_indices()
    srp = np.array([0, n], dtype=np.int64)
    p = gpu.cg_params(method=gpu.CG_PIPELINED, matrix_format=2)
    with pytest.raises(gpu.CfmError) as e:
        gpu.CgBatchPlan(srp, M.indptr, M.indices, M.data, p)
    assert e.value.status == gpu.ENOTSUP


def test_benchmark_batch_workload(gpu, orc):
    """The batched benchmark workload itself (256 systems, edges log-uniform in
    [32, 96], seed 2024) in one launch: every member reaches 1e-8 by its true
    residual, and three members (first, middle, last) equal the oracle's solo
    solve bit for bit."""
    systems = _systems(gpu, 2024, 256, 32, 96)
    srp, rp, col, val, b, p = _batch(gpu, systems, tol=1e-8, max_iter=20000)
    with gpu.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve(hist_cap=20001)
        x = plan.x()
    assert st == gpu.OK and np.all(stat == gpu.OK)
    for i, (rp_, col_, val_, f_) in enumerate(systems):
        r = f_ - sp.csr_matrix((val_, col_, rp_)) @ x[srp[i]:srp[i + 1]]
        assert np.linalg.norm(r) <= 1.2e-8 * np.linalg.norm(f_), i
    for i in (0, 128, 255):
        rp_, col_, val_, f_ = systems[i]
        ost, ox, oh, oit = orc.cg(orc.Csr(rp_, col_, val_), f_, tol=1e-8, max_iter=20000, variant="pipelined")
        assert iters[i] == oit
        assert np.array_equal(x[srp[i]:srp[i + 1]], ox)
        assert np.array_equal(hist[i, : oit + 1], oh)
