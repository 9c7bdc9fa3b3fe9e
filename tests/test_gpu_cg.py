"""CG parity on the GPU, through the C ABI.

The GPU kernels and the oracle use the same deterministic dot-product tree
and the same fused multiply-adds, so residual histories, iteration counts and
solutions must agree bit for bit (well inside the 1e-10 relative bar set for
the reference's residual history).  The discrete (host-driven) schedule and
any grid size must give the same bits as the persistent kernel.  At the full
128^3 size the whole solve to 1e-8 is compared bit for bit with the oracle
(history, iterations, solution) and the converged solution is checked
against the true residual b - A x.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _poisson(gpu, dims, seed=0):
    rp, col, val = gpu.poisson3d(*dims)
    b = gpu.uniform(seed, 0, rp.size - 1)
    return rp, col, val, b


def _oracle(orc, rp, col, val, b, method, tol, max_iter, quirk=True, base=0):
    A = orc.Csr(rp, col, val, base)
    if method == "ir":
        st, x, hist, outer, inner = orc.cg_ir(A, b, tol=tol, inner_max=max_iter)
        return st, x, hist, inner
    return orc.cg(A, b, tol=tol, max_iter=max_iter, variant=method, quirk=quirk)


def _gpu(gpu, rp, col, val, b, method, mode, tol, max_iter, quirk=True, base=0, grid=0):
    m = {"standard": gpu.CG_STANDARD, "pipelined": gpu.CG_PIPELINED, "ir": gpu.CG_IR_F32}[method]
    return gpu.cg(rp, col, val, b, index_base=base, method=m, mode=mode, tol=tol, max_iter=max_iter,
                  restart_quirk=1 if quirk else 0, grid_blocks=grid)


def _same(g, o):
    st, x, hist, it, t = g
    ost, ox, ohist, oit = o
    assert st == ost
    assert it == oit
    assert np.array_equal(hist, ohist), np.flatnonzero(hist != ohist)[:5]
    assert np.array_equal(x, ox)


DIMS = [(8, 8, 8), (16, 16, 16), (24, 20, 18), (33, 7, 5)]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("method", ["standard", "pipelined"])
@pytest.mark.parametrize("mode", [0, 1])
def test_poisson_bitexact(gpu, orc, dims, method, mode):
    rp, col, val, b = _poisson(gpu, dims, seed=sum(dims))
    g = _gpu(gpu, rp, col, val, b, method, mode, 1e-10, 2000)
    o = _oracle(orc, rp, col, val, b, method, 1e-10, 2000)
    assert g[0] == gpu.OK
    _same(g, o)


@pytest.mark.parametrize("quirk", [True, False])
def test_pipelined_restart_flag(gpu, orc, quirk):
    rp, col, val, b = _poisson(gpu, (20, 20, 20), 3)
    _same(_gpu(gpu, rp, col, val, b, "pipelined", 0, 1e-9, 1000, quirk=quirk),
          _oracle(orc, rp, col, val, b, "pipelined", 1e-9, 1000, quirk=quirk))


def test_fixed_iteration_mode(gpu, orc):
    # tol <= 0 runs exactly max_iter iterations, as the reference does
    rp, col, val, b = _poisson(gpu, (16, 16, 16), 1)
    for method in ("standard", "pipelined"):
        g = _gpu(gpu, rp, col, val, b, method, 0, 0.0, 37)
        assert g[0] == gpu.OK and g[3] == 37
        _same(g, _oracle(orc, rp, col, val, b, method, 0.0, 37))


@pytest.mark.parametrize("grid", [1, 2, 5])
def test_grid_size_does_not_change_bits(gpu, orc, grid):
    rp, col, val, b = _poisson(gpu, (24, 24, 24), 2)
    for method in ("standard", "pipelined"):
        _same(_gpu(gpu, rp, col, val, b, method, 0, 1e-10, 2000, grid=grid),
              _oracle(orc, rp, col, val, b, method, 1e-10, 2000))


def _irregular_spd(n, seed):
    # random symmetric pattern with 1..40 entries per row, diagonally dominant
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        k = rng.integers(0, 20)
        js = rng.integers(0, n, size=k)
        rows += [i] * k
        cols += js.tolist()
    v = rng.uniform(-1, 0, size=len(rows))
    M = sp.coo_matrix((v, (rows, cols)), shape=(n, n)).tocsr()
    M = M + M.T
    d = np.asarray(np.abs(M).sum(axis=1)).ravel() + 1.0
    M = sp.csr_matrix(M + sp.diags(d))
    M.sort_indices()
    return M


@pytest.mark.parametrize("n", [700, 5000])
def test_irregular_rows_sell_padding(gpu, orc, n):
    M = _irregular_spd(n, n)
    rp, col, val = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data
    b = gpu.uniform(9, 0, n)
    for method in ("standard", "pipelined"):
        _same(_gpu(gpu, rp, col, val, b, method, 0, 1e-11, 500),
              _oracle(orc, rp, col, val, b, method, 1e-11, 500))


def test_one_based_csr(gpu, orc):
    rp, col, val, b = _poisson(gpu, (10, 9, 8), 4)
    g1 = _gpu(gpu, rp + 1, col + 1, val, b, "standard", 0, 1e-10, 1000, base=1)
    g0 = _gpu(gpu, rp, col, val, b, "standard", 0, 1e-10, 1000)
    assert np.array_equal(g1[1], g0[1]) and np.array_equal(g1[2], g0[2])


def test_nonzero_x0(gpu, orc):
    rp, col, val, b = _poisson(gpu, (12, 12, 12), 6)
    x0 = gpu.uniform(6, 1, b.size)
    g = gpu.cg(rp, col, val, b, x0=x0, tol=1e-10, max_iter=1000)
    o = orc.cg(orc.Csr(rp, col, val), b, x0=x0, tol=1e-10, max_iter=1000)
    _same(g, o)


def test_ir_bitexact(gpu, orc):
    rp, col, val, b = _poisson(gpu, (20, 18, 16), 8)
    st, x, hist, it, t = _gpu(gpu, rp, col, val, b, "ir", 0, 1e-10, 10000)
    ost, ox, ohist, oouter, oinner = orc.cg_ir(orc.Csr(rp, col, val), b, tol=1e-10, inner_max=10000)
    assert st == ost == gpu.OK
    assert t["outer_iterations"] == oouter and it == oinner
    assert np.array_equal(hist, ohist) and np.array_equal(x, ox)


def test_breakdown_and_notconv(gpu):
    n = 3000
    rp = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    val = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    for m in (gpu.CG_STANDARD, gpu.CG_PIPELINED):
        for mode in (0, 1):
            st, *_ = gpu.cg(rp, col, val, np.ones(n), method=m, mode=mode, tol=1e-12, max_iter=50)
            assert st == gpu.EBREAKDOWN
    rp, col, val, b = _poisson(gpu, (16, 16, 16))
    st, x, hist, it, t = gpu.cg(rp, col, val, b, tol=1e-12, max_iter=5)
    assert st == gpu.ENOTCONV and it == 5 and len(hist) == 6


def test_cg_barrier_watchdog(gpu):
    rp, col, val, b = _poisson(gpu, (64, 64, 64))
    with pytest.raises(gpu.CfmError) as e:
        gpu.cg(rp, col, val, b, tol=1e-10, max_iter=500, spin_limit_us=1)
    assert e.value.status == gpu.ETIMEOUT


@pytest.mark.parametrize("method", ["standard", "pipelined"])
def test_full_size_128(gpu, orc, method):
    rp, col, val, b = _poisson(gpu, (128, 128, 128), 0)
    # first 20 iterations bit for bit
    g = _gpu(gpu, rp, col, val, b, method, 0, 0.0, 20)
    o = _oracle(orc, rp, col, val, b, method, 0.0, 20)
    assert np.array_equal(g[2], o[2])
    # converged solve: true residual
    st, x, hist, it, t = _gpu(gpu, rp, col, val, b, method, 0, 1e-8, 10000)
    assert st == gpu.OK and 100 < it < 2000
    A = sp.csr_matrix((val, col, rp))
    r = b - A @ x
    assert np.sqrt(r @ r) <= 1.2e-8 * np.sqrt(b @ b)
    assert hist[-1] <= 1e-16 * (b @ b)


@pytest.mark.parametrize("method", ["standard", "pipelined", "ir"])
def test_full_size_128_equals_oracle(gpu, orc, method):
    """The CG benchmark configuration end to end (128^3 Poisson, uniform rhs,
    x0 = 0, to 1e-8): status, iteration count, residual history and solution
    equal the oracle's bit for bit (the required bar is 1e-10 relative on the
    history).  Pipelined runs on the index-free DIA storage here."""
    rp, col, val, b = _poisson(gpu, (128, 128, 128), 0)
    _same(_gpu(gpu, rp, col, val, b, method, 0, 1e-8, 10000), _oracle(orc, rp, col, val, b, method, 1e-8, 10000))


def test_full_size_128_ir(gpu):
    rp, col, val, b = _poisson(gpu, (128, 128, 128), 0)
    st, x, hist, it, t = _gpu(gpu, rp, col, val, b, "ir", 0, 1e-8, 10000)
    assert st == gpu.OK
    A = sp.csr_matrix((val, col, rp))
    r = b - A @ x
    assert np.sqrt(r @ r) <= 1e-8 * np.sqrt(b @ b)


@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("method", ["standard", "pipelined", "ir"])
def test_dia_and_sell_storage_equal_oracle(gpu, orc, fmt, method):
    """The index-free DIA storage (banded CSR) and SELL-64 give the same bits."""
    rp, col, val, b = _poisson(gpu, (20, 13, 9), seed=4)
    m = {"standard": gpu.CG_STANDARD, "pipelined": gpu.CG_PIPELINED, "ir": gpu.CG_IR_F32}[method]
    g = gpu.cg(rp, col, val, b, method=m, tol=1e-10, max_iter=3000, matrix_format=fmt)
    _same(g, _oracle(orc, rp, col, val, b, method, 1e-10, 3000))


def test_dia_on_varcoef_and_one_based(gpu, orc):
    rp, col, val, f, dims = gpu.varcoef3d(5, 2, 7, 15)
    g = gpu.cg(rp + 1, col + 1, val, f, index_base=1, method=gpu.CG_PIPELINED, tol=1e-9, max_iter=3000,
               matrix_format=2)
    _same(g, _oracle(orc, rp + 1, col + 1, val, f, "pipelined", 1e-9, 3000, base=1))


def test_dia_rejects_unbanded(gpu):
    n = 3000
    rng = np.random.default_rng(2)
    M = sp.random(n, n, density=0.004, random_state=3, format="csr")
    M = (M + M.T + sp.eye(n) * 50).tocsr()
    M.sort_indices()
    with pytest.raises(gpu.CfmError) as e:
        gpu.cg(M.indptr, M.indices, M.data, rng.standard_normal(n), matrix_format=2)
    assert e.value.status == gpu.ENOTSUP
