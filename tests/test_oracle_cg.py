"""Pins of the CG CPU oracle (oracle/cg_oracle.c) and of the generators.

No reference golden vectors exist for CG (the reference's own CPU path never
stores its SpMV result, CG/src/common.cu:784-798), so the oracle is pinned by:
  * SplitMix64's published known-answer values (state 0);
  * the 7-point Poisson CSR rebuilt independently from Kronecker products;
  * exact solves: tiny SPD systems converge to numpy.linalg.solve;
  * scipy.sparse.linalg.cg (same Hestenes-Stiefel recurrence, other code)
    reproduces the residual history of the Standard oracle;
  * Pipelined with the restart fixed (k > 0) equals Standard in exact
    arithmetic, and the reference's k > 1 restart only delays convergence;
  * the recursively updated residual tracks b - A x;
  * the dot tree equals the exactly rounded sum to within its error bound.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def poisson_kron(nx, ny, nz):
    def lap1(n):
        return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity
    A = (sp.kron(I(nz), sp.kron(I(ny), lap1(nx))) + sp.kron(I(nz), sp.kron(lap1(ny), I(nx)))
         + sp.kron(lap1(nz), sp.kron(I(ny), I(nx))))
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def test_splitmix64_known_answers(orc):
    # splitmix64.c (Vigna) from state 0
    assert orc.splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_stream(orc):
    u = orc.uniform(0, 0, 1000)
    assert u.min() >= 0.0 and u.max() < 1.0
    st = orc.lib().orc_stream_state(0, 0)
    first = orc.splitmix64(st, 1)[0]
    assert u[0] == (first >> 11) * 2.0 ** -53
    assert not np.array_equal(u, orc.uniform(0, 1, 1000))
    assert abs(u.mean() - 0.5) < 0.03


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 3, 7), (1, 1, 6)])
def test_poisson_csr_matches_kron(orc, dims):
    A = orc.poisson3d(*dims)
    K = poisson_kron(*dims)
    assert A.n == K.shape[0]
    assert np.array_equal(A.rowptr, K.indptr)
    assert np.array_equal(A.col, K.indices)
    assert np.array_equal(A.val, K.data)
    nx, ny, nz = dims
    if dims == (4, 4, 4):
        assert A.nnz == 7 * 64 - 6 * 16 == 352


def test_spmv_matches_scipy(orc):
    A = orc.poisson3d(6, 5, 4)
    x = orc.uniform(3, 0, A.n)
    K = sp.csr_matrix((A.val, A.col, A.rowptr))
    assert np.allclose(orc.spmv(A, x), K @ x, rtol=1e-15, atol=1e-15)
    # integer-valued data makes every partial sum exact: bitwise
    xi = np.arange(A.n, dtype=np.float64)
    assert np.array_equal(orc.spmv(A, xi), K @ xi)


def test_spmv_one_based(orc):
    A = orc.poisson3d(3, 3, 3)
    B = orc.Csr(A.rowptr + 1, A.col + 1, A.val, base=1)
    x = orc.uniform(1, 0, A.n)
    assert np.array_equal(orc.spmv(A, x), orc.spmv(B, x))


@pytest.mark.parametrize("n", [1, 100, 1024, 1025, 5000, 70000])
def test_dot_tree_error_bound(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n)
    b = rng.standard_normal(n)
    exact = math.fsum(float(x) * float(y) for x, y in zip(a, b))
    bound = 1.1 * n * 2.0 ** -53 * float(np.abs(a) @ np.abs(b)) + 1e-300
    for tree in (orc.TREE_GPU, orc.TREE_SEQUENTIAL):
        assert abs(orc.dot(a, b, tree) - exact) <= bound


def test_dot_tree_is_the_chunk_tree(orc):
    # Chunk 0 holds 1e16 (row 0) and 1.0 (row 256, same thread): the 1.0 is
    # absorbed, partial = 1e16.  Chunk 1 holds -1e16 (row 1024, lane 0) and 1.0
    # (row 1025, lane 1): the butterfly absorbs the 1.0, partial = -1e16.  The
    # chunk tree therefore returns exactly 0.0, while the sequential order
    # ((1e16 + 1) - 1e16) + 1 returns 1.0: the value identifies the tree.
    n = 2048
    a = np.zeros(n)
    a[0], a[256], a[1024], a[1025] = 1e16, 1.0, -1e16, 1.0
    ones = np.ones(n)
    assert orc.dot(a, ones, orc.TREE_GPU) == 0.0
    assert orc.dot(a, ones, orc.TREE_SEQUENTIAL) == 1.0


def _dense(A):
    return sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n)).toarray()


@pytest.mark.parametrize("variant", ["standard", "pipelined"])
@pytest.mark.parametrize("dims", [(4, 4, 4), (3, 5, 2)])
def test_cg_solves_small_spd(orc, variant, dims):
    A = orc.poisson3d(*dims)
    b = orc.uniform(7, 0, A.n)
    st, x, hist, it = orc.cg(A, b, tol=1e-13, max_iter=500, variant=variant)
    assert st == orc.OK
    xs = np.linalg.solve(_dense(A), b)
    assert np.max(np.abs(x - xs)) <= 1e-11 * np.max(np.abs(xs))
    assert hist[-1] <= 1e-26 * (b @ b)
    assert it <= A.n + 5


def test_cg_history_matches_scipy(orc):
    A = orc.poisson3d(10, 9, 8)
    b = orc.uniform(11, 0, A.n)
    K = sp.csr_matrix((A.val, A.col, A.rowptr))
    xs = []
    spla.cg(K, b, x0=np.zeros(A.n), rtol=1e-12, maxiter=40, callback=lambda xk: xs.append(xk.copy()))
    st, x, hist, it = orc.cg(A, b, tol=0.0, max_iter=40, variant="standard")
    assert it == 40 and st == orc.OK
    for k in range(1, 30):
        r = b - K @ xs[k - 1]
        assert hist[k] == pytest.approx(r @ r, rel=1e-8)


def test_pipelined_equals_standard_without_quirk(orc):
    A = orc.poisson3d(12, 10, 9)
    b = orc.uniform(5, 0, A.n)
    _, xs_, hs, its = orc.cg(A, b, tol=1e-9, variant="standard")
    _, xp, hp, itp = orc.cg(A, b, tol=1e-9, variant="pipelined", quirk=False)
    m = min(40, len(hs), len(hp))
    assert np.allclose(hp[:m], hs[:m], rtol=1e-8)
    assert abs(itp - its) <= 2
    _, xq, hq, itq = orc.cg(A, b, tol=1e-9, variant="pipelined", quirk=True)
    # the k > 1 restart repeats a steepest-descent step: same first two residuals,
    # then a different but still convergent history
    assert hq[0] == hp[0] and hq[1] == hp[1] and hq[2] != hp[2]
    assert itq >= itp - 2
    assert np.max(np.abs(xq - xs_)) < 1e-7


def test_recursive_residual_tracks_true_residual(orc):
    A = orc.poisson3d(16, 16, 16)
    b = orc.uniform(0, 0, A.n)
    st, x, hist, it = orc.cg(A, b, tol=1e-10, variant="standard")
    r = b - orc.spmv(A, x)
    assert st == orc.OK
    assert abs(math.sqrt(r @ r) - math.sqrt(hist[-1])) < 1e-12 * math.sqrt(b @ b)


def test_cg_breakdown_and_notconv(orc):
    # indefinite matrix: diag(1, -1, 1, ...)
    n = 8
    A = orc.Csr(np.arange(n + 1), np.arange(n), np.array([1.0, -1.0] * 4))
    b = np.ones(n)
    for variant in ("standard", "pipelined"):
        st, *_ = orc.cg(A, b, tol=1e-12, max_iter=50, variant=variant)
        assert st == orc.EBREAKDOWN
    B = orc.poisson3d(8, 8, 8)
    st, x, hist, it = orc.cg(B, orc.uniform(0, 0, B.n), tol=1e-12, max_iter=5)
    assert st == orc.ENOTCONV and it == 5 and len(hist) == 6


def test_ir_reaches_fp64_tolerance(orc):
    A = orc.poisson3d(20, 20, 20)
    b = orc.uniform(0, 0, A.n)
    st, x, hist, outer, inner = orc.cg_ir(A, b, tol=1e-10, inner_tol=1e-4)
    assert st == orc.OK and outer >= 2
    r = b - orc.spmv(A, x)
    assert r @ r <= 1e-20 * (b @ b)
    # each refinement gains about the inner tolerance
    assert all(hist[k + 1] < 1e-5 * hist[k] for k in range(outer - 1))


def test_varcoef_generator_properties(orc):
    # the definition is the build's own; pin its mathematics: dims in range,
    # exact symmetry (harmonic means are symmetric bit for bit), positive
    # definiteness, zero row sums away from the box faces, k = exp(N(0,1)).
    logs = []
    for sid in range(6):
        A, f, dims = orc.varcoef3d(11, sid, 5, 9)
        assert all(5 <= d <= 9 for d in dims)
        n = A.n
        assert n == dims[0] * dims[1] * dims[2] and f.size == n
        M = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(n, n))
        assert (M != M.T).nnz == 0
        dense = M.toarray()
        assert np.linalg.eigvalsh(dense).min() > 0
        nx, ny, nz = dims
        iz, iy, ix = np.unravel_index(np.arange(n), (nz, ny, nx))
        inner = (ix > 0) & (ix < nx - 1) & (iy > 0) & (iy < ny - 1) & (iz > 0) & (iz < nz - 1)
        rows = np.asarray(M.sum(axis=1)).ravel()
        assert np.allclose(rows[inner], 0.0, atol=1e-12 * np.abs(M.diagonal()).max())
        assert np.all(rows[~inner] > 0)
        assert f.min() >= -1.0 and f.max() < 1.0
    # a 1x1x1 box: every face is a boundary face, diagonal = 6 k
    A, f, dims = orc.varcoef3d(3, 0, 1, 1)
    assert dims == (1, 1, 1) and A.nnz == 1 and A.val[0] > 0
    # log k ~ N(0,1): recover k from a corner cell's diagonal on large boxes
    ks = []
    for sid in range(30):
        A, f, dims = orc.varcoef3d(5, sid, 1, 1)
        ks.append(np.log(A.val[0] / 6.0))
    ks = np.array(ks)
    assert abs(ks.mean()) < 0.6 and 0.5 < ks.std() < 1.6


def test_varcoef_dims_are_log_uniform(orc):
    # dims only (no matrix): the size query draws only the three lengths
    import ctypes
    L = orc.lib()
    d = np.zeros(3, dtype=np.int32)
    draws = []
    for sid in range(600):
        L.orc_varcoef3d(0, sid, 32, 96, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None, None, None)
        draws += d.tolist()
    draws = np.array(draws)
    assert draws.min() >= 32 and draws.max() <= 96
    # log-uniform: the median is near sqrt(32*97) = 55.7
    assert 50 <= np.median(draws) <= 62


def test_jacobi_smooth_pins(orc):
    import scipy.sparse as sps
    A = orc.poisson3d(6, 5, 4)
    M = sps.csr_matrix((A.val, A.col, A.rowptr))
    rng = np.random.default_rng(8)
    xi = rng.integers(-5, 6, A.n).astype(np.float64)
    b = M @ xi  # integer arithmetic: exact, so xi is a fixed point bit for bit
    assert np.array_equal(orc.jacobi_smooth(A, b, xi, 3, 0.8), xi)
    # one sweep equals the textbook damped Jacobi update
    x0 = rng.standard_normal(A.n)
    b = rng.standard_normal(A.n)
    ref = x0 + 0.7 * (b - M @ x0) / M.diagonal()
    assert np.allclose(orc.jacobi_smooth(A, b, x0, 1, 0.7), ref, rtol=1e-13, atol=1e-13)
    # two sweeps compose
    one = orc.jacobi_smooth(A, b, x0, 1, 0.7)
    assert np.array_equal(orc.jacobi_smooth(A, b, one, 1, 0.7), orc.jacobi_smooth(A, b, x0, 2, 0.7))
    # sweeps reduce the error of an SPD system (omega < 1)
    xs = np.linalg.solve(M.toarray(), b)
    e0 = np.linalg.norm(x0 - xs)
    assert np.linalg.norm(orc.jacobi_smooth(A, b, x0, 20, 0.8) - xs) < 0.5 * e0


def test_ir_last_refinement_stops_at_half_target(orc):
    """The inner tolerance is floored at half the outer target: with a loose
    outer tolerance one refinement suffices and lands between tol/2-ish and tol."""
    A = orc.poisson3d(16, 16, 16)
    b = orc.uniform(1, 0, A.n)
    tol = 1e-3
    st, x, hist, outer, inner = orc.cg_ir(A, b, tol=tol, inner_tol=1e-6)
    assert st == orc.OK and outer == 1
    rel = np.sqrt(hist[1] / hist[0])
    assert rel <= tol
    # without the floor the same inner tolerance would solve to ~1e-6
    assert rel > 1e-5
