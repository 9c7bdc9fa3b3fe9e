"""The reference's CG inputs on the CPU: the default random tridiagonal matrix
(genTridiag, CG/src/common.cu:46-76) and the Matrix Market reader
(loadMMSparseMatrix, CG/src/mmio_wrapper.cpp:170-369; base test
CG/src/main.cu:180-186), as the library's cfm_gen_tridiag and cfm_mm_read.

Pins:
* cfm_gen_tridiag against the C library's own rand() called through ctypes
  (an independent source of the values) in the reference's call order, plus
  the structure (symmetric, tridiagonal, nnz = 3n - 2, diagonal dominance);
* cfm_mm_read against the reference's own reader compiled from its sources
  (oracle/_ref/libmmref.so, when built), against the outputs that reader gave
  for the committed fixtures (tests/golden/mm_expected.json), and against
  scipy.io.mmread for the matrix itself.
"""
import ctypes
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from conftest import REPO

GOLDEN = os.path.join(REPO, "tests", "golden")
MM_DIR = os.path.join(GOLDEN, "mm")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libmmref.so")
RAND_MAX = 2147483647


def _libc():
    c = ctypes.CDLL("libc.so.6")
    c.srand.argtypes = [ctypes.c_uint]
    c.rand.restype = ctypes.c_int
    return c


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 1), (1000, 1), (777, 12345)])
def test_tridiag_values_follow_libc_rand(cfm, n, seed):
    libc = _libc()
    libc.srand(seed)
    r = [libc.rand() / RAND_MAX for _ in range(2 * n - 1)]
    # reference order: a00, a01, then per row i >= 1: a(i,i), a(i,i+1) below the last row
    diag = np.empty(n)
    off = np.empty(n - 1)
    diag[0], off[0] = r[0] + 10.0, r[1]
    k = 2
    for i in range(1, n):
        diag[i] = r[k] + 10.0
        k += 1
        if i < n - 1:
            off[i] = r[k]
            k += 1
    want = sp.diags([off, diag, off], [-1, 0, 1], format="csr")
    want.sort_indices()
    rp, col, val = cfm.tridiag(n, seed)
    assert rp.size == n + 1 and col.size == 3 * n - 2
    assert np.array_equal(rp, want.indptr) and np.array_equal(col, want.indices)
    assert np.array_equal(val, want.data)


def test_tridiag_reference_size_structure(cfm):
    n = 20_971_520  # CG/src/main.cu:118-129
    rp, col, val = cfm.tridiag(n)
    assert col.size == (n - 2) * 3 + 4
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    assert (A - A.T).nnz == 0
    d = A.diagonal()
    assert d.min() >= 10.0 and d.max() <= 11.0
    offsum = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert (d - offsum).min() > 8.0  # strictly diagonally dominant, hence SPD


def test_tridiag_leaves_the_process_rand_state_alone(cfm):
    libc = _libc()
    libc.srand(5)
    a = [libc.rand() for _ in range(4)]
    libc.srand(5)
    libc.rand()
    cfm.tridiag(1000)
    b = [libc.rand() for _ in range(3)]
    assert a[1:] == b


def _ref_reader():
    if not os.path.exists(REF_LIB):
        return None
    L = ctypes.CDLL(REF_LIB)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.ref_load_mm.argtypes = [ctypes.c_char_p, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(ip),
                              ctypes.POINTER(ip)]

    def read(path):
        m, n, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        v, rp, cp = dp(), ip(), ip()
        if L.ref_load_mm(path.encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(v),
                         ctypes.byref(rp), ctypes.byref(cp)) != 0:
            return None
        k = nnz.value
        return (np.array([rp[i] for i in range(m.value + 1)], np.int32), np.array([cp[i] for i in range(k)], np.int32),
                np.array([v[i] for i in range(k)]))
    return read


EXPECTED = json.load(open(os.path.join(GOLDEN, "mm_expected.json")))
FILES = sorted(EXPECTED)


def _read(cfm, path):
    try:
        return cfm.mm_read(path)
    except cfm.CfmError as e:
        return e


@pytest.mark.parametrize("name", FILES)
def test_mm_reader_matches_the_reference_reader(cfm, name):
    path = os.path.join(MM_DIR, name)
    got = _read(cfm, path)
    exp = EXPECTED[name]
    if not exp["accepted"]:
        assert isinstance(got, cfm.CfmError) and got.status == cfm.EINVAL
        assert cfm.last_error()
        return
    assert not isinstance(got, Exception), got
    rp, col, val, base = got
    assert np.array_equal(rp, exp["rowptr"]) and np.array_equal(col, exp["col"])
    assert np.array_equal(val, np.array(exp["val"]))
    assert base == exp["rowptr"][0]
    ref = _ref_reader()
    if ref is not None:  # the reference's reader, live
        r = ref(path)
        assert r is not None
        assert np.array_equal(rp, r[0]) and np.array_equal(col, r[1]) and np.array_equal(val, r[2])


@pytest.mark.parametrize("name", [f for f in FILES if EXPECTED[f]["accepted"] and f != "no_last_index.mtx"])
def test_mm_reader_matches_scipy(cfm, name):
    path = os.path.join(MM_DIR, name)
    rp, col, val, base = cfm.mm_read(path)
    n = rp.size - 1
    mine = sp.csr_matrix((val, col - base, rp - base), shape=(n, n)).toarray()
    if name.startswith("base0"):  # scipy reads only 1-based files: shift this one first
        lines = open(path).read().splitlines()
        body = [ln for ln in lines if not ln.startswith("%")][1:]
        ent = np.array([[float(t) for t in ln.split()] for ln in body if ln.strip()])
        want = np.zeros((n, n))
        want[ent[:, 0].astype(int), ent[:, 1].astype(int)] = ent[:, 2]
    else:
        want = scipy.io.mmread(path)
        want = want.toarray() if sp.issparse(want) else np.asarray(want)
    assert np.array_equal(mine, want.astype(np.float64))


def test_mm_reader_base_test_of_the_reference(cfm):
    # 1-based file that never uses index n: the reference's test picks base 0,
    # which shifts every row down by one (row 0 empty); cfm_mm_read keeps that
    rp, col, val, base = cfm.mm_read(os.path.join(MM_DIR, "no_last_index.mtx"))
    assert base == 0 and rp[0] == 0 and rp[1] == 0


def test_malformed_csr_refused_before_any_device_work(cfm):
    # decreasing rowptr, column outside the matrix, rowptr[n] != nnz: every builder
    # (DIA via matrix_format 2, SELL-64) is protected by one up-front check
    n = 4
    good_rp = np.array([0, 2, 4, 6, 8], np.int32)
    col = np.array([0, 1, 0, 1, 2, 3, 2, 3], np.int32)
    val = np.full(8, 2.0)
    cases = [(np.array([0, 2, 1, 6, 8], np.int32), col), (good_rp, np.array([0, 1, 0, 1, 2, 9, 2, 3], np.int32)),
             (np.array([0, 2, 4, 6, 7], np.int32), col), (good_rp, np.array([0, 1, 0, 1, 2, -1, 2, 3], np.int32))]
    for rp, c in cases:
        for fmt in (1, 2):
            p = cfm.cg_params(method=cfm.CG_PIPELINED, matrix_format=fmt)
            h = ctypes.c_void_p()
            st = cfm.lib.cfm_cg_plan_create(n, 8, rp.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p),
                                            val.ctypes.data_as(ctypes.c_void_p), 0, 0, ctypes.byref(p), ctypes.byref(h))
            assert st == cfm.EINVAL and not h.value
            assert "cg:" in cfm.last_error()
