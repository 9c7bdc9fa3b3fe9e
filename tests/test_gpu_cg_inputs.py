"""The reference's own CG calling convention and inputs on the GPU.

* Device-resident CSR handed to the solver, as CG/src/main.cu:188-197 does
  (cudaMalloc + copy, then init(rowptr_d, col_d, val_d, ...)): the plan is
  created with csr_on_device = 1 from buffers the test allocated, 0- and
  1-based, and the solve equals the oracle bit for bit.
* The reference's default matrix (genTridiag at n = 20,971,520,
  CG/src/main.cu:118-129) with its b = 1 (CG/src/common.cu:133-140), Standard
  and Pipelined, against the oracle bit for bit; plus the `cg -tridiag` CLI.
* A Matrix Market file read by the reference's rules (1-based, symmetric
  storage), solved through index_base = 1, against the oracle; plus the
  `cg -matrix_path` CLI on it.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu
CG_BIN = os.path.join(PKG, "bin", "cg")


def _same(g, o):
    st, x, hist, it = g
    ost, ox, ohist, oit = o
    assert st == ost and it == oit
    assert np.array_equal(hist, ohist), np.flatnonzero(hist != ohist)[:5]
    assert np.array_equal(x, ox)


def _solve_device(gpu, rp, col, val, b, base, method, tol, max_iter):
    n, nnz = rp.size - 1, col.size
    p = gpu.cg_params(method=method, tol=tol, max_iter=max_iter)
    with gpu.DeviceBuffer.of(rp.astype(np.int32)) as drp, gpu.DeviceBuffer.of(col.astype(np.int32)) as dcol, \
            gpu.DeviceBuffer.of(val.astype(np.float64)) as dval:
        with gpu.CgPlan(drp, dcol, dval, p, index_base=base, n=n, nnz=nnz) as plan:
            plan.set_rhs(b)
            plan.set_x0(None)
            st, hist, it, t = plan.solve()
            x = plan.x()
    assert t["cooperative"] in (0, 1)
    return st, x, hist, it


@pytest.mark.parametrize("base", [0, 1])
@pytest.mark.parametrize("method", ["standard", "pipelined"])
def test_device_csr_equals_oracle(gpu, orc, base, method):
    rp, col, val = gpu.poisson3d(40, 36, 30)
    b = gpu.uniform(11, 0, rp.size - 1)
    rp, col = rp + base, col + base
    m = gpu.CG_STANDARD if method == "standard" else gpu.CG_PIPELINED
    g = _solve_device(gpu, rp, col, val, b, base, m, 1e-10, 3000)
    assert g[0] == gpu.OK
    _same(g, orc.cg(orc.Csr(rp, col, val, base), b, tol=1e-10, max_iter=3000, variant=method))


def test_device_csr_full_size_128(gpu, orc):
    # the benchmark system (128^3 Poisson, uniform b), handed over as device CSR; Pipelined picks DIA
    rp, col, val = gpu.poisson3d(128, 128, 128)
    b = gpu.uniform(0, 0, rp.size - 1)
    g = _solve_device(gpu, rp, col, val, b, 0, gpu.CG_PIPELINED, 1e-8, 10000)
    assert g[0] == gpu.OK
    _same(g, orc.cg(orc.Csr(rp, col, val), b, tol=1e-8, max_iter=10000, variant="pipelined"))


def test_device_csr_malformed_is_refused(gpu):
    rp = np.array([0, 2, 1, 4], np.int32)  # decreasing row pointer
    col = np.array([0, 1, 1, 2], np.int32)
    val = np.full(4, 2.0)
    p = gpu.cg_params()
    with gpu.DeviceBuffer.of(rp) as drp, gpu.DeviceBuffer.of(col) as dcol, gpu.DeviceBuffer.of(val) as dval:
        with pytest.raises(gpu.CfmError) as e:
            gpu.CgPlan(drp, dcol, dval, p, n=3, nnz=4)
    assert e.value.status == gpu.EINVAL and "rowptr" in str(e.value)


@pytest.mark.parametrize("method", ["standard", "pipelined"])
def test_reference_default_tridiag_equals_oracle(gpu, orc, method):
    n = 20_971_520  # the reference's default CG input (CG/src/main.cu:118-129)
    rp, col, val = gpu.tridiag(n)
    b = np.ones(n)  # the reference's right-hand side (r = 1, x = 0)
    m = gpu.CG_STANDARD if method == "standard" else gpu.CG_PIPELINED
    st, x, hist, it, t = gpu.cg(rp, col, val, b, method=m, tol=1e-8, max_iter=1000)
    assert st == gpu.OK and t["cooperative"] in (0, 1)
    _same((st, x, hist, it), orc.cg(orc.Csr(rp, col, val), b, tol=1e-8, max_iter=1000, variant=method))


def test_reference_tridiag_fixed_iterations(gpu, orc):
    # the reference runs a fixed count (no stopping test): tol 0, 6 iterations
    rp, col, val = gpu.tridiag(1_000_003)
    b = np.ones(rp.size - 1)
    for m, name in ((gpu.CG_STANDARD, "standard"), (gpu.CG_PIPELINED, "pipelined")):
        st, x, hist, it, t = gpu.cg(rp, col, val, b, method=m, tol=0.0, max_iter=6)
        assert it == 6
        _same((st, x, hist, it), orc.cg(orc.Csr(rp, col, val), b, tol=0.0, max_iter=6, variant=name))


def _write_mm_symmetric(path, rp, col, val):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = col <= rows  # lower triangle, 1-based, shuffled
    r, c, v = rows[keep] + 1, col[keep] + 1, val[keep]
    order = np.random.default_rng(3).permutation(r.size)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% 3D Poisson, lower triangle\n")
        f.write(f"{n} {n} {r.size}\n")
        for k in order:
            f.write(f"{r[k]} {c[k]} {float(v[k])!r}\n")


def test_matrix_market_file_equals_oracle(gpu, orc, tmp_path):
    rp0, col0, val0 = gpu.poisson3d(20, 18, 16)
    path = str(tmp_path / "poisson.mtx")
    _write_mm_symmetric(path, rp0, col0, val0)
    rp, col, val, base = gpu.mm_read(path)
    assert base == 1 and np.array_equal(rp - 1, rp0) and np.array_equal(col - 1, col0)
    b = gpu.uniform(2, 0, rp.size - 1)
    for m, name in ((gpu.CG_STANDARD, "standard"), (gpu.CG_PIPELINED, "pipelined")):
        st, x, hist, it, t = gpu.cg(rp, col, val, b, index_base=1, method=m, tol=1e-10, max_iter=2000)
        assert st == gpu.OK
        _same((st, x, hist, it), orc.cg(orc.Csr(rp, col, val, 1), b, tol=1e-10, max_iter=2000, variant=name))


def _cli(args):
    r = subprocess.run([CG_BIN, "-s", "-json"] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")], r.stdout


def test_cli_tridiag_and_matrix_path(gpu, tmp_path):
    recs, out = _cli(["-tridiag", "200000", "-rhs", "ones", "-v", "2,3", "-compare"])
    assert len(recs) == 2 and "Execution time" in out
    for r in recs:
        assert r["status"] == "ok" and r["rel_residual"] <= 1e-8 and r["n"] == 200000 and r["nnz"] == 599998
        assert r["cooperative"] in (0, 1)
    rp0, col0, val0 = gpu.poisson3d(12, 12, 12)
    path = str(tmp_path / "p12.mtx")
    _write_mm_symmetric(path, rp0, col0, val0)
    recs, out = _cli(["-matrix_path", path, "-v", "0,2,3,10", "-compare"])
    assert len(recs) == 4
    for r in recs:
        assert r["status"] == "ok" and r["rel_residual"] <= 1e-8 and r["n"] == 1728
    bad = str(tmp_path / "pattern.mtx")
    with open(bad, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n2 2\n")
    r = subprocess.run([CG_BIN, "-s", "-matrix_path", bad], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "pattern" in r.stderr
