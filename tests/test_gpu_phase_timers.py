"""In-place time breakdown of the persistent kernels (phase_timers).

With timers on, every workgroup of the persistent kernel reads its clock at
each phase boundary (SpMV, updates, reductions, barrier waits, set-up).  The
timed build must give the same bits as the default one (and the oracle), and
each workgroup's phases add up to its time in the kernel, so their mean must
match the kernel's HIP-event time within 5 %.  The discrete schedule with
timers times each operation on the host (the reference's profiling versions,
CG/src/profiling/discrete-standard.cu:222-264).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def _solve(gpu, rp, col, val, b, method, mode, timers):
    p = gpu.cg_params(method=method, mode=mode, tol=1e-8, max_iter=10000, phase_timers=timers)
    with gpu.CgPlan(rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, hist, it, t = plan.solve()
        x = plan.x()
        ph = plan.phase_times() if timers else None
    return st, x, hist, it, t, ph


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("method", ["standard", "pipelined", "ir"])
def test_cg_timers_bitexact_and_sum_to_kernel_time(gpu, orc, n, method):
    rp, col, val = gpu.poisson3d(n, n, n)
    b = gpu.uniform(0, 0, rp.size - 1)
    m = {"standard": gpu.CG_STANDARD, "pipelined": gpu.CG_PIPELINED, "ir": gpu.CG_IR_F32}[method]
    st0, x0, h0, it0, t0, _ = _solve(gpu, rp, col, val, b, m, gpu.CG_PERSISTENT, 0)
    st1, x1, h1, it1, t1, ph = _solve(gpu, rp, col, val, b, m, gpu.CG_PERSISTENT, 1)
    assert st0 == st1 == gpu.OK and it0 == it1
    assert np.array_equal(h0, h1) and np.array_equal(x0, x1)
    if method != "ir":
        ost, ox, ohist, oit = orc.cg(orc.Csr(rp, col, val), b, tol=1e-8, max_iter=10000, variant=method)
        assert oit == it1 and np.array_equal(ox, x1)
    assert ph["workgroups"] == t1["grid_blocks"] and not ph["host_timed"]
    phases = sum(ph["ms"].values())
    assert abs(phases - ph["total_ms"]) <= 1e-6 * ph["total_ms"] + 1e-6  # a workgroup's phases tile its span
    assert abs(phases - t1["kernel_ms"]) <= 0.05 * t1["kernel_ms"], (phases, t1["kernel_ms"])
    assert ph["ms"]["barrier"] > 0 and ph["ms"]["reduce"] > 0
    assert (ph["ms"]["spmv"] if method != "ir" else ph["ms"]["spmv"] + ph["ms"]["update"]) > 0


@pytest.mark.parametrize("method", ["standard", "pipelined"])
def test_cg_discrete_host_timed_operations(gpu, orc, method):
    rp, col, val = gpu.poisson3d(40, 40, 40)
    b = gpu.uniform(1, 0, rp.size - 1)
    m = gpu.CG_STANDARD if method == "standard" else gpu.CG_PIPELINED
    st, x, hist, it, t, ph = _solve(gpu, rp, col, val, b, m, gpu.CG_DISCRETE, 1)
    ost, ox, ohist, oit = orc.cg(orc.Csr(rp, col, val), b, tol=1e-8, max_iter=10000, variant=method)
    assert it == oit and np.array_equal(x, ox) and np.array_equal(hist, ohist)
    assert ph["host_timed"] and ph["ms"]["spmv"] > 0 and ph["ms"]["reduce"] > 0
    assert sum(ph["ms"].values()) <= ph["total_ms"] * 1.001


def test_cg_plan_without_timers_has_no_breakdown(gpu):
    rp, col, val = gpu.poisson3d(8, 8, 8)
    with gpu.CgPlan(rp, col, val, gpu.cg_params()) as plan:
        with pytest.raises(gpu.CfmError):
            plan.phase_times()


@pytest.mark.parametrize("shape,prec", [((256, 256, 256), "f64"), ((4096, 4096), "f64"), ((160, 96, 600), "f64"),
                                        ((256, 256, 256), "f32")])
def test_jacobi_timers_bitexact_and_sum_to_kernel_time(gpu, shape, prec):
    nx, ny = shape[-1], shape[-2]
    nz = shape[0] if len(shape) == 3 else 1
    pr = gpu.F64 if prec == "f64" else gpu.F32
    init = gpu.init_field(nx, ny, nz, prec=pr)
    out = {}
    for timers in (0, 1):
        p = gpu.jacobi_params(nx, ny, nz, niter=200, norm_every=50, prec=pr, bc=gpu.BC_DIRICHLET,
                              phase_timers=timers)
        with gpu.JacobiPlan(p) as plan:
            plan.load(init)
            norms, t = plan.run()
            out[timers] = (plan.download(), norms, t.as_dict(), plan.phase_times() if timers else None)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    ph, t = out[1][3], out[1][2]
    assert ph["workgroups"] == t["grid_blocks"]
    phases = sum(ph["ms"].values())
    assert abs(phases - ph["total_ms"]) <= 1e-6 * ph["total_ms"] + 1e-6
    assert abs(phases - t["kernel_ms"]) <= 0.05 * t["kernel_ms"], (phases, t["kernel_ms"])
    assert ph["ms"]["spmv"] > 0 and ph["ms"]["barrier"] > 0
    if len(shape) == 3:  # (a 4096^2 plane takes ~50 us per pass: there the barrier wait can be the larger)
        assert ph["ms"]["spmv"] > ph["ms"]["barrier"]


def test_cli_phase_timers_json(gpu):
    r = subprocess.run([os.path.join(PKG, "bin", "cg"), "-s", "-json", "-phase_timers", "-nx", "48", "-v", "2,3,6,10"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    recs = [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(recs) == 4 and r.stdout.count("phases (ms") == 4
    for rec in recs:
        ph = rec["phases"]
        assert rec["status"] == "ok" and ph["total_ms"] > 0 and ph["spmv_ms"] > 0
    assert recs[2]["phases"]["host_timed"] == 1 and recs[0]["phases"]["host_timed"] == 0
