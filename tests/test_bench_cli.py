"""The benchmark drivers over the command-line binaries: parsing of the
binaries' output contract (version table, ``Execution time``), size
arguments, and the kernel -> operation attribution of the breakdown (CPU);
end-to-end sweeps on the GPU."""
import csv
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cpu-free-model_amd"))
import bench_cli as bc  # noqa: E402

TABLE = ("Versions (select with -v):\n0:\tBaseline Copy\n1:\tBaseline Overlap\n5:\tPERKS\n\n"
         "Running PERKS\n\nSingle GPU jacobi relaxation: 10 iterations\nExecution time:   0.0123 s\n"
         "Non-persistent kernel - 1 GPU:   0.0456 s, 1 GPUs:   0.0123 s, speedup:     3.71, efficiency:   370.73\n")


def test_parse_version_table_and_times():
    assert bc.parse_versions(TABLE) == [(0, "Baseline Copy"), (1, "Baseline Overlap"), (5, "PERKS")]
    assert bc.parse_exec_times(TABLE) == [0.0123]
    assert bc.parse_exec_times("Execution time:  12.5000 s\nExecution time:   1.0000 s\n") == [12.5, 1.0]


def test_sizes_and_points():
    assert bc.parse_sizes("512", 3) == [[512, 512, 512]]
    assert bc.parse_sizes("64x32,128", 2) == [[64, 32], [128, 128]]
    with pytest.raises(ValueError):
        bc.parse_sizes("4x4x4", 2)
    assert bc.size_args([7, 8, 9]) == ["-nx", "7", "-ny", "8", "-nz", "9"]
    assert bc.interior_points([512, 512, 512]) == 510 ** 3
    assert bc.interior_points([2, 10]) == 0


def test_operation_attribution():
    stats = ('"Name","Calls","TotalDurationNs","AverageNs"\n'
             '"void cfm::k_std_spmv_dot<4>(cfm::CgArgs)",448,16731487,37347.0\n'
             '"void cfm::k_std_update_xr<4>(cfm::CgArgs, double)",448,7900646,17635.3\n'
             '"void cfm::k_reduce<4>(cfm::CgArgs, int, int, double*)",1346,7164000,5322.4\n'
             '"__amd_rocclr_copyBuffer",1356,5374573,3963.5\n'
             '"(anonymous namespace)::copy_probe_kernel(double const*)",6,2300000,383333.3\n')
    ops = bc.aggregate_stats(stats)
    assert ops == {"SpMV s = A p + <p,s> partials": 16731487.0,
                   "x += alpha p, r -= alpha s + <r,r> partials": 7900646.0,
                   "dot-product reduction (one workgroup)": 7164000.0,
                   "scalar read-back to the host": 5374573.0}
    assert bc.operation_of("void cfm::cg_standard_kernel<4>(cfm::CgArgs)") is None


def test_write_table(tmp_path):
    p = tmp_path / "t.csv"
    bc.write_table(str(p), "Version", ["A", "B"], ["64", "128"],
                   {("A", "64"): 1.5, ("A", "128"): float("nan"), ("B", "64"): 2.0, ("B", "128"): 3.25})
    rows = list(csv.reader(open(p)))
    assert rows == [["Version", "64", "128"], ["A", "1.5", ""], ["B", "2", "3.25"]]


@pytest.mark.gpu
def test_stencil_and_cg_sweeps(tmp_path):
    out = tmp_path / "stencil.csv"
    bc.main(["stencil", "--dims", "2", "--sizes", "66x34,130", "--versions", "3,5", "--niter", "8", "--repeat", "1",
             "--out", str(out), "--extra", "-bc", "dirichlet"])
    rows = list(csv.reader(open(out)))
    assert rows[0] == ["Version", "66x34", "130x130"]
    assert [r[0] for r in rows[1:]] == ["Design 1", "PERKS"]
    assert all(float(v) > 0 for r in rows[1:] for v in r[1:])
    glups = list(csv.reader(open(tmp_path / "stencil_glups.csv")))
    assert all(float(v) > 0 for r in glups[1:] for v in r[1:])
    out = tmp_path / "cg.csv"
    bc.main(["cg", "--poisson", "12,16", "--versions", "2,3", "--repeat", "1", "--out", str(out)])
    rows = list(csv.reader(open(out)))
    assert rows[0] == ["Version", "poisson12", "poisson16"]
    assert [r[0] for r in rows[1:]] == ["Single Stream Standard", "Single Stream Pipelined"]
    assert all(float(v) > 0 for r in rows[1:] for v in r[1:])
