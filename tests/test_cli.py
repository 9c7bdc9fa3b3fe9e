"""The command-line drivers jacobi2D, jacobi3D and cg: the reference's version
tables, flags and output lines (Stencil/jacobi2D/src/main.cu:24-69,
Stencil/jacobi3D/src/main.cu:30-73, CG/src/main.cu:23-278), run as separate
processes.  CPU tests check the tables and argument errors; GPU tests run
every version on small problems with the reference's -compare check."""
import json
import os
import re
import subprocess

import pytest

from conftest import PKG

BIN = os.path.join(PKG, "bin")
JACOBI_VERSIONS = ["Baseline Copy", "Baseline Overlap", "Baseline P2P", "Design 1", "Design 2", "PERKS",
                   "Baseline Copy (No computation)", "Baseline Overlap (No Computation)",
                   "Baseline P2P (No Computation)", "Design 1 (No Computation)", "Design 2 (No Computation)"]
CG_VERSIONS = ["Baseline Discrete Standard", "Baseline Discrete Pipelined", "Single Stream Standard",
               "Single Stream Pipelined", "Single Stream Pipelined Multi Overlap",
               "Single Stream Standard Saxpy Overlap", "Profiling Discrete Standard",
               "Profiling Discrete Pipelined", "Single GPU Discrete Standard", "Single Stream Pipelined Gather",
               "Single Stream Mixed-Precision Refinement"]


def _run(name, *args, timeout=120):
    return subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def _table(stdout):
    return re.findall(r"^(\d+):\t(.*)$", stdout, flags=re.M)


@pytest.mark.parametrize("name,versions", [("jacobi2D", JACOBI_VERSIONS), ("jacobi3D", JACOBI_VERSIONS),
                                           ("cg", CG_VERSIONS)])
def test_version_table_and_bad_version(name, versions):
    r = _run(name, "-v", 99)
    assert r.returncode == 1 and "no such version" in r.stderr
    # without a GPU the run fails after printing the table (with one it succeeds)
    r = _run(name, "-v", 0, "-niter", 1, "-nx", 16, "-ny", 16, "-nz", 16)
    table = _table(r.stdout)
    assert [int(i) for i, _ in table] == list(range(len(versions)))
    assert [t for _, t in table] == versions
    assert "Running " + versions[0] in r.stdout


def test_silent_flag_hides_the_table():
    r = _run("jacobi3D", "-s", "-v", 4, "-niter", 1, "-nx", 16, "-ny", 16, "-nz", 16)
    assert "Versions (select with -v):" not in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("v", range(11))
@pytest.mark.parametrize("name,shape", [("jacobi2D", ("-nx", 130, "-ny", 67)),
                                        ("jacobi3D", ("-nx", 66, "-ny", 37, "-nz", 35))])
def test_jacobi_versions_compare(gpu, name, shape, v):
    r = _run(name, "-v", v, *shape, "-niter", 9, "-compare")
    assert r.returncode == 0, r.stderr
    assert "ERROR" not in r.stderr
    assert re.search(r"^Execution time: +[0-9.]+ s$", r.stdout, flags=re.M)
    assert "Non-persistent kernel" in r.stdout


@pytest.mark.gpu
def test_persistent_run_while_another_process_holds_cooperative_launch(gpu):
    """ROCm grants cooperative launches to one process at a time: once this
    process has made one, the CLI's persistent runs go out as ordinary
    launches of the same co-resident grid and must still succeed."""
    import numpy as np
    out, _, _ = gpu.jacobi(np.ones((9, 70), dtype=np.float64), 3, variant=gpu.PERSISTENT)
    assert np.all(out == 1.0)
    for name, shape in (("jacobi2D", ("-nx", 130, "-ny", 67)), ("jacobi3D", ("-nx", 66, "-ny", 37, "-nz", 35))):
        for v in (3, 5):
            r = _run(name, "-v", v, *shape, "-niter", 9, "-compare")
            assert r.returncode == 0, r.stderr
    for v in (2, 3):  # Single Stream Standard / Pipelined: persistent CG
        r = _run("cg", "-v", v, "-nx", 12, "-ny", 12, "-nz", 12, "-compare-single-gpu")
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [("-bc", "dirichlet", "-init", "sin", "-tblock", 2),
                                   ("-bc", "dirichlet", "-prec", "f32", "-norm", 3),
                                   ("-bc", "periodic", "-norm", 2)])
def test_jacobi3d_options_and_json(gpu, extra):
    r = _run("jacobi3D", "-v", 5, "-nx", 64, "-ny", 40, "-nz", 24, "-niter", 10, "-compare", "-json", *extra)
    assert r.returncode == 0, r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
    d = json.loads(line)
    assert d["version"] == "PERKS" and d["glups"] > 0 and d["block"] > 0
    if "-norm" in extra:
        k = extra[extra.index("-norm") + 1]
        assert len(re.findall(r"^ *\d+, [0-9.e+-]+$", r.stdout, flags=re.M)) == (10 + k - 1) // k


@pytest.mark.gpu
@pytest.mark.parametrize("v", range(11))
def test_cg_versions_compare(gpu, v):
    r = _run("cg", "-v", v, "-nx", 20, "-ny", 18, "-nz", 16, "-compare-single-gpu")
    assert r.returncode == 0, r.stderr
    assert "ERROR" not in r.stderr
    assert re.search(r"^Execution time: +[0-9.]+ s$", r.stdout, flags=re.M)
    m = re.search(r"relative residual: ([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(1)) <= 1e-8


@pytest.mark.gpu
def test_cg_reference_tridiagonal_fixed_iterations(gpu):
    # the reference's default problem: genTridiag with glibc rand(), b = 1, -niter iterations
    r = _run("cg", "-v", 2, "-tridiag", 10000, "-rhs", "ones", "-tol", 0, "-niter", 50, "-compare-single-gpu")
    assert r.returncode == 0, r.stderr
    assert re.search(r"iterations: 50 ", r.stdout)
