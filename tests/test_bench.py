"""The benchmark entry point's contract: one JSON line with the metric,
roofline and CPU-baseline objects (GPU), and a clear refusal of more than
one GPU (CPU)."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.gpu
def test_bench_json_line():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--n", "64", "--niter", "20", "--steps", "2",
                        "--warmup", "1", "--no-cg", "--no-batch", "--no-warmstart", "--cpu-seconds", "0.5"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["unit"] == "GLUP/s" and d["value"] > 0
    assert d["config"]["workload"] == "jacobi3d_7pt_persistent" and d["config"]["nx"] == 64
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    assert rf["traffic"] is None  # counted bytes exist only for the 512^3 x 1000-sweep workload
    assert d["config"]["launch"] == "cooperative"  # a fallback launch is refused, not reported
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]
    assert cb["host"]["nproc"] >= 1 and "model" in cb["host"]
    c1 = cb["config1_2d_1024"]
    assert set(c1) == {f"threads_{cb['cores']}", "threads_1"} or cb["cores"] == 1
    for v in c1.values():
        assert v["norms"] == 100 and v["glups"] > 0
    cg = cb["cg_128_standard"]
    assert cg["iterations"] > 100 and cg["rel_residual"] <= 1e-8
