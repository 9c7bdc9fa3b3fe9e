"""Benchmark drivers over the jacobi2D / jacobi3D / cg command-line binaries.

The reference drives its binaries from Python and scrapes their stdout
(Scripts/bench.py:80-169, CG/scripts/measure_runtime.py, and
CG/scripts/measure_operation_breakdown.py for the per-operation split of the
"Profiling Discrete" versions).  This module does the same for one MI355X,
with the same output contract of the binaries: the version table
(``N:\\tName`` lines), ``Execution time: %8.4f s``.

    python bench_cli.py stencil --dims 3 --sizes 256x256x256,512x512x512 --versions 3,5 --out stencil.csv
    python bench_cli.py cg --poisson 64,128 [--matrix A.mtx ...] --versions 0,1,2,3,10 --out cg.csv
    python bench_cli.py breakdown --poisson 128 --versions 6,7 --out breakdown.csv

``stencil`` writes seconds per (version, size) and, beside it, the same table
in GLUP/s (the reference's "gcell" table: interior points x iterations / s).
``cg`` writes seconds per (version, matrix).  ``breakdown`` runs each CG
version once under ``rocprofv3 --kernel-trace --stats`` and attributes the
kernel time to the solver's operations: the discrete schedules launch one
kernel per operation, so the per-kernel totals are the operation breakdown.
The binaries are run as child processes; nothing here touches the GPU itself.
"""
import argparse
import csv
import math
import os
import re
import subprocess
import sys
import tempfile

_HERE = os.path.dirname(os.path.abspath(__file__))
BIN_DIR = os.path.join(_HERE, "bin")

VERSION_RE = re.compile(r"^(\d+):\s+(.*?)\s*$", re.MULTILINE)
EXEC_TIME_RE = re.compile(r"^Execution time:\s+([0-9.]+) s$", re.MULTILINE)

# kernel name (prefix after "cfm::") -> operation, for the discrete CG schedules
OPERATIONS = (
    ("k_std_spmv_dot", "SpMV s = A p + <p,s> partials"),
    ("k_std_update_xr", "x += alpha p, r -= alpha s + <r,r> partials"),
    ("k_std_update_p", "p = r + beta p"),
    ("k_std_init", "initial residual"),
    ("k_pipe_iterate", "q = A w + fused z,s,p,x,r,w updates + <r,r>,<r,w> partials"),
    ("k_pipe_init_r", "initial residual"),
    ("k_pipe_init_w", "w = A r + initial dots"),
    ("k_reduce", "dot-product reduction (one workgroup)"),
    ("__amd_rocclr_copyBuffer", "scalar read-back to the host"),
)


def parse_versions(text):
    """[(index, name)] from a binary's version table."""
    return [(int(i), name) for i, name in VERSION_RE.findall(text)]


def parse_exec_times(text):
    """Every ``Execution time`` value printed (seconds)."""
    return [float(v) for v in EXEC_TIME_RE.findall(text)]


def parse_sizes(spec, dims):
    out = []
    for item in spec.split(","):
        parts = [int(p) for p in item.lower().split("x")]
        if len(parts) == 1:
            parts = parts * dims
        if len(parts) != dims:
            raise ValueError(f"size {item!r} does not have {dims} dimensions")
        out.append(parts)
    return out


def size_args(size):
    return [a for flag, v in zip(("-nx", "-ny", "-nz"), size) for a in (flag, str(v))]


def interior_points(size):
    return math.prod(max(v - 2, 0) for v in size)


def _run(args, timeout):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def version_table(binary, probe_args, timeout=120):
    """The binary's version table, printed by a tiny run of version 0."""
    _, out, _ = _run([binary, "-v", "0", *probe_args], timeout)
    table = parse_versions(out)
    if not table:
        raise RuntimeError(f"{binary} printed no version table")
    return dict(table)


def time_run(args, repeat, timeout):
    """Best-of-`repeat` execution time (seconds); NaN when every run fails,
    so one failing configuration does not end the sweep."""
    best = math.nan
    for _ in range(repeat):
        code, out, _ = _run(args, timeout)
        t = parse_exec_times(out)
        if code == 0 and t:
            best = t[-1] if math.isnan(best) else min(best, t[-1])
    return best


def write_table(path, corner, rows, columns, cells):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow([corner, *columns])
        for r in rows:
            w.writerow([r, *[("" if math.isnan(cells[(r, c)]) else f"{cells[(r, c)]:.6g}") for c in columns]])


def stencil(args):
    binary = os.path.join(BIN_DIR, f"jacobi{args.dims}D")
    sizes = parse_sizes(args.sizes, args.dims)
    names = version_table(binary, ["-niter", "1", *size_args([16] * args.dims)])
    versions = [int(v) for v in args.versions.split(",")] if args.versions else sorted(names)
    cols = ["x".join(map(str, s)) for s in sizes]
    secs, glups = {}, {}
    for v in versions:
        for s, c in zip(sizes, cols):
            cmd = [binary, "-s", "-v", str(v), "-niter", str(args.niter), *size_args(s), *args.extra]
            t = time_run(cmd, args.repeat, args.timeout)
            secs[(names[v], c)] = t
            glups[(names[v], c)] = interior_points(s) * args.niter / t / 1e9 if t == t and t > 0 else math.nan
            print(f"{names[v]} {c}: {t:.4f} s ({glups[(names[v], c)]:.1f} GLUP/s)", file=sys.stderr, flush=True)
    rows = [names[v] for v in versions]
    write_table(args.out, "Version", rows, cols, secs)
    root, ext = os.path.splitext(args.out)
    write_table(f"{root}_glups{ext or '.csv'}", "Version", rows, cols, glups)


def _cg_inputs(args):
    inputs = [(f"poisson{n}", ["-nx", str(n)]) for n in (int(v) for v in args.poisson.split(","))] \
        if args.poisson else []
    inputs += [(os.path.splitext(os.path.basename(m))[0], ["-matrix_path", m]) for m in args.matrix]
    if not inputs:
        raise ValueError("no input: give --poisson and/or --matrix")
    return inputs


def cg(args):
    binary = os.path.join(BIN_DIR, "cg")
    names = version_table(binary, ["-nx", "4", "-niter", "1"])
    versions = [int(v) for v in args.versions.split(",")] if args.versions else sorted(names)
    inputs = _cg_inputs(args)
    cols = [n for n, _ in inputs]
    secs = {}
    for v in versions:
        for name, inp in inputs:
            cmd = [binary, "-s", "-v", str(v), "-niter", str(args.niter), "-tol", str(args.tol), *inp, *args.extra]
            secs[(names[v], name)] = time_run(cmd, args.repeat, args.timeout)
            print(f"{names[v]} {name}: {secs[(names[v], name)]:.4f} s", file=sys.stderr, flush=True)
    write_table(args.out, "Version", [names[v] for v in versions], cols, secs)


def operation_of(kernel_name):
    short = kernel_name.split("(")[0].replace("void ", "").replace("cfm::", "").split("<")[0].strip()
    for prefix, op in OPERATIONS:
        if short == prefix:
            return op
    return None


def aggregate_stats(stats_csv_text):
    """{operation: total ns} from a rocprofv3 kernel-stats CSV."""
    out = {}
    for row in csv.DictReader(stats_csv_text.splitlines()):
        op = operation_of(row["Name"])
        if op is not None:
            out[op] = out.get(op, 0.0) + float(row["TotalDurationNs"])
    return out


def breakdown(args):
    binary = os.path.join(BIN_DIR, "cg")
    names = version_table(binary, ["-nx", "4", "-niter", "1"])
    versions = [int(v) for v in args.versions.split(",")]
    inputs = _cg_inputs(args)
    table = {}
    for v in versions:
        for name, inp in inputs:
            with tempfile.TemporaryDirectory() as d:
                cmd = ["rocprofv3", "--kernel-trace", "--stats", "-d", d, "-o", "run", "--output-format", "csv",
                       "--", binary, "-s", "-v", str(v), "-niter", str(args.niter), "-tol", str(args.tol), *inp]
                # rocprofv3 may end with a signal after the program has finished; the CSV files are complete
                _run(cmd, args.timeout)
                path = None
                for root, _, files in os.walk(d):
                    for f in files:
                        if f.endswith("kernel_stats.csv"):
                            path = os.path.join(root, f)
                if path is None:
                    raise RuntimeError(f"no kernel statistics for version {v} on {name}")
                with open(path) as f:
                    ops = aggregate_stats(f.read())
            total = sum(ops.values())
            for op, ns in ops.items():
                table[(names[v], name, op)] = (ns / 1e6, ns / total if total else math.nan)
            print(f"{names[v]} {name}: {total / 1e6:.3f} ms in {len(ops)} operations", file=sys.stderr, flush=True)
    with open(args.out, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Version", "Input", "Operation", "ms", "fraction"])
        for (ver, inp, op), (ms, frac) in table.items():
            w.writerow([ver, inp, op, f"{ms:.4f}", f"{frac:.4f}"])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stencil")
    s.add_argument("--dims", type=int, choices=(2, 3), default=3)
    s.add_argument("--sizes", default="512")
    s.add_argument("--niter", type=int, default=1000)
    c = sub.add_parser("cg")
    b = sub.add_parser("breakdown")
    for p in (c, b):
        p.add_argument("--poisson", default="128", help="comma-separated grid edges of 7-point Poisson matrices")
        p.add_argument("--matrix", action="append", default=[], help="Matrix Market file (repeatable)")
        p.add_argument("--niter", type=int, default=10000)
        p.add_argument("--tol", type=float, default=1e-8)
    for p, out in ((s, "stencil.csv"), (c, "cg.csv"), (b, "breakdown.csv")):
        p.add_argument("--versions", default="6,7" if p is b else None,
                       help="comma-separated version indices (default: all; breakdown: 6,7)")
        p.add_argument("--repeat", type=int, default=3)
        p.add_argument("--timeout", type=float, default=600.0)
        p.add_argument("--out", default=out)
        p.add_argument("--extra", nargs=argparse.REMAINDER, default=[], help="further flags for the binary")
    args = ap.parse_args(argv)
    {"stencil": stencil, "cg": cg, "breakdown": breakdown}[args.cmd](args)


if __name__ == "__main__":
    main()
