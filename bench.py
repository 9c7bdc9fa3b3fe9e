"""Benchmark: 3D 7-point FP64 Jacobi, 512^3, 1000 sweeps per step, one
persistent cooperative kernel per step (the single-GPU Jacobi configuration),
plus CG time-to-1e-8 on the 128^3 Poisson system and the batched
variable-coefficient CG (256 systems, one launch) against plain GPU CG as
side figures.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-cg] [--no-batch]

Prints ONE JSON line.  `value` is GLUP/s (interior lattice updates per second)
over the K timed steps, with the field resident in HBM before timing starts.
The sweep kernel makes two sweeps per pass over HBM (the intermediate sweep
stays on chip), so `roofline.achieved` is the algorithmic HBM traffic of one
launch -- 16 B per interior point per pass (read a, write a_new), i.e. 8 B
per lattice update -- divided by the mean HIP-event duration of the launch.
`roofline.effective_gbs` restates the rate at 16 B per lattice update (the
traffic of one sweep per pass), the per-sweep roofline.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cpu-free-model_amd"))

PEAK_HBM_GBS = 8000.0
BYTES_PER_LUP_F64 = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--niter", type=int, default=1000)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--chunks", type=int, default=0)
    ap.add_argument("--time-block", type=int, default=0, help="sweeps per HBM pass (3D): 0 auto, 1, 2")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0 auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--no-warmstart", action="store_true")
    ap.add_argument("--warmstart-systems", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 1 GiB streaming-copy probe (counter calibration) and exit")
    ap.add_argument("--profile-batch", action="store_true",
                    help="profiling run: the copy probe, then one batched solve of the 256-system "
                         "variable-coefficient workload (default shape, one dispatch); prints its figures and exits")
    ap.add_argument("--profile-unet", action="store_true",
                    help="profiling run of the warm-start network at 32 x 64^3: forward passes and training "
                         "steps (for rocprofv3 --kernel-trace); prints their wall times and exits")
    ap.add_argument("--profile-cg", action="store_true",
                    help="profiling run: the copy probe, then one 128^3 Poisson solve per CG method and "
                         "schedule (one dispatch each); prints the iteration counts and exits")
    return ap.parse_args()


def host_description():
    """lscpu fields of the host whose cores time the CPU baseline."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    want = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for ln in out.splitlines():
        k, _, v = ln.partition(":")
        if k.strip() in want:
            v = v.strip()
            info[want[k.strip()]] = int(v) if v.isdigit() else v
    return info


def cpu_baseline(n, target_s):
    """The oracle (C, OpenMP; the CPU restatement of the reference's host loops)
    on the host cores, as a reported baseline:
      * value: a bounded sample of the benchmark workload itself, 3D 7-point
        FP64 sweeps of the n^3 field, all host threads (the fixed cost of a call
        -- array copies, first touch -- removed by differencing two runs);
      * config1_2d_1024: the reference's CPU-runnable case in full, 2D 1024^2
        FP64, 10000 sweeps, update norm every 100, on all threads and on 1;
      * cg_128_standard: Standard CG to 1e-8 on the 128^3 Poisson system."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import numpy as np
    cores = oracle.lib().orc_get_threads()
    init = oracle.init3d(n, n, n)

    def timed(k):
        t0 = time.perf_counter()
        oracle.jacobi(init, k, oracle.BC_DIRICHLET)
        return time.perf_counter() - t0

    t1, t3 = timed(1), timed(3)
    per_sweep = max((t3 - t1) / 2, 1e-4)
    sweeps = int(max(4, min(400, target_s / per_sweep)))
    tn = timed(sweeps)
    dt = max(tn - t1, 1e-9)
    lups = (n - 2) ** 3 * (sweeps - 1)
    del init
    res = {"value": round(lups / dt / 1e9, 3), "unit": "GLUP/s", "cores": cores, "kind": "port",
           "sample": f"{sweeps} vs 1 sweeps of the {n}^3 FP64 field (C oracle, OpenMP {cores} threads): "
                     f"{dt:.1f} s for {sweeps - 1} sweeps",
           "host": host_description()}
    # config 1: 2D 1024^2 FP64, 10000 sweeps, norms every 100 (the oracle's host loop, timed whole)
    init2 = oracle.init2d(1024, 1024)
    c1 = {}
    for threads in (cores, 1):
        prev = oracle.set_threads(threads)
        t0 = time.perf_counter()
        _, norms = oracle.jacobi(init2, 10000, oracle.BC_DIRICHLET, 100)
        el = time.perf_counter() - t0
        oracle.set_threads(prev)
        lup = 1022 * 1022 * 10000
        c1[f"threads_{threads}"] = {"seconds": round(el, 3), "glups": round(lup / el / 1e9, 3),
                                    "gbs_at_16B": round(16 * lup / el / 1e9, 1), "norms": len(norms)}
    res["config1_2d_1024"] = c1
    # CG: the oracle's Standard CG (CG/src/common.cu:399-538 restated) on the 128^3 Poisson system
    A = oracle.poisson3d(128, 128, 128)
    b = oracle.uniform(0, 0, A.n)
    t0 = time.perf_counter()
    st, x, hist, it = oracle.cg(A, b, tol=1e-8, max_iter=10000, variant="standard")
    el = time.perf_counter() - t0
    res["cg_128_standard"] = {"seconds_to_1e-8": round(el, 3), "iterations": int(it),
                              "ms_per_iter": round(1000 * el / max(it, 1), 3), "threads": cores,
                              "rel_residual": float(f"{np.sqrt(hist[it] / hist[0]):.3e}")}
    return res


def cg_figures(cfm):
    """Time-to-1e-8 of the 128^3 Poisson system per method, and the FP32-in-FP64
    refinement's solution error against the FP64 Standard CG solution (which
    equals the oracle's bit for bit) with both true residuals."""
    import numpy as np
    import scipy.sparse as sps
    out = {}
    rp, col, val = cfm.poisson3d(128, 128, 128)
    b = cfm.uniform(0, 0, rp.size - 1)
    M = sps.csr_matrix((val, col, rp))
    xs = {}
    for name, method, mode in (("standard", cfm.CG_STANDARD, cfm.CG_PERSISTENT),
                               ("pipelined", cfm.CG_PIPELINED, cfm.CG_PERSISTENT),
                               ("ir_f32", cfm.CG_IR_F32, cfm.CG_PERSISTENT),
                               ("standard_discrete", cfm.CG_STANDARD, cfm.CG_DISCRETE),
                               ("pipelined_discrete", cfm.CG_PIPELINED, cfm.CG_DISCRETE)):
        p = cfm.cg_params(method=method, mode=mode, tol=1e-8, max_iter=10000)
        with cfm.CgPlan(rp, col, val, p) as plan:
            plan.set_rhs(b)
            plan.set_x0(None)
            plan.solve()  # warm-up
            best = None
            for _ in range(3):
                st, hist, it, t = plan.solve()
                if best is None or t["kernel_ms"] < best[1]["kernel_ms"]:
                    best = (it, t, st)
            it, t, st = best
            out[name] = {"ms_to_1e-8": round(t["kernel_ms"], 3), "iterations": it,
                         "outer": t["outer_iterations"], "status": cfm.status_str(st),
                         "us_per_iter": round(1000 * t["kernel_ms"] / max(it, 1), 2)}
            if name in ("standard", "ir_f32"):
                xs[name] = plan.x()
    x64, x32 = xs["standard"], xs["ir_f32"]
    nb = float(np.linalg.norm(b))
    out["ir_f32"]["rel_err_vs_f64"] = float(f"{np.linalg.norm(x32 - x64) / np.linalg.norm(x64):.3e}")
    out["ir_f32"]["true_rel_residual"] = float(f"{np.linalg.norm(b - M @ x32) / nb:.3e}")
    out["standard"]["true_rel_residual"] = float(f"{np.linalg.norm(b - M @ x64) / nb:.3e}")
    return out


def cg_profile_run(cfm):
    """One solve per CG method on the 128^3 Poisson system, for counter runs:
    the persistent kernels are one dispatch per solve, the discrete schedule one
    dispatch per phase.  Returns the iterations and the matrix storage used."""
    rp, col, val = cfm.poisson3d(128, 128, 128)
    b = cfm.uniform(0, 0, rp.size - 1)
    out = {"n": int(rp.size - 1), "nnz": int(rp[-1])}
    for name, method, mode in (("standard", cfm.CG_STANDARD, cfm.CG_PERSISTENT),
                               ("pipelined", cfm.CG_PIPELINED, cfm.CG_PERSISTENT),
                               ("ir_f32", cfm.CG_IR_F32, cfm.CG_PERSISTENT),
                               ("standard_discrete", cfm.CG_STANDARD, cfm.CG_DISCRETE),
                               ("pipelined_discrete", cfm.CG_PIPELINED, cfm.CG_DISCRETE)):
        p = cfm.cg_params(method=method, mode=mode, tol=1e-8, max_iter=10000)
        with cfm.CgPlan(rp, col, val, p) as plan:
            plan.set_rhs(b)
            plan.set_x0(None)
            st, hist, it, t = plan.solve()
            out[name] = {"iterations": it, "outer": t["outer_iterations"], "ms": round(t["kernel_ms"], 3),
                         "status": cfm.status_str(st)}
    return out


def jacobi_f32_figures(cfm, shape, niter):
    """FP32 sweep (FP64 norms) against the FP64 sweep, whose field equals the
    oracle's bit for bit: rate and the field's error after niter sweeps, on one
    field shape ((nz, ny, nx) 3D or (ny, nx) 2D), Dirichlet, default kernels."""
    import numpy as np
    nx, ny = shape[-1], shape[-2]
    nz = shape[0] if len(shape) == 3 else 1
    interior = (nx - 2) * (ny - 2) * (nz - 2 if nz > 1 else 1)
    res = {}
    fields = {}
    for prec, name in ((cfm.F64, "f64"), (cfm.F32, "f32")):
        p = cfm.jacobi_params(nx, ny, nz, niter=niter, prec=prec, bc=cfm.BC_DIRICHLET, variant=cfm.PERSISTENT)
        with cfm.JacobiPlan(p) as plan:
            plan.load(cfm.init_field(nx, ny, nz, cfm.INIT_SIN, prec))
            plan.run()
            plan.load(cfm.init_field(nx, ny, nz, cfm.INIT_SIN, prec))
            _, t = plan.run()
            fields[name] = plan.download()
        res[name] = {"glups": round(interior * niter / t.kernel_ms / 1e6, 1), "ms": round(t.kernel_ms, 2),
                     "passes": t.outer_iterations}
    d = fields["f32"].astype(np.float64) - fields["f64"]
    res["f32"]["max_abs_err_vs_f64"] = float(f"{np.abs(d).max():.3e}")
    res["f32"]["rel_l2_err_vs_f64"] = float(f"{np.linalg.norm(d) / np.linalg.norm(fields['f64']):.3e}")
    return res


def cg_batch_figures(cfm, nsys, lmin=32, lmax=96, seed=2024):
    """Batched CG workload: nsys variable-coefficient systems -div(k grad u) = f
    (edge lengths log-uniform in [lmin, lmax] per axis, k lognormal(0,1)), all
    solved to 1e-8 in one persistent launch, against plain GPU CG solving them
    one launch each (same Pipelined method; kernel time summed)."""
    import numpy as np
    systems = [cfm.varcoef3d(seed, sid, lmin, lmax)[:4] for sid in range(nsys)]
    srp, rp, col, val = cfm.block_diag([s_[:3] for s_ in systems])
    b = np.concatenate([s_[3] for s_ in systems])
    p = cfm.cg_params(method=cfm.CG_PIPELINED, tol=1e-8, max_iter=20000)
    with cfm.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        blk, grid = plan.tune(trial_iters=20)
        fmt = plan.matrix_format()
        plan.solve()
        st, iters, stat, hist, t = plan.solve()
    batched_ms = t["kernel_ms"]
    plain_ms, plain_iters = 0.0, 0
    for s_ in systems:
        with cfm.CgPlan(s_[0], s_[1], s_[2], cfm.cg_params(method=cfm.CG_PIPELINED, tol=1e-8,
                                                           max_iter=20000)) as sp:
            sp.set_rhs(s_[3])
            sp.set_x0(None)
            sp.solve()
            st2, h2, it2, t2 = sp.solve()
            plain_ms += t2["kernel_ms"]
            plain_iters += it2
    return {"systems": nsys, "rows": int(srp[-1]), "edge": [lmin, lmax], "status": cfm.status_str(st),
            "batched_ms_to_1e-8": round(batched_ms, 2), "plain_ms_to_1e-8": round(plain_ms, 2),
            "speedup": round(plain_ms / batched_ms, 3), "max_iters": int(iters.max()),
            "mean_iters": round(float(iters.mean()), 1), "tuned_block": blk, "tuned_grid": grid,
            "tuned_format": fmt}


def cg_batch_profile_run(cfm, nsys, lmin=32, lmax=96, seed=2024):
    """One batched solve of the cg_batch_figures workload with the default
    workgroup shape (a single dispatch, for counter runs).  Returns the rows,
    the stored entries and the per-system iteration counts' sum, from which the
    algorithmic traffic of the solve follows."""
    import numpy as np
    systems = [cfm.varcoef3d(seed, sid, lmin, lmax)[:4] for sid in range(nsys)]
    srp, rp, col, val = cfm.block_diag([s_[:3] for s_ in systems])
    b = np.concatenate([s_[3] for s_ in systems])
    p = cfm.cg_params(method=cfm.CG_PIPELINED, tol=1e-8, max_iter=20000)
    with cfm.CgBatchPlan(srp, rp, col, val, p) as plan:
        plan.set_rhs(b)
        plan.set_x0(None)
        st, iters, stat, hist, t = plan.solve()
    rows = np.diff(srp)
    nnz = np.array([s_[0][-1] for s_ in systems], dtype=np.int64)
    return {"systems": nsys, "rows": int(srp[-1]), "nnz": int(nnz.sum()), "status": cfm.status_str(st),
            "ms": round(t["kernel_ms"], 3), "max_iters": int(iters.max()), "sum_iters": int(iters.sum()),
            "row_iterations": int((rows * iters).sum()), "nnz_iterations": int((nnz * iters).sum()),
            "grid": t["grid_blocks"], "block": t["block_threads"]}


def unet_profile_run(cfm, n=64, batch=32, reps=5):
    """The warm-start network's forward pass and Adam training step on a
    batch of 32 random 64^3 inputs (He-uniform weights): wall time per call,
    host transfers of the input, output and target included; the kernel
    times come from a kernel trace of this run."""
    import numpy as np
    rng = np.random.default_rng(5)
    x = rng.standard_normal((batch, 2, n ** 3)).astype(np.float32)
    y = rng.standard_normal((batch, n ** 3)).astype(np.float32)
    out = {"n": n, "batch": batch}
    with cfm.UNet(n, batch, params=cfm.unet_init_params(1)) as net:
        for name, call in (("forward", lambda: net.forward(x)), ("train_step", lambda: net.train_step(x, y, 1e-4))):
            call()
            cfm.lib.cfm_device_synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            cfm.lib.cfm_device_synchronize()
            out[f"{name}_wall_ms"] = round(1000 * (time.perf_counter() - t0) / reps, 2)
    return out


def hbm_ceiling_sweep(cfm, kinds=("copy", "read", "write"), grids=(256, 1024, 4096, 16384, 65536),
                      unrolls=(1, 4), nts=(False, True), nbytes=1 << 30):
    """HBM ceiling over probe shapes (1 GiB per buffer, best of 5 launches each): streaming copy
    (read + write bytes), read only and write only; workgroups of 256 lanes, grid-strided 16-byte
    vectors, 1 or 4 in flight per lane, plain or non-temporal.  Returns every shape's GB/s and
    the best per kind."""
    shapes, best = [], {}
    for kind in kinds:
        for grid in grids:
            for unroll in unrolls:
                for nt in nts:
                    gbs, _ = cfm.bandwidth_probe_shape(kind, grid, unroll, nt, nbytes, 5)
                    shapes.append({"kind": kind, "grid": grid, "unroll": unroll, "nt": nt, "gbs": round(gbs, 1)})
                    if gbs > best.get(kind, {"gbs": 0.0})["gbs"]:
                        best[kind] = shapes[-1]
    return {"hbm_ceiling_bytes": nbytes, "best": best, "shapes": shapes}


def warmstart_figures(nsys):
    """Learned warm start on held-out 64^3 variable-coefficient systems: one
    batched plan solved from x0 = 0 and from the trained network's guess
    (inference + 2 damped Jacobi sweeps + guard), both to 1e-8, plus plain
    single-system GPU CG on a 32-system sample."""
    import warmstart as ws
    params, meta = ws.load()
    r = ws.evaluate(params, meta, systems=nsys, group=nsys, log=lambda *_: None)
    r["trained_on_systems"] = meta["systems"]
    return r


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU runs are not part of this benchmark)", file=sys.stderr)
        sys.exit(2)
    import cfm

    if args.profile_batch:
        gbs, ms = cfm.bandwidth_probe(1 << 30, 5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "copy_gbs": round(gbs, 1)}), flush=True)
        print(json.dumps(cg_batch_profile_run(cfm, args.batch_systems)), flush=True)
        return
    if args.profile_unet:
        print(json.dumps(unet_profile_run(cfm)), flush=True)
        return
    if args.profile_cg:
        gbs, ms = cfm.bandwidth_probe(1 << 30, 5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "copy_gbs": round(gbs, 1)}), flush=True)
        print(json.dumps(cg_profile_run(cfm)), flush=True)
        return
    if args.copy_probe:
        gbs, ms = cfm.bandwidth_probe(1 << 30, 5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "copy_gbs": round(gbs, 1), "best_ms": round(ms, 4)}),
              flush=True)
        print(json.dumps(hbm_ceiling_sweep(cfm)), flush=True)
        return
    n, niter = args.n, args.niter
    dev = cfm.device_info(0)
    init = cfm.init_field(n, n, n, cfm.INIT_SIN, cfm.F64)
    p = cfm.jacobi_params(n, n, n, niter=niter, prec=cfm.F64, bc=cfm.BC_DIRICHLET, variant=cfm.PERSISTENT,
                          tile_rows=args.tile_rows, grid_blocks=args.grid, chunks=args.chunks,
                          time_block=args.time_block, block_threads=args.block)
    plan = cfm.JacobiPlan(p)
    plan.load(init)
    del init
    for _ in range(args.warmup):
        plan.run()
    cfm.lib.cfm_device_synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, t = plan.run()
        kernel_ms.append(t.kernel_ms)
    cfm.lib.cfm_device_synchronize()
    wall = time.perf_counter() - t0
    grid = t.grid_blocks
    block = t.block_threads
    plan.close()
    if t.cooperative != 1:
        # an ordinary launch does not guarantee co-residency: not a valid headline
        print("bench.py: the persistent kernel went out as an ordinary launch (another process held the "
              "cooperative-launch grant); no figure is reported for this run", file=sys.stderr)
        sys.exit(3)

    interior = (n - 2) ** 3
    lups = interior * niter * args.steps
    passes = t.outer_iterations  # HBM passes per launch (two sweeps each when time-blocked)
    value = lups / wall / 1e9
    mean_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    achieved = BYTES_PER_LUP_F64 * interior * passes / mean_kernel_s / 1e9
    effective = BYTES_PER_LUP_F64 * interior * niter / mean_kernel_s / 1e9

    traffic = None  # counted HBM bytes per launch, from the committed counter runs of this exact workload
    tpath = os.path.join(REPO, "profiles", "jacobi3d_t2_f64_512_traffic.json" if passes < niter
                         else "jacobi3d_f64_512_traffic.json")
    if n == 512 and niter == 1000 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    line = {
        "metric": "3D Jacobi GLUP/s + HBM %peak; CG time-to-1e-8 (learned/batched vs plain)",
        "value": round(value, 2),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (sin*sin Dirichlet x faces, zero interior)",
        "config": {"workload": "jacobi3d_7pt_persistent", "nx": n, "ny": n, "nz": n, "sweeps_per_step": niter,
                   "bc": "dirichlet", "grid_blocks": grid, "block_threads": block, "device": dev["name"],
                   "parallelism": "1 GPU", "launch": "cooperative"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": ("cfm::jacobi3d_t2_kernel (FP64, 7-point, two sweeps per HBM pass)" if passes < niter
                                else "cfm::jacobi3d_wg_kernel (FP64, 7-point, one sweep per pass)"),
                     "sweeps_per_launch": niter, "passes_per_launch": passes,
                     "bytes_per_launch": BYTES_PER_LUP_F64 * interior * passes,
                     "effective_gbs": round(effective, 1), "effective_frac": round(effective / PEAK_HBM_GBS, 4),
                     "mean_launch_ms": round(1000 * mean_kernel_s, 3)},
    }
    if not args.no_cg:
        line["cg_128"] = cg_figures(cfm)
    if not args.no_cg:
        line["jacobi_precision"] = {"3d_512": jacobi_f32_figures(cfm, (n, n, n), niter),
                                    "2d_16384": jacobi_f32_figures(cfm, (16384, 16384), niter)}
    if not args.no_batch:
        line["cg_batch"] = cg_batch_figures(cfm, args.batch_systems)
    if not args.no_warmstart:
        line["cg_warmstart"] = warmstart_figures(args.warmstart_systems)
    gbs, _ = cfm.bandwidth_probe(1 << 30, 5)
    line["roofline"]["copy_ceiling_gbs"] = round(gbs, 1)
    # the best streaming copy over a few probe shapes (the full sweep: --copy-probe)
    best = hbm_ceiling_sweep(cfm, kinds=("copy",), grids=(1024, 4096, 65536), unrolls=(1,))["best"]["copy"]
    line["roofline"]["copy_ceiling_best_gbs"] = best["gbs"]
    line["roofline"]["copy_ceiling_best_shape"] = {k: best[k] for k in ("grid", "unroll", "nt")}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
