"""Learned warm start for batched variable-coefficient CG.

A 3D U-Net (include/cfm.h, cfm_unet_*) maps the per-cell fields (f, log k)
of a -div(k grad u) = f system on an n^3 grid to an initial guess u0; the
batched CG plan then starts from u0 instead of 0, with the in-kernel guard
(x0_guard) restarting any system whose guess is worse than zero.  The
network is trained on the GPU by the library's own forward/backward kernels
on solutions that plain CG (x0 = 0, tol 1e-8) computes for generated systems.
No counterpart in the reference: its CG always starts from zero
(CG/src/single-stream/pipelined.cu:240-250).

Data: system `sid` of seed `seed` is cfm.varcoef3d(seed, sid, n, n) (k per
cell lognormal(0,1), harmonic-mean faces, f uniform[-1,1)); network input
channel 0 = f * sqrt(3) (unit variance), channel 1 = ln k; target u / u_scale.

    python warmstart.py train --systems 20000 --out weights/unet_varcoef64.npz
    python warmstart.py evaluate --weights weights/unet_varcoef64.npz --systems 1000
"""
import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)
import cfm  # noqa: E402

N = 64
F_GAIN = math.sqrt(3.0)
TRAIN_SEED = 20240601
EVAL_SEED = 77003311
DEFAULT_WEIGHTS = os.path.join(_HERE, "weights", "unet_varcoef64.npz")
# The network's error has a high-frequency part that inflates ||b - A u0||
# (the guard's test) while costing CG little; two damped Jacobi sweeps on
# the guess remove it before the guard checks the guess.
SMOOTH_SWEEPS = 2
SMOOTH_OMEGA = 0.8


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def gen_system(seed, sid, n=N):
    """(rowptr, col, val, f, features[2, n^3] float32) of system sid."""
    rowptr, col, val, f, dims = cfm.varcoef3d(seed, sid, n, n)
    k = cfm.varcoef3d_coef(seed, sid, n, n).reshape(-1)
    feat = np.empty((2, f.size), dtype=np.float32)
    feat[0] = f * F_GAIN
    feat[1] = np.log(k)
    return rowptr, col, val, f, feat


def generate(seed, ids, n=N, pool=None):
    own = pool is None
    pool = pool or cf.ThreadPoolExecutor(_workers())
    try:
        return list(pool.map(lambda s: gen_system(seed, s, n), ids))
    finally:
        if own:
            pool.shutdown()


class Batch:
    """A batched CG plan over generated systems, with their features loaded."""

    def __init__(self, systems, tol=1e-8, max_iter=20000, guard=True):
        srp, rp, col, val = cfm.block_diag([s[:3] for s in systems])
        self.n_rows = systems[0][3].size
        self.b = np.concatenate([s[3] for s in systems])
        p = cfm.cg_params(method=cfm.CG_PIPELINED, tol=tol, max_iter=max_iter, x0_guard=1 if guard else 0)
        self.plan = cfm.CgBatchPlan(srp, rp, col, val, p)
        del rp, col, val
        self.plan.set_rhs(self.b)
        self.plan.load_features(np.concatenate([s[4].reshape(-1) for s in systems]))
        self.nsys = len(systems)

    def solve_plain(self, hist_cap=0):
        self.plan.set_x0(None)
        return self.plan.solve(hist_cap)

    def solve_learned(self, net, u_scale, smooth=SMOOTH_SWEEPS, hist_cap=0):
        """Inference into x0, damped Jacobi smoothing, then CG with the guard.
        Returns (inference+smoothing device ms, solve result)."""
        tw = self.plan.warm_start(net, u_scale)
        ts = self.plan.smooth_x0(smooth, SMOOTH_OMEGA) if smooth > 0 else {"kernel_ms": 0.0}
        res = self.plan.solve(hist_cap)
        return tw["kernel_ms"] + ts["kernel_ms"], res

    def solutions(self):
        return self.plan.x().reshape(self.nsys, self.n_rows)

    def close(self):
        self.plan.close()


def _lr(step, total, lr0, lr_min=1e-5, warmup=100):
    if step < warmup:
        return lr0 * (step + 1) / warmup
    q = min(1.0, (step - warmup) / max(1, total - warmup))
    return lr_min + 0.5 * (lr0 - lr_min) * (1.0 + math.cos(math.pi * q))


def _labelled_chunk(seed, ids, n):
    """Host side of one training chunk: generate the systems and build their
    batched plan (runs on a worker thread while the GPU trains)."""
    sysl = generate(seed, ids, n)
    feats = np.stack([s[4] for s in sysl])
    return Batch(sysl, guard=False), feats


def train(systems=20000, chunk=128, batch=4, passes=2, lr=1e-3, seed=TRAIN_SEED, out=DEFAULT_WEIGHTS, n=N,
          init_seed=1, log=print):
    """Stream `systems` generated systems in chunks: label each chunk with
    batched plain CG (x0 = 0, 1e-8), then take `passes` shuffled passes of
    Adam steps over it.  The next chunk is generated and its plan built on a
    host thread while the GPU labels and trains on the current one."""
    pool = cf.ThreadPoolExecutor(1)
    rng = np.random.default_rng(seed)
    net = cfm.UNet(n, batch, seed=init_seed)
    nchunks = (systems + chunk - 1) // chunk
    total_steps = sum(passes * ((min(chunk, systems - c * chunk)) // batch) for c in range(nchunks))
    step, u_scale = 0, None
    t_start = time.perf_counter()
    history = []
    nxt = pool.submit(_labelled_chunk, seed, range(0, min(chunk, systems)), n)
    for c in range(nchunks):
        t0 = time.perf_counter()
        bt, feats = nxt.result()
        lo = (c + 1) * chunk
        if lo < systems:
            nxt = pool.submit(_labelled_chunk, seed, range(lo, min(lo + chunk, systems)), n)
        t1 = time.perf_counter()
        st, iters, stat, _, tm = bt.solve_plain()
        if st != cfm.OK:
            raise RuntimeError(f"label solve failed: {cfm.status_str(st)}")
        u = bt.solutions()
        bt.close()
        if u_scale is None:
            u_scale = float(np.sqrt(np.mean(u[: min(32, len(u))] ** 2)))
        tgt = (u / u_scale).astype(np.float32)
        t2 = time.perf_counter()
        losses = []
        m = (len(tgt) // batch) * batch
        for _ in range(passes):
            order = rng.permutation(len(tgt))[:m]
            for g in range(0, m, batch):
                idx = order[g: g + batch]
                losses.append(net.train_step(feats[idx], tgt[idx], _lr(step, total_steps, lr)))
                step += 1
        t3 = time.perf_counter()
        rec = {"chunk": c, "systems_done": min(lo, systems), "label_iters": round(float(iters.mean()), 1),
               "label_ms": round(tm["kernel_ms"], 1), "wait_s": round(t1 - t0, 2), "label_s": round(t2 - t1, 2),
               "train_s": round(t3 - t2, 2), "loss": round(float(np.mean(losses)), 5), "steps": step,
               "elapsed_s": round(time.perf_counter() - t_start, 1)}
        history.append(rec)
        log(json.dumps(rec))
        sys.stdout.flush()
    pool.shutdown()
    params = net.params()
    net.close()
    meta = {"n": n, "u_scale": u_scale, "f_gain": F_GAIN, "train_seed": seed, "systems": systems,
            "batch": batch, "passes": passes, "lr": lr, "steps": step, "final_loss": history[-1]["loss"],
            "train_seconds": round(time.perf_counter() - t_start, 1)}
    if out:
        save(out, params, meta)
    return params, meta, history


def save(path, params, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32), meta=np.frombuffer(json.dumps(meta).encode(),
                                                                                  dtype=np.uint8))


def load(path=DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        meta = json.loads(bytes(z["meta"]).decode())
    return params, meta


def evaluate(params, meta, systems=1000, group=250, seed=EVAL_SEED, net_batch=8, smooth=SMOOTH_SWEEPS, log=print):
    """Held-out systems: plain batched CG from x0 = 0 against the learned
    start (inference + CG), both to 1e-8, both timed by HIP events with the
    systems already resident on the device.  Also times plain single-system
    GPU CG (one persistent launch per system) on the first group."""
    n = int(meta["n"])
    net = cfm.UNet(n, net_batch, params=params)
    pool = cf.ThreadPoolExecutor(_workers())
    tot = {"plain_ms": 0.0, "learned_cg_ms": 0.0, "infer_ms": 0.0, "plain_iters": [], "learned_iters": [],
           "rejected": 0, "r0_rel": [], "solo_ms": 0.0, "solo_systems": 0}
    nxt = pool.submit(generate, seed, range(0, min(group, systems)), n, None)
    for g0 in range(0, systems, group):
        sysl = nxt.result()
        g1 = g0 + group
        if g1 < systems:
            nxt = pool.submit(generate, seed, range(g1, min(g1 + group, systems)), n, None)
        bt = Batch(sysl, guard=True)
        bt.solve_plain()  # warm-up launch
        st0, it0, _, _, t0 = bt.solve_plain()
        bt.solve_learned(net, meta["u_scale"], smooth)
        tw, (st1, it1, _, h1, t1) = bt.solve_learned(net, meta["u_scale"], smooth, hist_cap=1)
        if st0 != cfm.OK or st1 != cfm.OK:
            raise RuntimeError(f"solve failed: plain {cfm.status_str(st0)}, learned {cfm.status_str(st1)}")
        bb = np.array([float(s[3] @ s[3]) for s in sysl])
        tot["r0_rel"] += list(np.sqrt(h1[:, 0] / bb))
        tot["rejected"] += int(bt.plan.x0_rejected().sum())
        tot["plain_ms"] += t0["kernel_ms"]
        tot["learned_cg_ms"] += t1["kernel_ms"]
        tot["infer_ms"] += tw
        tot["plain_iters"] += list(it0)
        tot["learned_iters"] += list(it1)
        bt.close()
        if g0 == 0:  # plain GPU CG, one system per launch, on a sample
            for s in sysl[:32]:
                with cfm.CgPlan(s[0], s[1], s[2], cfm.cg_params(method=cfm.CG_PIPELINED, tol=1e-8,
                                                                max_iter=20000)) as sp:
                    sp.set_rhs(s[3])
                    sp.set_x0(None)
                    sp.solve()
                    _, _, _, ts = sp.solve()
                    tot["solo_ms"] += ts["kernel_ms"]
                    tot["solo_systems"] += 1
        log(json.dumps({"evaluated": min(g1, systems), "plain_ms": round(tot["plain_ms"], 1),
                        "learned_ms": round(tot["learned_cg_ms"] + tot["infer_ms"], 1)}))
        sys.stdout.flush()
        del sysl
    pool.shutdown()
    net.close()
    solo_per_sys = tot["solo_ms"] / max(1, tot["solo_systems"])
    learned_ms = tot["learned_cg_ms"] + tot["infer_ms"]
    pi, li = np.array(tot["plain_iters"]), np.array(tot["learned_iters"])
    return {
        "systems": systems, "grid": n, "seed": seed,
        "plain_mean_iters": round(float(pi.mean()), 1), "learned_mean_iters": round(float(li.mean()), 1),
        "iter_reduction": round(1.0 - float(li.sum()) / float(pi.sum()), 4),
        "median_r0_rel": round(float(np.median(tot["r0_rel"])), 4),
        "rejected_guesses": tot["rejected"],
        "batched_plain_ms": round(tot["plain_ms"], 2),
        "batched_learned_ms": round(learned_ms, 2), "inference_and_smoothing_ms": round(tot["infer_ms"], 2),
        "smoothing_sweeps": smooth,
        "solo_plain_ms_per_system": round(solo_per_sys, 3),
        "solo_plain_ms_extrapolated": round(solo_per_sys * systems, 1),
        "learned_vs_batched_plain": round(tot["plain_ms"] / learned_ms, 3),
        "learned_vs_solo_plain": round(solo_per_sys * systems / learned_ms, 3),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--systems", type=int, default=20000)
    t.add_argument("--chunk", type=int, default=128)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--passes", type=int, default=2)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--n", type=int, default=N)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--systems", type=int, default=1000)
    e.add_argument("--group", type=int, default=250)
    e.add_argument("--smooth", type=int, default=SMOOTH_SWEEPS)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        _, meta, _ = train(a.systems, a.chunk, a.batch, a.passes, a.lr, out=a.out, n=a.n)
        print(json.dumps(meta))
    else:
        params, meta = load(a.weights)
        print(json.dumps(evaluate(params, meta, a.systems, a.group, smooth=a.smooth)))


if __name__ == "__main__":
    main()
