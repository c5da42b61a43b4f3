"""Benchmark: FW iterations/s of the node-relaxation hot path on one MI355X,
plus time-to-proven-optimum on the tractable binary-QP sizes.

One step = one gb_relax_batch launch over a fixed batch of open B&B nodes of
a seeded n=200 binary QP (A 300x200 ~ N(0,1), Q = A'A/300 + 0.1 I,
c ~ U(-1,1)), reference default algorithm (DICG + secant line search, 4
secant iterations), per-node FW tolerance 1e-3 * 0.8^depth, iteration budget
per step.  value = total FW iterations executed / timed wall time (inputs
resident in HBM).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--nodes 1024] [--budget 200]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP64_TFLOPS = 78.6
MI355X_HBM_GBS = 8000.0


def node_batch(n, N, seed):
    """N open nodes of a binary problem: node 0 is the root, the others carry
    random fixings of up to ~30% of the variables (depth = #fixings)."""
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    depth = np.zeros(N, dtype=np.int64)
    for i in range(1, N):
        fix = rng.random(n) < 0.3 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
        depth[i] = int(fix.sum())
    eps = np.maximum(1e-6, 1e-3 * 0.8 ** depth)
    return LB, UB, eps


def gpu_relax_throughput(args, Q, c, LB, UB, EPS, form):
    """Times `steps` launches over the same node batch.  form 'plain': the
    FP64 per-node-workgroup kernel with the reference defaults (DICG +
    secant); form 'batched': the persistent wave64 kernel (DICG, exact step,
    incremental products), no cutoff so every node relaxes to its tolerance
    or the budget."""
    import torch

    import gpuboscia_amd as gb
    n = Q.shape[0]
    N = LB.shape[0]
    cuda = torch.device("cuda:0")
    dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).to(cuda)
    dc = torch.from_numpy(c).to(cuda)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    lmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
    alg = gb.algorithm()  # DICG + secant(4, 1e-10, 1e-14): reference defaults
    x = torch.zeros((N, n), dtype=torch.float64, device=cuda)
    lb = torch.from_numpy(LB).to(cuda)
    ub = torch.from_numpy(UB).to(cuda)
    eps = torch.from_numpy(EPS).to(cuda)
    it = torch.zeros(N, dtype=torch.int32, device=cuda)
    st = torch.zeros(N, dtype=torch.uint8, device=cuda)
    gap = torch.zeros(N, dtype=torch.float64, device=cuda)
    val = torch.zeros(N, dtype=torch.float64, device=cuda)
    bv = torch.zeros((N, n), dtype=torch.float64, device=cuda)
    bvv = torch.zeros(N, dtype=torch.float64, device=cuda)
    flags = torch.ones(n, dtype=torch.uint8, device=cuda)
    stats = torch.zeros((N, 2), dtype=torch.int64, device=cuda)
    stream = torch.cuda.Stream(device=cuda)

    def launch():
        if form == "plain":
            gb.relax_batch(obj, alg, lmo, N, args.budget, x, lb, ub, eps, it, st, gap, val, bv, bvv,
                           flags, 1e-6, 0, stream.cuda_stream)
        else:
            gb.relax_batch_fast(obj, alg, lmo, N, args.budget, x, lb, ub, eps, it, st, gap, val, bv,
                                bvv, flags, 1e-6, None, 64, 4, 0, stream.cuda_stream, stats)

    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            it.zero_()
            st.zero_()
            launch()
        stream.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            it.zero_()
            st.zero_()
            starts[k].record(stream)
            launch()
            ends[k].record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    iters_per_step = int(it.sum().item())
    kernel_ms = [s_.elapsed_time(e_) for s_, e_ in zip(starts, ends)]
    out = {"wall_s": wall, "total_iters": iters_per_step * args.steps,
           "iters_per_step": iters_per_step, "kernel_ms_avg": float(np.mean(kernel_ms)),
           "status_hist": np.bincount(st.cpu().numpy(), minlength=5).tolist(),
           "value": iters_per_step * args.steps / wall}
    if form == "batched":
        sv = stats.cpu().numpy()
        out["column_updates"] = int(sv[:, 0].sum())
        out["direct_products"] = int(sv[:, 1].sum())
    return out


def cpu_baseline(args, Q, c, LB, UB, EPS):
    """The CPU oracle (FP64 restatement of the reference algorithm) on a
    bounded sample of the same node batch, node-parallel on host threads."""
    from oracle import pyoracle as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    sample = min(len(LB), max(threads, args.cpu_nodes))
    obj = orc.dense_objective(Q, c, 2.0, 0.5, 0.0)
    a = orc.alg()
    X = LB[:sample].copy()
    t0 = time.perf_counter()
    r = orc.relax_batch(obj, a, orc.box_lmo(), X, LB[:sample].copy(), UB[:sample].copy(),
                        EPS[:sample].copy(), np.zeros(sample, np.int32), np.zeros(sample, np.uint8),
                        flags=np.ones(Q.shape[0], np.uint8), max_iters=args.budget, lanes=256,
                        threads=threads)
    dt = time.perf_counter() - t0
    iters = int(r["iterations"].sum())
    return {"value": iters / dt, "unit": "FW iterations/s", "cores": threads, "kind": "port",
            "sample": f"{sample} of the {len(LB)} nodes, budget {args.budget}, {iters} iterations "
                      f"in {dt:.2f} s"}


def time_to_optimum(args):
    """gb_solve to the proven optimum (abs = rel gap 1e-6), plain form
    (reference defaults) and batched form, on the tractable binary-QP sizes."""
    import torch

    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    out = []
    for n, rows, seeds in ((20, 30, (1, 2, 3)), (40, 60, (1, 2, 3)), (60, 90, (1, 2, 3))):
        for seed in seeds:
            Q, c, sha = instances.binary_qp(n, seed=seed, rows=rows)
            dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
            dc = torch.from_numpy(c).cuda()
            obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
            row = {"n": n, "seed": seed, "sha256": sha[:16]}
            for form in (0, 1):
                s = gb.default_settings()
                s.abs_gap_tol = 1e-6
                s.rel_gap_tol = 1e-6
                s.time_limit_seconds = 120.0
                s.form = form
                r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
                key = "plain" if form == 0 else "batched"
                row[key] = {"seconds": round(r["elapsed_seconds"], 5),
                            "relax_seconds": round(r["relax_seconds"], 5), "nodes": r["nodes_processed"],
                            "fw_iterations": r["fw_iterations"], "objective": r["objective"],
                            "dual_bound": r["dual_bound"], "proven": not r["reached_time_limit"]}
            fp, fb = row["plain"]["objective"], row["batched"]["objective"]
            row["same_optimum"] = abs(fp - fb) <= 1e-9 * max(1.0, abs(fp))
            row["speedup"] = round(row["plain"]["seconds"] / max(row["batched"]["seconds"], 1e-9), 2)
            out.append(row)
            print(f"[bench] time-to-optimum n={n} seed={seed}: plain {row['plain']['seconds']:.4f}s "
                  f"batched {row['batched']['seconds']:.4f}s nodes {row['plain']['nodes']}/"
                  f"{row['batched']['nodes']}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nodes", type=int, default=1024)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--budget", type=int, default=200)
    ap.add_argument("--cpu-nodes", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tto", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(B&B open nodes would shard across GPUs with a min-allreduce of incumbent "
              "and bound per round)", file=sys.stderr)
        sys.exit(2)
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    gb.load()
    n = args.n
    Q, c, sha = instances.binary_qp(n, seed=1)
    LB, UB, EPS = node_batch(n, args.nodes, seed=2)
    print(f"[bench] n={n} nodes={args.nodes} budget={args.budget} steps={args.steps}",
          file=sys.stderr, flush=True)
    plain = gpu_relax_throughput(args, Q, c, LB, UB, EPS, "plain")
    print(f"[bench] plain: {plain['value']:.4g} FW it/s, kernel {plain['kernel_ms_avg']:.3f} ms",
          file=sys.stderr, flush=True)
    fast = gpu_relax_throughput(args, Q, c, LB, UB, EPS, "batched")
    print(f"[bench] batched: {fast['value']:.4g} FW it/s, kernel {fast['kernel_ms_avg']:.3f} ms",
          file=sys.stderr, flush=True)
    kernel_s = fast["kernel_ms_avg"] / 1e3
    # FP64 flops the batched kernel executed on matrix work: 2n per Q-column
    # update, 2n^2 per direct product (counted by the kernel itself).
    flops = (fast["column_updates"] * 2.0 * n + fast["direct_products"] * 2.0 * n * n) / args.steps
    achieved = flops / kernel_s / 1e12
    plain_flops = plain["iters_per_step"] * 2.0 * n * n
    result = {
        "metric": "FW iterations/s of the node relaxation (batched form) on n=200 binary-QP nodes; "
                  "time-to-proven-optimum @gap 1e-6 plain vs batched in time_to_optimum",
        "value": fast["value"], "unit": "FW iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fast["wall_s"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"binary QP n={n} (A {int(1.5*n)}x{n}), {args.nodes} open B&B nodes, "
                               f"DICG, budget {args.budget} iterations/node/step",
                   "instance_sha256": sha[:16], "form": "batched (persistent wave64, FP64)"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MI355X_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MI355X_FP64_TFLOPS, "traffic": None,
                     "kernel": "relax_fast_kernel<4>", "kernel_ms": fast["kernel_ms_avg"],
                     "basis": "FP64 matrix flops executed (2n per column update + 2n^2 per "
                              "direct product, kernel-counted) / kernel time; FP64 peak"},
        "fw_iters_per_step": fast["iters_per_step"], "node_status_hist": fast["status_hist"],
        "plain": {"value": plain["value"], "kernel_ms": plain["kernel_ms_avg"],
                  "fw_iters_per_step": plain["iters_per_step"],
                  "roofline_achieved_tflops": plain_flops / (plain["kernel_ms_avg"] / 1e3) / 1e12,
                  "basis": "2 n^2 per FW iteration (one gradient), DICG + secant reference defaults",
                  "node_status_hist": plain["status_hist"]},
        "speedup_fw_iters_per_s": fast["value"] / plain["value"],
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, Q, c, LB, UB, EPS)
    if not args.no_tto:
        tto = time_to_optimum(args)
        result["time_to_optimum"] = tto
        sp = [r["speedup"] for r in tto if r["n"] >= 40]
        result["time_to_optimum_speedup_geomean_n40_60"] = float(np.exp(np.mean(np.log(sp))))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
