"""Benchmark of the node-relaxation hot path on one MI355X.

Workload (one step): one relaxation launch over a fixed batch of 1024 open
B&B nodes of a seeded n=200 binary QP (A 300x200 ~ N(0,1), Q = A'A/300 +
0.1 I, c ~ U(-1,1); nodes = the root plus random fixings), DICG, per-node FW
tolerance 1e-3 * 0.8^depth, 200 iterations per node per step.  Inputs are
resident in HBM; each step resets the per-node iteration counters and
statuses (fill kernels) and relaxes the batch again.

  value      FW iterations/s of the mixed-precision form as the searches run
             it (three bf16 planes of Q on MFMA, FP32-class products, + the
             FP64 certification pass), whole batch, HIP events on the launch
             stream bracket the timed steps
  mixed      the same with one bf16 plane of Q
  plain      the reference algorithm (FP64, DICG + secant), same batch
  batched    the FP64 batched form, same batch
  scale      the mixed forms on 16384 nodes
  roofline   the BF16 MFMA kernel of the headline form (relax_mixed_kernel,
             three planes): algorithmic flops 2 n^2 per FW iteration / its
             launch time, against 2.5 PFLOP/s; HBM bytes per launch and the
             MFMA-busy and VALU-issue fractions of its SIMD-cycles from the
             committed rocprofv3 counter passes under profiles/
  heterogeneous_batch  256 binary QPs, n log-uniform in [32, 512], box or
             bounded-simplex LMO by seed, 1024 open nodes each, one
             gb_relax_instances call (and its captured replay)
  cpu_baseline  the CPU oracle (FP64 restatement of the plain form) on the
             same 1024 nodes and budget, on the host CPUs this process may
             use (capped by OMP_NUM_THREADS), with nproc and the CPU model
  time_to_optimum  gb_solve to the proven optimum (abs = rel gap 1e-6) on
             binary QPs of n = 20, 40, 60, 80, 100 (three seeds each): plain
             vs batched vs mixed vs mixed with the device-resident search
             (mixed_device) vs learned + batched vs learned + mixed, and the
             geometric-mean speedups over plain per n
  multiblock  one node of a dense quadratic at n = 8192 .. 65536 (Q up to
             32 GiB) relaxed by 256 workgroups (MultiBlock): FW iterations/s and
             the bytes of Q streamed per second against 8 TB/s, with the
             one-workgroup kernel at n = 8192 for comparison
  time_to_target_n200  n = 200 held-out binary QPs: the time each form
             needs to reach the certified relative gap plain reaches in
             --tt200-seconds (python bench.py --only-tt200 runs only this)

Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--nodes 1024] [--budget 200]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP64_TFLOPS = 78.6
MI355X_BF16_TFLOPS = 2500.0
SCORER = os.path.join(ROOT, "gpuboscia_amd", "scorer_binary_qp_n40.npz")


def say(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


class Batch:
    """Device buffers of one node batch and the launch of each form."""

    def __init__(self, Q, c, LB, UB, EPS):
        import torch

        import gpuboscia_amd as gb
        self.gb = gb
        self.torch = torch
        n = Q.shape[0]
        N = LB.shape[0]
        self.n, self.N = n, N
        cuda = torch.device("cuda:0")
        f64 = dict(dtype=torch.float64, device=cuda)
        self.dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).to(cuda)
        self.dc = torch.from_numpy(c).to(cuda)
        self.obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, self.dQt.data_ptr(), self.dc.data_ptr(), 2.0, 0.5, 0.0, 0,
                                    0, 0)
        self.lmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
        self.alg = gb.algorithm()  # DICG + secant(4, 1e-10, 1e-14): reference defaults
        self.x = torch.zeros((N, n), **f64)
        self.lb = torch.from_numpy(LB).to(cuda)
        self.ub = torch.from_numpy(UB).to(cuda)
        self.eps = torch.from_numpy(EPS).to(cuda)
        self.it = torch.zeros(N, dtype=torch.int32, device=cuda)
        self.st = torch.zeros(N, dtype=torch.uint8, device=cuda)
        self.gap = torch.zeros(N, **f64)
        self.val = torch.zeros(N, **f64)
        self.bv = torch.zeros((N, n), **f64)
        self.bvv = torch.full((N,), 1e300, **f64)
        self.flags = torch.ones(n, dtype=torch.uint8, device=cuda)
        self.mit = torch.zeros(N, dtype=torch.int64, device=cuda)
        self.stats = torch.zeros((N, 2), dtype=torch.int64, device=cuda)
        self.stream = torch.cuda.Stream(device=cuda)
        self.mq = {}

    def reset(self):
        self.it.zero_()
        self.st.zero_()
        self.bvv.fill_(1e300)

    def launch(self, form, budget):
        gb, s = self.gb, self.stream.cuda_stream
        if form == "plain":
            gb.relax_batch(self.obj, self.alg, self.lmo, self.N, budget, self.x, self.lb, self.ub, self.eps,
                           self.it, self.st, self.gap, self.val, self.bv, self.bvv, self.flags, 1e-6, 0, s)
        elif form == "batched":
            gb.relax_batch_fast(self.obj, self.alg, self.lmo, self.N, budget, self.x, self.lb, self.ub,
                                self.eps, self.it, self.st, self.gap, self.val, self.bv, self.bvv, self.flags,
                                1e-6, None, 64, 0, 0, s, self.stats)
        else:
            planes = 3 if form.startswith("mixed_fp32") else 1
            if self.mq.get(planes) is None:
                self.mq[planes] = gb.MixedQ(self.obj, planes, s)
            gb.relax_batch_mixed(self.mq[planes], self.obj, self.alg, self.lmo, self.N, budget, self.x, self.lb,
                                 self.ub, self.eps, self.it, self.st, self.gap, self.val, self.bv, self.bvv,
                                 self.flags, 1e-6, None, fp64_max_iters=-1 if form.endswith("_pass1") else 0,
                                 mixed_iterations=self.mit, stream=s)

    def measure(self, form, budget, steps, warmup):
        """Times `steps` launches; HIP events on the launch stream give the
        average launch duration, wall time the throughput."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for _ in range(warmup):
                self.reset()
                self.launch(form, budget)
            self.stream.synchronize()
            starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
            ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(steps):
                self.reset()
                starts[k].record(self.stream)
                self.launch(form, budget)
                ends[k].record(self.stream)
            self.stream.synchronize()
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        iters = int(self.it.sum().item())
        out = {"value": iters * steps / wall, "wall_s": wall, "fw_iters_per_step": iters,
               "launch_ms": float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)])),
               "node_status_hist": np.bincount(self.st.cpu().numpy() & 0x7F, minlength=5)[:5].tolist()}
        if form.startswith("mixed"):
            out["mixed_iters_per_step"] = int(self.mit.sum().item())
        if form == "batched":
            sv = self.stats.cpu().numpy()
            out["column_updates"] = int(sv[:, 0].sum())
            out["direct_products"] = int(sv[:, 1].sum())
        return out


def traffic_bytes(planes):
    """HBM bytes per launch of relax_mixed_kernel with `planes` bf16 planes
    on this workload, from the committed rocprofv3 counter passes
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known
    size); None if not recorded."""
    name = "relax_mixed_kernel_n200_1024nodes_traffic.json" if planes == 1 else \
        "relax_mixed_kernel_3planes_n200_1024nodes_traffic.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def kernel_counters(planes):
    """Issue fractions of relax_mixed_kernel with `planes` bf16 planes on
    this workload, from the committed rocprofv3 counter passes (MFMA-busy and
    VALU-issue SIMD-cycles over the launch's SIMD-cycles); None if not
    recorded."""
    path = os.path.join(ROOT, "profiles", "relax_mixed_kernel_n200_1024nodes_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(f"{planes}_planes")
    if k is None:
        return None
    return {key: k[key] for key in ("kernel", "trace_avg_ms", "mfma_busy_frac", "valu_issue_frac",
                                    "wait_frac_of_wave_cycles", "per_wave_pass")}


def cpu_baseline(args, Q, c, LB, UB, EPS):
    """The CPU oracle (FP64 restatement of the plain form) on the node batch
    (or its first --cpu-nodes nodes), node-parallel on host threads."""
    from oracle import pyoracle as orc
    threads = host_threads()
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
                      f"in {dt:.2f} s (plain-form algorithm, FP64)",
            "nproc": os.cpu_count(), "cpus_allowed": len(os.sched_getaffinity(0)), "cpu_model": cpu_model(),
            "threads_basis": "the CPUs this process may use, capped by OMP_NUM_THREADS (the host share "
                             "given to one GPU)"}


def host_threads():
    """Worker threads of the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when it is set (the host's share per GPU)."""
    allowed = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        allowed = min(allowed, int(cap))
    return max(1, allowed)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def hetero_batches(args):
    """Heterogeneous batch at its configured shape: args.c3_instances seeded
    binary QPs with n log-uniform in [32, 512] (instances.batched_sizes),
    each with a box or a bounded-simplex LMO (sum x <= ceil(n / 4)) chosen
    by seed (instances.batched_lmos) and args.c3_nodes open nodes
    (instances.open_nodes), relaxed by ONE gb_relax_instances call over 8
    streams: box instances in the mixed form (args.c3_planes bf16 planes of
    Q on MFMA with FP32 accumulation, then the FP64 certification pass) for
    every instance, the simplex's greedy vertices picked by one wave per
    node column; DICG with the exact
    step, budget args.budget iterations per node.  HIP events on the calling
    stream bracket each call; the same call captured as one HIP graph is
    replayed on refilled inputs."""
    import torch

    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    sizes = instances.batched_sizes(args.c3_instances, seed=1)
    lmos = instances.batched_lmos(args.c3_instances, seed=1, sizes=sizes)
    N = args.c3_nodes
    f64 = dict(dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    entries, host = [], []
    for k, (n, (kind, K)) in enumerate(zip(sizes, lmos)):
        n = int(n)
        Q, c, _ = instances.binary_qp(n, seed=100000 + k)
        LB, UB, EPS = instances.open_nodes(n, N, 200000 + k, card=K)
        dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
        dc = torch.from_numpy(c).cuda()
        flags = torch.ones(n, dtype=torch.uint8, device="cuda")
        obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        lmo = (gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0) if kind == "box"
               else gb.LmoDesc(gb.LMO_SIMPLEX, float(K), 0, flags.data_ptr(), 0))
        e = {"obj": obj, "lmo": lmo, "keep": (dQt, dc), "kind": kind,
             "mixed": gb.MixedQ(obj, args.c3_planes, stream.cuda_stream),
             "x": torch.from_numpy(LB).cuda(), "lb": torch.from_numpy(LB).cuda(), "ub": torch.from_numpy(UB).cuda(),
             "eps": torch.from_numpy(EPS).cuda(), "iterations": torch.zeros(N, dtype=torch.int32, device="cuda"),
             "status": torch.zeros(N, dtype=torch.uint8, device="cuda"), "gap": torch.zeros(N, **f64),
             "value": torch.zeros(N, **f64), "best_vertex": torch.zeros((N, n), **f64),
             "best_vertex_val": torch.full((N,), 1e300, **f64), "flags": flags, "cutoff": None}
        entries.append(e)
        host.append(LB)
    alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)

    def reset():
        for e, LB in zip(entries, host):
            e["x"].copy_(torch.from_numpy(LB))
            e["iterations"].zero_()
            e["status"].zero_()
            e["best_vertex_val"].fill_(1e300)

    def count_iters(kind=None):
        return int(sum(int(e["iterations"].sum().item()) for e in entries if kind in (None, e["kind"])))

    times = []
    with torch.cuda.stream(stream):
        for step in range(args.c3_steps + 1):
            reset()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            gb.relax_instances(entries, alg, args.budget, concurrency=8, stream=stream.cuda_stream)
            b.record(stream)
            stream.synchronize()
            if step > 0:  # the first call is the warm-up
                times.append(a.elapsed_time(b) / 1e3)
        iters = count_iters()
        iters_box, iters_simplex = count_iters("box"), count_iters("simplex")
        nodes = len(entries) * N
        # the same call captured once as a HIP graph, replayed on refilled inputs
        g = torch.cuda.CUDAGraph()
        reset()
        stream.synchronize()
        with torch.cuda.graph(g, stream=stream):
            gb.relax_instances(entries, alg, args.budget, concurrency=8, stream=stream.cuda_stream)
        stream.synchronize()
        replay = []
        for step in range(args.c3_steps):
            reset()
            stream.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            g.replay()
            b.record(stream)
            stream.synchronize()
            replay.append(a.elapsed_time(b) / 1e3)
        iters_replay = count_iters()
        del g
        # a captured call keeps the one-stream tile width (graph replays
        # overlap the streams less than eager calls): it equals an eager call
        # with concurrency 1, timed here too
        reset()
        stream.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        gb.relax_instances(entries, alg, args.budget, concurrency=1, stream=stream.cuda_stream)
        b.record(stream)
        stream.synchronize()
        one_stream_ms = a.elapsed_time(b)
        iters_one = count_iters()
        # statuses 1 and 4: fractional / integral, converged to the node tolerance
        converged = int(sum(int(((e["status"] == 1) | (e["status"] == 4)).sum().item()) for e in entries))
    t = float(np.mean(times))
    kinds = [e["kind"] for e in entries]
    return {"instances": len(entries), "box_instances": kinds.count("box"),
            "simplex_instances": kinds.count("simplex"), "nodes_per_instance": N, "n_min": int(sizes.min()),
            "n_max": int(sizes.max()),
            "precision": f"box and bounded simplex: {args.c3_planes} bf16 plane(s) of Q on MFMA (FP32 "
                         "accumulate, FP32 iterates) + FP64 certification pass",
            "algorithm": "DICG, exact quadratic step", "budget": args.budget,
            "call_ms": t * 1e3, "captured_replay_ms": float(np.mean(replay)) * 1e3,
            "replay_iterations_equal_one_stream_call": iters_replay == iters_one,
            "one_stream_call_ms": one_stream_ms,
            "fw_iterations_per_call": iters, "fw_iterations_box": iters_box, "fw_iterations_simplex": iters_simplex,
            "value": iters / t, "unit": "FW iterations/s", "nodes_per_s": nodes / t, "converged_nodes": converged}


def multiblock_large_n(args):
    """MultiBlock leg: one node of a dense quadratic over the 0/1 box at n
    beyond the one-workgroup kernel's LDS limit, relaxed by K = 256
    workgroups (gb_relax_batch_multiblock), FW open loop for a fixed count of
    iterations (tolerance 0), so the products are counted exactly: per node
    2 iterations + 4 passes over Q (gradient and best-vertex value per
    iteration; final gradient, vertex value, value and rounded value), each
    reading 8 n^2 bytes of Q.  HIP events on the launch stream time each
    call.  At n = 8192 the one-workgroup kernel (gb_relax_batch) runs the
    same node for comparison."""
    import torch
    import gpuboscia_amd as gb
    f64 = dict(dtype=torch.float64, device="cuda")
    rows = []
    stream = torch.cuda.Stream()
    for n, K, iters, nodes in [(8192, 0, 4, 1), (8192, 64, 20, 1), (16384, 256, 20, 1), (16384, 64, 20, 4),
                               (32768, 256, 10, 1), (65536, 256, 4, 1)]:
        if n > args.mb_max_n:
            continue
        gen = torch.Generator(device="cuda")
        gen.manual_seed(n)
        with torch.cuda.stream(stream):
            A = torch.randn(n, n, generator=gen, **f64) / np.sqrt(n)
            Qt = (A + A.T) * 0.5
            del A
            Qt.diagonal().add_(2.0)
            cvec = torch.rand(n, generator=gen, **f64) * 2 - 1
        stream.synchronize()
        obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, Qt.data_ptr(), cvec.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        lmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
        alg = gb.algorithm(dicg=False, line_search=gb.LS_OPEN_LOOP)
        fl = torch.ones(n, dtype=torch.uint8, device="cuda")
        times = []
        for rep in range(3):
            with torch.cuda.stream(stream):
                x = torch.zeros((nodes, n), **f64)
                lb, ub = torch.zeros((nodes, n), **f64), torch.ones((nodes, n), **f64)
                e = torch.zeros(nodes, **f64)
                it = torch.zeros(nodes, dtype=torch.int32, device="cuda")
                st = torch.zeros(nodes, dtype=torch.uint8, device="cuda")
                gap, val = torch.zeros(nodes, **f64), torch.zeros(nodes, **f64)
                bv, bvv = torch.zeros((nodes, n), **f64), torch.full((nodes,), 1e300, **f64)
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                if K == 0:
                    gb.relax_batch(obj, alg, lmo, nodes, iters, x, lb, ub, e, it, st, gap, val, bv, bvv, fl, 1e-6,
                                   256, stream=stream.cuda_stream)
                else:
                    gb.relax_batch_multiblock(obj, alg, lmo, nodes, iters, x, lb, ub, e, it, st, gap, val, bv, bvv,
                                              fl, 1e-6, 256, K, stream=stream.cuda_stream)
                b.record(stream)
            stream.synchronize()
            if rep > 0 or K == 0:
                times.append(a.elapsed_time(b) / 1e3)
        assert int(it.sum().item()) == nodes * iters and int((st.cpu() == 0xFF).sum()) == 0
        t = float(np.mean(times))
        passes = nodes * (2 * iters + 4)
        gbs = passes * 8.0 * n * n / t / 1e9
        row = {"n": n, "nodes": nodes, "blocks_per_node": K, "fw_iterations": nodes * iters, "call_ms": t * 1e3,
               "fw_iters_per_s": nodes * iters / t, "q_passes": passes, "q_bytes_per_s": gbs * 1e9,
               "hbm_frac": gbs / 8000.0}
        rows.append(row)
        say(f"multiblock n={n} nodes={nodes} K={K}: {t * 1e3:.2f} ms per call, {gbs:.0f} GB/s of Q")
        del Qt
        torch.cuda.empty_cache()
    return {"workload": "dense quadratic (Q = (A + A')/2 + 2 I, A ~ N(0, 1/n)) over the 0/1 box, one node at "
                        "the root, FW open loop, tolerance 0; K = 0: one workgroup (gb_relax_batch)",
            "basis": "q_bytes_per_s = passes over Q x 8 n^2 bytes / call time (HIP events on the launch stream); "
                     "hbm_frac against 8 TB/s; several nodes of one Q share its reads in L2",
            "rows": rows}


def qap_root_proof(args):
    """Birkhoff QAP (planted: B = Pi A Pi', optimum 0 at P = Pi'): wall time of
    gb_solve to the proven optimum with the natural lower bound 0, FP32 form
    (form 2: FP32 iterates and assignment costs, FP64 certificate) at every
    size in --qap-sizes, and the plain FP64 form (form 0, the reference's
    BirkhoffLMO arithmetic) at the sizes in --qap-plain-sizes.  DICG with the
    exact step in both; gap tolerances 1e-6."""
    import torch
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    rows = []
    plain = {int(v) for v in args.qap_plain_sizes.split(",") if v}
    for m in [int(v) for v in args.qap_sizes.split(",") if v]:
        A, B, perm, sha = instances.qap_instance(m, seed=m)
        n = m * m
        dA = torch.from_numpy(A).cuda()
        dB = torch.from_numpy(B).cuda()
        obj = gb.ObjectiveDesc(gb.OBJ_QAP, n, 0, 0, 1.0, 1.0, 0.0, m, dA.data_ptr(), dB.data_ptr())
        lmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, m)
        row = {"m": m, "variables": n, "sha256": sha[:16]}
        for form in ([2, 0] if m in plain else [2]):
            s = gb.default_settings()
            s.form = form
            s.precision = gb.PREC_FP32 if form == 2 else gb.PREC_FP64
            s.abs_gap_tol = s.rel_gap_tol = 1e-6
            s.natural_lower_bound = 0.0
            s.time_limit_seconds = args.qap_seconds
            s.algorithm.line_search = gb.LS_QUADRATIC_EXACT
            t0 = time.perf_counter()
            r = gb.solve(n, [0.0] * n, [1.0] * n, list(range(n)), obj, lmo, s)
            t = time.perf_counter() - t0
            X = np.asarray(r["x"]).reshape(m, m)
            key = "fp32" if form == 2 else "plain"
            row[key] = {"seconds": t, "objective": r["objective"], "dual_bound": r["dual_bound"],
                        "nodes": r["nodes_processed"], "fw_iterations": r["fw_iterations"],
                        "proven": bool(r["objective"] - r["dual_bound"] <= 1e-6) and not r["reached_time_limit"],
                        "planted_found": bool(r["objective"] == 0.0 and np.allclose(A @ X, X @ B))}
            say(f"qap m={m} {key}: {t:.3f} s, objective {r['objective']}, {r['fw_iterations']} FW iterations")
        if "plain" in row:
            row["speedup_fp32_vs_plain"] = row["plain"]["seconds"] / row["fp32"]["seconds"]
        rows.append(row)
        del dA, dB
    return {"workload": "planted Birkhoff QAP (symmetric ER(0.3) A, B = Pi A Pi'), gb_solve to the proven "
                        "optimum 0 with natural_lower_bound 0, DICG + exact step, gap 1e-6",
            "rows": rows}


def solve_row(gb, Q, c, n, form, learned=None, max_nodes=0, time_limit=120.0, rel_gap=1e-6, device=False, card=0):
    """card > 0: over the bounded simplex sum x <= card (0/1 variables)."""
    import torch
    dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
    dc = torch.from_numpy(c).cuda()
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    ii = torch.ones(n, dtype=torch.uint8, device="cuda")
    lmo = (gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, ii.data_ptr(), 0) if card > 0
           else gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0))
    s = gb.default_settings()
    s.abs_gap_tol = 1e-6
    s.rel_gap_tol = rel_gap
    s.time_limit_seconds = time_limit
    s.max_nodes = max_nodes
    s.form = form
    if form == 2:
        # three bf16 planes (FP32-class products): fewer FP64 polish
        # iterations in the search than one plane, at nearly the same pass cost
        s.precision = gb.PREC_FP32
    if learned is not None:
        s.branch_rule = gb.BRANCH_LEARNED
        s.learned_model = learned.handle
    # the device-resident search: frontier, pruning and branching on the GPU
    # (the host driver's one-stream schedule)
    s.device_frontier = 1 if device else 0
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, lmo, s)
    return {"seconds": round(r["elapsed_seconds"], 5), "relax_seconds": round(r["relax_seconds"], 5),
            "nodes": r["nodes_processed"], "fw_iterations": r["fw_iterations"], "objective": r["objective"],
            "dual_bound": r["dual_bound"], "proven": not r["reached_time_limit"]}


def gap_relative(bound, inc):
    """gap_utils.h:34-42: (incumbent - bound) / min(|bound|, |incumbent|)."""
    if not (np.isfinite(bound) and np.isfinite(inc)):
        return float("inf")
    scale = min(abs(bound), abs(inc))
    return float("inf") if scale == 0.0 else (inc - bound) / scale


def time_to_optimum(args):
    """gb_solve to the proven optimum (abs = rel gap 1e-6): plain form
    (reference defaults), batched FP64 form, mixed form (BF16 MFMA + FP64
    certification), and the learned rule with the batched and with the mixed
    form, on the tractable binary-QP sizes."""
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    scorer = gb.Brancher.from_file(SCORER) if os.path.exists(SCORER) else None
    forms = [("plain", 0, None, False), ("batched", 1, None, False), ("mixed", 2, None, False),
             ("mixed_device", 2, None, True)]
    if scorer is not None:
        forms.append(("learned_batched", 1, scorer, False))
        forms.append(("learned_mixed", 2, scorer, False))
    keep = set(args.tto_forms.split(","))
    forms = [f for f in forms if f[0] in keep]
    sizes = [int(v) for v in args.tto_sizes.split(",")]
    out = []
    for n in sizes:
        rows, seeds = 3 * n // 2, (1, 2, 3)
        for seed in seeds:
            Q, c, sha = instances.binary_qp(n, seed=seed, rows=rows)
            # Every form, plain included, first runs the same untimed,
            # node-capped search on another instance of this size: the first
            # launches of a kernel variant (some are reached only by some
            # batch sizes) and of each pipeline slot are one-time costs of
            # the process (about 6 ms), not of the search.
            Qw, cw, _ = instances.binary_qp(n, seed=100 + seed, rows=rows)
            row = {"n": n, "seed": seed, "sha256": sha[:16]}
            for key, form, learned, device in forms:
                solve_row(gb, Qw, cw, n, form, learned, max_nodes=16384, device=device)
                row[key] = solve_row(gb, Q, c, n, form, learned, device=device)
            if "plain" in row:
                fp = row["plain"]["objective"]
                row["same_optimum"] = all(abs(row[k]["objective"] - fp) <= 1e-9 * max(1.0, abs(fp))
                                          for k, _, _, _ in forms)
                for key, _, _, _ in forms[1:]:
                    row[f"speedup_{key}"] = round(row["plain"]["seconds"] / max(row[key]["seconds"], 1e-9), 2)
            out.append(row)
            say(f"time-to-optimum n={n} seed={seed}: " +
                " ".join(f"{k} {row[k]['seconds']:.4f}s/{row[k]['nodes']}n" for k, _, _, _ in forms))
    return out


def simplex_time_to_optimum(args):
    """gb_solve to the proven optimum (abs = rel gap 1e-6) over the bounded
    simplex sum x <= ceil(n / 10) (binary QPs built as the heterogeneous
    batch's, A of 1.5 n x n; at ceil(n / 4) the box optimum is inside the
    simplex and the searches are the box's): plain form (reference defaults),
    batched FP64 form and mixed form (three bf16 planes + FP64
    certification), each after an untimed node-capped run on another
    instance of the size."""
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    forms = [("plain", 0), ("batched", 1), ("mixed", 2)]
    out = []
    for n in [int(v) for v in args.sx_sizes.split(",") if v]:
        card = (n + 9) // 10
        for seed in (1, 2):
            Q, c, sha = instances.binary_qp(n, seed=seed)
            Qw, cw, _ = instances.binary_qp(n, seed=100 + seed)
            row = {"n": n, "seed": seed, "card": card, "sha256": sha[:16]}
            for key, form in forms:
                solve_row(gb, Qw, cw, n, form, max_nodes=4096, card=card)
                row[key] = solve_row(gb, Q, c, n, form, card=card, time_limit=args.sx_seconds)
            fp = row["plain"]["objective"]
            row["same_optimum"] = all(abs(row[k]["objective"] - fp) <= 1e-6 * max(1.0, abs(fp)) for k, _ in forms)
            for key, _ in forms[1:]:
                row[f"speedup_{key}"] = round(row["plain"]["seconds"] / max(row[key]["seconds"], 1e-9), 2)
            out.append(row)
            say(f"simplex time-to-optimum n={n} seed={seed} card={card}: " +
                " ".join(f"{k} {row[k]['seconds']:.4f}s/{row[k]['nodes']}n" for k, _ in forms))
    return out


def time_to_target(args):
    """n = 200 binary QPs of the learned-branching configuration's held-out
    set (A 300 x 200, seeds 50000 + k): the searches cannot be proven within
    minutes, so each form is timed to the certified relative gap the plain
    form reaches in args.tt200_seconds.  Plain runs once with that time
    limit (its final gap G); then every form, plain included, runs with
    rel_gap_tol = G and stops at the first moment its FP64-certified gap is
    within G (the library's own termination rule), so each time is a
    time-to-G.  Every form first runs an untimed 2048-node search on another
    instance."""
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    scorer = gb.Brancher.from_file(SCORER) if os.path.exists(SCORER) else None
    forms = [("plain", 0, None, False), ("batched", 1, None, False), ("mixed", 2, None, False),
             ("mixed_device", 2, None, True)]
    if scorer is not None:
        forms += [("learned_batched", 1, scorer, False), ("learned_mixed", 2, scorer, False)]
    n, T = 200, args.tt200_seconds
    Qw, cw, _ = instances.binary_qp(n, seed=49999)
    for key, form, learned, device in forms:
        solve_row(gb, Qw, cw, n, form, learned, max_nodes=2048, device=device)
    out = []
    for k in range(args.tt200_instances):
        seed = args.tt200_first_seed + k
        Q, c, sha = instances.binary_qp(n, seed=seed)
        first = solve_row(gb, Q, c, n, 0, None, time_limit=T)
        G = gap_relative(first["dual_bound"], first["objective"])
        row = {"n": n, "seed": seed, "sha256": sha[:16], "plain_time_limit_s": T, "target_rel_gap": G,
               "plain_first_run": first}
        for key, form, learned, device in forms:
            r = solve_row(gb, Q, c, n, form, learned, time_limit=max(2 * T, 10.0), rel_gap=G, device=device)
            r["reached_target"] = gap_relative(r["dual_bound"], r["objective"]) <= G
            row[key] = r
        for key, _, _, _ in forms[1:]:
            row[f"speedup_{key}"] = round(row["plain"]["seconds"] / max(row[key]["seconds"], 1e-9), 2)
        out.append(row)
        say(f"time-to-target n=200 seed={seed} gap {G:.4g}: " +
            " ".join(f"{kk} {row[kk]['seconds']:.3f}s/{row[kk]['nodes']}n" for kk, _, _, _ in forms))
    gm = {}
    for key, _, _, _ in forms[1:]:
        sp = [r[f"speedup_{key}"] for r in out if r[key]["reached_target"] and r["plain"]["reached_target"]]
        gm[key] = round(float(np.exp(np.mean(np.log(sp)))), 2) if sp else None
    return {"instances": out, "speedup_geomean": gm,
            "all_reached_target": all(r[kk]["reached_target"] for r in out for kk, _, _, _ in forms)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nodes", type=int, default=1024)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--budget", type=int, default=200)
    ap.add_argument("--cpu-nodes", type=int, default=1024)
    ap.add_argument("--scale-nodes", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tto", action="store_true")
    ap.add_argument("--tto-sizes", default="20,40,60,80,100", help="binary-QP sizes of time_to_optimum")
    ap.add_argument("--tto-forms", default="plain,batched,mixed,mixed_device,learned_batched,learned_mixed",
                    help="forms of time_to_optimum (speedups need plain)")
    ap.add_argument("--no-tt200", action="store_true", help="skip the n=200 time-to-target leg")
    ap.add_argument("--only-tt200", action="store_true", help="run only the n=200 time-to-target leg")
    ap.add_argument("--tt200-instances", type=int, default=2)
    ap.add_argument("--tt200-first-seed", type=int, default=50000)
    ap.add_argument("--tt200-seconds", type=float, default=8.0,
                    help="time limit of plain's first run; its final gap is every form's target")
    ap.add_argument("--no-scale", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-instances", type=int, default=256)
    ap.add_argument("--c3-nodes", type=int, default=1024)
    ap.add_argument("--c3-steps", type=int, default=2)
    ap.add_argument("--c3-planes", type=int, default=3, choices=(1, 3))
    ap.add_argument("--forms", default="plain,batched,mixed,mixed_fp32,mixed_pass1,mixed_fp32_pass1",
                    help="forms to time (the profiled runs time one)")
    ap.add_argument("--no-multiblock", action="store_true", help="skip the MultiBlock large-n leg")
    ap.add_argument("--only-multiblock", action="store_true", help="run only the MultiBlock large-n leg")
    ap.add_argument("--mb-max-n", type=int, default=65536, help="largest n of the MultiBlock leg")
    ap.add_argument("--no-qap", action="store_true", help="skip the Birkhoff QAP root-proof leg")
    ap.add_argument("--only-qap", action="store_true", help="run only the Birkhoff QAP root-proof leg")
    ap.add_argument("--no-sx", action="store_true", help="skip the bounded-simplex time-to-optimum leg")
    ap.add_argument("--only-sx", action="store_true", help="run only the bounded-simplex time-to-optimum leg")
    ap.add_argument("--sx-sizes", default="40,60,80,100", help="binary-QP sizes of the bounded-simplex leg")
    ap.add_argument("--sx-seconds", type=float, default=60.0, help="time limit of each bounded-simplex solve")
    ap.add_argument("--qap-sizes", default="64,128,200", help="QAP orders of the FP32 form")
    ap.add_argument("--qap-plain-sizes", default="64", help="QAP orders also solved by the plain FP64 form")
    ap.add_argument("--qap-seconds", type=float, default=120.0, help="time limit of each QAP solve")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(B&B open nodes would shard across GPUs with a min-allreduce of incumbent "
              "and bound per round)", file=sys.stderr)
        sys.exit(2)
    import gpuboscia_amd as gb
    from gpuboscia_amd import instances
    gb.load()
    if args.only_tt200:
        print(json.dumps({"time_to_target_n200": time_to_target(args)}), flush=True)
        return
    if args.only_multiblock:
        print(json.dumps({"multiblock": multiblock_large_n(args)}), flush=True)
        return
    if args.only_qap:
        print(json.dumps({"qap_root_proof": qap_root_proof(args)}), flush=True)
        return
    if args.only_sx:
        print(json.dumps({"simplex_time_to_optimum": simplex_time_to_optimum(args)}), flush=True)
        return
    n = args.n
    Q, c, sha = instances.binary_qp(n, seed=1)
    LB, UB, EPS = node_batch(n, args.nodes, seed=2)
    say(f"n={n} nodes={args.nodes} budget={args.budget} steps={args.steps}")
    batch = Batch(Q, c, LB, UB, EPS)
    forms = args.forms.split(",")
    res = {}
    for form in forms:
        res[form] = batch.measure(form, args.budget, args.steps, args.warmup)
        say(f"{form}: {res[form]['value']:.4g} FW it/s, launch {res[form]['launch_ms']:.3f} ms")
    if args.copy_probe:
        import torch
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        say("copy probe: 268435456 bytes copied device to device")
        del src, dst
    main_form = "mixed_fp32" if "mixed_fp32" in res else ("mixed" if "mixed" in res else forms[0])
    m = res[main_form]
    planes = 3 if main_form.startswith("mixed_fp32") else 1
    result = {
        "metric": "FW iterations/s of the node relaxation on n=200 binary-QP nodes (mixed-precision form as the "
                  "searches run it: three bf16 planes of Q on MFMA + FP64 certification); time-to-proven-optimum "
                  "@gap 1e-6 plain vs batched / mixed / learned+batched / learned+mixed in time_to_optimum "
                  "(n <= 100) and time to plain's certified gap at n = 200 in time_to_target_n200",
        "value": m["value"], "unit": "FW iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["wall_s"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"binary QP n={n} (A {int(1.5 * n)}x{n}), {args.nodes} open B&B nodes, DICG, "
                               f"budget {args.budget} iterations/node/step",
                   "instance_sha256": sha[:16], "form": main_form,
                   "precision": f"BF16 MFMA products over {planes} bf16 plane(s) of Q (FP32 accumulate), FP32 "
                                "iterates, FP64 steps; FP64 certificate of gap, value and incumbent"},
        "forms": res,
    }
    if "plain" in res:
        result["speedup_vs_plain"] = m["value"] / res["plain"]["value"]
    pass1 = f"{main_form}_pass1" if f"{main_form}_pass1" in res else ("mixed_pass1" if "mixed_pass1" in res else None)
    if pass1 is not None:
        p1 = res[pass1]
        kplanes = 3 if pass1.startswith("mixed_fp32") else 1
        kernel_s = p1["launch_ms"] / 1e3
        alg_flops = 2.0 * n * n * p1["mixed_iters_per_step"]
        n_pad = 32 * ((n + 31) // 32)
        # Q v and Q a per iteration, each over kplanes bf16 planes of Q
        column_flops = 2 * kplanes * 2.0 * n_pad * n_pad * p1["mixed_iters_per_step"]
        achieved = alg_flops / kernel_s / 1e12
        result["roofline"] = {
            "bound": "mfma", "achieved": achieved, "peak": MI355X_BF16_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / MI355X_BF16_TFLOPS, "traffic": traffic_bytes(kplanes), "kernel": "relax_mixed_kernel",
            "planes": kplanes, "kernel_ms": p1["launch_ms"],
            "basis": "algorithmic 2 n^2 flops per FW iteration x iterations of one launch / launch time "
                     "(HIP events on the launch stream); bf16 dense MFMA peak",
            "column_product_tflops": column_flops / kernel_s / 1e12,
            "column_product_basis": "2 products (Q v, Q a) x planes x 2 n_pad^2 per column-iteration, "
                                    "n_pad = 32 ceil(n/32) (a 16-column MFMA tile carries 4 node columns "
                                    "of v and 4 of a at 1024 nodes)",
            "counters": kernel_counters(kplanes)}
    if not args.no_scale and "mixed" in forms:
        LBs, UBs, EPSs = node_batch(n, args.scale_nodes, seed=3)
        sb = Batch(Q, c, LBs, UBs, EPSs)
        steps = max(2, args.steps // 2)
        sc = {f: sb.measure(f, args.budget, steps, 1) for f in ("mixed_fp32", "mixed", "batched")}
        result["scale"] = {"nodes": args.scale_nodes, "steps": steps,
                           **{f: {"value": v["value"], "launch_ms": v["launch_ms"]} for f, v in sc.items()}}
        say(f"scale {args.scale_nodes}: " + " ".join(f"{f} {v['value']:.4g}" for f, v in sc.items()))
        del sb
    if not args.no_c3:
        result["heterogeneous_batch"] = hetero_batches(args)
        h = result["heterogeneous_batch"]
        say(f"heterogeneous batch: {h['instances']} instances x {h['nodes_per_instance']} nodes, n {h['n_min']}.."
            f"{h['n_max']}: {h['call_ms']:.1f} ms per call, {h['value']:.4g} FW it/s, {h['nodes_per_s']:.4g} nodes/s")
    if not args.no_multiblock:
        result["multiblock"] = multiblock_large_n(args)
    if not args.no_qap:
        result["qap_root_proof"] = qap_root_proof(args)
    if not args.no_sx:
        result["simplex_time_to_optimum"] = simplex_time_to_optimum(args)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, Q, c, LB, UB, EPS)
    if not args.no_tto:
        tto = time_to_optimum(args)
        result["time_to_optimum"] = tto
        gm = {}
        for key in ("batched", "mixed", "mixed_device", "learned_batched", "learned_mixed"):
            per_n = {}
            for nn in sorted({r["n"] for r in tto}):
                sp = [r[f"speedup_{key}"] for r in tto if f"speedup_{key}" in r and r["n"] == nn]
                if sp:
                    per_n[str(nn)] = round(float(np.exp(np.mean(np.log(sp)))), 2)
            sp = [r[f"speedup_{key}"] for r in tto if f"speedup_{key}" in r and r["n"] >= 40]
            if sp:
                per_n["n40_100"] = round(float(np.exp(np.mean(np.log(sp)))), 2)
            gm[key] = per_n
        result["time_to_optimum_speedup_geomean"] = gm
        result["time_to_optimum_all_same_optimum"] = all(r.get("same_optimum", True) for r in tto)
    if not args.no_tt200:
        result["time_to_target_n200"] = time_to_target(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
