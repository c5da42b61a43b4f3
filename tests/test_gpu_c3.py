"""The heterogeneous batch at its configured shape: 256 binary QPs with n
log-uniform in [32, 512], each with a box or a bounded-simplex LMO
(sum x <= ceil(n / 4)) chosen by seed and 1024 open nodes, relaxed by one
gb_relax_instances call (every instance in the mixed form: 3 bf16 planes of
Q on MFMA, then the FP64 certification pass; over a simplex pass 1 picks the
greedy vertices with one wave per node column).  Every node's outputs are checked for what must hold whatever the
arithmetic (valid status, feasible iterate, FP64 bound f - gap no greater
than the best vertex found), and four sampled instances are compared bit
for bit with the oracle on a slice of their nodes."""
import math
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import assert_bit_equal

pytestmark = pytest.mark.gpu

COUNT, NODES, BUDGET = 256, 1024, 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(count=COUNT, nodes=NODES):
    import torch
    sizes = instances.batched_sizes(count, seed=1)
    lmos = instances.batched_lmos(count, seed=1, sizes=sizes)
    f64 = dict(dtype=torch.float64, device="cuda")
    entries = []
    for k, (n, (kind, K)) in enumerate(zip(sizes, lmos)):
        n = int(n)
        Q, c, _ = instances.binary_qp(n, seed=100000 + k)
        LB, UB, EPS = instances.open_nodes(n, nodes, 200000 + k, card=K)
        dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
        dc = torch.from_numpy(c).cuda()
        obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        flags = torch.ones(n, dtype=torch.uint8, device="cuda")
        if kind == "box":
            lmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
        else:
            lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(K), 0, flags.data_ptr(), 0)
        e = {"k": k, "n": n, "kind": kind, "K": K, "Q": Q, "c": c, "LB": LB, "UB": UB, "EPS": EPS,
             "obj": obj, "lmo": lmo, "keep": (dQt, dc),
             "mixed": gb.MixedQ(obj, 3),
             "x": torch.from_numpy(LB).cuda(), "lb": torch.from_numpy(LB).cuda(), "ub": torch.from_numpy(UB).cuda(),
             "eps": torch.from_numpy(EPS).cuda(), "iterations": torch.zeros(nodes, dtype=torch.int32, device="cuda"),
             "status": torch.zeros(nodes, dtype=torch.uint8, device="cuda"), "gap": torch.zeros(nodes, **f64),
             "value": torch.zeros(nodes, **f64), "best_vertex": torch.zeros((nodes, n), **f64),
             "best_vertex_val": torch.full((nodes,), 1e300, **f64), "flags": flags, "cutoff": None}
        entries.append(e)
    return entries


def results(e, rows=slice(None)):
    return {"x": e["x"][rows].cpu().numpy(), "gap": e["gap"][rows].cpu().numpy(),
            "value": e["value"][rows].cpu().numpy(), "iterations": e["iterations"][rows].cpu().numpy(),
            "status": e["status"][rows].cpu().numpy(), "best_vertex": e["best_vertex"][rows].cpu().numpy(),
            "best_vertex_val": e["best_vertex_val"][rows].cpu().numpy()}


def check_properties(e):
    r = results(e)
    LB, UB, K = e["LB"], e["UB"], e["K"]
    st = r["status"]
    assert np.all(st <= 4), f"instance {e['k']}: status codes"
    infeasible = st == 2
    if e["kind"] == "box":
        assert not infeasible.any(), f"instance {e['k']}: a box node is always feasible"
    else:
        # a node with more than K variables fixed at 1 is the only infeasible kind here
        assert np.array_equal(infeasible, LB.sum(axis=1) > K), f"instance {e['k']}: infeasible nodes"
    ok = ~infeasible
    x = r["x"][ok]
    tol = 1e-9
    assert np.all(x >= LB[ok] - tol) and np.all(x <= UB[ok] + tol), f"instance {e['k']}: x within bounds"
    if e["kind"] == "simplex":
        assert np.all(x.sum(axis=1) <= K + 1e-8), f"instance {e['k']}: sum x <= K"
    gap, val, bvv = r["gap"][ok], r["value"][ok], r["best_vertex_val"][ok]
    assert np.all(np.isfinite(gap)) and np.all(gap >= -1e-9), f"instance {e['k']}: FW gap"
    # f(x) from the returned iterate, in FP64 on the host
    Q, c = e["Q"], e["c"]
    f = np.einsum("ij,jk,ik->i", x, Q, x) + x @ c
    assert np.allclose(val, f, rtol=1e-9, atol=1e-9), f"instance {e['k']}: value at x"
    # the certified bound f - gap is below every feasible point, the best vertex included
    has = bvv < 1e299
    assert np.all(val[has] - gap[has] <= bvv[has] + 1e-9 * (1.0 + np.abs(bvv[has]))), f"instance {e['k']}: bound"
    bv = r["best_vertex"][ok][has]
    assert np.all((bv == 0.0) | (bv == 1.0)), f"instance {e['k']}: integral best vertices"
    fb = np.einsum("ij,jk,ik->i", bv, Q, bv) + bv @ c
    assert np.allclose(bvv[has], fb, rtol=1e-9, atol=1e-9), f"instance {e['k']}: best vertex value"
    return int(r["iterations"].sum()), int(np.isin(st, (1, 4)).sum())


def in_call_columns(N, k):
    """Node columns per tile of a mixed entry in a k-entry gb_relax_instances
    call: min(16, max(ceil(N / 256), ceil(N k / CUs)))."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return min(16, max(min(16, max(1, -(-N // 256))), -(-N * k // cus)))


def compare_with_oracle(e, alg, rows, columns):
    """A slice of one instance's nodes against the oracle: the mixed form
    through its pass-1 output (a solo call of pass 1 only with the in-call
    tile width, which equals the in-call pass 1) and the oracle's FP64 pass
    from it (box or simplex)."""
    import torch
    from oracle import pyoracle as orc
    n = e["n"]
    N = rows.stop - rows.start
    LB, UB, EPS = e["LB"][rows], e["UB"][rows], e["EPS"][rows]
    got = results(e, rows)
    ones = np.ones(n, np.uint8)
    f64 = dict(dtype=torch.float64, device="cuda")
    x = torch.from_numpy(LB.copy()).cuda()
    it = torch.zeros(N, dtype=torch.int32, device="cuda")
    st = torch.zeros(N, dtype=torch.uint8, device="cuda")
    gap, val = torch.zeros(N, **f64), torch.zeros(N, **f64)
    bv, bvv = torch.zeros((N, n), **f64), torch.full((N,), 1e300, **f64)
    gb.relax_batch_mixed(e["mixed"], e["obj"], alg, e["lmo"], N, BUDGET, x, torch.from_numpy(LB).cuda(),
                         torch.from_numpy(UB).cuda(), torch.from_numpy(EPS).cuda(), it, st, gap, val, bv, bvv,
                         e["flags"], 1e-6, None, fp64_max_iters=-1, columns=columns)
    torch.cuda.synchronize()
    o = orc.relax_fast_batch(orc.dense_objective(e["Q"], e["c"], 2.0, 0.5, 0.0), True, x.cpu().numpy(), LB.copy(),
                             UB.copy(), EPS.copy(), it.cpu().numpy(), st.cpu().numpy(), flags=ones,
                             max_iters=max(1, BUDGET // 16), refresh=64, threads=8,
                             best_vertex=bv.cpu().numpy(), best_vertex_val=bvv.cpu().numpy(),
                             lmo=orc.simplex_lmo(e["K"], ones) if e["kind"] == "simplex" else None)
    assert_bit_equal(got, o, f"{e['kind']} instance {e['k']} n={n}")


def test_c3_full_shape_one_call():
    import torch
    say(f"[c3] building {COUNT} instances x {NODES} nodes")
    entries = build()
    kinds = [e["kind"] for e in entries]
    assert 0 < kinds.count("simplex") < COUNT
    alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)
    say("[c3] one gb_relax_instances call, concurrency 8")
    gb.relax_instances(entries, alg, BUDGET, concurrency=8)
    torch.cuda.synchronize()
    total_it, total_conv = 0, 0
    for e in entries:
        it, conv = check_properties(e)
        total_it += it
        total_conv += conv
    say(f"[c3] {total_it} FW iterations, {total_conv} of {COUNT * NODES} nodes converged")
    assert total_conv > 0
    # two box and two simplex instances, the smallest and the largest n of each kind
    for kind in ("box", "simplex"):
        ks = sorted((e for e in entries if e["kind"] == kind), key=lambda e: e["n"])
        for e in (ks[0], ks[-1]):
            say(f"[c3] oracle comparison: {kind} instance {e['k']} n={e['n']}")
            compare_with_oracle(e, alg, slice(0, 48), in_call_columns(NODES, min(8, COUNT)))
