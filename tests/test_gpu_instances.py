"""Heterogeneous instance batches (gb_relax_instances): batches of different
problems and sizes relaxed in one call give, entry by entry, bit for bit
what each entry's solo call gives."""
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import assert_bit_equal, dev

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_entry(n, N, seed, mixed_planes, card=0):
    """card > 0: a bounded-simplex entry (sum x <= card), relaxed in the batched FP64 form."""
    import torch
    Q, c, _ = instances.binary_qp(n, seed=seed)
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    if card > 0:
        LB, UB, eps = instances.open_nodes(n, N, seed, card=card)
        f64 = dict(dtype=torch.float64, device="cuda")
        flags = dev(np.ones(n, np.uint8))
        return {"obj": obj, "lmo": gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, flags.data_ptr(), 0),
                "keep": (dQt, dc), "mixed": None,
                "x": dev(LB.copy()), "lb": dev(LB), "ub": dev(UB), "eps": dev(eps),
                "iterations": torch.zeros(N, dtype=torch.int32, device="cuda"),
                "status": torch.zeros(N, dtype=torch.uint8, device="cuda"),
                "gap": torch.zeros(N, **f64), "value": torch.zeros(N, **f64),
                "best_vertex": torch.zeros((N, n), **f64), "best_vertex_val": torch.full((N,), 1e300, **f64),
                "flags": flags, "cutoff": None}
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.3 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    eps = np.maximum(1e-6, 1e-3 * 0.8 ** rng.integers(0, 10, N))
    f64 = dict(dtype=torch.float64, device="cuda")
    e = {"obj": obj, "lmo": gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), "keep": (dQt, dc),
         "mixed": gb.MixedQ(obj, mixed_planes) if mixed_planes else None,
         "x": dev(LB.copy()), "lb": dev(LB), "ub": dev(UB), "eps": dev(eps),
         "iterations": torch.zeros(N, dtype=torch.int32, device="cuda"),
         "status": torch.zeros(N, dtype=torch.uint8, device="cuda"),
         "gap": torch.zeros(N, **f64), "value": torch.zeros(N, **f64), "best_vertex": torch.zeros((N, n), **f64),
         "best_vertex_val": torch.full((N,), 1e300, **f64), "flags": dev(np.ones(n, np.uint8)),
         "cutoff": dev(np.full(N, -1.0)) if seed % 2 else None}
    return e


def results(e):
    return {k: e[k].cpu().numpy() for k in ("x", "gap", "value", "iterations", "status", "best_vertex",
                                              "best_vertex_val")}


@pytest.mark.parametrize("concurrency", [1, 4])
def test_instances_equal_solo_runs(concurrency):
    import torch
    sizes = [(32, 40, 1, 0), (45, 17, 3, 0), (64, 100, 3, 0), (70, 30, 0, 18), (100, 64, 1, 0), (150, 33, 3, 0),
             (200, 80, 1, 0), (256, 20, 3, 0), (260, 24, 0, 65), (300, 50, 1, 0), (420, 25, 3, 0), (512, 16, 0, 0)]
    alg = gb.algorithm(dicg=True)
    budget = 250
    together = [make_entry(n, N, 100 + k, planes, card) for k, (n, N, planes, card) in enumerate(sizes)]
    solo = [make_entry(n, N, 100 + k, planes, card) for k, (n, N, planes, card) in enumerate(sizes)]
    say(f"[instances] {len(sizes)} instances, n from 32 to 512, concurrency {concurrency}")
    gb.relax_instances(together, alg, budget, concurrency=concurrency)
    for e in solo:
        if e["mixed"] is not None:
            gb.relax_batch_mixed(e["mixed"], e["obj"], alg, e["lmo"], e["x"].shape[0], budget, e["x"], e["lb"],
                                 e["ub"], e["eps"], e["iterations"], e["status"], e["gap"], e["value"],
                                 e["best_vertex"], e["best_vertex_val"], e["flags"], 1e-6, e["cutoff"])
        else:
            gb.relax_batch_fast(e["obj"], alg, e["lmo"], e["x"].shape[0], budget, e["x"], e["lb"], e["ub"],
                                e["eps"], e["iterations"], e["status"], e["gap"], e["value"], e["best_vertex"],
                                e["best_vertex_val"], e["flags"], 1e-6, e["cutoff"])
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(together, solo)):
        assert_bit_equal(results(a), results(b), f"instance {k} n={sizes[k][0]}")
        assert int(a["iterations"].sum().item()) > 0


def test_instances_widen_mixed_tiles():
    """Large mixed entries sharing the chip run wider tiles (up to the MFMA's
    16 node columns): each equals its solo call with
    columns = min(16, max(ceil(N / 256), ceil(N k / CUs))), k = min(concurrency, count)."""
    import torch
    specs = [(100, 1024, 0), (300, 700, 75), (60, 1024, 15), (450, 512, 0)]
    alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)
    budget, conc = 60, 4
    cus = torch.cuda.get_device_properties(0).multi_processor_count

    def entry(k, n, N, card):
        Q, c, _ = instances.binary_qp(n, seed=900 + k)
        dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
        obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        LB, UB, eps = instances.open_nodes(n, N, 950 + k, card=card)
        flags = dev(np.ones(n, np.uint8))
        lmo = (gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, flags.data_ptr(), 0) if card
               else gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0))
        f64 = dict(dtype=torch.float64, device="cuda")
        return {"obj": obj, "lmo": lmo, "keep": (dQt, dc), "mixed": gb.MixedQ(obj, 3),
                "x": dev(LB.copy()), "lb": dev(LB), "ub": dev(UB), "eps": dev(eps),
                "iterations": torch.zeros(N, dtype=torch.int32, device="cuda"),
                "status": torch.zeros(N, dtype=torch.uint8, device="cuda"),
                "gap": torch.zeros(N, **f64), "value": torch.zeros(N, **f64),
                "best_vertex": torch.zeros((N, n), **f64), "best_vertex_val": torch.full((N,), 1e300, **f64),
                "flags": flags, "cutoff": None}

    together = [entry(k, *sp) for k, sp in enumerate(specs)]
    solo = [entry(k, *sp) for k, sp in enumerate(specs)]
    k = min(conc, len(specs))
    cols = [min(16, max(min(16, max(1, -(-N // 256))), -(-N * k // cus))) for _, N, _ in specs]
    say(f"[instances] wide mixed tiles: columns {cols} on {cus} CUs")
    assert max(cols) == 16 and min(cols) < 16
    gb.relax_instances(together, alg, budget, concurrency=conc)
    for e, cc in zip(solo, cols):
        gb.relax_batch_mixed(e["mixed"], e["obj"], alg, e["lmo"], e["x"].shape[0], budget, e["x"], e["lb"],
                             e["ub"], e["eps"], e["iterations"], e["status"], e["gap"], e["value"],
                             e["best_vertex"], e["best_vertex_val"], e["flags"], 1e-6, e["cutoff"], columns=cc)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(together, solo)):
        assert_bit_equal(results(a), results(b), f"entry {i} n={specs[i][0]} columns={cols[i]}")
        assert int(a["iterations"].sum().item()) > 0

    # under capture the call keeps the one-stream width: a replay equals an
    # eager call with concurrency 1
    one = [entry(k, *sp) for k, sp in enumerate(specs)]
    gb.relax_instances(one, alg, budget, concurrency=1)
    torch.cuda.synchronize()
    cap = [entry(k, *sp) for k, sp in enumerate(specs)]
    init = [(e["x"].clone(), e["iterations"].clone(), e["status"].clone()) for e in cap]

    def refill():
        for e, (x0, it0, st0) in zip(cap, init):
            e["x"].copy_(x0)
            e["iterations"].copy_(it0)
            e["status"].copy_(st0)
            e["best_vertex_val"].fill_(1e300)

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        gb.relax_instances(cap, alg, budget, concurrency=conc, stream=s.cuda_stream)  # allocations
        refill()
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gb.relax_instances(cap, alg, budget, concurrency=conc, stream=s.cuda_stream)
    s.synchronize()
    for rep in range(2):
        with torch.cuda.stream(s):
            refill()
        s.synchronize()
        g.replay()
        torch.cuda.synchronize()
        say(f"[instances] wide tiles: captured replay {rep}")
        for i, (a, b) in enumerate(zip(cap, one)):
            assert_bit_equal(results(a), results(b), f"replay {rep} entry {i} n={specs[i][0]}")
    del g


def test_instances_reject_bad_entries():
    e = make_entry(40, 8, 1, 0)
    e["obj"] = gb.ObjectiveDesc(gb.OBJ_DENSE, 40, 0, 0, 2.0, 0.5, 0.0, 0, 0, 0)  # no Qt
    with pytest.raises(gb.InvalidArgument):
        gb.relax_instances([e], gb.algorithm(), 10)


def test_instances_captured_call_replays_match_eager():
    """One gb_relax_instances call captured in a HIP graph (after a warm-up
    call has made its allocations): each replay, with the inputs refilled by
    eager work and a synchronisation in between, equals the eager call."""
    import torch
    sizes = [(48, 40, 1), (150, 33, 3), (300, 20, 1), (420, 16, 0)]
    alg = gb.algorithm(dicg=True)
    budget = 150
    entries = [make_entry(n, N, 300 + k, planes) for k, (n, N, planes) in enumerate(sizes)]
    init = [(e["x"].clone(), e["iterations"].clone(), e["status"].clone()) for e in entries]

    def refill():
        for e, (x0, it0, st0) in zip(entries, init):
            e["x"].copy_(x0)
            e["iterations"].copy_(it0)
            e["status"].copy_(st0)
            e["best_vertex_val"].fill_(1e300)

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        refill()
        gb.relax_instances(entries, alg, budget, concurrency=4, stream=s.cuda_stream)
    s.synchronize()
    eager = [results(e) for e in entries]
    with torch.cuda.stream(s):
        refill()
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gb.relax_instances(entries, alg, budget, concurrency=4, stream=s.cuda_stream)
    s.synchronize()
    for rep in range(3):
        with torch.cuda.stream(s):
            refill()
        s.synchronize()
        g.replay()
        torch.cuda.synchronize()
        say(f"[instances] captured replay {rep}")
        for k, e in enumerate(entries):
            assert_bit_equal(results(e), eager[k], f"replay {rep} instance {k} n={sizes[k][0]}")


def test_instances_captured_call_with_simplex_entries():
    """Box entries (mixed and batched forms) and bounded-simplex entries
    (batched FP64 form) in one captured gb_relax_instances call: every replay,
    with refills in between, equals the eager call."""
    import torch
    specs = [(48, 40, 1, 0), (150, 33, 3, 38), (300, 20, 0, 0), (64, 25, 0, 16), (420, 16, 3, 0)]
    alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)
    budget = 150
    entries = [make_entry(n, N, 500 + k, planes, card) for k, (n, N, planes, card) in enumerate(specs)]
    init = [(e["x"].clone(), e["iterations"].clone(), e["status"].clone()) for e in entries]

    def refill():
        for e, (x0, it0, st0) in zip(entries, init):
            e["x"].copy_(x0)
            e["iterations"].copy_(it0)
            e["status"].copy_(st0)
            e["best_vertex_val"].fill_(1e300)

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        refill()
        gb.relax_instances(entries, alg, budget, concurrency=3, stream=s.cuda_stream)
    s.synchronize()
    eager = [results(e) for e in entries]
    with torch.cuda.stream(s):
        refill()
    s.synchronize()
    say("[instances] capturing a call with bounded-simplex entries")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gb.relax_instances(entries, alg, budget, concurrency=3, stream=s.cuda_stream)
    s.synchronize()
    for rep in range(3):
        with torch.cuda.stream(s):
            refill()
        s.synchronize()
        g.replay()
        torch.cuda.synchronize()
        # eager calls between replays use counters of their own
        gb.relax_instances(entries[:1], alg, budget, concurrency=1)
        torch.cuda.synchronize()
        say(f"[instances] captured replay {rep} with simplex entries")
        for k, e in enumerate(entries[1:], start=1):
            assert_bit_equal(results(e), eager[k], f"replay {rep} instance {k} n={specs[k][0]}")
