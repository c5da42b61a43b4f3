"""Mixed-precision form over the bounded simplex (0/1 variables, integral
rhs): pass 1 (BF16 MFMA products, FP32 iterates, the greedy vertices chosen
by one wave per node column) and the FP64 pass of the batched simplex
kernel, which starts from pass 1's iterate scaled onto the mass limit.

Checks: pass 1's outputs are feasible (iterate within the bounds and the
mass limit up to FP32 rounding, best vertices 0/1 points of the set with
their FP32-tracked value close to the exact one); the FP64 pass from pass
1's output equals the oracle's batched simplex run from the same output,
bit for bit; the full relaxation reaches the batched FP64 form's values and
certifies bounds below its best vertices; the mixed search over a simplex
proves the enumerated optimum; heterogeneous batches with mixed simplex
entries equal their solo calls."""
import itertools
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


def nodes01(n, N, card, seed, equality=False):
    """0/1 node bounds (node 0 the root); node N-1 infeasible (sum lb > card)
    and, for an equality, node N-2 too (sum ub < card)."""
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.25 * rng.random()
        val = rng.random(n) < 0.3
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
        while LB[i].sum() > card:
            LB[i, np.flatnonzero(LB[i])[0]] = 0.0
    LB[N - 1, : card + 1] = 1.0
    UB[N - 1, : card + 1] = 1.0
    if equality:
        LB[N - 2, :] = 0.0
        UB[N - 2, :] = 0.0
        UB[N - 2, : card - 1] = 1.0
    return LB, UB


class Case:
    def __init__(self, n, card, equality=False, planes=3, seed=7):
        import torch
        self.n, self.card, self.equality = n, card, equality
        self.Q, self.c, _ = instances.binary_qp(n, seed=seed)
        self.dQt, self.dc = dev(np.ascontiguousarray(self.Q.T)), dev(self.c)
        self.obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, self.dQt.data_ptr(), self.dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        self.ii = dev(np.ones(n, np.uint8))
        self.flags = self.ii
        self.lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 1 if equality else 0, self.ii.data_ptr(), 0)
        self.mq = gb.MixedQ(self.obj, planes)
        self.alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)
        self.torch = torch

    def buffers(self, LB, UB, eps):
        torch = self.torch
        N, n = LB.shape
        f64 = dict(dtype=torch.float64, device="cuda")
        return {"x": dev(LB.copy()), "lb": dev(LB), "ub": dev(UB), "eps": dev(eps),
                "iterations": torch.zeros(N, dtype=torch.int32, device="cuda"),
                "status": torch.zeros(N, dtype=torch.uint8, device="cuda"), "gap": torch.zeros(N, **f64),
                "value": torch.zeros(N, **f64), "best_vertex": torch.zeros((N, n), **f64),
                "best_vertex_val": torch.full((N,), 1e300, **f64)}

    def mixed(self, b, budget, fp64_max_iters=0, columns=0):
        N = b["x"].shape[0]
        gb.relax_batch_mixed(self.mq, self.obj, self.alg, self.lmo, N, budget, b["x"], b["lb"], b["ub"], b["eps"],
                             b["iterations"], b["status"], b["gap"], b["value"], b["best_vertex"],
                             b["best_vertex_val"], self.flags, 1e-6, None, fp64_max_iters=fp64_max_iters,
                             columns=columns)
        self.torch.cuda.synchronize()

    def fast(self, b, budget):
        N = b["x"].shape[0]
        gb.relax_batch_fast(self.obj, self.alg, self.lmo, N, budget, b["x"], b["lb"], b["ub"], b["eps"],
                            b["iterations"], b["status"], b["gap"], b["value"], b["best_vertex"],
                            b["best_vertex_val"], self.flags, 1e-6, None, 64, 4)
        self.torch.cuda.synchronize()


def host(b):
    return {k: v.cpu().numpy() for k, v in b.items() if k not in ("lb", "ub", "eps")}


@pytest.mark.parametrize("n,planes,equality,columns", [
    (20, 3, False, 0), (100, 1, False, 0), (100, 3, True, 8), (200, 3, False, 0), (200, 1, False, 16),
    (300, 3, False, 0), (512, 3, True, 0), (512, 1, False, 8)])
def test_mixed_simplex_pass1_and_fp64_pass(n, planes, equality, columns):
    from oracle import pyoracle as orc
    card = max(2, (n + 3) // 4)
    C = Case(n, card, equality, planes, seed=n + 3)
    N = 48
    LB, UB = nodes01(n, N, card, seed=n, equality=equality)
    infeasible = [N - 1] + ([N - 2] if equality else [])
    eps = np.full(N, 1e-6)
    say(f"[mixed-simplex] n={n} planes={planes} equality={equality} columns={columns} card={card}")
    b = C.buffers(LB, UB, eps)
    C.mixed(b, 200, fp64_max_iters=-1, columns=columns)
    p1 = host(b)
    # pass 1: feasible iterates and vertices; infeasible nodes untouched
    for i in range(N):
        x = p1["x"][i]
        if i in infeasible:  # left to the FP64 pass
            assert p1["status"][i] == 0 and p1["iterations"][i] == 0
            continue
        assert p1["iterations"][i] > 0 and p1["status"][i] == gb.NODE_INITIAL_POINT_SUPPLIED
        assert np.all(x >= LB[i] - 1e-6) and np.all(x <= UB[i] + 1e-6), i
        assert x.sum() <= card + 1e-3 * (1 + card), i
        if equality:
            assert abs(x.sum() - card) <= 1e-3 * (1 + card), i
        if p1["best_vertex_val"][i] < 1e299:
            v = p1["best_vertex"][i]
            assert np.all((v == 0.0) | (v == 1.0)) and np.all(v >= LB[i]) and np.all(v <= UB[i])
            assert v.sum() <= card and (not equality or v.sum() == card)
            fv = v @ C.Q @ v + C.c @ v
            tol = 1e-2 if planes == 1 else 1e-4  # one bf16 plane of Q: about 2^-8 relative per product
            assert abs(p1["best_vertex_val"][i] - fv) <= tol * (1 + abs(fv)), i
    # the whole call (pass 1 again, then the FP64 pass, which re-scores pass
    # 1's vertex) equals the oracle's FP64 pass from pass 1's output, bit for bit
    fp64_iters = 200 // 16
    b2 = C.buffers(LB, UB, eps)
    C.mixed(b2, 200, columns=columns)
    got = host(b2)
    o = orc.relax_fast_batch(orc.dense_objective(C.Q, C.c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].copy(), p1["status"].copy(),
                             flags=np.ones(n, np.uint8), max_iters=fp64_iters, refresh=64, threads=8,
                             best_vertex=p1["best_vertex"].copy(), best_vertex_val=p1["best_vertex_val"].copy(),
                             lmo=orc.simplex_lmo(float(card), np.ones(n, np.uint8), equality))
    assert_bit_equal(got, o, f"FP64 pass n={n}")
    ok = np.isin(got["status"], (0, 1, 3, 4))
    assert all(got["status"][i] == gb.NODE_INFEASIBLE for i in infeasible)
    assert np.all(got["x"][ok].sum(axis=1) <= card + 1e-9)


@pytest.mark.parametrize("n,equality", [(60, False), (200, False), (200, True), (420, False)])
def test_mixed_simplex_full_relaxation_reaches_fp64_values(n, equality):
    card = (n + 3) // 4
    C = Case(n, card, equality, 3, seed=n + 11)
    N = 64
    LB, UB = nodes01(n, N, card, seed=2 * n, equality=equality)
    eps = np.full(N, 1e-7)
    say(f"[mixed-simplex] full relaxation n={n} equality={equality}")
    bm = C.buffers(LB, UB, eps)
    C.mixed(bm, 4000, fp64_max_iters=4000)
    bf = C.buffers(LB, UB, eps)
    C.fast(bf, 4000)
    m, f = host(bm), host(bf)
    assert np.array_equal(m["status"] == 2, f["status"] == 2)
    conv = np.isin(m["status"], (1, 4)) & np.isin(f["status"], (1, 4))
    say(f"[mixed-simplex] converged: mixed {np.isin(m['status'], (1, 4)).sum()}, batched "
        f"{np.isin(f['status'], (1, 4)).sum()}, both {conv.sum()} of {N}; iterations mixed "
        f"{int(m['iterations'].sum())} batched {int(f['iterations'].sum())}")
    assert conv.sum() >= N // 2
    assert np.allclose(m["value"][conv], f["value"][conv], rtol=1e-5, atol=1e-5)
    has = m["best_vertex_val"] < 1e299
    ok = np.isin(m["status"], (0, 1, 3, 4)) & has
    assert np.all(m["value"][ok] - m["gap"][ok] <= m["best_vertex_val"][ok] + 1e-9 * (1 + np.abs(m["best_vertex_val"][ok])))


@pytest.mark.parametrize("n,seed,card", [(14, 6, 4), (40, 2, 10), (70, 4, 18)])
def test_mixed_simplex_bnb_same_optimum(n, seed, card):
    Q, c, sha = instances.binary_qp(n, seed=seed)
    say(f"[mixed-simplex-bnb] n={n} seed={seed} card={card} sha={sha[:12]}")
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    ii = dev(np.ones(n, np.uint8))
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, ii.data_ptr(), 0)
    res = {}
    for form in (2, 1, 0):
        s = gb.default_settings()
        s.abs_gap_tol = s.rel_gap_tol = 1e-6
        s.form = form
        if form == 2:
            s.precision = gb.PREC_BF16
        res[form] = gb.solve(n, [0] * n, [1] * n, range(n), obj, lmo, s)
        assert res[form]["reached_tolerance"] or res[form]["objective"] - res[form]["dual_bound"] <= 1e-6 * (
            1 + abs(res[form]["objective"]))
    for form in (1, 0):
        assert abs(res[2]["objective"] - res[form]["objective"]) <= 1e-6 * max(1.0, abs(res[form]["objective"]))
    assert sum(res[2]["x"]) <= card
    if n <= 16:
        best = min(np.array(b) @ Q @ np.array(b) + c @ np.array(b)
                   for b in itertools.product((0.0, 1.0), repeat=n) if sum(b) <= card)
        assert abs(res[2]["objective"] - best) <= 1e-9 * max(1.0, abs(best))


def test_instances_with_mixed_simplex_entries_equal_solo_calls():
    import torch
    specs = [(48, 40, 12), (150, 33, 38), (300, 20, 75), (70, 25, 0), (420, 16, 0)]
    alg = gb.algorithm(dicg=True, line_search=gb.LS_QUADRATIC_EXACT)
    budget = 150

    def entry(k, n, N, card):
        Q, c, _ = instances.binary_qp(n, seed=700 + k)
        dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
        obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
        LB, UB, eps = instances.open_nodes(n, N, 800 + k, card=card)
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
    say("[mixed-simplex] heterogeneous call with mixed simplex entries")
    gb.relax_instances(together, alg, budget, concurrency=3)
    for e in solo:
        gb.relax_batch_mixed(e["mixed"], e["obj"], alg, e["lmo"], e["x"].shape[0], budget, e["x"], e["lb"],
                             e["ub"], e["eps"], e["iterations"], e["status"], e["gap"], e["value"],
                             e["best_vertex"], e["best_vertex_val"], e["flags"], 1e-6, e["cutoff"])
    torch.cuda.synchronize()
    keys = ("x", "gap", "value", "iterations", "status", "best_vertex", "best_vertex_val")
    for k, (a, b) in enumerate(zip(together, solo)):
        assert_bit_equal({q: a[q].cpu().numpy() for q in keys}, {q: b[q].cpu().numpy() for q in keys},
                         f"entry {k}")


@pytest.mark.parametrize("n,rw,columns,planes", [
    (100, 1, 2, 3), (100, 2, 8, 3), (100, 3, 16, 1), (100, 4, 4, 3), (150, 3, 4, 3), (150, 2, 16, 1),
    (300, 4, 2, 3), (300, 4, 16, 1)])
def test_mixed_simplex_tile_shapes(n, rw, columns, planes):
    """Every tile shape (row tiles per wave, node columns per tile) computes
    the same algorithm: pass 1 stays feasible and the whole call equals the
    oracle's FP64 pass from pass 1's output."""
    import torch
    from oracle import pyoracle as orc
    card = (n + 3) // 4
    C = Case(n, card, False, planes, seed=n + 29)
    N = 40
    LB, UB = nodes01(n, N, card, seed=3 * n + rw)
    eps = np.full(N, 1e-6)
    say(f"[mixed-simplex] shape n={n} rw={rw} columns={columns} planes={planes}")

    def run(fp64):
        b = C.buffers(LB, UB, eps)
        gb.relax_batch_mixed(C.mq, C.obj, C.alg, C.lmo, N, 150, b["x"], b["lb"], b["ub"], b["eps"], b["iterations"],
                             b["status"], b["gap"], b["value"], b["best_vertex"], b["best_vertex_val"], C.flags, 1e-6,
                             None, fp64_max_iters=fp64, row_tiles_per_wave=rw, columns=columns)
        torch.cuda.synchronize()
        return host(b)

    p1 = run(-1)
    for i in range(N - 1):
        x = p1["x"][i]
        assert p1["iterations"][i] > 0, i
        assert np.all(x >= LB[i] - 1e-6) and np.all(x <= UB[i] + 1e-6) and x.sum() <= card + 1e-3 * (1 + card), i
    got = run(0)
    o = orc.relax_fast_batch(orc.dense_objective(C.Q, C.c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].copy(), p1["status"].copy(),
                             flags=np.ones(n, np.uint8), max_iters=150 // 16, refresh=64, threads=8,
                             best_vertex=p1["best_vertex"].copy(), best_vertex_val=p1["best_vertex_val"].copy(),
                             lmo=orc.simplex_lmo(float(card), np.ones(n, np.uint8)))
    assert_bit_equal(got, o, f"shape n={n} rw={rw} columns={columns}")
