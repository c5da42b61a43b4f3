"""Batched-form relaxation over the bounded simplex (0/1 variables, integral
rhs; gb_relax_batch_fast with a simplex LMO) against its CPU oracle mirror,
bit for bit, and the batched-form B&B over the simplex against the oracle's
search and the enumerated optimum."""
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


def simplex_nodes(n, N, card, seed, warm=False):
    """0/1 node bounds; node N-1 infeasible (sum lb > card), node N-2 outside
    the 0/1 scope (a bound of 0.5)."""
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.2 * rng.random()
        val = rng.random(n) < 0.3
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
        while LB[i].sum() > card:
            LB[i, np.flatnonzero(LB[i])[0]] = 0.0
    LB[N - 1, : card + 1] = 1.0
    LB[N - 2, 0], UB[N - 2, 0] = 0.0, 0.5
    X = LB.copy()
    ST = np.zeros(N, np.uint8)
    if warm:
        # supplied feasible points on half of the nodes: lb plus mass on free
        # coordinates within the right side
        for i in range(0, N - 2, 2):
            free = np.flatnonzero(UB[i] > LB[i])
            room = card - LB[i].sum()
            w = rng.random(len(free))
            if w.sum() > 0 and room > 0:
                X[i, free] = np.minimum(w / w.sum() * room * 0.9, 1.0)
            ST[i] = gb.NODE_INITIAL_POINT_SUPPLIED
    return LB, UB, X, ST


def run_pair(Q, c, dicg, LB, UB, X, ST, card, equality, eps, max_iters, is_int=None, wpb=4):
    import torch
    from oracle import pyoracle as orc
    N, n = LB.shape
    flags = np.ones(n, np.uint8)
    ii = np.ones(n, np.uint8) if is_int is None else np.asarray(is_int, np.uint8)
    dii = dev(ii)
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 1 if equality else 0, dii.data_ptr(), 0)
    x, lb, ub, e = dev(X), dev(LB), dev(UB), dev(eps)
    it = dev(np.zeros(N, np.int32))
    st = dev(ST)
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    gb.relax_batch_fast(obj, gb.algorithm(dicg=dicg), lmo, N, max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv,
                        dev(flags), 1e-6, None, 64, wpb)
    torch.cuda.synchronize()
    g = {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
         "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
         "best_vertex_val": bvv.cpu().numpy()}
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), dicg, X.copy(), LB.copy(), UB.copy(),
                             np.ascontiguousarray(eps, dtype=np.float64), np.zeros(N, np.int32), ST.copy(),
                             flags=flags, max_iters=max_iters, refresh=64, threads=8,
                             lmo=orc.simplex_lmo(float(card), ii, equality))
    return g, o


@pytest.mark.parametrize("n,dicg,equality,warm,wpb", [
    (20, True, False, False, 4), (20, False, False, False, 4), (48, True, True, False, 1),
    (64, True, False, True, 4), (100, True, False, False, 0), (100, False, True, False, 4),
    (128, False, False, True, 4),
    (200, True, False, True, 4), (300, True, False, False, 2), (512, True, True, False, 4),
    # wpb 0: one node per workgroup with its rows over the waves (the batch fits in flight)
    (80, True, True, True, 0), (128, False, False, False, 0), (200, True, False, True, 0),
    (256, True, True, False, 0), (300, True, False, True, 0), (512, True, False, False, 0)])
def test_fast_simplex_kernel_parity(n, dicg, equality, warm, wpb):
    Q, c, _ = instances.binary_qp(n, seed=17)
    N = 40
    card = max(2, (n + 3) // 4)
    LB, UB, X, ST = simplex_nodes(n, N, card, seed=n + 1, warm=warm)
    say(f"[fast-simplex] n={n} dicg={dicg} equality={equality} warm={warm} N={N} card={card} wpb={wpb}")
    g, o = run_pair(Q, c, dicg, LB, UB, X, ST, card, equality, np.full(N, 1e-7), 300, wpb=wpb)
    assert_bit_equal(g, o, f"fast simplex n={n}")
    assert g["status"][N - 1] == gb.NODE_INFEASIBLE and g["status"][N - 2] == 0xFF
    ok = np.isin(g["status"], (0, 1, 3, 4))  # relaxed nodes
    xs = g["x"][ok]
    assert np.all(xs.sum(axis=1) <= card + 1e-9)
    if equality:
        assert np.all(np.abs(xs.sum(axis=1) - card) <= 1e-9)


def test_fast_simplex_ties_and_signed_zeros():
    """A linear objective with integer costs (and -0.0 / +0.0): the gradient
    has many ties, so the selection's index order decides every vertex."""
    n, N, card = 150, 24, 37
    rng = np.random.default_rng(4)
    Q = np.zeros((n, n))
    c = rng.integers(-3, 3, n).astype(np.float64)
    c[rng.random(n) < 0.2] = -0.0
    LB, UB, X, ST = simplex_nodes(n, N, card, seed=8)
    say("[fast-simplex] ties")
    for equality in (False, True):
        g, o = run_pair(Q, c, True, LB, UB, X, ST, card, equality, np.full(N, 1e-9), 50)
        assert_bit_equal(g, o, f"ties equality={equality}")


def test_fast_simplex_not_integer_is_outside_scope():
    n, N, card = 40, 4, 10
    Q, c, _ = instances.binary_qp(n, seed=5)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    ii = np.ones(n, np.uint8)
    ii[7] = 0
    say("[fast-simplex] is_int scope")
    g, o = run_pair(Q, c, True, LB, UB, LB.copy(), np.zeros(N, np.uint8), card, False, np.full(N, 1e-7), 50,
                    is_int=ii)
    assert_bit_equal(g, o, "scope")
    assert np.all(g["status"] == 0xFF)


@pytest.mark.parametrize("n,seed,card", [(14, 6, 4), (40, 2, 10), (60, 3, 15)])
def test_fast_simplex_bnb_matches_oracle(n, seed, card):
    from oracle import pyoracle as orc
    Q, c, sha = instances.binary_qp(n, seed=seed)
    say(f"[fast-simplex-bnb] n={n} seed={seed} card={card} sha={sha[:12]}")
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    ii = dev(np.ones(n, np.uint8))
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, ii.data_ptr(), 0)
    s = gb.default_settings()
    s.abs_gap_tol = s.rel_gap_tol = 1e-6
    s.form = 1
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, lmo, s)
    o = orc.bnb_solve(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), orc.alg(),
                      orc.simplex_lmo(float(card), np.ones(n, np.uint8)), [0] * n, [1] * n, list(range(n)),
                      abs_gap_tol=1e-6, rel_gap_tol=1e-6, threads=8, form=1)
    assert r["nodes_processed"] == o["nodes_processed"]
    assert r["objective"] == o["objective"] and r["dual_bound"] == o["dual_bound"]
    assert r["x"] == list(o["x"])
    assert sum(r["x"]) <= card
    # the plain form proves the same optimum
    s.form = 0
    p = gb.solve(n, [0] * n, [1] * n, range(n), obj, lmo, s)
    assert abs(p["objective"] - r["objective"]) <= 1e-6 * max(1.0, abs(r["objective"]))
    if n <= 16:
        best = min(np.array(b) @ Q @ np.array(b) + c @ np.array(b)
                   for b in itertools.product((0.0, 1.0), repeat=n) if sum(b) <= card)
        assert abs(r["objective"] - best) <= 1e-9 * max(1.0, abs(best))
