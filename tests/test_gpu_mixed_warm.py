"""Warm start of the mixed-precision pass at supplied points
(gb_relax_batch_mixed with status GB_NODE_INITIAL_POINT_SUPPLIED, iterations
0): the FP64 products Q x that start it, bit for bit against exact
k-ordered FMA chains; a node supplied at its relaxation optimum stops after
one iteration; and, for the shapes that start at x and for those that start
at the rounded vertex (3 row tiles per wave, n > 256), the FP64 pass that
follows is bit for bit the oracle's from the pass-1 output."""
import sys
from fractions import Fraction

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import assert_bit_equal, dev

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fma(a, b, c):
    """Correctly rounded a * b + c (exact rational arithmetic, one rounding)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def fma_chain_qx(Q, x):
    """(Q x)_j = fma(Q[j, n-1], x[n-1], ... fma(Q[j, 0], x[0], 0.0)), k ascending."""
    n = len(x)
    out = np.zeros(n)
    for j in range(n):
        acc = 0.0
        for k in range(n):
            acc = fma(float(Q[j, k]), float(x[k]), acc)
        out[j] = acc
    return out


def dense(Q, c):
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, Q.shape[0], dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    return obj, (dQt, dc)


@pytest.mark.parametrize("n", [24, 130])
def test_warm_products_are_k_ordered_fma_chains(n):
    import torch
    Q, c, _ = instances.binary_qp(n, seed=31)
    obj, keep = dense(Q, c)
    rng = np.random.default_rng(n)
    N = 5
    X = rng.random((N, n))
    ST = np.array([gb.NODE_INITIAL_POINT_SUPPLIED, 0, gb.NODE_INITIAL_POINT_SUPPLIED,
                   gb.NODE_INITIAL_POINT_SUPPLIED, gb.NODE_INITIAL_POINT_SUPPLIED], np.uint8)
    IT = np.array([0, 0, 0, 3, 0], np.int32)  # node 1: no point; node 3: a continuation
    out = torch.full((N, n), -7.0, dtype=torch.float64, device="cuda")
    say(f"[warm] products n={n}")
    gb.mixed_warm_products(obj, N, dev(X), dev(ST), dev(IT), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for i in range(N):
        if ST[i] and IT[i] == 0:
            assert np.array_equal(got[i], fma_chain_qx(Q, X[i])), f"node {i}"
        else:
            assert np.all(got[i] == -7.0), f"node {i} must be left as it was"


def run_mixed(mq, obj, LB, UB, X, ST, eps, budget, fp64_iters, rw=0):
    import torch
    N, n = LB.shape
    x = dev(X)
    it = torch.zeros(N, dtype=torch.int32, device="cuda")
    st = dev(ST.astype(np.uint8))
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    mit = torch.zeros(N, dtype=torch.int64, device="cuda")
    gb.relax_batch_mixed(mq, obj, gb.algorithm(dicg=True), gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), N, budget, x,
                         dev(LB), dev(UB), dev(eps), it, st, gap, val, bv, bvv, dev(np.ones(n, np.uint8)), 1e-6,
                         None, fp64_max_iters=fp64_iters, row_tiles_per_wave=rw, mixed_iterations=mit)
    torch.cuda.synchronize()
    return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
            "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
            "best_vertex_val": bvv.cpu().numpy(), "mixed_iterations": mit.cpu().numpy()}


def boxes(n, N, seed):
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.3
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    return LB, UB


@pytest.mark.parametrize("n,rw,starts_at_x", [(60, 0, True), (150, 2, True), (200, 3, False), (300, 0, False)])
def test_supplied_optimum(n, rw, starts_at_x):
    """A node supplied at its relaxation optimum: the shapes that start at x
    stop after one iteration (their Q x is exact, so the gap at x is below
    eps); the others start at the vertex nearest to x and need more."""
    from oracle import pyoracle as orc
    N = 12
    Q, c, _ = instances.binary_qp(n, seed=8)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, 3)
    LB, UB = boxes(n, N, 2)
    # relaxation optima of the nodes, from the oracle's FP64 batched form
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, LB.copy(), LB.copy(), UB.copy(),
                             np.full(N, 1e-12), np.zeros(N, np.int32), np.zeros(N, np.uint8),
                             flags=np.ones(n, np.uint8), max_iters=20000, refresh=64, threads=8)
    assert np.all(o["gap"] < 1e-5)  # well below the pass-1 tolerance 1e-3
    xs = o["x"]
    eps = np.full(N, 1e-3)
    say(f"[warm] supplied optimum, n={n} rw={rw}")
    warm = run_mixed(mq, obj, LB, UB, xs.copy(), np.full(N, gb.NODE_INITIAL_POINT_SUPPLIED), eps, 500, -1, rw)
    if starts_at_x:
        assert np.all(warm["mixed_iterations"] == 1), warm["mixed_iterations"]
        cold = run_mixed(mq, obj, LB, UB, LB.copy(), np.zeros(N), eps, 500, -1, rw)
        assert warm["mixed_iterations"].sum() < cold["mixed_iterations"].sum()
    else:
        assert warm["mixed_iterations"].max() > 1, warm["mixed_iterations"]
    mq.close()


@pytest.mark.parametrize("n,rw", [(60, 0), (150, 2), (200, 3), (300, 0)])
def test_fp64_pass_after_supplied_points_matches_oracle(n, rw):
    from oracle import pyoracle as orc
    N, budget = 40, 400
    Q, c, _ = instances.binary_qp(n, seed=17)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, 3)
    LB, UB = boxes(n, N, 5)
    rng = np.random.default_rng(n)
    X = np.clip(rng.random((N, n)), LB, UB)
    ST = np.full(N, gb.NODE_INITIAL_POINT_SUPPLIED)
    eps = np.full(N, 1e-5)
    say(f"[warm] supplied points n={n} rw={rw}")
    p1 = run_mixed(mq, obj, LB, UB, X.copy(), ST, eps, budget, -1, rw)
    assert np.all(p1["mixed_iterations"] >= 1)
    assert np.all(p1["x"] >= LB) and np.all(p1["x"] <= UB)
    full = run_mixed(mq, obj, LB, UB, X.copy(), ST, eps, budget, 0, rw)
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].astype(np.int32).copy(), p1["status"].copy(),
                             flags=np.ones(n, np.uint8), max_iters=max(1, budget // 16), refresh=64, threads=8,
                             best_vertex=p1["best_vertex"].copy(), best_vertex_val=p1["best_vertex_val"].copy())
    assert_bit_equal({k: v for k, v in full.items() if k != "mixed_iterations"}, o, f"warm n={n} rw={rw}")
    mq.close()
