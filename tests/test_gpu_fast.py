"""Batched-form relaxation kernel (gb_relax_batch_fast) against its CPU
oracle mirror, bit for bit, and the batched-form B&B against the oracle and
the exhaustive optimum."""
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


def nodes(n, N, seed, warm=False):
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.3 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    X = np.clip(rng.random((N, n)), LB, UB) if warm else LB.copy()
    st = np.full(N, gb.NODE_INITIAL_POINT_SUPPLIED if warm else 0, dtype=np.uint8)
    return LB, UB, X, st


def run_pair(Q, c, dicg, LB, UB, X, ST, eps, max_iters, cutoff=None, refresh=64, wpb=4):
    import torch
    from oracle import pyoracle as orc
    N, n = LB.shape
    flags = np.ones(n, np.uint8)
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    x, lb, ub, e = dev(X), dev(LB), dev(UB), dev(eps)
    it = dev(np.zeros(N, np.int32))
    st = dev(ST)
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    cut = None if cutoff is None else dev(cutoff)
    gb.relax_batch_fast(obj, gb.algorithm(dicg=dicg), gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), N,
                        max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv, dev(flags), 1e-6, cut,
                        refresh, wpb)
    torch.cuda.synchronize()
    g = {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
         "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
         "best_vertex_val": bvv.cpu().numpy()}
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), dicg, X.copy(), LB.copy(),
                             UB.copy(), np.ascontiguousarray(eps, dtype=np.float64),
                             np.zeros(N, np.int32), ST.copy(), flags=flags, max_iters=max_iters,
                             cutoff=cutoff, refresh=refresh, threads=8)
    return g, o


@pytest.mark.parametrize("n,dicg,warm", [(20, True, False), (20, False, False), (60, True, True),
                                         (100, True, False), (200, True, True), (200, False, True),
                                         (300, True, False)])
def test_fast_kernel_parity(n, dicg, warm):
    Q, c, _ = instances.binary_qp(n, seed=11)
    N = 48
    LB, UB, X, ST = nodes(n, N, seed=n, warm=warm)
    say(f"[fast] n={n} dicg={dicg} warm={warm} N={N}")
    g, o = run_pair(Q, c, dicg, LB, UB, X, ST, np.full(N, 1e-6), 300)
    assert_bit_equal(g, o, f"fast n={n}")


def test_fast_kernel_cutoff_and_refresh():
    n, N = 60, 32
    Q, c, _ = instances.binary_qp(n, seed=3)
    LB, UB, X, ST = nodes(n, N, seed=9)
    cutoff = np.linspace(-3.0, 0.5, N)
    for refresh in (0, 7):
        say(f"[fast] cutoff refresh={refresh}")
        g, o = run_pair(Q, c, True, LB, UB, X, ST, np.full(N, 1e-7), 500, cutoff=cutoff, refresh=refresh)
        assert_bit_equal(g, o, "cutoff")
    # the certificate is a valid lower bound: f(x) - gap <= every feasible value
    # scored on the node (best vertex)
    lbnd = g["value"] - g["gap"]
    ok = g["best_vertex_val"] < 1e299
    assert np.all(lbnd[ok] <= g["best_vertex_val"][ok] + 1e-9)


def test_fast_kernel_single_wave_blocks():
    n, N = 128, 40
    Q, c, _ = instances.binary_qp(n, seed=5)
    LB, UB, X, ST = nodes(n, N, seed=2)
    g1, o = run_pair(Q, c, True, LB, UB, X, ST, np.full(N, 1e-6), 200, wpb=1)
    assert_bit_equal(g1, o, "wpb1")


@pytest.mark.parametrize("n,seed", [(20, 1), (20, 3), (40, 2), (60, 1), (80, 2)])
def test_fast_bnb_matches_oracle(n, seed):
    # n = 80: full 1024-node batches (byte bounds on the bus, best vertices
    # fetched on incumbent improvements, sibling families in the frontier)
    from oracle import pyoracle as orc
    Q, c, sha = instances.binary_qp(n, seed=seed, rows=30 if n == 20 else None)
    say(f"[fast-bnb] n={n} seed={seed} sha={sha[:12]}")
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    s = gb.default_settings()
    s.abs_gap_tol = s.rel_gap_tol = 1e-6
    s.form = 1
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
    o = orc.bnb_solve(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), orc.alg(), orc.box_lmo(), [0] * n,
                      [1] * n, list(range(n)), abs_gap_tol=1e-6, rel_gap_tol=1e-6, threads=8, form=1)
    assert r["nodes_processed"] == o["nodes_processed"]
    assert r["objective"] == o["objective"] and r["dual_bound"] == o["dual_bound"]
    assert r["x"] == list(o["x"])
    # the plain form proves the same optimum
    s.form = 0
    p = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
    assert p["x"] == r["x"]
    assert abs(p["objective"] - r["objective"]) <= 1e-9 * max(1.0, abs(r["objective"]))
    if n == 20:
        f, xe = orc.enumerate_binary_qp(Q, c)
        assert r["x"] == list(xe)
