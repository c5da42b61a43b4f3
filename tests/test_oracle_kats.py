"""Pins the CPU oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, brute force, enumeration).

Every expected value below is either copied from a reference test
(file:line cited) or computed independently of the oracle's formulas.
"""
import ctypes
import itertools
import math

import numpy as np
import pytest

from gpuboscia_amd import instances


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --- gap_utils.h KATs (tests/test_cpu.cpp:41-158) -----------------------------

def test_certified_bound(orc):
    lib = orc.load()
    assert abs(lib.orc_certified_fw_lower_bound(-7.0, 10.0, 3.0) - 7.0) < 1e-12
    assert lib.orc_certified_fw_lower_bound(-7.0, 10.0, -1e-12) == -7.0
    assert lib.orc_certified_fw_lower_bound(-7.0, 10.0, float("nan")) == -7.0
    assert lib.orc_certified_fw_lower_bound(-7.0, float("inf"), 1.0) == -7.0


def test_gap_helpers(orc):
    lib = orc.load()
    assert abs(lib.orc_gap_absolute(9.0, 10.0) - 1.0) < 1e-12
    assert abs(lib.orc_gap_relative(9.0, 10.0) - 1.0 / 9.0) < 1e-12
    assert abs(lib.orc_gap_relative(0.25, 0.75) - 2.0) < 1e-12
    assert abs(lib.orc_gap_relative(-10.0, -9.0) - 1.0 / 9.0) < 1e-12
    assert math.isinf(lib.orc_gap_relative(0.0, 0.0))
    assert math.isinf(lib.orc_gap_relative(0.0, 1.0))
    assert math.isinf(lib.orc_gap_relative(-1.0, 0.0))
    assert lib.orc_gap_within_tolerance(0.0, 0.0, 0.0, 0.01)
    assert not lib.orc_gap_within_tolerance(9.0, 10.0, 0.0, 0.105)
    assert not lib.orc_gap_within_tolerance(0.0, 1.0, 0.01, 0.01)
    assert not lib.orc_gap_within_tolerance(-1.0, 0.0, 0.01, 0.01)
    for bad in (float("inf"), float("-inf"), float("nan")):
        assert math.isinf(lib.orc_gap_relative(bad, 1.0)) and math.isinf(lib.orc_gap_relative(1.0, bad))
        assert not lib.orc_gap_within_tolerance(bad, 1.0, 0.01, 0.01)
        assert not lib.orc_gap_within_tolerance(1.0, bad, 0.01, 0.01)
    assert lib.orc_gap_within_tolerance(9.99, 10.0, 0.02, 1e-12)
    assert lib.orc_gap_within_tolerance(99.0, 100.0, 1e-12, 0.011)
    assert not lib.orc_gap_within_tolerance(-1.0, 1.0, 0.1, 0.1)
    assert abs(lib.orc_depth_dependent_tolerance(1e-3, 0.8, 1e-6, 0) - 1e-3) < 1e-15
    assert abs(lib.orc_depth_dependent_tolerance(1e-3, 0.8, 1e-6, 2) - 6.4e-4) < 1e-15
    assert abs(lib.orc_depth_dependent_tolerance(1e-3, 0.8, 1e-6, 1000) - 1e-6) < 1e-15


# --- branching.h KATs (tests/test_cpu.cpp:311-370) -------------------------

def topk(orc, x, iv, k, tol=1e-6):
    lib = orc.load()
    out = (ctypes.c_int * max(1, k))()
    xa = c(x)
    iva = np.ascontiguousarray(iv, dtype=np.int32)
    m = lib.orc_top_k_infeasible(orc.p(xa), orc.p(iva), len(iv), k, ctypes.c_double(tol), out)
    return [out[i] for i in range(m)]


def test_top_k(orc):
    assert topk(orc, [1.1, 2.5, 3.0], [0, 1, 2], 2) == [1, 0]
    assert len(topk(orc, [1.1, 2.5, 3.0], [0, 1, 2], 5)) == 2
    assert topk(orc, [1.0, 2.0], [0, 1], 3) == []
    assert topk(orc, [1.0 + 5e-5], [0], 1, 1e-4) == []
    assert topk(orc, [1.0 + 5e-5], [0], 1, 1e-6) == [0]
    lib = orc.load()
    xa = c([2.5, 1.1, 3.0])
    iva = np.array([0, 1, 2], dtype=np.int32)
    assert lib.orc_most_infeasible(orc.p(xa), orc.p(iva), 3) == 0
    xa = c([1.0, 2.0, 3.0])
    assert lib.orc_most_infeasible(orc.p(xa), orc.p(iva), 3) == 0


# --- lmo_dicg.h KATs (tests/test_cpu.cpp:399-517) --------------------------

def test_box_dicg_kat(orc):
    lib = orc.load()
    lb, ub = c([0, -2, 1, 0]), c([4, 3, 5, 2])
    x, g = c([0, 3, 3, 1]), c([-2, 4, 1, -1])
    v = np.where(g > 0, lb, ub)
    assert v[0] == 4 and v[1] == -2
    a = np.zeros(4)
    lib.orc_box_inface_lmo_cpu(4, orc.p(c(-g)), orc.p(x), orc.p(lb), orc.p(ub), orc.p(a),
                               ctypes.c_double(1e-6), ctypes.c_double(1e-4))
    assert list(a) == [0.0, 3.0, 5.0, 0.0]
    d = c(a - v)
    gm = lib.orc_box_dicg_maximum_step_cpu(4, orc.p(d), orc.p(x), orc.p(lb), orc.p(ub))
    assert abs(gm - 0.5) < 1e-12
    nxt = x - gm * d
    assert np.all(nxt >= lb - 1e-12) and np.all(nxt <= ub + 1e-12)
    # Device-semantics in-face (lmo.cuh:278-295) agrees on this case.
    a2 = np.zeros(4)
    lib.orc_compute_inface(ctypes.byref(orc.box_lmo()), 4, orc.p(c(-g)), orc.p(x), orc.p(lb),
                           orc.p(ub), orc.p(a2), 64)
    assert list(a2) == list(a)
    gm2 = lib.orc_max_step(ctypes.byref(orc.box_lmo()), 4, orc.p(d), orc.p(x), orc.p(lb),
                           orc.p(ub), 64)
    assert abs(gm2 - 0.5) < 1e-12


def test_simplex_dicg_kat(orc):
    lib = orc.load()
    is_int = np.array([1, 1, 0, 0], dtype=np.uint8)
    lb, ub = c([1, 0, 0, 0]), c([3, 4, 5, 5])
    x = c([1, 2, 0, 1])
    away = c([2, -1, -3, 4])
    a = np.zeros(4)
    lib.orc_unit_simplex_inface_lmo_cpu(4, ctypes.c_double(5.0), orc.p(is_int), orc.p(away),
                                        orc.p(x), orc.p(lb), orc.p(ub), orc.p(a),
                                        ctypes.c_double(1e-6), ctypes.c_double(1e-4), 0)
    assert list(a) == [1.0, 4.0, 0.0, 0.0]
    fwd = c([3, 0, 0, 2])
    d = c(a - fwd)
    gm = lib.orc_unit_simplex_dicg_maximum_step_cpu(4, ctypes.c_double(5.0), orc.p(is_int),
                                                    orc.p(d), orc.p(x), orc.p(lb), orc.p(ub),
                                                    ctypes.c_double(1e-6), 0)
    assert abs(gm - 0.5) < 1e-12
    # Equality simplex (test_cpu.cpp:491-517)
    xe = c([1, 2, 0, 2])
    a = np.zeros(4)
    lib.orc_unit_simplex_inface_lmo_cpu(4, ctypes.c_double(5.0), orc.p(is_int), orc.p(c([4, 3, 2, 1])),
                                        orc.p(xe), orc.p(lb), orc.p(ub), orc.p(a),
                                        ctypes.c_double(1e-6), ctypes.c_double(1e-4), 1)
    assert abs(a.sum() - 5.0) < 1e-12
    gm = lib.orc_unit_simplex_dicg_maximum_step_cpu(4, ctypes.c_double(5.0), orc.p(is_int),
                                                    orc.p(c([0, 0, 1, 0])), orc.p(xe), orc.p(lb),
                                                    orc.p(ub), ctypes.c_double(1e-6), 1)
    assert gm == 0.0


def test_quadratic_exact_step(orc):
    # tests/test_cpu.cpp:543-552
    q = orc.load().orc_quadratic_exact_step
    assert abs(q(3, 6, 1) - 0.5) < 1e-14
    assert q(3, 2, 0.25) == 0.25
    assert q(-3, 6, 1) == 0
    assert q(3, 0, 1) == 1
    assert q(0, 0, 1) == 0
    assert q(-1, -4, 1) == 1
    assert q(3, 6, 0) == 0
    assert abs(q(3e-100, 6e-100, 1) - 0.5) < 1e-14


# --- reductions ---------------------------------------------------------------

def test_block_sum_order(orc):
    lib = orc.load()
    lib.orc_block_sum.restype = ctypes.c_double
    vals = c(np.arange(128, dtype=np.float64))
    assert lib.orc_block_sum(orc.p(vals), 128) == 128 * 127 / 2
    # 1e16 + 1 loses the 1 unless the tree pairs lane 0 with lane 32 first.
    v = np.zeros(64)
    v[0], v[32], v[1] = 1e16, -1e16, 1.0
    assert lib.orc_block_sum(orc.p(c(v)), 64) == 1.0


# --- LMO vertices against brute force --------------------------------------

def test_box_lmo_closed_form(orc):
    rng = np.random.default_rng(3)
    n = 37
    g = c(rng.standard_normal(n))
    g[::5] = 0.0
    g[3] = -0.0
    lb, ub = c(-rng.random(n)), c(rng.random(n) + 0.5)
    v = np.zeros(n)
    orc.load().orc_compute_v(ctypes.byref(orc.box_lmo()), n, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), 64)
    # argmin of a separable linear function over a box, ties (g == 0) -> ub
    expect = np.where(g > 0, lb, ub)
    assert np.array_equal(v, expect)
    assert np.all(g * v <= np.minimum(g * lb, g * ub))


def enumerate_simplex(lb, ub, is_int, rhs, equality):
    """All integer points with integer coordinates in [lb,ub], sum <= rhs (or ==)."""
    n = len(lb)
    ranges = [range(int(math.ceil(lb[j])), int(math.floor(ub[j])) + 1) for j in range(n)]
    for pt in itertools.product(*ranges):
        s = sum(pt)
        if (equality and s == rhs) or (not equality and s <= rhs):
            yield np.array(pt, dtype=np.float64)


@pytest.mark.parametrize("equality", [False, True])
def test_simplex_lmo_integer_enumeration(orc, equality):
    # tests/test_pipeline.cu:526-645: with all-integer variables the greedy
    # vertex attains the minimum over the integer points of the set.
    rng = np.random.default_rng(11 + equality)
    for trial in range(40):
        n = 5
        lb = c(rng.integers(0, 2, n))
        ub = c(lb + rng.integers(0, 3, n))
        rhs = float(lb.sum() + rng.integers(0, 4))
        if equality and ub.sum() < rhs:
            rhs = float(ub.sum())
        g = c(rng.integers(-3, 4, n).astype(float))
        is_int = np.ones(n, dtype=np.uint8)
        lmo = orc.simplex_lmo(rhs, is_int, equality)
        v = np.zeros(n)
        orc.load().orc_compute_v(ctypes.byref(lmo), n, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), 64)
        best = min(g @ pt for pt in enumerate_simplex(lb, ub, is_int, rhs, equality))
        assert g @ v == best, (trial, g, v)
        assert np.all(v >= lb) and np.all(v <= ub)
        assert (v.sum() == rhs) if equality else (v.sum() <= rhs)


def test_simplex_lmo_continuous_lp(orc):
    # Continuous unit simplex {x >= 0, sum x <= 1}: the minimum of g'x is
    # min(0, min_j g_j), attained at the vertex e_argmin (lowest index on ties).
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = 33
        g = c(np.round(rng.standard_normal(n), 1))
        lmo = orc.simplex_lmo(1.0, np.zeros(n, dtype=np.uint8))
        v = np.zeros(n)
        z = np.zeros(n)
        orc.load().orc_compute_v(ctypes.byref(lmo), n, orc.p(g), orc.p(z), orc.p(c(np.ones(n))), orc.p(v), 64)
        j = int(np.argmin(g))
        expect = np.zeros(n)
        if g[j] <= 0:
            expect[j] = 1.0
        assert np.array_equal(v, expect)


def brute_assignment(cost, m, lb=None, ub=None):
    best, arg = math.inf, None
    for perm in itertools.permutations(range(m)):
        ok = True
        for r in range(m):
            for cc in range(m):
                e = r * m + cc
                on = perm[r] == cc
                if lb is not None and lb[e] > 0.5 and not on:
                    ok = False
                if ub is not None and ub[e] < 0.5 and on:
                    ok = False
        if not ok:
            continue
        val = sum(cost[r * m + perm[r]] for r in range(m))
        if val < best:
            best, arg = val, perm
    return best, arg


def test_birkhoff_lmo_bruteforce(orc):
    # tests/test_pipeline.cu:165-299: exhaustive permutations, order 5..6.
    rng = np.random.default_rng(7)
    lib = orc.load()
    for trial in range(25):
        m = 5 if trial % 2 else 6
        n = m * m
        cost = c(rng.integers(-5, 6, n).astype(float) if trial % 3 == 0 else rng.standard_normal(n))
        lb, ub = np.zeros(n), np.ones(n)
        if trial % 4 == 1:  # force one edge, forbid another
            lb[0 * m + 2] = 1.0
            ub[1 * m + 1] = 0.0
        lb, ub = c(lb), c(ub)
        v = np.zeros(n)
        lib.orc_compute_v(ctypes.byref(orc.birkhoff_lmo(m)), n, orc.p(cost), orc.p(lb), orc.p(ub), orc.p(v), 64)
        best, _ = brute_assignment(cost, m, lb, ub)
        P = v.reshape(m, m)
        assert np.all(P.sum(0) == 1) and np.all(P.sum(1) == 1)
        assert abs(cost @ v - best) < 1e-10, (trial, cost @ v, best)


def test_birkhoff_infeasible_forcing(orc):
    lib = orc.load()
    m, n = 3, 9
    lb, ub = np.zeros(n), np.ones(n)
    lb[0], lb[1] = 1.0, 1.0   # two forced edges in row 0
    assert lib.orc_node_feasible(ctypes.byref(orc.birkhoff_lmo(m)), n, orc.p(c(lb)), orc.p(c(ub)), None, 64) == 0


# --- node relaxation KATs (tests/test_pipeline.cu:1211-1441) ---------------

def ils_objective(orc, Q, target):
    """ILSGlobalCachedObjective: s_g=2, s_l=-1, kappa=|t|^2 -> ||x - t||^2 for Q=I."""
    t = c(target)
    return orc.dense_objective(Q, t, 2.0, -1.0, float(t @ t))


def run_one(orc, obj, a, lmo, x0, lb, ub, flags, budget, iters=0, status=0, eps=1e-12, lanes=64):
    x = c(np.atleast_2d(x0))
    r = orc.relax_batch(obj, a, lmo, x, c(np.atleast_2d(lb)), c(np.atleast_2d(ub)), c([eps]),
                        np.array([iters], dtype=np.int32), np.array([status], dtype=np.uint8),
                        flags=flags, max_iters=budget, candidate_tol=1e-6, lanes=lanes)
    return r


@pytest.mark.parametrize("dicg", [False, True])
def test_continuation_two_thirds(orc, dicg):
    obj = ils_objective(orc, np.eye(2), [0.3, 0.4])
    a = orc.alg(dicg=dicg, ls=orc.LS_OPEN_LOOP)
    flags = np.array([1, 1], dtype=np.uint8)
    split = run_one(orc, obj, a, orc.box_lmo(), [0, 0], [0, 0], [1, 1], flags, 1)
    assert split["status"][0] == 3  # integral, uncertified
    split2 = run_one(orc, obj, a, orc.box_lmo(), split["x"][0], [0, 0], [1, 1], flags, 1,
                     iters=int(split["iterations"][0]))
    whole = run_one(orc, obj, a, orc.box_lmo(), [0, 0], [0, 0], [1, 1], flags, 2)
    assert split2["iterations"][0] == 2
    assert abs(split2["x"][0][0] - 2.0 / 3.0) < 1e-12
    assert np.allclose(split2["x"], whole["x"], atol=1e-12, rtol=0)
    assert abs(split2["gap"][0] - whole["gap"][0]) < 1e-12


def test_rounded_candidate_kats(orc):
    obj = ils_objective(orc, np.eye(2), [0.3, 0.4])
    flags = np.array([1, 0], dtype=np.uint8)
    for dicg in (False, True):
        normal = run_one(orc, obj, orc.alg(dicg=dicg, ls=orc.LS_OPEN_LOOP), orc.box_lmo(),
                         [0.6, 0.4], [0, 0], [1, 1], flags, 0, iters=1)
        assert abs(normal["best_vertex_val"][0] - 0.45) < 1e-12
        floor = run_one(orc, obj, orc.alg(dicg=dicg, ls=orc.LS_OPEN_LOOP, heuristic=1),
                        orc.box_lmo(), [0.6, 0.4], [0, 0], [1, 1], flags, 0, iters=1)
        assert abs(floor["best_vertex_val"][0] - 0.09) < 1e-12
        bad = run_one(orc, obj, orc.alg(dicg=dicg, ls=orc.LS_OPEN_LOOP, heuristic=1,
                                        heuristic_offset=2.0),
                      orc.box_lmo(), [0.6, 0.4], [0, 0], [1, 1], flags, 0, iters=1)
        assert abs(bad["best_vertex_val"][0] - 0.45) < 1e-12
    mixed = run_one(orc, obj, orc.alg(dicg=False, ls=orc.LS_OPEN_LOOP), orc.box_lmo(),
                    [0.3, 0.4], [0, 0], [1, 1], flags, 0, iters=5)
    assert list(mixed["best_vertex"][0]) == [0.0, 0.4]
    assert abs(mixed["best_vertex_val"][0] - 0.09) < 1e-12
    eq = run_one(orc, obj, orc.alg(dicg=False, ls=orc.LS_OPEN_LOOP),
                 orc.simplex_lmo(1.0, np.ones(2, dtype=np.uint8), True),
                 [0.5, 0.5], [0, 0], [1, 1], np.array([1, 1], dtype=np.uint8), 0, iters=5)
    assert abs(eq["best_vertex"][0].sum() - 1.0) < 1e-12


def test_quadratic_exact_interior_one_step(orc):
    # tests/test_pipeline.cu:1240-1270: on x0+x1=1 one exact step reaches (0.25, 0.75).
    obj = ils_objective(orc, np.eye(2), [0.25, 0.75])
    lmo = orc.simplex_lmo(1.0, np.zeros(2, dtype=np.uint8), True)
    for dicg in (False, True):
        r = run_one(orc, obj, orc.alg(dicg=dicg, ls=orc.LS_QUADRATIC_EXACT), lmo, [0, 0], [0, 0],
                    [1, 1], None, 1, eps=1e-10)
        assert r["iterations"][0] == 1
        # converged (1), reported as integral-converged (4) since no
        # coordinate is marked integer
        assert r["status"][0] in (1, 4) and abs(r["gap"][0]) < 1e-10
        assert np.allclose(r["x"][0], [0.25, 0.75], atol=1e-15)


def test_curvature_uses_scaled_hessian(orc):
    # tests/test_pipeline.cu:1211-1228: Hessian 2 diag(2,3), d=(1,-2): d'Hd = 28,
    # gap 7 -> gamma = 1/4; the linear term must not enter.
    Q = np.diag([2.0, 3.0])
    obj = orc.dense_objective(Q, [100.0, -200.0], 2.0, 1.0, 0.0)
    lib = orc.load()
    lib.orc_curvature.restype = ctypes.c_double
    d = c([1.0, -2.0])
    curv = lib.orc_curvature(ctypes.byref(obj), orc.p(d), 64)
    assert curv == 28.0
    assert lib.orc_quadratic_exact_step(7.0, curv, 1.0) == 0.25


def test_infeasible_node_skips_objective(orc):
    # Null objective data: an infeasible node must not touch it
    # (tests/test_pipeline.cu:1496-1574).
    obj = orc.Objective(orc.OBJ_DENSE, 2, None, None, 1.0, 1.0, 0.0, 0, None, None)
    r = run_one(orc, obj, orc.alg(), orc.box_lmo(), [0, 0], [0.2, 0], [0.8, 1],
                np.array([1, 1], dtype=np.uint8), 5)
    assert r["status"][0] == 2 and math.isinf(r["value"][0]) and r["gap"][0] == 0.0


# --- branch-and-bound -----------------------------------------------------

def test_minimal_example(orc):
    # examples/minimal.cu:49-50: x = [0, 1], objective 0.18.
    obj = ils_objective(orc, np.eye(2), [0.3, 0.7])
    r = orc.bnb_solve(obj, orc.alg(ls=orc.LS_QUADRATIC_EXACT), orc.box_lmo(), [0, 0], [1, 1], [0, 1],
                      natural_lower_bound=0.0)
    assert list(r["x"]) == [0.0, 1.0]
    assert abs(r["objective"] - 0.18) < 1e-8


@pytest.mark.parametrize("n,cvec,lo,hi", [
    (2, [1.3, 1.6], 0.0, 3.0),
    (4, [0.7, 1.6, -0.3, 2.8], -1.0, 4.0),
    (8, [0.7, 1.6, -0.3, 2.8, 0.1, -1.4, 3.6, 0.9], -2.0, 4.0)])
def test_separable_analytic_optimum(orc, n, cvec, lo, hi):
    # tests/test_pipeline.cu:449-463, :657-708: optimum of sum (x_j - c_j)^2 over
    # integers is per-coordinate rounding clamped to the bounds.
    t = c(cvec)
    xs = np.clip(np.round(t), lo, hi)
    f_true = float(((xs - t) ** 2).sum())
    obj = ils_objective(orc, np.eye(n), t)
    for a in (orc.alg(dicg=False, ls=orc.LS_OPEN_LOOP), orc.alg(), orc.alg(ls=orc.LS_QUADRATIC_EXACT)):
        r = orc.bnb_solve(obj, a, orc.box_lmo(), [lo] * n, [hi] * n, list(range(n)),
                          batch_size=16, max_iters=50000, eps=1e-6, time_limit_seconds=60)
        assert abs(r["objective"] - f_true) < 1e-9
        assert np.array_equal(r["x"], xs)


def test_c0_matches_exhaustive_enumeration(orc):
    # Config C0: n=20 binary QP; the B&B optimum equals the 2^20 enumeration.
    Q, cv, _ = instances.binary_qp(20, seed=1, rows=30)
    f_enum, x_enum = orc.enumerate_binary_qp(Q, cv, threads=8)
    obj = orc.dense_objective(Q, cv, 2.0, 0.5, 0.0)
    r = orc.bnb_solve(obj, orc.alg(), orc.box_lmo(), [0] * 20, [1] * 20, list(range(20)),
                      abs_gap_tol=1e-6, rel_gap_tol=1e-6, threads=8)
    assert np.array_equal(r["x"], x_enum)
    assert abs(r["objective"] - f_enum) <= 1e-9 * max(1.0, abs(f_enum))
    assert r["dual_bound"] <= r["objective"] + 1e-12
    # brute-force cross-check of the enumerator itself on a smaller case
    Q8, c8, _ = instances.binary_qp(8, seed=3, rows=12)
    best = min((np.array(b) @ Q8 @ np.array(b) + c8 @ np.array(b), b)
               for b in itertools.product([0.0, 1.0], repeat=8))
    f8, x8 = orc.enumerate_binary_qp(Q8, c8, threads=2)
    assert abs(f8 - best[0]) < 1e-12 and list(x8) == list(best[1])


def test_qap_objective_gradient_and_optimum(orc):
    m = 5
    A, B, perm, _ = instances.qap_instance(m, seed=2)
    obj = orc.qap_objective(A, B)
    lib = orc.load()
    lib.orc_value.restype = ctypes.c_double
    Pt = np.zeros((m, m))
    Pt[perm, np.arange(m)] = 1.0  # Pi^T
    assert lib.orc_value(ctypes.byref(obj), orc.p(c(Pt.ravel())), 64) == 0.0
    rng = np.random.default_rng(0)
    X = c(rng.random(m * m))
    g = np.zeros(m * m)
    lib.orc_grad(ctypes.byref(obj), orc.p(X), orc.p(g))
    h = 1e-6
    for e in range(0, m * m, 3):
        xp, xm = X.copy(), X.copy()
        xp[e] += h
        xm[e] -= h
        fd = (lib.orc_value(ctypes.byref(obj), orc.p(c(xp)), 64)
              - lib.orc_value(ctypes.byref(obj), orc.p(c(xm)), 64)) / (2 * h)
        assert abs(fd - g[e]) < 1e-5 * max(1, abs(g[e]))
