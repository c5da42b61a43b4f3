"""Pins of the oracle's line search and LMOs beyond the closed forms of
test_oracle_kats.py:

* the reference's secant stall regression (tests/test_pipeline.cu:1133-1205)
  and equality-simplex convergence matrix (:1062-1131);
* the secant step against the exact quadratic step, which it must equal on a
  quadratic (the derivative along the direction is linear there);
* the reference's LMO benchmark cases (benchmarks/benchmark_lmo.cu:173-214)
  against LP and assignment optima computed by scipy
  (tests/golden/lmo_fixtures.json, made by tests/golden/make_lmo_fixtures.py);
* the reference's order-7 Birkhoff case against all 5040 permutations
  (tests/test_pipeline.cu:193-299);
* the Hungarian's tie order on all-zero costs, derived by hand below.
"""
import ctypes
import itertools
import json
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_lmo_fixtures as lmofix  # noqa: E402


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --- secant line search ---------------------------------------------------------

def stall_probe(orc, iters, tol, denom):
    """tests/test_pipeline.cu:1142-1181: f = 2e-4 x^2 - 2e-4 x on [0, 1]
    (Q = 4e-4, c = -2e-4, s_g = s_l = 1), DICG + secant, 64 iterations,
    eps 1e-6, fresh root (its first vertex is x = 1, where the gap is 2e-4)."""
    obj = orc.dense_objective(np.array([[4e-4]]), [-2e-4], 1.0, 1.0, 0.0)
    a = orc.alg(dicg=True, ls=orc.LS_SECANT, secant_iters=iters, secant_tol=tol, secant_denom_eps=denom)
    return orc.relax_batch(obj, a, orc.box_lmo(), np.zeros((1, 1)), np.zeros((1, 1)), np.ones((1, 1)),
                           np.array([1e-6]), np.zeros(1, np.int32), np.zeros(1, np.uint8),
                           flags=np.ones(1, np.uint8), max_iters=64)


def converged(status):
    return int(status) in (1, 4)


def test_secant_stall_regression(orc):
    sparse = stall_probe(orc, 8, 1e-3, 1e-4)
    tight = stall_probe(orc, 16, 1e-10, 1e-14)
    # test_pipeline.cu:1198-1204
    assert not converged(sparse["status"][0]) and sparse["iterations"][0] >= 64 and sparse["gap"][0] > 1e-6
    assert converged(tight["status"][0]) and tight["iterations"][0] < 64 and tight["gap"][0] < 1e-6
    # the exact minimiser is x = 1/2
    assert abs(tight["x"][0, 0] - 0.5) < 1e-9


@pytest.mark.parametrize("seed", range(6))
def test_secant_equals_exact_step_on_quadratics(orc, seed):
    """Along d the derivative phi'(t) = phi0 + t d'Hd is linear, so the first
    secant point from the bracket [0, max] is the exact minimiser gap / curv,
    which the best-value safeguard then returns; beyond max the step is max."""
    rng = np.random.default_rng(seed)
    n = 9 + 7 * seed
    M = rng.standard_normal((n, n))
    Q = M.T @ M / n + 0.2 * np.eye(n)
    c = rng.uniform(-1, 1, n)
    obj = orc.dense_objective(Q, c, 2.0, 0.5, 0.3)
    x = rng.random(n)
    g = 2.0 * Q @ x + c
    d = -g / np.linalg.norm(g)  # a descent direction
    gap = float(-g @ d)
    curv = float(2.0 * d @ Q @ d)
    exact = gap / curv
    a = orc.alg(dicg=True, ls=orc.LS_SECANT)
    # interior minimiser
    t = orc.secant_search(obj, a, x, d, gap, 2.0 * exact)
    assert abs(t - exact) <= 1e-10 * (1.0 + exact), (t, exact)
    assert abs(t - orc.load().orc_quadratic_exact_step(gap, curv, 2.0 * exact)) <= 1e-10 * (1.0 + exact)
    # minimiser beyond the maximum step: the maximum
    t = orc.secant_search(obj, a, x, d, gap, 0.5 * exact)
    assert t == 0.5 * exact
    # no descent along d: zero step
    t = orc.secant_search(obj, a, x, -d, -gap, 1.0)
    assert t == 0.0


def test_secant_linear_objective_takes_the_maximum(orc):
    n = 5
    obj = orc.dense_objective(np.zeros((n, n)), np.arange(1.0, n + 1), 1.0, 1.0, 0.0)
    x = np.full(n, 0.5)
    d = -np.arange(1.0, n + 1)
    gap = float(np.arange(1.0, n + 1) @ -d)
    assert orc.secant_search(obj, orc.alg(), x, d, gap, 0.7) == 0.7


# --- equality-simplex convergence matrix (test_pipeline.cu:1062-1131) ----------

MATRIX = [
    # (name, dicg, line search, secant iterations, must converge)
    ("FW + open-loop", False, 0, 4, False),
    ("FW + secant", False, 3, 16, False),
    ("DICG + open-loop", True, 0, 4, False),
    ("DICG + scaled open-loop", True, 1, 4, False),
    ("DICG + safeguarded scaled", True, 2, 4, True),
    ("DICG + secant", True, 3, 16, True),
]


def matrix_instance(orc):
    n, m = 120, 200
    Qt, lin, kappa = orc.least_squares_instance(0x5EED1234, m, n)
    return n, Qt, lin, kappa


def test_least_squares_instance_is_the_normal_equations(orc):
    n, Qt, lin, kappa = matrix_instance(orc)
    assert np.array_equal(Qt, Qt.T)
    assert np.linalg.eigvalsh(Qt).min() > 0.0
    assert kappa > 0.0
    # 0.5 ||A x - b||^2 >= 0 everywhere: the unconstrained minimum is >= 0
    xs = np.linalg.solve(Qt, -lin)
    assert 0.5 * xs @ Qt @ xs + lin @ xs + kappa >= -1e-9


@pytest.mark.parametrize("name,dicg,ls,si,must", MATRIX)
def test_equality_simplex_convergence_matrix(orc, name, dicg, ls, si, must):
    n, Qt, lin, kappa = matrix_instance(orc)
    obj = orc.dense_objective(Qt.T, lin, 1.0, 1.0, kappa)
    lmo = orc.simplex_lmo(200.0, np.ones(n, np.uint8), equality=True)
    a = orc.alg(dicg=dicg, ls=ls, secant_iters=si, secant_tol=1e-10, secant_denom_eps=1e-14)
    r = orc.relax_batch(obj, a, lmo, np.zeros((1, n)), np.zeros((1, n)), np.full((1, n), 8.0), np.array([1e-4]),
                        np.zeros(1, np.int32), np.zeros(1, np.uint8), flags=np.ones(n, np.uint8),
                        max_iters=20000, lanes=128)
    gap, it = r["gap"][0], r["iterations"][0]
    assert math.isfinite(gap) and gap >= 0.0, name
    assert 0 < it <= 20000, name
    if must:
        assert converged(r["status"][0]), (name, it, gap)
    x = r["x"][0]
    assert abs(x.sum() - 200.0) < 1e-8 and x.min() >= -1e-12 and x.max() <= 8.0 + 1e-12


# --- LMO benchmark fixtures (benchmark_lmo.cu:173-214) ----------------------

def fixture():
    with open(os.path.join(HERE, "golden", "lmo_fixtures.json")) as f:
        return json.load(f)


def orc_lmo_for(orc, case):
    if case["kind"] == "simplex":
        return orc.simplex_lmo(case["rhs"], case["is_int"], equality=case["equality"])
    return orc.birkhoff_lmo(case["order"])


def check_vertex(case, node, v, optimum):
    n = case["n"]
    sl = slice(node * n, (node + 1) * n)
    g, lb, ub = case["g"][sl], case["lb"][sl], case["ub"][sl]
    if optimum is None:
        assert np.all(v == 0.0), f"{case['label']} node {node}: infeasible case returns zeros"
        return
    if case["kind"] == "simplex":
        # integer coordinates within [lb, ub]; continuous ones >= 0 (the
        # reference's simplex LMO gives them the remaining mass unbounded)
        ints = np.asarray(case["is_int"], bool)
        assert np.all(v[ints] >= lb[ints]) and np.all(v[ints] <= ub[ints]), f"{case['label']} node {node}: bounds"
        assert np.all(v[~ints] >= 0.0), f"{case['label']} node {node}: bounds"
        s = v.sum()
        assert (s == case["rhs"]) if case["equality"] else (s <= case["rhs"]), f"{case['label']} node {node}: mass"
        assert np.all(v[ints] == np.round(v[ints])), f"{case['label']} node {node}: integers"
    else:
        assert np.all(v >= lb) and np.all(v <= ub), f"{case['label']} node {node}: bounds"
        m = case["order"]
        P = v.reshape(m, m)
        assert np.all((P == 0) | (P == 1)) and np.all(P.sum(0) == 1) and np.all(P.sum(1) == 1)
    assert float(g @ v) == optimum, f"{case['label']} node {node}: {g @ v} vs optimum {optimum}"


def test_mt19937_reproduces_std():
    assert lmofix.mt_self_check()


def test_lmo_fixture_inputs_match():
    fx = fixture()
    cases = lmofix.build_cases()
    assert [c["label"] for c in cases] == [c["label"] for c in fx["cases"]]
    for c, f in zip(cases, fx["cases"]):
        assert lmofix.digest(c) == f["sha256"], c["label"]


def test_oracle_lmo_on_benchmark_fixtures(orc):
    lib = orc.load()
    for case, f in zip(lmofix.build_cases(), fixture()["cases"]):
        n = case["n"]
        lmo = orc_lmo_for(orc, case)
        lanes = max(64, min(256, (n + 63) // 64 * 64))
        for node in range(lmofix.CASES):
            sl = slice(node * n, (node + 1) * n)
            v = np.zeros(n)
            lib.orc_compute_v(ctypes.byref(lmo), n, orc.p(f64(case["g"][sl])), orc.p(f64(case["lb"][sl])),
                              orc.p(f64(case["ub"][sl])), orc.p(v), lanes)
            check_vertex(case, node, v, f["optimum"][node])


# --- Birkhoff ---------------------------------------------------------------

def order7_cases():
    """tests/test_pipeline.cu:193-226: 20 cases of order 7, each keeping a
    cyclic perfect matching; other arcs forbidden by a fixed rule and one
    matching arc forced in two of every three cases."""
    order, cases = 7, 20
    nv = order * order
    cost = np.zeros(cases * nv)
    lb = np.zeros(cases * nv)
    ub = np.ones(cases * nv)
    for s in range(cases):
        for r in range(order):
            gcol = (r + s) % order
            for cc in range(order):
                e = r * order + cc
                pos = s * nv + e
                cost[pos] = math.sin(0.37 * (s + 1) * (e + 2)) + 0.013 * e + 0.001 * s
                if cc != gcol and (r * 11 + cc * 7 + s * 3) % 5 == 0:
                    ub[pos] = 0.0
        if s % 3 != 0:
            fr = s % order
            lb[s * nv + fr * order + (fr + s) % order] = 1.0
    return order, cases, cost, lb, ub


def exhaustive_assignment(order, cost, lb, ub):
    best = math.inf
    for perm in itertools.permutations(range(order)):
        ok = all(ub[r * order + perm[r]] >= 0.5 for r in range(order))
        ok = ok and all(not (lb[r * order + cc] > 0.5 and perm[r] != cc) for r in range(order) for cc in range(order))
        if ok:
            best = min(best, sum(cost[r * order + perm[r]] for r in range(order)))
    return best


def test_birkhoff_order7_against_all_permutations(orc):
    order, cases, cost, lb, ub = order7_cases()
    nv = order * order
    lib = orc.load()
    for s in range(cases):
        sl = slice(s * nv, (s + 1) * nv)
        v = np.zeros(nv)
        lib.orc_compute_v(ctypes.byref(orc.birkhoff_lmo(order)), nv, orc.p(f64(cost[sl])), orc.p(f64(lb[sl])),
                          orc.p(f64(ub[sl])), orc.p(v), 64)
        P = v.reshape(order, order)
        assert np.all(P.sum(0) == 1) and np.all(P.sum(1) == 1)
        assert np.all(v >= lb[sl]) and np.all(v <= ub[sl])
        assert abs(cost[sl] @ v - exhaustive_assignment(order, cost[sl], lb[sl], ub[sl])) <= 1e-10


@pytest.mark.parametrize("order", [2, 3, 6])
def test_birkhoff_zero_costs_give_identity(orc, order):
    """Tie order of the shortest-augmenting-path Hungarian (lmo.cuh:1111-1171):
    the next column is the unused one of least reduced cost, the lowest index
    among equals (a strict < scan), and a column's path link changes only on
    a strictly smaller value.  With all costs zero, row i starts at column 0
    with every reduced cost 0: the scan takes column 1, then (strict <: no
    link changes) columns 2, 3, ... held by rows 2, 3, ..., until column i,
    which is free; its link is still column 0, so row i is matched to column
    i.  The result is the identity.  Taking the highest index among equals
    instead matches row 1 to column m and ends at the anti-diagonal for
    m = 2."""
    n = order * order
    v = np.zeros(n)
    g, lb, ub = np.zeros(n), np.zeros(n), np.ones(n)
    orc.load().orc_compute_v(ctypes.byref(orc.birkhoff_lmo(order)), n, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), 64)
    assert np.array_equal(v.reshape(order, order), np.eye(order))
