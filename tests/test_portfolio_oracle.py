"""The cardinality-portfolio LMO of the oracle against LP optima computed by
scipy (committed fixtures, tests/golden/make_portfolio_fixtures.py): the
returned vertex is feasible, attains the LP optimum, and infeasible nodes are
detected."""
import json
import os

import numpy as np

from oracle import pyoracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    with open(os.path.join(HERE, "golden", "portfolio_lmo.json")) as f:
        return json.load(f)["cases"]


def test_vertex_attains_lp_optimum():
    lib = orc.load()
    checked = 0
    for case in cases():
        m = case["m"]
        n = 2 * m
        l = orc.portfolio_lmo(m, case["cap"], case["card"], case["rhs"])
        lb = np.concatenate([np.zeros(m), case["lz"]])
        ub = np.concatenate([np.ones(m), case["uz"]])
        feas = lib.orc_node_feasible(orc.ctypes.byref(l), n, orc.p(lb), orc.p(ub), None, 64)
        assert bool(feas) == case["feasible"], case
        if not case["feasible"]:
            continue
        g = np.asarray(case["g"], dtype=np.float64)
        v = np.zeros(n)
        lib.orc_compute_v(orc.ctypes.byref(l), n, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), 64)
        x, z = v[:m], v[m:]
        tol = 1e-9
        assert abs(x.sum() - case["rhs"]) <= tol
        assert np.all(x >= -tol) and np.all(x <= case["cap"] * z + tol)
        assert np.all(z >= np.asarray(case["lz"]) - tol) and np.all(z <= np.asarray(case["uz"]) + tol)
        assert z.sum() <= case["card"] + tol
        assert abs(g @ v - case["optimum"]) <= 1e-7 * max(1.0, abs(case["optimum"])), (case, g @ v)
        checked += 1
    assert checked >= 20


def test_cardinality_pool_binds_in_the_fixtures():
    # Without the cardinality pool the greedy vertex breaks sum z <= card on some
    # fixtures: the fixtures exercise the binding case.
    lib = orc.load()
    misses = 0
    for case in cases():
        if not case["feasible"]:
            continue
        m = case["m"]
        g = np.asarray(case["g"])
        lb = np.concatenate([np.zeros(m), case["lz"]])
        ub = np.concatenate([np.ones(m), case["uz"]])
        big = orc.portfolio_lmo(m, case["cap"], 1e9, case["rhs"])   # pool never binds
        v = np.zeros(2 * m)
        lib.orc_compute_v(orc.ctypes.byref(big), 2 * m, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), 64)
        if v[m:].sum() > case["card"] + 1e-9:
            misses += 1
    assert misses > 0
