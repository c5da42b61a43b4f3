"""The reference's line-search and LMO known-answer cases on the HIP path,
through the C-ABI, each also compared bit for bit with the oracle:

* secant stall regression (tests/test_pipeline.cu:1133-1205);
* equality-simplex convergence matrix (tests/test_pipeline.cu:1062-1131);
* LMO benchmark cases (benchmarks/benchmark_lmo.cu:173-214) against the
  scipy optima of tests/golden/lmo_fixtures.json;
* order-7 Birkhoff cases against all permutations (tests/test_pipeline.cu:193-299);
* the Hungarian's tie order on all-zero costs (identity).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from tests.gpu_util import Problem, assert_bit_equal, dev

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_lmo_fixtures as lmofix  # noqa: E402
from test_oracle_pins import MATRIX, check_vertex, converged, exhaustive_assignment, fixture, order7_cases  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relax_both(prob, alg, n, max_iters, eps, flags, T):
    """One fresh root node [lb, ub] relaxed by gb_relax_batch and by the oracle."""
    from oracle import pyoracle as orc
    import torch
    LB = prob.lb.reshape(1, n)
    UB = prob.ub.reshape(1, n)
    args = (LB.copy(), LB, UB, np.array([eps]), np.zeros(1, np.int32), np.zeros(1, np.uint8))
    x, lb, ub, e = dev(args[0]), dev(LB), dev(UB), dev(args[3])
    it = torch.zeros(1, dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.uint8, device="cuda")
    gap = torch.zeros(1, dtype=torch.float64, device="cuda")
    val = torch.zeros(1, dtype=torch.float64, device="cuda")
    bv = torch.zeros((1, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((1,), 1e300, dtype=torch.float64, device="cuda")
    fl = dev(np.asarray(flags, dtype=np.uint8))
    gb.relax_batch(prob.gobj, alg, prob.glmo, 1, max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv, fl, 1e-6, T)
    torch.cuda.synchronize()
    g = {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(), "iterations": it.cpu().numpy(),
         "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(), "best_vertex_val": bvv.cpu().numpy()}
    oa = orc.alg(dicg=alg.algorithm == gb.ALG_DICG, ls=alg.line_search, secant_iters=alg.secant_iters,
                 secant_tol=alg.secant_tol, secant_denom_eps=alg.secant_denom_eps)
    o = orc.relax_batch(prob.oobj, oa, prob.olmo, LB.copy(), LB.copy(), UB.copy(), np.array([eps]),
                        np.zeros(1, np.int32), np.zeros(1, np.uint8), flags=np.asarray(flags, np.uint8),
                        max_iters=max_iters, lanes=T)
    return g, o


@pytest.mark.parametrize("iters,tol,denom,stalls", [(8, 1e-3, 1e-4, True), (16, 1e-10, 1e-14, False)])
def test_gpu_secant_stall_regression(iters, tol, denom, stalls):
    prob = Problem("dense", np.array([[4e-4]]), [-2e-4], 1.0, 1.0, 0.0)
    prob.lb, prob.ub = np.zeros(1), np.ones(1)
    alg = gb.algorithm(dicg=True, line_search=gb.LS_SECANT, secant_iters=iters, secant_tol=tol,
                       secant_denom_eps=denom)
    say(f"[pins] secant stall probe SecantLS{{{iters},{tol},{denom}}}")
    g, o = relax_both(prob, alg, 1, 64, 1e-6, [1], 64)
    assert_bit_equal(g, o, "secant stall probe")
    if stalls:  # test_pipeline.cu:1198-1200
        assert not converged(g["status"][0]) and g["iterations"][0] >= 64 and g["gap"][0] > 1e-6
    else:  # :1201-1202
        assert converged(g["status"][0]) and g["iterations"][0] < 64 and g["gap"][0] < 1e-6


@pytest.mark.parametrize("name,dicg,ls,si,must", MATRIX)
def test_gpu_equality_simplex_convergence_matrix(name, dicg, ls, si, must):
    from oracle import pyoracle as orc
    n = 120
    Qt, lin, kappa = orc.least_squares_instance(0x5EED1234, 200, n)
    prob = Problem("dense", Qt.T, lin, 1.0, 1.0, kappa, lmo="simplex", rhs=200.0, equality=True,
                   is_int=np.ones(n, np.uint8))
    prob.lb, prob.ub = np.zeros(n), np.full(n, 8.0)
    alg = gb.algorithm(dicg=dicg, line_search=ls, secant_iters=si, secant_tol=1e-10, secant_denom_eps=1e-14)
    say(f"[pins] convergence matrix: {name}")
    g, o = relax_both(prob, alg, n, 20000, 1e-4, np.ones(n, np.uint8), 128)
    assert_bit_equal(g, o, name)
    assert np.isfinite(g["gap"][0]) and g["gap"][0] >= 0.0
    assert 0 < g["iterations"][0] <= 20000
    if must:
        assert converged(g["status"][0]), (name, g["iterations"][0], g["gap"][0])


def gpu_vertices(lmo, n, N, G, LB, UB):
    import torch
    v = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    dg, dl, du = dev(G.reshape(N, n)), dev(LB.reshape(N, n)), dev(UB.reshape(N, n))
    gb.lmo_compute_v(lmo, n, N, dg, dl, du, v)
    torch.cuda.synchronize()
    return v.cpu().numpy()


def oracle_vertices(olmo, n, N, G, LB, UB, lanes):
    from oracle import pyoracle as orc
    out = np.zeros((N, n))
    lib = orc.load()
    for i in range(N):
        g = np.ascontiguousarray(G[i * n:(i + 1) * n])
        lb = np.ascontiguousarray(LB[i * n:(i + 1) * n])
        ub = np.ascontiguousarray(UB[i * n:(i + 1) * n])
        v = np.zeros(n)
        lib.orc_compute_v(ctypes.byref(olmo), n, orc.p(g), orc.p(lb), orc.p(ub), orc.p(v), lanes)
        out[i] = v
    return out


def lmo_pair(case):
    from oracle import pyoracle as orc
    if case["kind"] == "simplex":
        ii = dev(np.asarray(case["is_int"], np.uint8))
        g = gb.LmoDesc(gb.LMO_SIMPLEX, float(case["rhs"]), 1 if case["equality"] else 0, ii.data_ptr(), 0)
        g._keep = ii
        return g, orc.simplex_lmo(case["rhs"], case["is_int"], equality=case["equality"])
    return gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, case["order"]), orc.birkhoff_lmo(case["order"])


def test_gpu_lmo_benchmark_fixtures():
    for case, f in zip(lmofix.build_cases(), fixture()["cases"]):
        n, N = case["n"], lmofix.CASES
        assert lmofix.digest(case) == f["sha256"]
        glmo, olmo = lmo_pair(case)
        lanes = max(64, min(256, (n + 63) // 64 * 64))
        say(f"[pins] LMO fixture {case['label']} n={n}")
        vg = gpu_vertices(glmo, n, N, case["g"], case["lb"], case["ub"])
        vo = oracle_vertices(olmo, n, N, case["g"], case["lb"], case["ub"], lanes)
        assert np.array_equal(vg, vo), f"{case['label']}: GPU vertices differ from the oracle's"
        for node in range(N):
            check_vertex(case, node, vg[node], f["optimum"][node])


def test_gpu_birkhoff_order7_against_all_permutations():
    from oracle import pyoracle as orc
    order, cases, cost, lb, ub = order7_cases()
    nv = order * order
    say("[pins] Birkhoff order 7, 20 cases")
    glmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, order)
    vg = gpu_vertices(glmo, nv, cases, cost, lb, ub)
    vo = oracle_vertices(orc.birkhoff_lmo(order), nv, cases, cost, lb, ub, 64)
    assert np.array_equal(vg, vo)
    for s in range(cases):
        sl = slice(s * nv, (s + 1) * nv)
        P = vg[s].reshape(order, order)
        assert np.all(P.sum(0) == 1) and np.all(P.sum(1) == 1)
        assert np.all(vg[s] >= lb[sl]) and np.all(vg[s] <= ub[sl])
        assert abs(cost[sl] @ vg[s] - exhaustive_assignment(order, cost[sl], lb[sl], ub[sl])) <= 1e-10


@pytest.mark.parametrize("order", [2, 3, 6, 40])
def test_gpu_birkhoff_zero_costs_give_identity(order):
    n = order * order
    say(f"[pins] Birkhoff zero costs order {order}")
    glmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, order)
    v = gpu_vertices(glmo, n, 1, np.zeros(n), np.zeros(n), np.ones(n))
    assert np.array_equal(v[0].reshape(order, order), np.eye(order))
