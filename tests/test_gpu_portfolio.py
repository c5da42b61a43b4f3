"""Cardinality portfolio (GB_LMO_PORTFOLIO) on the GPU: node relaxations bit
for bit against the CPU oracle (whose LMO is pinned by the scipy LP
fixtures), and the branch-and-bound optimum against the oracle's search and
against enumeration of the asset subsets."""
import itertools
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import Problem, assert_bit_equal

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def portfolio_nodes(m, N, card, seed):
    """Random z fixings (z_i = 0 or 1) within the cardinality budget."""
    rng = np.random.default_rng(seed)
    n = 2 * m
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        ones = rng.choice(m, size=int(rng.integers(0, min(card, m) + 1)), replace=False)
        zeros = rng.choice(m, size=int(rng.integers(0, m // 3 + 1)), replace=False)
        UB[i, m + zeros] = 0.0
        LB[i, m + ones] = 1.0
        UB[i, m + ones] = 1.0
    return LB, UB


@pytest.mark.parametrize("m,ls", [(12, gb.LS_SECANT), (12, gb.LS_QUADRATIC_EXACT), (100, gb.LS_QUADRATIC_EXACT),
                                  (100, gb.LS_OPEN_LOOP)])
def test_portfolio_relaxation_parity(m, ls):
    Q, c, _, _, sha = instances.portfolio(m, seed=3, cap=0.2 if m >= 10 else 0.5, card=min(10, m))
    cap, card = (0.2, 10) if m >= 10 else (0.5, 3)
    prob = Problem("dense", Q, c, 1.0, 1.0, 0.0, lmo="portfolio", cap=cap, card=card)
    N = 40
    LB, UB = portfolio_nodes(m, N, card, seed=m)
    flags = np.concatenate([np.zeros(m, np.uint8), np.ones(m, np.uint8)])
    say(f"[portfolio] parity m={m} ls={ls} sha={sha[:12]}")
    alg = {"dicg": False, "ls": ls}
    args = (LB.copy(), LB, UB, np.full(N, 1e-6), np.zeros(N, np.int32), np.zeros(N, np.uint8), flags, 300)
    g = prob.run_gpu(alg, *args, T=128)
    o = prob.run_orc(alg, *args, T=128)
    assert_bit_equal(g, o, "portfolio")
    # every returned iterate is in the polytope of its node
    for i in range(N):
        if g["status"][i] == gb.NODE_INFEASIBLE:
            continue
        x, z = g["x"][i, :m], g["x"][i, m:]
        assert abs(x.sum() - 1.0) <= 1e-9 and np.all(x >= -1e-12) and np.all(x <= cap * z + 1e-9)
        assert z.sum() <= card + 1e-9 and np.all(z >= LB[i, m:] - 1e-12) and np.all(z <= UB[i, m:] + 1e-12)


def subset_optimum(Sigma, mu, cap, card):
    """min x'Sigma x - mu'x over sum x = 1, 0 <= x <= cap, support of size <= card."""
    from scipy.optimize import minimize
    m = len(mu)
    best = np.inf
    for k in range(1, card + 1):
        if k * cap < 1.0 - 1e-12:
            continue
        for S in itertools.combinations(range(m), k):
            S = list(S)
            A = Sigma[np.ix_(S, S)]
            b = mu[S]
            r = minimize(lambda y: y @ A @ y - b @ y, np.full(k, 1.0 / k), jac=lambda y: 2 * A @ y - b,
                         bounds=[(0.0, cap)] * k, constraints=[{"type": "eq", "fun": lambda y: y.sum() - 1.0,
                                                                "jac": lambda y: np.ones(k)}],
                         method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
            best = min(best, float(r.fun))
    return best


@pytest.mark.parametrize("seed", [1, 2])
def test_portfolio_bnb_optimum(seed):
    from oracle import pyoracle as orc
    m, cap, card = 8, 0.4, 4   # cap 0.4: an optimum needs a fractional third weight, so the search branches
    Q, c, Sigma, mu, sha = instances.portfolio(m, seed=seed, periods=120, cap=cap, card=card)
    # a larger return weight makes the support choice non-trivial
    c[:m] *= 20.0
    prob = Problem("dense", Q, c, 1.0, 1.0, 0.0, lmo="portfolio", cap=cap, card=card)
    n = 2 * m
    lb, ub = [0.0] * n, [1.0] * n
    iv = list(range(m, n))
    s = gb.default_settings()
    s.abs_gap_tol = 1e-7
    s.rel_gap_tol = 1e-7
    s.algorithm.algorithm = gb.ALG_FRANK_WOLFE
    s.algorithm.line_search = gb.LS_QUADRATIC_EXACT
    s.time_limit_seconds = 60.0
    # Leaves keep a continuous x: the per-node FW tolerance (reference
    # schedule max(1e-6, 1e-3 0.8^depth)) is what bounds their accuracy, so a
    # comparison at 1e-6 needs a fixed 1e-8 per node.
    s.gpu_eps = 1e-8
    s.gpu_eps_min = 1e-8
    s.gpu_max_iters = 100000
    r = gb.solve(n, lb, ub, iv, prob.gobj, prob.glmo, s)
    o = orc.bnb_solve(prob.oobj, orc.alg(dicg=False, ls=orc.LS_QUADRATIC_EXACT), prob.olmo, lb, ub, iv,
                      abs_gap_tol=1e-7, rel_gap_tol=1e-7, threads=8, eps=1e-8, eps_min=1e-8, max_iters=100000)
    best = subset_optimum(Sigma, 20.0 * mu, cap, card)
    say(f"[portfolio] bnb seed={seed} sha={sha[:12]}: gpu {r['objective']:.10f} oracle {o['objective']:.10f} "
        f"subsets {best:.10f} nodes {r['nodes_processed']}/{o['nodes_processed']}")
    assert not r["reached_time_limit"]
    assert r["nodes_processed"] > 1
    assert r["nodes_processed"] == o["nodes_processed"]
    assert r["objective"] == o["objective"] and r["dual_bound"] == o["dual_bound"]
    assert abs(r["objective"] - best) <= 1e-6 * max(1.0, abs(best))
    z = np.array(r["x"][m:])
    assert np.all((z == 0) | (z == 1)) and z.sum() <= card


def test_portfolio_rejects_dicg():
    Q, c, _, _, _ = instances.portfolio(6, seed=1, periods=60)
    prob = Problem("dense", Q, c, 1.0, 1.0, 0.0, lmo="portfolio", cap=0.5, card=3)
    s = gb.default_settings()  # DICG by default
    with pytest.raises(gb.InvalidArgument):
        gb.solve(12, [0.0] * 12, [1.0] * 12, range(6, 12), prob.gobj, prob.glmo, s)
