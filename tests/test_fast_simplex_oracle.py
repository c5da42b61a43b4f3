"""Oracle of the batched form's bounded simplex (0/1 variables, integral rhs).

The batched form (one wave per node) selects the simplex vertex without a
sort; the oracle restates it as a sort of the eligible coordinates by
(direction, index).  Pins:
  * forward and in-face vertices equal the plain form's greedy
    (orc_compute_v / orc_compute_inface, restating lmo.cuh:517-618 and
    :814-863, themselves pinned by the reference's fixtures) on 0/1 data
    with heavy ties, -0.0 and +0.0, rhs at and beyond the candidate count;
  * the forward vertex minimises h'v over the 0/1 points of the set
    (enumeration);
  * a batched relaxation converges to the plain form's relaxation value and
    certifies a bound below it; nodes outside the 0/1 scope get status 0xFF,
    infeasible ones status 2.
"""
import ctypes
import itertools

import numpy as np
import pytest


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def random_case(rng, n, ties):
    lb = np.zeros(n)
    ub = np.ones(n)
    r = rng.random(n)
    lb[r < 0.15] = 1.0           # fixed at 1
    ub[(r >= 0.15) & (r < 0.3)] = 0.0  # fixed at 0
    if ties:
        h = rng.integers(-3, 3, n).astype(np.float64)
        h[rng.random(n) < 0.2] = -0.0
    else:
        h = rng.standard_normal(n)
    return f64(lb), f64(ub), f64(h)


@pytest.mark.parametrize("equality", [False, True])
def test_forward_vertex_equals_plain_greedy(orc, equality):
    lib = orc.load()
    rng = np.random.default_rng(7 + equality)
    for trial in range(300):
        n = int(rng.integers(1, 200))
        lb, ub, h = random_case(rng, n, ties=trial % 2 == 0)
        free = int((ub - lb).sum())
        rhs = float(lb.sum() + rng.integers(0, free + 3))
        if equality:
            rhs = float(lb.sum() + rng.integers(0, free + 1))
        lmo = orc.simplex_lmo(rhs, np.ones(n, dtype=np.uint8), equality)
        v_plain = np.zeros(n)
        lib.orc_compute_v(ctypes.byref(lmo), n, orc.p(h), orc.p(lb), orc.p(ub), orc.p(v_plain), 64)
        v = np.zeros(n)
        lib.orc_fast_simplex_vertex(n, orc.p(h), orc.p(lb), orc.p(ub), ctypes.c_double(rhs), int(equality),
                                    orc.p(v))
        assert np.array_equal(v, v_plain), (trial, n, rhs)


def test_forward_vertex_minimises_over_01_points(orc):
    lib = orc.load()
    rng = np.random.default_rng(3)
    for trial in range(60):
        n = 9
        lb, ub, h = random_case(rng, n, ties=trial % 3 == 0)
        rhs = float(lb.sum() + rng.integers(0, 5))
        v = np.zeros(n)
        lib.orc_fast_simplex_vertex(n, orc.p(h), orc.p(lb), orc.p(ub), ctypes.c_double(rhs), 0, orc.p(v))
        pts = [np.array(p) for p in itertools.product(*[range(int(lb[j]), int(ub[j]) + 1) for j in range(n)])
               if sum(p) <= rhs]
        assert h @ v == min(h @ p for p in pts)
        assert v.sum() <= rhs and np.all(v >= lb) and np.all(v <= ub)


@pytest.mark.parametrize("equality", [False, True])
def test_inface_vertex_equals_plain_greedy(orc, equality):
    lib = orc.load()
    rng = np.random.default_rng(21 + equality)
    for trial in range(300):
        n = int(rng.integers(2, 150))
        lb, ub, h = random_case(rng, n, ties=trial % 2 == 1)
        free = np.flatnonzero(ub > lb)
        # x on the face: some free coordinates at 0 or 1, some interior;
        # half the cases with the full mass sum x == rhs
        x = lb.copy()
        for j in free:
            u = rng.random()
            x[j] = 0.0 if u < 0.3 else (1.0 if u < 0.5 else rng.uniform(0.05, 0.95))
        rhs = float(np.ceil(x.sum()) + (0 if trial % 2 else rng.integers(0, 3)))
        if trial % 3 == 0:
            # dyadic interior pairs (0.25 + 0.75, 0.5 + 0.5) so that
            # sum x == rhs exactly: the face keeps the full mass
            inner = [j for j in free if 0.0 < x[j] < 1.0]
            for k in range(0, len(inner) - 1, 2):
                q = (0.25, 0.5)[k // 2 % 2]
                x[inner[k]], x[inner[k + 1]] = q, 1.0 - q
            if len(inner) % 2:
                x[inner[-1]] = 1.0
            rhs = float(x.sum())
        lmo = orc.simplex_lmo(rhs, np.ones(n, dtype=np.uint8), equality)
        a_plain = np.zeros(n)
        lib.orc_compute_inface(ctypes.byref(lmo), n, orc.p(h), orc.p(x), orc.p(lb), orc.p(ub), orc.p(a_plain), 64)
        a = np.zeros(n)
        lib.orc_fast_simplex_inface(n, orc.p(h), orc.p(x), orc.p(lb), orc.p(ub), ctypes.c_double(rhs),
                                    int(equality), orc.p(a))
        assert np.array_equal(a, a_plain), (trial, n, rhs)


def simplex_problem(orc, n, seed, card):
    from gpuboscia_amd import instances
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj = orc.dense_objective(Q, c)
    lmo = orc.simplex_lmo(float(card), np.ones(n, dtype=np.uint8))
    return Q, c, obj, lmo


def node_batch01(rng, n, N, card):
    LB = np.zeros((N, n))
    UB = np.ones((N, n))
    for i in range(1, N):
        k = rng.integers(0, 6)
        idx = rng.choice(n, k, replace=False)
        for j in idx:
            if rng.random() < 0.5:
                UB[i, j] = 0.0
            else:
                LB[i, j] = 1.0
    # keep sum lb <= card on all but the last node (infeasible on purpose)
    for i in range(N - 1):
        while LB[i].sum() > card:
            LB[i, np.flatnonzero(LB[i])[0]] = 0.0
    LB[N - 1, : card + 1] = 1.0
    UB[N - 1, : card + 1] = 1.0
    return f64(LB), f64(UB)


@pytest.mark.parametrize("dicg", [True, False])
def test_batched_simplex_relaxation_matches_plain_form(orc, dicg):
    n, N, card = 30, 12, 7
    _, _, obj, lmo = simplex_problem(orc, n, seed=4, card=card)
    rng = np.random.default_rng(9)
    LB, UB = node_batch01(rng, n, N, card)
    flags = np.ones(n, dtype=np.uint8)
    X = np.zeros((N, n))
    r = orc.relax_fast_batch(obj, dicg, X, LB, UB, np.full(N, 1e-9), np.zeros(N, np.int32), np.zeros(N, np.uint8),
                             flags=flags, max_iters=20000, lmo=lmo)
    a = orc.alg(dicg=dicg, ls=orc.LS_QUADRATIC_EXACT)
    Xp = np.zeros((N, n))
    rp = orc.relax_batch(obj, a, lmo, Xp, LB.copy(), UB.copy(), np.full(N, 1e-9), np.zeros(N, np.int32),
                         np.zeros(N, np.uint8), flags=flags, max_iters=20000)
    assert r["status"][N - 1] == 2 and rp["status"][N - 1] == 2
    for i in range(N - 1):
        x = r["x"][i]
        assert x.sum() <= card + 1e-9 and np.all(x >= LB[i] - 1e-12) and np.all(x <= UB[i] + 1e-12)
        # both forms converge to the same relaxation value; each certificate
        # value - gap is a lower bound on it
        assert abs(r["value"][i] - rp["value"][i]) <= 1e-6 * (1 + abs(rp["value"][i])), i
        assert r["value"][i] - r["gap"][i] <= rp["value"][i] + 1e-9
        if dicg:  # plain FW zig-zags and need not reach gap 1e-9 within the budget
            assert r["status"][i] in (1, 4)


def test_batched_simplex_scope_and_supplied_point(orc):
    n, card = 20, 5
    _, _, obj, lmo = simplex_problem(orc, n, seed=2, card=card)
    N = 3
    LB = np.zeros((N, n))
    UB = np.ones((N, n))
    LB[1, 3] = 0.5          # not 0/1: outside the scope
    X = np.zeros((N, n))
    X[2, :card] = 1.0       # supplied feasible start
    st = np.array([0, 0, 0x80], np.uint8)
    r = orc.relax_fast_batch(obj, True, X, f64(LB), f64(UB), np.full(N, 1e-9), np.zeros(N, np.int32), st,
                             flags=np.ones(n, np.uint8), max_iters=5000, lmo=lmo)
    assert r["status"][1] == 0xFF and np.isnan(r["value"][1])
    assert r["status"][0] in (1, 4) and r["status"][2] in (1, 4)
    assert abs(r["value"][0] - r["value"][2]) <= 1e-6 * (1 + abs(r["value"][0]))
    # a variable not marked integer is outside the scope too
    ii = np.ones(n, np.uint8)
    ii[0] = 0
    lmo2 = orc.simplex_lmo(float(card), ii)
    r2 = orc.relax_fast_batch(obj, True, np.zeros((1, n)), f64(np.zeros((1, n))), f64(np.ones((1, n))),
                              np.full(1, 1e-9), np.zeros(1, np.int32), np.zeros(1, np.uint8),
                              flags=np.ones(n, np.uint8), max_iters=10, lmo=lmo2)
    assert r2["status"][0] == 0xFF


def test_batched_simplex_bnb_reaches_enumerated_optimum(orc):
    """Form 1 (batched) B&B over the simplex ends at the optimum of the
    0/1 points with sum x <= card, found by enumeration."""
    n, card = 14, 4
    Q, c, obj, lmo = simplex_problem(orc, n, seed=6, card=card)
    best = np.inf
    for bits in itertools.product((0.0, 1.0), repeat=n):
        x = np.array(bits)
        if x.sum() <= card:
            best = min(best, x @ Q @ x + c @ x)
    a = orc.alg(dicg=True, ls=orc.LS_QUADRATIC_EXACT)
    res = orc.bnb_solve(obj, a, lmo, np.zeros(n), np.ones(n), np.arange(n, dtype=np.int32), batch_size=16,
                        eps=1e-7, eps_min=1e-9, abs_gap_tol=1e-9, rel_gap_tol=1e-9, form=1)
    assert abs(res["objective"] - best) <= 1e-9 * (1 + abs(best))
    assert res["x"].sum() <= card


@pytest.mark.parametrize("equality,free_value", [(False, 0.6), (True, 0.6), (True, 0.4), (False, 0.4)])
def test_supplied_point_is_put_on_the_faces_and_the_mass_limit(orc, equality, free_value):
    """A supplied point (an FP32 iterate of the mixed pass, say) near bounds
    and off the mass limit: coordinates within the face tolerance of a bound
    land on it, and the free ones absorb the excess (or, for an equality, the
    shortfall) in proportion to their room; a point within the limit of an
    inequality keeps its free coordinates."""
    n, card = 40, 10
    _, _, obj, _ = simplex_problem(orc, n, seed=8, card=card)
    lmo = orc.simplex_lmo(float(card), np.ones(n, dtype=np.uint8), equality)
    x0 = np.zeros(n)
    x0[:6] = 1.0 - 3e-7                # within the face tolerance of 1
    x0[6:12] = 2e-8                    # within the face tolerance of 0
    x0[12:20] = free_value             # free: 6 + 8 * 0.6 = 10.8 > 10, or 6 + 3.2 = 9.2 < 10
    x0[20] = -1e-3                     # below its bound: clipped to 0
    X = x0[None, :].copy()
    r = orc.relax_fast_batch(obj, True, X, f64(np.zeros((1, n))), f64(np.ones((1, n))), np.full(1, 1e-9),
                             np.zeros(1, np.int32), np.array([0x80], np.uint8), flags=np.ones(n, np.uint8),
                             max_iters=0, lmo=lmo)
    x = r["x"][0]
    assert np.all(x[:6] == 1.0) and np.all(x[6:12] == 0.0) and x[20] == 0.0 and np.all(x[21:] == 0.0)
    free = x[12:20]
    if free_value > 0.5:       # excess: each free coordinate gives up 0.8 / 4.8 of its value
        assert np.allclose(free, 0.6 - 0.6 * (0.8 / 4.8), rtol=0, atol=1e-15)
        assert abs(x.sum() - card) <= 1e-12
    elif equality:             # shortfall: each takes 0.8 / 4.8 of its room up to 1
        assert np.allclose(free, 0.4 + 0.6 * (0.8 / 4.8), rtol=0, atol=1e-15)
        assert abs(x.sum() - card) <= 1e-12
    else:                      # within the limit: unchanged
        assert np.all(free == 0.4)
