"""Fixtures of the reference's LMO benchmark cases (benchmarks/benchmark_lmo.cu:173-214):
tie-heavy simplex LMOs at n = 17 and 257 (inequality and equality forms) and
Birkhoff LMOs at orders 5, 32 and 129 with forced and forbidden edges.

The inputs are rebuilt here exactly as that file builds them, from raw
std::mt19937 outputs (MT19937 below, checked against the C++ standard's
value for the 10000th output of a default-seeded engine).  The expected
outputs are independent of the oracle:

* simplex: the minimum of g.x over {sum x <= rhs (== rhs), lb_j <= x_j <= ub_j
  for integer j, x_j >= 0 for continuous j}, a linear program solved by
  scipy's HiGHS (the data are integers, so the optimum is an integer the LP
  solver returns exactly).  Continuous coordinates have no upper bound and
  ignore lb because the reference's simplex LMO starts them at 0 and lets
  the chosen one take all remaining mass (lmo.cuh:399-450, :528-618);
* Birkhoff: the minimum-cost perfect matching with forced edges (lb > 1/2)
  and forbidden edges (ub < 1/2), from scipy.optimize.linear_sum_assignment,
  or "infeasible" when no perfect matching exists.

Run:  python tests/golden/make_lmo_fixtures.py  (writes lmo_fixtures.json).
The tests import build_cases() from this file to regenerate the inputs and
check their SHA-256 against the fixture.
"""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lmo_fixtures.json")
CASES = 3  # benchmark_lmo.cu:9


class MT19937:
    """The 32-bit Mersenne Twister of std::mt19937 (seed by init_genrand)."""

    def __init__(self, seed=5489):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.index = 624

    def __call__(self):
        if self.index >= 624:
            for i in range(624):
                y = (self.mt[i] & 0x80000000) | (self.mt[(i + 1) % 624] & 0x7FFFFFFF)
                self.mt[i] = self.mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.index = 0
        y = self.mt[self.index]
        self.index += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def mt_self_check():
    r = MT19937()
    for _ in range(9999):
        r()
    return r() == 4123659995  # [rand.predef]: the 10000th output of std::mt19937


def simplex_cases(N, equality):
    """benchmark_lmo.cu:173-194: (label, rhs, is_int, g, lb, ub) per case set."""
    is_int = np.array([1 if j % 3 != 0 else 0 for j in range(N)], dtype=np.uint8)
    rng = MT19937(42)
    g = np.zeros(CASES * N)
    lb = np.zeros(CASES * N)
    ub = np.full(CASES * N, 3.0)
    for c in range(CASES):
        for j in range(N):
            e = c * N + j
            if c == 1:
                g[e] = 1.0 + rng() % 4
            else:
                g[e] = float(int(rng() % 9) - 4)
            if c == 2 and j < 3 and is_int[j]:
                lb[e] = 1.0
            if j % 5 == 0 and is_int[j]:
                ub[e] = lb[e]
    out = [("equality_simplex" if equality else "simplex", 7.0, is_int, g, lb, ub)]
    out.append(("equality_integer" if equality else "simplex_integer", float(N), np.ones(N, np.uint8), g.copy(),
                np.zeros(CASES * N), np.ones(CASES * N)))
    return out


def birkhoff_cases(order):
    """benchmark_lmo.cu:196-214."""
    n = order * order
    rng = MT19937(123)
    g = np.zeros(CASES * n)
    lb = np.zeros(CASES * n)
    ub = np.ones(CASES * n)
    for c in range(CASES):
        for j in range(n):
            g[c * n + j] = 0.0 if c == 1 else float(int(rng() % 17) - 8)
    for r in range(order):
        for cc in range(order):
            if r != cc and (r + cc) % 3 == 0:
                ub[2 * n + r * order + cc] = 0.0
    lb[2 * n] = 1.0
    out = [("birkhoff", g.copy(), lb.copy(), ub.copy())]
    ub2 = ub.copy()
    for c in range(CASES):
        for j in range(order):
            ub2[c * n + j] = 0.0
    out.append(("birkhoff_infeasible", g.copy(), np.zeros(CASES * n), ub2))
    return out


def build_cases():
    """Every fixture case: dict(kind, label, n, order, rhs, equality, is_int, g, lb, ub), CASES nodes each."""
    cases = []
    for N in (17, 257):
        for equality in (False, True):
            for label, rhs, is_int, g, lb, ub in simplex_cases(N, equality):
                cases.append(dict(kind="simplex", label=f"{label}<{N}>", n=N, order=0, rhs=rhs,
                                  equality=equality, is_int=is_int, g=g, lb=lb, ub=ub))
    for order in (5, 32, 129):
        for label, g, lb, ub in birkhoff_cases(order):
            cases.append(dict(kind="birkhoff", label=f"{label}<{order}>", n=order * order, order=order, rhs=1.0,
                              equality=True, is_int=None, g=g, lb=lb, ub=ub))
    return cases


def digest(case):
    h = hashlib.sha256()
    for k in ("g", "lb", "ub"):
        h.update(np.ascontiguousarray(case[k], dtype=np.float64).tobytes())
    if case["is_int"] is not None:
        h.update(np.ascontiguousarray(case["is_int"], dtype=np.uint8).tobytes())
    return h.hexdigest()


def simplex_optimum(g, lb, ub, rhs, equality, is_int):
    from scipy.optimize import linprog
    n = len(g)
    A = np.ones((1, n))
    kw = dict(A_eq=A, b_eq=[rhs]) if equality else dict(A_ub=A, b_ub=[rhs])
    bounds = [(lb[j], ub[j]) if is_int[j] else (0.0, None) for j in range(n)]
    res = linprog(g, bounds=bounds, method="highs", **kw)
    if res.status == 2:
        return None
    assert res.status == 0, res.message
    return float(res.fun)


def birkhoff_optimum(g, lb, ub, order):
    from scipy.optimize import linear_sum_assignment
    C = g.reshape(order, order).astype(np.float64).copy()
    allowed = ub.reshape(order, order) >= 0.5
    forced = lb.reshape(order, order) > 0.5
    for r, cc in zip(*np.nonzero(forced)):
        row, col = allowed[r, cc], True
        allowed[r, :] = False
        allowed[:, cc] = False
        allowed[r, cc] = row and col
    big = 1e9
    C = np.where(allowed, C, big)
    rows, cols = linear_sum_assignment(C)
    if np.any(~allowed[rows, cols]):
        return None
    return float(g.reshape(order, order)[rows, cols].sum())


def main():
    assert mt_self_check(), "MT19937 does not reproduce std::mt19937"
    out = {"source": "benchmarks/benchmark_lmo.cu:173-214", "nodes_per_case": CASES, "cases": []}
    for case in build_cases():
        n = case["n"]
        opts = []
        for c in range(CASES):
            sl = slice(c * n, (c + 1) * n)
            if case["kind"] == "simplex":
                opts.append(simplex_optimum(case["g"][sl], case["lb"][sl], case["ub"][sl], case["rhs"],
                                            case["equality"], case["is_int"]))
            else:
                opts.append(birkhoff_optimum(case["g"][sl], case["lb"][sl], case["ub"][sl], case["order"]))
        out["cases"].append({"label": case["label"], "kind": case["kind"], "n": n, "sha256": digest(case),
                             "optimum": opts})
        print(case["label"], opts)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
