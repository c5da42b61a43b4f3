"""Writes tests/golden/portfolio_lmo.json: LP optima of the cardinality-
portfolio LMO on tiny random instances, solved by scipy's linprog (HiGHS).
Run in the build container (scipy present); the fixture is committed.

Polytope over v = (x, z), x, z in R^m:
  sum x = rhs, 0 <= x_i <= cap z_i, lz_i <= z_i <= uz_i, sum z <= card.
Objective min g'v with gz >= 0 (the LMO's stated domain).
"""
import json
import os

import numpy as np
from scipy.optimize import linprog


def solve(g, lz, uz, cap, card, rhs):
    m = len(lz)
    c = np.asarray(g, dtype=float)
    A_ub, b_ub = [], []
    for i in range(m):                     # x_i - cap z_i <= 0
        row = np.zeros(2 * m)
        row[i] = 1.0
        row[m + i] = -cap
        A_ub.append(row)
        b_ub.append(0.0)
    row = np.zeros(2 * m)                  # sum z <= card
    row[m:] = 1.0
    A_ub.append(row)
    b_ub.append(card)
    A_eq = [np.concatenate([np.ones(m), np.zeros(m)])]
    bounds = [(0.0, None)] * m + list(zip(lz, uz))
    r = linprog(c, A_ub=np.array(A_ub), b_ub=b_ub, A_eq=np.array(A_eq), b_eq=[rhs], bounds=bounds,
                method="highs")
    return (float(r.fun), r.x.tolist()) if r.status == 0 else (None, None)


def main():
    rng = np.random.default_rng(20261203)
    cases = []
    while len(cases) < 60:
        m = int(rng.integers(2, 7))
        cap = float(rng.choice([0.2, 0.35, 0.5, 1.0]))
        card = float(rng.integers(1, m + 1)) if rng.random() < 0.8 else float(rng.uniform(0.5, m))
        rhs = 1.0
        lz = np.zeros(m)
        uz = np.ones(m)
        for i in range(m):
            u = rng.random()
            if u < 0.2:
                lz[i] = 1.0
            elif u < 0.35:
                uz[i] = 0.0
            elif u < 0.45:
                lz[i] = float(np.round(rng.uniform(0, 0.5), 3))
                uz[i] = float(np.round(rng.uniform(lz[i], 1.0), 3))
        gx = np.round(rng.normal(size=m), 4)
        if rng.random() < 0.3:                   # ties
            gx[rng.integers(0, m)] = gx[0]
        gz = np.zeros(m) if rng.random() < 0.5 else np.round(rng.uniform(0, 1, size=m), 4)
        g = np.concatenate([gx, gz])
        f, v = solve(g, lz, uz, cap, card, rhs)
        cases.append({"m": m, "cap": cap, "card": card, "rhs": rhs, "g": g.tolist(), "lz": lz.tolist(),
                      "uz": uz.tolist(), "feasible": f is not None, "optimum": f, "lp_vertex": v})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "portfolio_lmo.json")
    with open(out, "w") as fh:
        json.dump({"generator": "tests/golden/make_portfolio_fixtures.py (scipy linprog, HiGHS)",
                   "cases": cases}, fh, indent=0)
    print(f"wrote {len(cases)} cases ({sum(c['feasible'] for c in cases)} feasible) to {out}")


if __name__ == "__main__":
    main()
