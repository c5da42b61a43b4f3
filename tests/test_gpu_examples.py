"""The reference's example programs over the drop-in header
(gpuboscia_amd/examples, built by __graft_entry__.build()), run end to end
on the reference's own instance files (tests/golden/mat):

* graph_iso (examples/main.cu): a MAT graph against a seeded relabelling of
  itself, through the QAP's FP32 form: the isomorphism is found (objective
  0, B X = X A checked by the program) on the triangular graph T(10), the
  rook's graph K_4 x K_4 and a 3-regular graph (K_8 x K_8, whose symmetric
  relaxation branches for about 130 000 nodes, takes about 110 s);
* weighted_clique (examples/weighted_clique.cu): the optimum equals the
  enumeration of all C(20, 5) vertex sets within the search's 1e-3 relative
  tolerance;
* sparse_sharpe (examples/sparse_sharpe.cu, LinkedCardinalityLMO and its
  objective as device plug-ins): on FF25EU (25 assets, 3 chosen) the
  incumbent is feasible and within the search's 1e-2 relative tolerance of
  the optimum found by enumerating subsets and active sets.
"""
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "gpuboscia_amd", "examples")
MAT = os.path.join(ROOT, "tests", "golden", "mat")


def run(prog, *args, timeout=300):
    exe = os.path.join(EX, prog)
    assert os.path.exists(exe), f"{exe} not built: run __graft_entry__.build()"
    print(f"[examples] {prog} {' '.join(args)}", flush=True)
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    print(out.stdout[-3000:], flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.parametrize("instance", ["Triangular_10_45.mat", "Lattice_4_16.mat", "CHH_cc_1_1_22_1.mat"])
def test_graph_isomorphism_example(instance):
    out = run("graph_iso", "--instance", os.path.join(MAT, instance), "--seed", "3", "--time-limit", "120")
    assert "Hello:)" in out and "Solution Statistics." in out
    assert re.search(r"Isomorphism found:\s+yes", out), out
    assert re.search(r"Objective:\s+0\n", out), out


def test_graph_isomorphism_example_rejects_bad_options():
    exe = os.path.join(EX, "graph_iso")
    out = subprocess.run([exe, "--algorithm", "pgd"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "--algorithm must be either 'fw' or 'dicg'." in out.stderr
    out = subprocess.run([exe, "--instance", os.path.join(MAT, "FF25EUnew.mat")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "no matrix named 'A'" in out.stderr


def _uniforms(seed):
    M = (1 << 64) - 1
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M
        z = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        yield (z >> 11) * 2.0 ** -53


def test_weighted_clique_example():
    V, C, seed = 20, 5, 5
    out = run("weighted_clique", "--seed", str(seed), "--time-limit", "120")
    u = _uniforms(seed)
    w = np.zeros((V, V))
    for i in range(V):
        for j in range(i + 1, V):
            w[i, j] = w[j, i] = next(u)
    comp = 1.0 - w
    best = min(2.0 * sum(comp[a, b] for a, b in itertools.combinations(S, 2))
               for S in itertools.combinations(range(V), C))
    obj = float(re.search(r"GPUBoscia minimization objective: (\S+)", out).group(1))
    sel = [int(v) - 1 for v in re.search(r"selected vertices:(.*)", out).group(1).split()]
    assert len(sel) == C
    assert abs(2.0 * sum(comp[a, b] for a, b in itertools.combinations(sel, 2)) - obj) < 1e-9
    assert best - 1e-9 <= obj <= best * (1 + 1e-3) + 1e-7, (obj, best)


def _subset_qp(H, mu, U, S):
    """min 0.5 v'Hv - mu'v, 0 <= v <= U on S: each coordinate at 0, at U or
    free; the KKT point of the positive definite H is the minimiser."""
    k = len(S)
    if k == 0:
        return 0.0
    best = np.inf
    Hs, ms, Us = H[np.ix_(S, S)], mu[S], U[S]
    for pat in itertools.product((0, 1, 2), repeat=k):
        st = np.array(pat)
        v = np.where(st == 1, Us, 0.0)
        F = np.where(st == 2)[0]
        if len(F):
            B = np.where(st != 2)[0]
            rhs = ms[F] - Hs[np.ix_(F, B)] @ v[B]
            v[F] = np.linalg.solve(Hs[np.ix_(F, F)], rhs)
            if np.any(v[F] < -1e-12) or np.any(v[F] > Us[F] + 1e-12):
                continue
        g = Hs @ v - ms
        if np.any(g[st == 0] < -1e-10) or np.any(g[st == 1] > 1e-10):
            continue
        best = min(best, 0.5 * v @ Hs @ v - ms @ v)
    return best


def test_sparse_sharpe_example():
    from scipy.io import loadmat
    assets, card, T, eps = 25, 3, 60, 1e-3
    out = run("sparse_sharpe", "--data", os.path.join(MAT, "FF25EUnew.mat"), "--assets", str(assets),
              "--cardinality", str(card), "--observations", str(T), "--time-limit", "120")
    prices = loadmat(os.path.join(MAT, "FF25EUnew.mat"))["data"][-T:, :assets]
    R = prices - 1.0
    mu = R.mean(axis=0)
    H = np.cov(R, rowvar=False, ddof=1) + eps * np.eye(assets)
    center = np.linalg.solve(H, mu)
    r2 = max(mu @ center, 0.0)
    U = np.maximum(center + np.sqrt(np.maximum(np.diag(np.linalg.inv(H)) * r2, 0.0)), 0.0)
    best = min(_subset_qp(H, mu, U, list(S)) for k in range(card + 1)
               for S in itertools.combinations(range(assets), k))
    m = re.search(r"Objective 0\.5 v'Hv - mu'v:\s+(\S+) \(solver (\S+)\)", out)
    f, solver = float(m.group(1)), float(m.group(2))
    assert abs(f - solver) <= 1e-9 * (1 + abs(f))
    weights = re.findall(r"asset (\d+)\s+v=(\S+)\s+w=(\S+)\s+y=(\d)", out)
    assert 1 <= len(weights) <= card and all(y == "1" for *_, y in weights)
    assert all(0.0 <= float(v) <= U[int(a) - 1] + 1e-9 for a, v, _, _ in weights)
    assert best - 1e-9 <= f <= best + 1e-2 * abs(best), (f, best)
