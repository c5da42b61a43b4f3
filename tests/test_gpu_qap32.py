"""The QAP's FP32 form (gb_qap32_*, relax_qap32.hip): FP32 iterates and
assignment costs, FP64 certificates.

* The warm-started assignment solver returns optimal assignments: on
  sequences of FP32 cost matrices (perturbed steps, ties, forbidden edges)
  every permutation's cost equals scipy's linear_sum_assignment optimum on
  the same costs, and the Lagrangian bound of its potentials lies at or
  below that optimum and within rounding of it.
* The relaxation's outputs are certified in FP64: value and gap recomputed
  on the host from the returned P (residual, gradient, the exact minimum of
  <G, V> by scipy) agree, and value - gap never exceeds the relaxation's
  optimum (0 for a planted instance; the plain FP64 form's value otherwise).
* Node bounds (forced and forbidden edges) are honoured by every vertex, and
  contradictory forced edges make a node infeasible.
* Branch-and-bound with form 2 on a QAP proves the planted optimum 0, and on
  a non-planted m = 6 instance finds the optimum of all 720 permutations.
"""
import itertools

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import dev, torch_mod

pytestmark = pytest.mark.gpu

scipy_opt = pytest.importorskip("scipy.optimize")


def say(msg):
    print(msg, flush=True)


def _lsa_cost(C):
    """Optimal assignment cost of C (float32, +inf forbidden) in FP64."""
    big = np.where(np.isfinite(C), C.astype(np.float64), 1e30)
    r, c = scipy_opt.linear_sum_assignment(big)
    return float(np.sum(C.astype(np.float64)[r, c]))


def _perm_cost(C, p):
    return float(np.sum(C.astype(np.float64)[np.arange(len(p)), p]))


@pytest.mark.parametrize("m", [5, 17, 64, 200])
def test_assignment_sequence_is_optimal(m):
    torch = torch_mod()
    rng = np.random.default_rng(m)
    seq = []
    C = rng.standard_normal((m, m)).astype(np.float32)
    seq.append(C.copy())
    for t in range(3):  # Frank-Wolfe-like steps: C <- (1 - g) C + g D
        g = [0.5, 0.05, 0.002][t]
        C = ((1 - g) * C + g * rng.standard_normal((m, m))).astype(np.float32)
        seq.append(C.copy())
    seq.append(rng.integers(0, 4, (m, m)).astype(np.float32))  # ties everywhere
    F = rng.standard_normal((m, m)).astype(np.float32)
    perm = rng.permutation(m)
    F[rng.random((m, m)) < 0.5] = np.inf
    F[np.arange(m), perm] = rng.standard_normal(m).astype(np.float32)  # a feasible assignment stays
    seq.append(F)
    T = len(seq)
    costs = dev(np.stack(seq), torch.float32)
    perms = torch.zeros((T, m), dtype=torch.int32, device="cuda")
    bounds = torch.zeros(T, dtype=torch.float64, device="cuda")
    say(f"[qap32] assignments m={m} T={T}")
    gb.qap32_assignments(costs, perms, bounds)
    P = perms.cpu().numpy()
    Bd = bounds.cpu().numpy()
    for t, Ct in enumerate(seq):
        p = P[t]
        assert sorted(p.tolist()) == list(range(m)), (t, p)
        assert np.all(np.isfinite(Ct[np.arange(m), p])), t
        opt = _lsa_cost(Ct)
        got = _perm_cost(Ct, p)
        scale = 1.0 + float(np.abs(Ct[np.isfinite(Ct)].astype(np.float64)).sum()) / m
        assert abs(got - opt) <= 1e-9 * scale, (t, got, opt)
        assert Bd[t] <= opt + 1e-9 * scale and Bd[t] >= opt - 1e-6 * scale, (t, Bd[t], opt)


def test_assignment_small_brute_force():
    torch = torch_mod()
    rng = np.random.default_rng(3)
    m = 6
    seq = [rng.integers(-3, 3, (m, m)).astype(np.float32) for _ in range(4)]
    costs = dev(np.stack(seq), torch.float32)
    perms = torch.zeros((4, m), dtype=torch.int32, device="cuda")
    bounds = torch.zeros(4, dtype=torch.float64, device="cuda")
    gb.qap32_assignments(costs, perms, bounds)
    for t, Ct in enumerate(seq):
        best = min(sum(float(Ct[i, q[i]]) for i in range(m)) for q in itertools.permutations(range(m)))
        assert _perm_cost(Ct, perms[t].cpu().numpy()) == best
        assert abs(bounds[t].item() - best) < 1e-9


def test_assignment_infeasible_forbidden_pattern():
    torch = torch_mod()
    m = 7
    C = np.ones((1, m, m), np.float32)
    C[0, :, 0] = np.inf  # column 0 cannot be covered
    costs = dev(C, torch.float32)
    perms = torch.zeros((1, m), dtype=torch.int32, device="cuda")
    bounds = torch.zeros(1, dtype=torch.float64, device="cuda")
    gb.qap32_assignments(costs, perms, bounds)
    assert (perms.cpu().numpy() == -1).all()
    assert bounds.item() == float("-inf")


def _nodes(m, N, fixes=None):
    n = m * m
    LB = np.zeros((N, n))
    UB = np.ones((N, n))
    for k, (kind, i, j) in (fixes or {}).items():
        if kind == "force":
            LB[k, i * m + j] = 1.0
        else:
            UB[k, i * m + j] = 0.0
    return LB, UB


def _run_qap32(A, B, LB, UB, max_iters=3000, eps=1e-7, stop=float("-inf"), x0=None, status0=None):
    torch = torch_mod()
    m = A.shape[0]
    N, n = LB.shape
    dA, dB = dev(A), dev(B)
    obj = gb.ObjectiveDesc(gb.OBJ_QAP, n, 0, 0, 1.0, 1.0, 0.0, m, dA.data_ptr(), dB.data_ptr())
    lmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, m)
    q = gb.Qap32(obj)
    x = dev(x0 if x0 is not None else np.zeros((N, n)))
    lb, ub = dev(LB), dev(UB)
    e = dev(np.full(N, eps))
    it = torch.zeros(N, dtype=torch.int32, device="cuda")
    st = dev(status0 if status0 is not None else np.zeros(N, np.uint8))
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    flags = dev(np.ones(n, np.uint8))
    q.relax(lmo, N, max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv, flags, 1e-6, stop)
    torch.cuda.synchronize()
    out = dict(x=x.cpu().numpy(), gap=gap.cpu().numpy(), value=val.cpu().numpy(), it=it.cpu().numpy(),
               st=st.cpu().numpy(), bv=bv.cpu().numpy(), bvv=bvv.cpu().numpy())
    q.close()
    del dA, dB
    return out


def _host_certificate(A, B, P, mask):
    R = A @ P - P @ B
    f = float(np.sum(R * R))
    G = 2.0 * (A.T @ R - R @ B.T)
    Gm = np.where(mask, G, 1e300)
    r, c = scipy_opt.linear_sum_assignment(Gm)
    gmin = float(np.sum(G[r, c]))
    return f, float(np.sum(G * P)) - gmin


def test_relaxation_certificate_planted():
    m = 12
    A, B, perm, _ = instances.qap_instance(m, seed=7)
    LB, UB = _nodes(m, 1)
    say(f"[qap32] relaxation m={m} planted")
    out = _run_qap32(A, B, LB, UB, max_iters=400)
    P = out["x"][0].reshape(m, m)
    f, gap_exact = _host_certificate(A, B, P, np.ones((m, m), bool))
    assert abs(out["value"][0] - f) <= 1e-9 * (1.0 + f)
    # the Lagrangian bound can only make the certified gap larger
    assert out["gap"][0] >= gap_exact - 1e-9 * (1.0 + abs(gap_exact))
    assert out["gap"][0] <= gap_exact + 1e-5 * (1.0 + abs(gap_exact))
    assert out["value"][0] - out["gap"][0] <= 1e-9  # the relaxation's optimum is 0
    assert np.allclose(P.sum(0), 1.0, atol=1e-4) and np.allclose(P.sum(1), 1.0, atol=1e-4)
    assert out["bvv"][0] >= 0.0


def test_relaxation_bound_against_plain_form():
    """A non-planted instance: value - gap of the FP32 form never exceeds the
    plain FP64 form's relaxation value (an upper bound of the optimum)."""
    from tests.gpu_util import Problem
    m = 8
    rng = np.random.default_rng(11)
    A = (rng.random((m, m)) < 0.4).astype(np.float64)
    A = np.triu(A, 1) + np.triu(A, 1).T
    Bm = (rng.random((m, m)) < 0.4).astype(np.float64)
    Bm = np.triu(Bm, 1) + np.triu(Bm, 1).T
    n = m * m
    prob = Problem("qap", A=A, B=Bm, lmo="birkhoff", order=m)
    alg = {"dicg": True, "ls": gb.LS_QUADRATIC_EXACT}
    LB, UB = _nodes(m, 1)
    g = prob.run_gpu(alg, LB.copy(), LB, UB, np.full(1, 1e-9), np.zeros(1, np.int32), np.zeros(1, np.uint8),
                     np.ones(n, np.uint8), 20000)
    plain_value = float(g["value"][0])
    out = _run_qap32(A, Bm, LB, UB, max_iters=3000, eps=1e-9)
    bound = out["value"][0] - out["gap"][0]
    say(f"[qap32] m={m} plain value {plain_value:.9f}, FP32 form bound {bound:.9f} value {out['value'][0]:.9f}")
    assert bound <= plain_value + 1e-9 * (1.0 + abs(plain_value))
    assert out["value"][0] >= bound
    assert abs(out["value"][0] - plain_value) <= 1e-3 * (1.0 + plain_value)


def test_node_bounds_and_infeasible_nodes():
    m = 10
    A, B, perm, _ = instances.qap_instance(m, seed=3)
    fixes = {0: ("force", 2, 5), 1: ("forbid", 0, 0), 2: ("force", 4, 4)}
    LB, UB = _nodes(m, 4, fixes)
    # node 3: two forced edges in one row
    LB[3, 1 * m + 2] = 1.0
    LB[3, 1 * m + 3] = 1.0
    out = _run_qap32(A, B, LB, UB, max_iters=200)
    assert out["st"][3] == gb.NODE_INFEASIBLE
    for k in range(3):
        V = out["bv"][k].reshape(m, m)
        assert out["bvv"][k] < 1e299
        assert np.allclose(V.sum(0), 1) and np.allclose(V.sum(1), 1)
        assert np.all(V[LB[k].reshape(m, m) > 0.5] == 1.0)
        assert np.all(V[UB[k].reshape(m, m) < 0.5] == 0.0)
        P = out["x"][k].reshape(m, m)
        assert np.all(P[UB[k].reshape(m, m) < 0.5] == 0.0)


def _solve_qap(A, B, natural=None, max_nodes=0):
    m = A.shape[0]
    n = m * m
    dA, dB = dev(A), dev(B)
    obj = gb.ObjectiveDesc(gb.OBJ_QAP, n, 0, 0, 1.0, 1.0, 0.0, m, dA.data_ptr(), dB.data_ptr())
    lmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, m)
    s = gb.default_settings()
    s.form = 2
    s.precision = gb.PREC_FP32
    s.abs_gap_tol = s.rel_gap_tol = 1e-6
    s.gpu_batch_size = 64
    s.algorithm.line_search = gb.LS_QUADRATIC_EXACT
    if natural is not None:
        s.natural_lower_bound = natural
    s.max_nodes = max_nodes
    r = gb.solve(n, [0.0] * n, [1.0] * n, list(range(n)), obj, lmo, s)
    del dA, dB
    return r


def test_search_proves_planted_optimum():
    m = 32
    A, B, perm, _ = instances.qap_instance(m, seed=m)
    say(f"[qap32] gb_solve form 2 m={m}")
    r = _solve_qap(A, B, natural=0.0)
    X = np.asarray(r["x"]).reshape(m, m)
    assert r["objective"] == 0.0, r["objective"]
    assert np.allclose(A @ X, X @ B)


def test_search_matches_enumeration():
    m = 6
    rng = np.random.default_rng(2)
    A = rng.integers(0, 3, (m, m)).astype(np.float64)
    B = rng.integers(0, 3, (m, m)).astype(np.float64)
    best = min(float(np.sum((A[:, np.argsort(q)] - B[list(q), :]) ** 2))
               for q in itertools.permutations(range(m)))
    r = _solve_qap(A, B)
    say(f"[qap32] m={m} enumeration {best}, search {r['objective']} ({r['nodes_processed']} nodes)")
    assert abs(r["objective"] - best) <= 1e-6 * (1.0 + best)
    assert r["dual_bound"] <= best + 1e-6 * (1.0 + best)
