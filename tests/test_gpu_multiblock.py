"""MultiBlock relaxation (gb_relax_batch_multiblock, K workgroups per node
meeting at the device-atomic group barrier): the same relaxation as the
one-workgroup kernel, so every output must equal gb_relax_batch's bit for
bit for any K, over every step rule of FW and DICG, with branched nodes, a
warm start and a continuation, over a box and over unit and equality
simplices (integer and continuous coordinates, fixed faces, an infeasible
node).  Beyond the one-workgroup kernel's LDS limit (n > 8192 over a box,
n > about 4000 over a simplex) the oracle is the reference; and gb_solve
with blocks_per_node must find the same tree as without it."""
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import Problem, alg_gpu, assert_bit_equal, auto_T, dev

pytestmark = pytest.mark.gpu

FW_KINDS = [gb.LS_OPEN_LOOP, gb.LS_SECANT, gb.LS_QUADRATIC_EXACT]
DICG_KINDS = [gb.LS_OPEN_LOOP, gb.LS_SCALED, gb.LS_SAFEGUARDED, gb.LS_SECANT, gb.LS_QUADRATIC_EXACT]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nodes(n, N, seed):
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.3 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    X0 = LB.copy()
    ST = np.zeros(N, np.uint8)
    IT = np.zeros(N, np.int32)
    # node 1: a supplied start point; node 2: a continuation
    X0[1] = np.clip(rng.random(n), LB[1], UB[1])
    ST[1] = gb.NODE_INITIAL_POINT_SUPPLIED
    X0[2] = np.clip(rng.random(n), LB[2], UB[2])
    IT[2] = 7
    return LB, UB, X0, ST, IT


def run(prob, alg, LB, UB, X0, ST, IT, eps, max_iters, T, K, flags=None):
    import torch
    N, n = LB.shape
    x, lb, ub, e = dev(X0), dev(LB), dev(UB), dev(np.full(N, eps))
    it, st = dev(IT), dev(ST)
    f64 = dict(dtype=torch.float64, device="cuda")
    gap, val = torch.zeros(N, **f64), torch.zeros(N, **f64)
    bv, bvv = torch.zeros((N, n), **f64), torch.full((N,), 1e300, **f64)
    fl = dev(np.ones(n, np.uint8) if flags is None else flags)
    if K == 0:
        gb.relax_batch(prob.gobj, alg_gpu(alg), prob.glmo, N, max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv,
                       fl, 1e-6, T)
    else:
        gb.relax_batch_multiblock(prob.gobj, alg_gpu(alg), prob.glmo, N, max_iters, x, lb, ub, e, it, st, gap,
                                  val, bv, bvv, fl, 1e-6, T, K)
    torch.cuda.synchronize()
    return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
            "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
            "best_vertex_val": bvv.cpu().numpy()}


@pytest.mark.parametrize("n", [100, 300, 1000])
@pytest.mark.parametrize("dicg,ls", [(False, k) for k in FW_KINDS] + [(True, k) for k in DICG_KINDS])
def test_multiblock_equals_one_workgroup(n, dicg, ls):
    Q, c, _ = instances.binary_qp(n, seed=11)
    prob = Problem("dense", Q, c)
    LB, UB, X0, ST, IT = nodes(n, 6, seed=n + 3)
    alg = {"dicg": dicg, "ls": ls}
    T = auto_T(n)
    say(f"[multiblock] n={n} dicg={dicg} ls={ls} one workgroup")
    ref = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 40, T, 0)
    for K in (1, 2, 4, 8):
        say(f"[multiblock] n={n} dicg={dicg} ls={ls} K={K}")
        got = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 40, T, K)
        assert_bit_equal(got, ref, f"n={n} K={K}")


@pytest.mark.parametrize("T", [64, 512, 1024])
def test_multiblock_follows_the_block_size_order(T):
    """The sums follow the one-workgroup launch of the same block size."""
    n = 700
    Q, c, _ = instances.binary_qp(n, seed=2)
    prob = Problem("dense", Q, c)
    LB, UB, X0, ST, IT = nodes(n, 4, seed=9)
    alg = {"dicg": True, "ls": gb.LS_SECANT}
    ref = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 30, T, 0)
    K = max(1, T // 256)
    say(f"[multiblock] T={T} K={K}")
    got = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 30, T, K)
    assert_bit_equal(got, ref, f"T={T}")


def test_multiblock_beyond_lds_matches_oracle():
    """n = 9000 exceeds the one-workgroup kernel's LDS-staged vector: the
    oracle (same algorithm, canonical 256-lane sums) is the reference."""
    from oracle import pyoracle as orc
    n = 9000
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n // 8, n)) / np.sqrt(n)
    Q = A.T @ A + 0.1 * np.eye(n)
    c = rng.uniform(-1, 1, n)
    prob = Problem("dense", Q, c)
    LB, UB, X0, ST, IT = nodes(n, 3, seed=5)
    alg = {"dicg": True, "ls": gb.LS_QUADRATIC_EXACT}
    say("[multiblock] n=9000 K=16 against the oracle")
    got = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 6, 256, 16)
    o = prob.run_orc(alg, X0, LB, UB, np.full(3, 1e-6), IT, ST, np.ones(n, np.uint8), 6, T=256, threads=3)
    assert_bit_equal(got, o, "n=9000")


def test_solve_with_blocks_per_node_is_the_same_search():
    n = 40
    Q, c, _ = instances.binary_qp(n, seed=3)
    prob = Problem("dense", Q, c)
    out = []
    for K in (0, 4):
        s = gb.default_settings()
        s.abs_gap_tol = 1e-6
        s.rel_gap_tol = 1e-6
        s.blocks_per_node = K
        say(f"[multiblock] gb_solve n={n} blocks_per_node={K}")
        out.append(gb.solve(n, [0] * n, [1] * n, range(n), prob.gobj, prob.glmo, s))
    a, b = out
    for key in ("objective", "dual_bound", "nodes_processed", "fw_iterations"):
        assert a[key] == b[key], (key, a[key], b[key])
    assert a["x"] == b["x"]


def simplex_problem(n, equality, seed):
    """Mixed integer/continuous simplex nodes after the simplex parity test:
    integer coordinates up to 3, continuous up to 5, rhs 4, branched nodes
    (upper bounds 0/1 on integers, a lower bound 1 on one integer), and one
    node whose fixed mass exceeds rhs (infeasible)."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n // 2, n))
    Q = M.T @ M / n + 0.05 * np.eye(n)
    c = rng.uniform(-1, 1, n)
    is_int = (rng.random(n) < 0.5).astype(np.uint8)
    N = 7
    LB = np.zeros((N, n))
    UB = np.where(is_int, 3.0, 5.0) * np.ones((N, n))
    ints = np.flatnonzero(is_int)
    for i in range(1, N - 1):
        j = rng.integers(0, n, 3)
        UB[i, j] = np.where(is_int[j], rng.integers(0, 2, 3), UB[i, j])
    LB[N - 2, ints[0]] = 1.0
    LB[N - 1, ints[:5]] = 1.0  # fixed mass 5 > rhs 4
    prob = Problem("dense", Q, c, lmo="simplex", rhs=4.0, equality=equality, is_int=is_int)
    return prob, LB, UB, is_int


@pytest.mark.parametrize("n", [100, 700, 2000])
@pytest.mark.parametrize("equality", [False, True])
@pytest.mark.parametrize("dicg,ls", [(True, gb.LS_SECANT), (True, gb.LS_OPEN_LOOP), (False, gb.LS_QUADRATIC_EXACT)])
def test_multiblock_simplex_equals_one_workgroup(n, equality, dicg, ls):
    prob, LB, UB, is_int = simplex_problem(n, equality, seed=n + equality)
    N = LB.shape[0]
    X0 = LB.copy()
    ST, IT = np.zeros(N, np.uint8), np.zeros(N, np.int32)
    alg = {"dicg": dicg, "ls": ls}
    T = auto_T(n)
    say(f"[multiblock] simplex n={n} eq={equality} dicg={dicg} ls={ls} one workgroup")
    ref = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 30, T, 0, flags=is_int)
    assert ref["status"][-1] == gb.NODE_INFEASIBLE
    for K in (1, 3, 8):
        say(f"[multiblock] simplex n={n} eq={equality} dicg={dicg} ls={ls} K={K}")
        got = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 30, T, K, flags=is_int)
        assert_bit_equal(got, ref, f"simplex n={n} K={K}")


@pytest.mark.parametrize("equality", [False, True])
def test_multiblock_simplex_beyond_lds_matches_oracle(equality):
    """n = 9000 over a simplex (past the one-workgroup kernel's LDS-staged
    ranks): the oracle is the reference."""
    n = 9000
    prob, LB, UB, is_int = simplex_problem(n, equality, seed=17 + equality)
    N = LB.shape[0]
    X0 = LB.copy()
    ST, IT = np.zeros(N, np.uint8), np.zeros(N, np.int32)
    alg = {"dicg": True, "ls": gb.LS_QUADRATIC_EXACT}
    say(f"[multiblock] simplex n=9000 eq={equality} K=32 against the oracle")
    got = run(prob, alg, LB, UB, X0, ST, IT, 1e-6, 6, 256, 32, flags=is_int)
    o = prob.run_orc(alg, X0, LB, UB, np.full(N, 1e-6), IT, ST, is_int, 6, T=256, threads=8)
    assert_bit_equal(got, o, f"simplex n=9000 eq={equality}")
