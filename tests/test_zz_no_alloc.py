"""The allocation-free relaxation (gb_relax_batch_noalloc) through the C-ABI:
bit for bit the output of gb_relax_batch on the same inputs, on two calls in
a row (the second with a smaller batch, so it runs on the work array the
first call left behind), and a clear error beyond the array's size."""
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import Problem, alg_gpu, assert_bit_equal, dev

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def node_batch(n, N, seed):
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.3 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    eps = np.maximum(1e-6, 1e-3 * 0.8 ** rng.integers(0, 10, N))
    return LB, UB, eps


def run(entry, prob, alg, LB, UB, eps, flags, max_iters):
    import torch
    N, n = LB.shape
    x, lb, ub, e = dev(LB.copy()), dev(LB), dev(UB), dev(eps)
    it = torch.zeros(N, dtype=torch.int32, device="cuda")
    st = torch.zeros(N, dtype=torch.uint8, device="cuda")
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    fl = dev(np.asarray(flags, dtype=np.uint8))
    entry(prob.gobj, alg_gpu(alg), prob.glmo, N, max_iters, x, lb, ub, e, it, st, gap, val, bv, bvv, fl, 1e-6, 0)
    torch.cuda.synchronize()
    return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
            "iterations": it.cpu().numpy(), "status": st.cpu().numpy(),
            "best_vertex": bv.cpu().numpy(), "best_vertex_val": bvv.cpu().numpy()}


def test_noalloc_equals_relax_batch_on_two_calls():
    cases = [
        # (label, n, N, seed, algorithm, budget): the second call is smaller
        ("n=160 N=200", 160, 200, 3, {"dicg": True, "ls": gb.LS_SECANT}, 120),
        ("n=48 N=24", 48, 24, 4, {"dicg": True, "ls": gb.LS_QUADRATIC_EXACT}, 300),
    ]
    inputs, refs = [], []
    for label, n, N, seed, alg, budget in cases:
        assert N * n * 48 <= gb_noalloc_bytes()
        Q, c, _ = instances.binary_qp(n, seed=seed)
        prob = Problem("dense", Q, c)
        LB, UB, eps = node_batch(n, N, seed)
        inputs.append((prob, alg, LB, UB, eps, np.ones(n, np.uint8), budget))
        say(f"[noalloc] gb_relax_batch {label} alg={alg} budget={budget}")
        refs.append(run(gb.relax_batch, *inputs[-1]))
        assert int(refs[-1]["iterations"].sum()) > 0
    # the two allocation-free calls back to back, larger batch first
    for (label, *_), args, ref in zip(cases, inputs, refs):
        say(f"[noalloc] gb_relax_batch_noalloc {label}")
        got = run(gb.relax_batch_noalloc, *args)
        assert_bit_equal(got, ref, f"noalloc {label}")


def test_noalloc_simplex_after_dense():
    """A simplex batch (its own LDS layout and work-vector use) after the
    dense calls above left their contents in the array."""
    rng = np.random.default_rng(9)
    n, N = 40, 12
    M = rng.standard_normal((n, n))
    Q = M.T @ M / n + 0.05 * np.eye(n)
    c = rng.uniform(-1, 1, n)
    is_int = (rng.random(n) < 0.5).astype(np.uint8)
    LB = np.zeros((N, n))
    UB = np.where(is_int, 3.0, 5.0) * np.ones((N, n))
    prob = Problem("dense", Q, c, lmo="simplex", rhs=4.0, equality=True, is_int=is_int)
    alg = {"dicg": True, "ls": gb.LS_SECANT}
    say("[noalloc] simplex n=40 N=12")
    ref = run(gb.relax_batch, prob, alg, LB, UB, np.full(N, 1e-6), is_int, 200)
    got = run(gb.relax_batch_noalloc, prob, alg, LB, UB, np.full(N, 1e-6), is_int, 200)
    assert_bit_equal(got, ref, "noalloc simplex")


def gb_noalloc_bytes():
    return 16 * 1024 * 1024


def test_noalloc_rejects_batches_beyond_the_array():
    import torch
    n, N = 100, 4096  # 4096 * 100 * 48 bytes > 16 MiB
    assert N * n * 48 > gb_noalloc_bytes()
    Q, c, _ = instances.binary_qp(n, seed=1)
    prob = Problem("dense", Q, c)
    small = torch.zeros(8, dtype=torch.float64, device="cuda")  # never read: rejected before launch
    it = torch.zeros(8, dtype=torch.int32, device="cuda")
    st = torch.zeros(8, dtype=torch.uint8, device="cuda")
    say("[noalloc] oversize batch is rejected")
    with pytest.raises(gb.InvalidArgument, match="GB_NOALLOC_WORK_BYTES"):
        gb.relax_batch_noalloc(prob.gobj, gb.algorithm(), prob.glmo, N, 10, small, small, small, small, it, st,
                               small, small, small, small)
