"""The device-resident search (gb_settings.device_frontier): frontier,
incumbent, pruning and branching on the GPU, one captured graph per round.
Pinned against the host driver with one stream slot (num_streams = 1),
which it restates: the same problem and settings must give the same tree,
so the objective, incumbent, dual bound, node count and FW iteration count
are compared exactly, on proven searches, on searches cut by the node cap,
with small batches (many rounds, frequent refills) and both branching rules
and both batched forms."""
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import dev

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense(Q, c):
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, Q.shape[0], dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    return obj, (dQt, dc)


BOX = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)


def settings(form, rule, batch, device, max_nodes=0, k=3):
    s = gb.default_settings()
    s.abs_gap_tol = 1e-6
    s.rel_gap_tol = 1e-6
    s.form = form
    if form == 2:
        s.precision = gb.PREC_FP32
    s.branch_rule = rule
    s.branch_k = k
    s.gpu_batch_size = batch
    s.num_streams = 1
    s.max_nodes = max_nodes
    s.time_limit_seconds = 120.0
    s.device_frontier = 1 if device else 0
    return s


def both(n, seed, form, rule, batch, max_nodes=0, k=3):
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj, keep = dense(Q, c)
    out = []
    for device in (False, True):
        say(f"[dsearch] n={n} seed={seed} form={form} rule={rule} batch={batch} cap={max_nodes} k={k} "
            f"{'device' if device else 'host'}")
        r = gb.solve(n, [0] * n, [1] * n, range(n), obj, BOX, settings(form, rule, batch, device, max_nodes, k))
        say(f"[dsearch]   objective {r['objective']!r} bound {r['dual_bound']!r} nodes {r['nodes_processed']} "
            f"iterations {r['fw_iterations']} {r['elapsed_seconds']:.3f}s")
        out.append(r)
    return out


def assert_same(h, d, what):
    for key in ("objective", "dual_bound", "nodes_processed", "fw_iterations", "reached_tolerance",
                "reached_time_limit"):
        assert h[key] == d[key], f"{what}: {key} host {h[key]!r} device {d[key]!r}"
    assert np.array_equal(np.array(h["x"]), np.array(d["x"]), equal_nan=True), f"{what}: incumbent"


@pytest.mark.parametrize("n,seed", [(20, 1), (40, 2), (60, 1)])
@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("rule", [gb.BRANCH_POW2, gb.BRANCH_MOST_FRACTIONAL])
def test_device_search_equals_host_search(n, seed, form, rule):
    h, d = both(n, seed, form, rule, 1024)
    assert_same(h, d, f"n={n} seed={seed} form={form} rule={rule}")
    assert d["reached_tolerance"] or d["dual_bound"] == d["objective"]


@pytest.mark.parametrize("batch", [8, 32, 128])
@pytest.mark.parametrize("form", [1, 2])
def test_small_batches_refill_like_the_host(batch, form):
    h, d = both(60, 2, form, gb.BRANCH_POW2, batch)
    assert_same(h, d, f"batch={batch} form={form}")


@pytest.mark.parametrize("cap", [1, 37, 300])
def test_node_cap_stops_where_the_host_stops(cap):
    h, d = both(80, 3, 2, gb.BRANCH_POW2, 64, max_nodes=cap)
    assert_same(h, d, f"cap={cap}")
    assert d["reached_time_limit"]


@pytest.mark.parametrize("k", [1, 2])
def test_branch_k(k):
    h, d = both(40, 5, 1, gb.BRANCH_POW2, 64, k=k)
    assert_same(h, d, f"k={k}")


def test_larger_tree():
    h, d = both(100, 1, 2, gb.BRANCH_POW2, 1024)
    assert_same(h, d, "n=100")


def test_unsupported_settings_are_rejected():
    Q, c, _ = instances.binary_qp(20, seed=1)
    obj, keep = dense(Q, c)
    s = settings(0, gb.BRANCH_POW2, 1024, True)
    with pytest.raises(RuntimeError, match="device_frontier"):
        gb.solve(20, [0] * 20, [1] * 20, range(20), obj, BOX, s)
    s = settings(1, gb.BRANCH_LEARNED, 1024, True)
    with pytest.raises(RuntimeError, match="device_frontier"):
        gb.solve(20, [0] * 20, [1] * 20, range(20), obj, BOX, s)


@pytest.mark.parametrize("n,seed,card", [(30, 1, 3), (60, 2, 6)])
@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("batch", [1024, 16])
def test_device_search_over_simplex_equals_host_search(n, seed, card, form, batch):
    """Over the bounded simplex sum x <= card (0/1 variables) children start
    at the LMO vertex in both drivers; the trees are the same."""
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj, keep = dense(Q, c)
    ii = dev(np.ones(n, np.uint8))
    lmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(card), 0, ii.data_ptr(), 0)
    out = []
    for device in (False, True):
        say(f"[dsearch] simplex n={n} seed={seed} card={card} form={form} batch={batch} "
            f"{'device' if device else 'host'}")
        s = settings(form, gb.BRANCH_POW2, batch, device)
        if form == 2:
            s.algorithm.algorithm = gb.ALG_DICG
        r = gb.solve(n, [0] * n, [1] * n, range(n), obj, lmo, s)
        say(f"[dsearch]   objective {r['objective']!r} nodes {r['nodes_processed']}")
        out.append(r)
    assert_same(out[0], out[1], f"simplex n={n} form={form} batch={batch}")
    assert sum(out[1]["x"]) <= card
