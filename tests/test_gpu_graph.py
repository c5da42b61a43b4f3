"""Captured batched-form launches (gb_relax_graph_*): every replay of one
graph, with eager work and synchronisations between replays and the inputs
refilled in place, gives bit for bit what the eager launch and the CPU
oracle give on the same inputs."""
import sys

import numpy as np
import pytest

import gpuboscia_amd as gb
from gpuboscia_amd import instances
from tests.gpu_util import assert_bit_equal, dev

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def node_set(n, N, seed):
    rng = np.random.default_rng(seed)
    LB, UB = np.zeros((N, n)), np.ones((N, n))
    for i in range(1, N):
        fix = rng.random(n) < 0.4 * rng.random()
        val = rng.random(n) < 0.5
        LB[i, fix & val] = 1.0
        UB[i, fix & ~val] = 0.0
    eps = np.maximum(1e-6, 1e-3 * 0.8 ** rng.integers(0, 12, N))
    return LB, UB, eps


@pytest.mark.parametrize("n", [40, 150])
def test_graph_replays_match_eager_and_oracle(n):
    import torch
    from oracle import pyoracle as orc
    N, budget = 96, 300
    Q, c, _ = instances.binary_qp(n, seed=11)
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    lmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
    alg = gb.algorithm(dicg=True)
    flags = dev(np.ones(n, np.uint8))
    f64 = dict(dtype=torch.float64, device="cuda")

    def buffers():
        return {"x": torch.zeros((N, n), **f64), "lb": torch.zeros((N, n), **f64),
                "ub": torch.zeros((N, n), **f64), "eps": torch.zeros(N, **f64),
                "it": torch.zeros(N, dtype=torch.int32, device="cuda"),
                "st": torch.zeros(N, dtype=torch.uint8, device="cuda"),
                "gap": torch.zeros(N, **f64), "val": torch.zeros(N, **f64),
                "bv": torch.zeros((N, n), **f64), "bvv": torch.zeros(N, **f64)}

    def fill(b, LB, UB, eps):
        b["x"].copy_(torch.from_numpy(LB))
        b["lb"].copy_(torch.from_numpy(LB))
        b["ub"].copy_(torch.from_numpy(UB))
        b["eps"].copy_(torch.from_numpy(eps))
        b["it"].zero_()
        b["st"].zero_()
        b["bvv"].fill_(1e300)

    def result(b):
        return {"x": b["x"].cpu().numpy(), "gap": b["gap"].cpu().numpy(), "value": b["val"].cpu().numpy(),
                "iterations": b["it"].cpu().numpy(), "status": b["st"].cpu().numpy(),
                "best_vertex": b["bv"].cpu().numpy(), "best_vertex_val": b["bvv"].cpu().numpy()}

    gbuf, ebuf = buffers(), buffers()
    graph = gb.RelaxGraph(obj, alg, lmo, N, budget, gbuf["x"], gbuf["lb"], gbuf["ub"], gbuf["eps"],
                          gbuf["it"], gbuf["st"], gbuf["gap"], gbuf["val"], gbuf["bv"], gbuf["bvv"],
                          flags, 1e-6)
    for replay, seed in enumerate((1, 2, 3, 2)):
        LB, UB, eps = node_set(n, N, seed)
        fill(gbuf, LB, UB, eps)
        fill(ebuf, LB, UB, eps)
        say(f"[graph] n={n} replay {replay} seed {seed}")
        graph.launch()
        torch.cuda.synchronize()
        # eager work between replays
        _ = (torch.arange(1 << 16, device="cuda", dtype=torch.float64) * 2).sum().item()
        gb.relax_batch_fast(obj, alg, lmo, N, budget, ebuf["x"], ebuf["lb"], ebuf["ub"], ebuf["eps"],
                            ebuf["it"], ebuf["st"], ebuf["gap"], ebuf["val"], ebuf["bv"], ebuf["bvv"],
                            flags, 1e-6)
        torch.cuda.synchronize()
        g, e = result(gbuf), result(ebuf)
        assert_bit_equal(g, e, f"graph replay {replay} vs eager")
        if replay == 0:
            o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, LB.copy(), LB.copy(),
                                     UB.copy(), eps.copy(), np.zeros(N, np.int32), np.zeros(N, np.uint8),
                                     flags=np.ones(n, np.uint8), max_iters=budget, refresh=64, threads=8)
            assert_bit_equal(g, o, "graph replay vs oracle")
    graph.close()


def test_graph_rejects_empty_batch():
    import torch
    n = 8
    z = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, n, z.data_ptr(), 0, 2.0, 0.5, 0.0, 0, 0, 0)
    with pytest.raises(gb.InvalidArgument):
        gb.RelaxGraph(obj, gb.algorithm(), gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), 0, 10, z, z, z, z, z, z, z,
                      z, z, z)
