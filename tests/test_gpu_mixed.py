"""Mixed-precision form (gb_mixed_*, gb_relax_batch_mixed): the BF16 MFMA
products against FP64 within the stated error bound, the FP64 pass bit for
bit against the oracle started from the mixed pass's output, and the
mixed-form B&B at the exhaustive / plain-form optimum."""
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


def dense(Q, c):
    dQt, dc = dev(np.ascontiguousarray(Q.T)), dev(c)
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, Q.shape[0], dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    return obj, (dQt, dc)


@pytest.mark.parametrize("n", [7, 40, 200, 256, 300, 512])
@pytest.mark.parametrize("planes", [1, 3])
def test_products_within_error_bound(n, planes):
    import torch
    Q, c, _ = instances.binary_qp(n, seed=5)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, planes)
    rng = np.random.default_rng(n)
    N = 37  # not a multiple of the 16-column tile
    V = rng.integers(0, 2, (N, n)).astype(np.float64)
    dY = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    mq.products(dev(V), dY)
    torch.cuda.synchronize()
    Y = dY.cpu().numpy()
    ref = V @ Q.T
    mag = np.abs(V) @ np.abs(Q).T
    # bf16 rounding of Q (2^-9 per entry for one plane, about 2^-25 for three)
    # plus FP32 accumulation over n terms
    rel = (2.0 ** -8 if planes == 1 else 2.0 ** -21) + n * 2.0 ** -23
    err = np.abs(Y - ref)
    say(f"[mixed] n={n} planes={planes}: max rel err {float((err / np.maximum(mag, 1e-300)).max()):.3e}")
    assert np.all(err <= rel * mag + 1e-300)
    if planes == 3:
        assert err.max() <= 1e-5 * mag.max()
    mq.close()


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


def run_mixed(mq, obj, dicg, LB, UB, X, ST, IT, eps, budget, fp64_iters, cutoff=None):
    import torch
    N, n = LB.shape
    x = dev(X)
    it = dev(IT.astype(np.int32))
    st = dev(ST.astype(np.uint8))
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    mit = torch.zeros(N, dtype=torch.int64, device="cuda")
    gb.relax_batch_mixed(mq, obj, gb.algorithm(dicg=dicg), gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), N, budget, x,
                         dev(LB), dev(UB), dev(eps), it, st, gap, val, bv, bvv, dev(np.ones(n, np.uint8)), 1e-6,
                         None if cutoff is None else dev(cutoff), fp64_max_iters=fp64_iters,
                         mixed_iterations=mit)
    torch.cuda.synchronize()
    return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
            "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
            "best_vertex_val": bvv.cpu().numpy(), "mixed_iterations": mit.cpu().numpy()}


@pytest.mark.parametrize("n,planes,dicg,cut", [(40, 3, True, False), (130, 1, True, True),
                                               (200, 3, False, False), (256, 3, True, True),
                                               (300, 1, True, False), (512, 3, True, True)])
def test_fp64_pass_matches_oracle_from_mixed_output(n, planes, dicg, cut):
    from oracle import pyoracle as orc
    N, budget = 70, 400
    Q, c, _ = instances.binary_qp(n, seed=21)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, planes)
    LB, UB, eps = node_set(n, N, 3)
    zeros = np.zeros(N)
    cutoff = np.full(N, -2.0) if cut else None
    say(f"[mixed] n={n} planes={planes} dicg={dicg} cutoff={cut}")
    p1 = run_mixed(mq, obj, dicg, LB, UB, LB.copy(), zeros, zeros, eps, budget, -1, cutoff)
    assert np.all(p1["status"] == gb.NODE_INITIAL_POINT_SUPPLIED)
    assert np.all(p1["x"] >= LB) and np.all(p1["x"] <= UB)
    assert np.all(p1["iterations"] == p1["mixed_iterations"]) and np.all(p1["iterations"] >= 1)
    full = run_mixed(mq, obj, dicg, LB, UB, LB.copy(), zeros, zeros, eps, budget, 0, cutoff)
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), dicg, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].astype(np.int32).copy(),
                             p1["status"].copy(), flags=np.ones(n, np.uint8), max_iters=max(1, budget // 16),
                             cutoff=cutoff, refresh=64, threads=8, best_vertex=p1["best_vertex"].copy(),
                             best_vertex_val=p1["best_vertex_val"].copy())
    # the FP64 pass runs max(1, budget / 16) iterations by default
    assert_bit_equal({k: v for k, v in full.items() if k != "mixed_iterations"}, o, "mixed + FP64 pass")
    conv = np.isin(full["status"], (1, 4))
    say(f"[mixed] converged {conv.sum()}/{N}, mixed iters {int(p1['iterations'].sum())}, "
        f"total {int(full['iterations'].sum())}")
    if dicg:  # plain FW converges sublinearly: no node need reach eps within the budget
        assert conv.sum() > 0


def test_out_of_scope_nodes_skip_the_mixed_pass():
    from oracle import pyoracle as orc
    n, N, budget = 50, 20, 300
    Q, c, _ = instances.binary_qp(n, seed=4)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, 3)
    LB, UB, eps = node_set(n, N, 5)
    UB[3, 7] = 2.0            # integer bound beyond 0/1
    LB[5, 2] = 0.25           # fractional bound
    X = LB.copy()
    IT = np.zeros(N)
    ST = np.zeros(N)
    IT[8] = 5                 # continuation
    X[8] = np.clip(0.5, LB[8], UB[8])
    ST[8] = gb.NODE_INITIAL_POINT_SUPPLIED
    p1 = run_mixed(mq, obj, True, LB, UB, X.copy(), ST, IT, eps, budget, -1)
    for i in (3, 5, 8):
        assert p1["mixed_iterations"][i] == 0 and p1["iterations"][i] == IT[i]
        assert p1["status"][i] == ST[i] and np.array_equal(p1["x"][i], X[i])
    full = run_mixed(mq, obj, True, LB, UB, X.copy(), ST, IT, eps, budget, 0)
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].astype(np.int32).copy(),
                             p1["status"].copy(), flags=np.ones(n, np.uint8), max_iters=max(1, budget // 16),
                             refresh=64, threads=8, best_vertex=p1["best_vertex"].copy(),
                             best_vertex_val=p1["best_vertex_val"].copy())
    assert_bit_equal({k: v for k, v in full.items() if k != "mixed_iterations"}, o, "out-of-scope nodes")


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_mixed_bnb_matches_enumeration(seed, precision):
    from oracle import pyoracle as orc
    n = 20
    Q, c, _ = instances.binary_qp(n, seed=seed, rows=30)
    obj, keep = dense(Q, c)
    s = gb.default_settings()
    s.abs_gap_tol = 1e-6
    s.rel_gap_tol = 1e-6
    s.form = 2
    s.precision = gb.PREC_BF16 if precision == "bf16" else gb.PREC_FP32
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
    f, xo = orc.enumerate_binary_qp(Q, c)
    say(f"[mixed] C0 seed={seed} {precision}: {r['objective']:.12f} vs {f:.12f}, nodes {r['nodes_processed']}")
    assert np.array_equal(np.array(r["x"]), xo)
    assert abs(r["objective"] - f) <= 1e-9 * max(1.0, abs(f))
    assert r["dual_bound"] <= f + 1e-9 * max(1.0, abs(f))


@pytest.mark.parametrize("n,seed", [(40, 1), (40, 2), (60, 1)])
def test_mixed_bnb_matches_plain_optimum(n, seed):
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj, keep = dense(Q, c)
    res = {}
    for form in (1, 2):
        s = gb.default_settings()
        s.abs_gap_tol = 1e-6
        s.rel_gap_tol = 1e-6
        s.form = form
        s.time_limit_seconds = 60.0
        if form == 2:
            s.precision = gb.PREC_FP32
        res[form] = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
        say(f"[mixed] n={n} seed={seed} form={form}: {res[form]['objective']:.12f} nodes "
            f"{res[form]['nodes_processed']} {res[form]['elapsed_seconds']:.4f}s")
    assert not res[2]["reached_time_limit"]
    assert abs(res[1]["objective"] - res[2]["objective"]) <= 1e-9 * max(1.0, abs(res[1]["objective"]))


def test_mixed_search_pipeline_slots():
    """The mixed form keeps up to three batches in flight (num_streams >= 3),
    two with num_streams = 2 and one with 1: every schedule proves the same
    optimum, and a repeated search with the same schedule is the same tree."""
    n, seed = 80, 2
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj, keep = dense(Q, c)
    res = {}
    for streams in (1, 2, 8, 8):
        s = gb.default_settings()
        s.abs_gap_tol = 1e-6
        s.rel_gap_tol = 1e-6
        s.form = 2
        s.precision = gb.PREC_FP32
        s.num_streams = streams
        s.time_limit_seconds = 60.0
        r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
        say(f"[mixed] streams={streams}: {r['objective']:.12f} nodes {r['nodes_processed']} "
            f"{r['elapsed_seconds']:.4f}s")
        assert not r["reached_time_limit"]
        if streams in res:
            for key in ("nodes_processed", "fw_iterations", "objective", "dual_bound"):
                assert r[key] == res[streams][key], key
        res[streams] = r
    ref = res[1]["objective"]
    for r in res.values():
        assert abs(r["objective"] - ref) <= 1e-9 * max(1.0, abs(ref))


@pytest.mark.parametrize("form", [0, 1, 2])
def test_byte_bounds_give_the_same_search(form):
    """0/1 problems keep node bounds as bytes; lower bounds of -0.0 (not
    eligible) keep them as doubles.  Both searches are the same tree."""
    n = 40
    Q, c, _ = instances.binary_qp(n, seed=3)
    obj, keep = dense(Q, c)
    res = []
    for lb0 in (0.0, -0.0):
        s = gb.default_settings()
        s.abs_gap_tol = 1e-6
        s.rel_gap_tol = 1e-6
        s.form = form
        s.time_limit_seconds = 60.0
        r = gb.solve(n, [lb0] * n, [1.0] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
        say(f"[bounds] form={form} lb={lb0!r}: {r['objective']:.12f} nodes {r['nodes_processed']}")
        res.append(r)
    a, b = res
    assert not a["reached_time_limit"] and not b["reached_time_limit"]
    for key in ("objective", "dual_bound", "nodes_processed", "fw_iterations"):
        assert a[key] == b[key], key
    assert np.array_equal(np.asarray(a["x"]), np.asarray(b["x"]))


def test_autotuner_picks_a_valid_shape_and_leaves_inputs_untouched():
    from oracle import pyoracle as orc
    import torch
    n, N, budget = 120, 256, 300
    Q, c, _ = instances.binary_qp(n, seed=31)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, 1)
    LB, UB, eps = node_set(n, N, 8)
    x, lb, ub, e = dev(LB.copy()), dev(LB), dev(UB), dev(eps)
    it = dev(np.zeros(N, np.int32))
    st = dev(np.zeros(N, np.uint8))
    flags = dev(np.ones(n, np.uint8))
    alg = gb.algorithm(dicg=True)
    box = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
    opt, sec = gb.mixed_autotune(mq, obj, alg, box, N, budget, x, lb, ub, e, it, st, flags)
    say(f"[mixed] autotune n={n} N={N}: rw={opt.row_tiles_per_wave} cols={opt.columns} "
        f"fp64={opt.fp64_max_iters} {sec * 1e3:.3f} ms")
    assert 1 <= opt.row_tiles_per_wave <= 4 and opt.columns in (2, 4, 8, 16)
    assert opt.fp64_max_iters in (max(1, budget // d) for d in (32, 16, 8))
    torch.cuda.synchronize()
    assert np.array_equal(x.cpu().numpy(), LB) and not it.cpu().numpy().any() and not st.cpu().numpy().any()
    # the tuned shape computes the same algorithm: its FP64 pass matches the
    # oracle started from its own mixed-pass output
    zeros = np.zeros(N)
    p1 = run_mixed_opts(mq, obj, LB, UB, eps, budget, -1, opt)
    full = run_mixed_opts(mq, obj, LB, UB, eps, budget, opt.fp64_max_iters, opt)
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].astype(np.int32).copy(), p1["status"].copy(),
                             flags=np.ones(n, np.uint8), max_iters=opt.fp64_max_iters, refresh=64, threads=8,
                             best_vertex=p1["best_vertex"].copy(), best_vertex_val=p1["best_vertex_val"].copy())
    assert_bit_equal(full, o, "tuned shape")
    del zeros


@pytest.mark.parametrize("n,cols,planes", [(300, 8, 1), (300, 16, 3), (450, 8, 3)])
def test_wide_tiles_fp64_pass_matches_oracle(n, cols, planes):
    """n > 256 (more than 16 row tiles: four per wave, fragments from L2) at
    two and four rows per lane: the FP64 pass equals the oracle started from
    the mixed pass's own output, and the mixed pass reaches the same
    objective values as the one-row-per-lane shape (FP32 sums in another
    order; values agree to the bf16 product error)."""
    from types import SimpleNamespace
    from oracle import pyoracle as orc
    N, budget = 60, 300
    Q, c, _ = instances.binary_qp(n, seed=41)
    obj, keep = dense(Q, c)
    mq = gb.MixedQ(obj, planes)
    LB, UB, eps = node_set(n, N, 9)
    opt = SimpleNamespace(row_tiles_per_wave=4, columns=cols, fp64_max_iters=budget // 16)
    say(f"[mixed] wide n={n} cols={cols} planes={planes}")
    p1 = run_mixed_opts(mq, obj, LB, UB, eps, budget, -1, opt)
    assert np.all(p1["x"] >= LB) and np.all(p1["x"] <= UB) and np.all(p1["iterations"] >= 1)
    full = run_mixed_opts(mq, obj, LB, UB, eps, budget, opt.fp64_max_iters, opt)
    o = orc.relax_fast_batch(orc.dense_objective(Q, c, 2.0, 0.5, 0.0), True, p1["x"].copy(), LB.copy(),
                             UB.copy(), eps.copy(), p1["iterations"].astype(np.int32).copy(), p1["status"].copy(),
                             flags=np.ones(n, np.uint8), max_iters=opt.fp64_max_iters, refresh=64, threads=8,
                             best_vertex=p1["best_vertex"].copy(), best_vertex_val=p1["best_vertex_val"].copy())
    assert_bit_equal(full, o, "wide tiles")
    ref = run_mixed_opts(mq, obj, LB, UB, eps, budget, -1, SimpleNamespace(row_tiles_per_wave=4, columns=4,
                                                                           fp64_max_iters=-1))

    def f(X):
        return np.einsum("ij,jk,ik->i", X, Q, X) + X @ c

    fw, fr = f(p1["x"]), f(ref["x"])
    dev_rel = float(np.max(np.abs(fw - fr) / (1.0 + np.abs(fr))))
    conv = np.isin(full["status"], (1, 4))
    say(f"[mixed] wide converged {conv.sum()}/{N}, objective vs one row per lane {dev_rel:.2e}")
    assert dev_rel < 1e-3
    if planes == 3:  # one bf16 plane limits the reachable gap at this budget
        assert conv.sum() > 0
    mq.close()


def run_mixed_opts(mq, obj, LB, UB, eps, budget, fp64_iters, opt):
    import torch
    N, n = LB.shape
    x = dev(LB.copy())
    it = dev(np.zeros(N, np.int32))
    st = dev(np.zeros(N, np.uint8))
    gap = torch.zeros(N, dtype=torch.float64, device="cuda")
    val = torch.zeros(N, dtype=torch.float64, device="cuda")
    bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
    gb.relax_batch_mixed(mq, obj, gb.algorithm(dicg=True), gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), N, budget, x,
                         dev(LB), dev(UB), dev(eps), it, st, gap, val, bv, bvv, dev(np.ones(n, np.uint8)), 1e-6,
                         None, fp64_max_iters=fp64_iters, row_tiles_per_wave=opt.row_tiles_per_wave,
                         columns=opt.columns)
    torch.cuda.synchronize()
    return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
            "iterations": it.cpu().numpy(), "status": st.cpu().numpy(), "best_vertex": bv.cpu().numpy(),
            "best_vertex_val": bvv.cpu().numpy()}


@pytest.mark.parametrize("seed", [1, 2])
def test_autotuned_bnb_reaches_the_same_optimum(seed):
    n = 60
    Q, c, _ = instances.binary_qp(n, seed=seed)
    obj, keep = dense(Q, c)
    res = {}
    for tune in (0, 1):
        s = gb.default_settings()
        s.abs_gap_tol = 1e-6
        s.rel_gap_tol = 1e-6
        s.form = 2
        s.precision = gb.PREC_BF16
        s.autotune = tune
        s.time_limit_seconds = 60.0
        res[tune] = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
        say(f"[mixed] autotune={tune} seed={seed}: {res[tune]['objective']:.12f} nodes "
            f"{res[tune]['nodes_processed']} {res[tune]['elapsed_seconds']:.4f}s")
    assert not res[1]["reached_time_limit"]
    assert abs(res[0]["objective"] - res[1]["objective"]) <= 1e-9 * max(1.0, abs(res[0]["objective"]))
