"""Learned branching: strong-branching labels from the solver, training of
the on-GPU MLP scorer, and its evaluation against most-fractional
branching at the same proven optimum.

All device work runs through the C-ABI: labels come from
gb_branching_dataset (batched child relaxations on the GPU), training from
gb_brancher_train (the library's forward / backward / Adam kernels), and the
scorer is used by gb_solve with branch_rule = GB_BRANCH_LEARNED.
"""
from __future__ import annotations

import time

import numpy as np

import gpuboscia_amd as gb
from gpuboscia_amd import instances


def _objective(Q, c):
    import torch
    dQt = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
    dc = torch.from_numpy(np.ascontiguousarray(c)).cuda()
    obj = gb.ObjectiveDesc(gb.OBJ_DENSE, Q.shape[0], dQt.data_ptr(), dc.data_ptr(), 2.0, 0.5, 0.0, 0, 0, 0)
    return obj, (dQt, dc)


def _settings(time_limit=60.0, max_nodes=0, form=1):
    s = gb.default_settings()
    s.abs_gap_tol = 1e-6
    s.rel_gap_tol = 1e-6
    s.form = form
    if form == 2:
        s.precision = gb.PREC_BF16
    s.time_limit_seconds = time_limit
    s.max_nodes = max_nodes
    return s


def generate(n, seeds, labeled_nodes=16, capacity_per_instance=4096, bnb_node_cap=5000, progress=False):
    """Strong-branching rows for binary QPs of size n (one instance per seed)."""
    Xs, Ts = [], []
    for k, seed in enumerate(seeds):
        if progress and k % 256 == 0:
            print(f"[learned] labelling instance {k}", flush=True)
        Q, c, _ = instances.binary_qp(n, seed=seed)
        obj, keep = _objective(Q, c)
        F, L = gb.branching_dataset(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0),
                                    _settings(max_nodes=bnb_node_cap), labeled_nodes, capacity_per_instance)
        Xs.append(F)
        Ts.append(L)
    return np.concatenate(Xs), np.concatenate(Ts)


def fit(X, T, epochs=30, batch=1024, lr=2e-3, seed=0, progress=False):
    """Trains a scorer on (X, T) with the library's GPU kernels (Adam state
    persists in the scorer, so training in chunks of epochs is the same as
    one call)."""
    import torch
    mu = X.mean(0)
    sigma = X.std(0)
    sigma = np.where(sigma > 1e-6, sigma, 1.0)
    b = gb.Brancher(seed)
    b.set_normalization(mu, sigma)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(T))
    dX = torch.from_numpy(np.ascontiguousarray(X[perm], dtype=np.float32)).cuda()
    dT = torch.from_numpy(np.ascontiguousarray(T[perm], dtype=np.float32)).cuda()
    loss0, _ = b.loss_grad(dX, dT)
    loss = loss0
    done = 0
    while done < epochs:
        step = min(10, epochs - done)
        loss = b.train(dX, dT, step, batch, lr)
        done += step
        if progress:
            print(f"[learned] epoch {done}: loss {loss:.5f}", flush=True)
    return b, {"initial_loss": loss0, "final_loss": loss, "rows": int(len(T)), "mu": mu.tolist(),
               "sigma": sigma.tolist()}


def solve_with(Q, c, rule, brancher=None, time_limit=120.0, form=2, max_nodes=0, branch_k=3):
    n = Q.shape[0]
    obj, keep = _objective(Q, c)
    s = _settings(time_limit, max_nodes=max_nodes, form=form)
    s.branch_rule = rule
    s.branch_k = branch_k
    if brancher is not None:
        s.learned_model = brancher.handle
    t0 = time.perf_counter()
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
    r["wall"] = time.perf_counter() - t0
    return r


BASELINES = (("mf", gb.BRANCH_MOST_FRACTIONAL), ("pow2", gb.BRANCH_POW2))


def evaluate(brancher, n, seeds, time_limit=120.0, form=2, learned_ks=(3,)):
    """Node count and time of the learned rule against the rules without it,
    on the same instances, all to the proven optimum (gap 1e-6): "mf" is
    most-fractional branching (one variable, two children; the comparison
    the learned-branching configuration names) and "pow2" the reference
    default (branch_pow2 on the 3 most fractional variables, the learned
    rule's own tree shape).  The learned rule splits on its branch_k
    best-scored candidates: key "learned" for k = 3, "learned_k<k>" for the
    other k in learned_ks."""
    rows = []
    for seed in seeds:
        Q, c, sha = instances.binary_qp(n, seed=seed)
        row = {"seed": seed, "sha256": sha[:16]}
        runs = {key: solve_with(Q, c, rule, None, time_limit, form) for key, rule in BASELINES}
        for k in learned_ks:
            runs["learned" if k == 3 else f"learned_k{k}"] = solve_with(Q, c, gb.BRANCH_LEARNED, brancher,
                                                                        time_limit, form, branch_k=k)
        if "learned" not in runs:
            runs["learned"] = solve_with(Q, c, gb.BRANCH_LEARNED, brancher, time_limit, form)
        for key, r in runs.items():
            row[f"{key}_nodes"] = r["nodes_processed"]
            row[f"{key}_seconds"] = r["wall"]
            row[f"{key}_objective"] = r["objective"]
        ref = runs["learned"]["objective"]
        row["same_optimum"] = all(abs(r["objective"] - ref) <= 1e-9 * max(1.0, abs(ref)) for r in runs.values())
        row["proven"] = not any(r["reached_time_limit"] for r in runs.values())
        rows.append(row)
    return rows


def evaluate_budget(brancher, n, seeds, max_nodes, time_limit=120.0, form=2):
    """For sizes whose searches do not finish: the learned rule and the rules
    without it (most-fractional "mf" and branch_pow2 "pow2") each get the same
    node budget on the same instances; the better rule ends with the higher
    global dual bound (all bounds are FP64-certified, so each is a valid
    lower bound on the optimum)."""
    rows = []
    for k, seed in enumerate(seeds):
        if k % 128 == 0:
            print(f"[learned] budget evaluation, instance {k}", flush=True)
        Q, c, sha = instances.binary_qp(n, seed=seed)
        row = {"seed": seed, "sha256": sha[:16]}
        runs = {key: solve_with(Q, c, rule, None, time_limit, form, max_nodes) for key, rule in BASELINES}
        runs["learned"] = solve_with(Q, c, gb.BRANCH_LEARNED, brancher, time_limit, form, max_nodes)
        for key, r in runs.items():
            row[f"{key}_nodes"] = r["nodes_processed"]
            row[f"{key}_seconds"] = r["wall"]
            row[f"{key}_bound"] = r["dual_bound"]
            row[f"{key}_objective"] = r["objective"]
        rows.append(row)
    return rows


def budget_summary(rows, baseline="mf"):
    """Bound comparison at equal node budgets against one baseline rule: the
    gap of each rule is measured against the better incumbent of the two
    runs."""
    mb = np.array([r[f"{baseline}_bound"] for r in rows])
    lb = np.array([r["learned_bound"] for r in rows])
    inc = np.minimum([r[f"{baseline}_objective"] for r in rows], [r["learned_objective"] for r in rows])
    gm, gl = inc - mb, inc - lb
    ok = (gm > 0) & (gl > 0)
    return {"baseline": baseline, "instances": len(rows), "learned_bound_higher": int((lb > mb).sum()),
            "learned_bound_lower": int((lb < mb).sum()), "mean_bound_gain": float(np.mean(lb - mb)),
            "gap_ratio_geomean": float(np.exp(np.mean(np.log(gl[ok] / gm[ok])))) if ok.any() else None,
            "baseline_seconds": float(sum(r[f"{baseline}_seconds"] for r in rows)),
            "learned_seconds": float(sum(r["learned_seconds"] for r in rows))}


def train_main(argv=None):
    """Trains a scorer on solver-generated strong-branching labels and writes it
    (weights + feature normalization, plain arrays) to --out:

      python -m gpuboscia_amd.learned --n 40 --train 3072 --out scorer.npz

    Sizes whose searches cannot be proven are compared at a node budget:

      python -m gpuboscia_amd.learned --n 200 --train 0 --scorer scorer.npz --test 1024 --node-budget 20000
    """
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=40)
    ap.add_argument("--train", type=int, default=3072, help="training instances (seeds 1000..)")
    ap.add_argument("--test", type=int, default=64, help="test instances (seeds 50000..)")
    ap.add_argument("--labeled-nodes", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--node-budget", type=int, default=0,
                    help="evaluate at this node budget per search (for sizes that cannot be proven)")
    ap.add_argument("--scorer", help="evaluate this saved scorer instead of training one (--train 0)")
    ap.add_argument("--out", help="where to write the trained scorer")
    a = ap.parse_args(argv)
    if a.train > 0 and not a.out:
        ap.error("--out is required when training")
    if a.train == 0 and not a.scorer:
        ap.error("--scorer is required with --train 0")
    if a.train == 0:
        b = gb.Brancher.from_file(a.scorer)
        if a.node_budget > 0:
            rows = evaluate_budget(b, a.n, range(50000, 50000 + a.test), a.node_budget)
            summary = {"n": a.n, "node_budget": a.node_budget, "vs_mf": budget_summary(rows, "mf"),
                       "vs_pow2": budget_summary(rows, "pow2")}
            print(json.dumps(summary), flush=True)
            return summary
        rows = evaluate(b, a.n, seeds=range(50000, 50000 + a.test))
        info = {"rows": 0, "initial_loss": None, "final_loss": None}
        t0 = t1 = t2 = time.perf_counter()
    else:
        t0 = time.perf_counter()
        X, T = generate(a.n, seeds=range(1000, 1000 + a.train), labeled_nodes=a.labeled_nodes, progress=True)
        t1 = time.perf_counter()
        b, info = fit(X, T, epochs=a.epochs, lr=a.lr, progress=True)
        t2 = time.perf_counter()
        b.save(a.out)
        if a.node_budget > 0 and a.test > 0:
            rows = evaluate_budget(b, a.n, range(50000, 50000 + a.test), a.node_budget)
            summary = {"n": a.n, "train_instances": a.train, "rows": info["rows"],
                       "label_seconds": round(t1 - t0, 2), "fit_seconds": round(t2 - t1, 2),
                       "final_loss": info["final_loss"], "node_budget": a.node_budget,
                       "vs_mf": budget_summary(rows, "mf"), "vs_pow2": budget_summary(rows, "pow2")}
            print(json.dumps(summary), flush=True)
            return summary
        rows = evaluate(b, a.n, seeds=range(50000, 50000 + a.test)) if a.test > 0 else []
    summary = {"n": a.n, "train_instances": a.train, "rows": info["rows"], "label_seconds": round(t1 - t0, 2),
               "fit_seconds": round(t2 - t1, 2), "initial_loss": info["initial_loss"],
               "final_loss": info["final_loss"], "test_instances": len(rows)}
    if rows:
        ln = np.array([r["learned_nodes"] for r in rows], dtype=float)
        summary.update({"learned_nodes": int(ln.sum()),
                        "learned_seconds": float(sum(r["learned_seconds"] for r in rows)),
                        "same_optimum": all(r["same_optimum"] for r in rows),
                        "proven": all(r["proven"] for r in rows)})
        for key, _ in BASELINES:
            bn = np.array([r[f"{key}_nodes"] for r in rows], dtype=float)
            summary[f"vs_{key}"] = {"nodes": int(bn.sum()), "node_ratio": float(ln.sum() / bn.sum()),
                                    "node_ratio_geomean": float(np.exp(np.mean(np.log(ln / bn)))),
                                    "seconds": float(sum(r[f"{key}_seconds"] for r in rows))}
    print(json.dumps(summary), flush=True)
    return summary


if __name__ == "__main__":
    train_main()
