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


def _settings(time_limit=60.0, max_nodes=0):
    s = gb.default_settings()
    s.abs_gap_tol = 1e-6
    s.rel_gap_tol = 1e-6
    s.form = 1
    s.time_limit_seconds = time_limit
    s.max_nodes = max_nodes
    return s


def generate(n, seeds, labeled_nodes=16, capacity_per_instance=4096, bnb_node_cap=5000):
    """Strong-branching rows for binary QPs of size n (one instance per seed)."""
    Xs, Ts = [], []
    for seed in seeds:
        Q, c, _ = instances.binary_qp(n, seed=seed)
        obj, keep = _objective(Q, c)
        F, L = gb.branching_dataset(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0),
                                    _settings(max_nodes=bnb_node_cap), labeled_nodes, capacity_per_instance)
        Xs.append(F)
        Ts.append(L)
    return np.concatenate(Xs), np.concatenate(Ts)


def fit(X, T, epochs=30, batch=1024, lr=2e-3, seed=0):
    """Trains a scorer on (X, T) with the library's GPU kernels."""
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
    loss = b.train(dX, dT, epochs, batch, lr)
    return b, {"initial_loss": loss0, "final_loss": loss, "rows": int(len(T)), "mu": mu.tolist(),
               "sigma": sigma.tolist()}


def solve_with(Q, c, rule, brancher=None, time_limit=120.0):
    n = Q.shape[0]
    obj, keep = _objective(Q, c)
    s = _settings(time_limit)
    s.branch_rule = rule
    if brancher is not None:
        s.learned_model = brancher.handle
    t0 = time.perf_counter()
    r = gb.solve(n, [0] * n, [1] * n, range(n), obj, gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0), s)
    r["wall"] = time.perf_counter() - t0
    return r


def evaluate(brancher, n, seeds, time_limit=120.0):
    """Node count and time of learned vs most-fractional branching on the
    same instances, both to the proven optimum (gap 1e-6)."""
    rows = []
    for seed in seeds:
        Q, c, sha = instances.binary_qp(n, seed=seed)
        mf = solve_with(Q, c, gb.BRANCH_MOST_FRACTIONAL, None, time_limit)
        ln = solve_with(Q, c, gb.BRANCH_LEARNED, brancher, time_limit)
        rows.append({"seed": seed, "sha256": sha[:16], "mf_nodes": mf["nodes_processed"],
                     "learned_nodes": ln["nodes_processed"], "mf_seconds": mf["wall"],
                     "learned_seconds": ln["wall"], "mf_objective": mf["objective"],
                     "learned_objective": ln["objective"],
                     "same_optimum": abs(mf["objective"] - ln["objective"]) <= 1e-9 * max(1.0, abs(mf["objective"])),
                     "proven": not (mf["reached_time_limit"] or ln["reached_time_limit"])})
    return rows
