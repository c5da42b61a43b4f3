"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

OBJ_DENSE, OBJ_QAP = 0, 1
LMO_BOX, LMO_SIMPLEX, LMO_BIRKHOFF, LMO_PORTFOLIO = 0, 1, 2, 3
LS_OPEN_LOOP, LS_SCALED, LS_SAFEGUARDED, LS_SECANT, LS_QUADRATIC_EXACT = 0, 1, 2, 3, 4

_dp = ctypes.POINTER(ctypes.c_double)


class Objective(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("Qt", ctypes.c_void_p),
                ("linear", ctypes.c_void_p), ("grad_scale", ctypes.c_double),
                ("linear_scale", ctypes.c_double), ("constant", ctypes.c_double),
                ("order", ctypes.c_int), ("A", ctypes.c_void_p), ("B", ctypes.c_void_p)]


class Lmo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("right_side", ctypes.c_double),
                ("equality", ctypes.c_int), ("is_int", ctypes.c_void_p), ("order", ctypes.c_int),
                ("assets", ctypes.c_int), ("cap", ctypes.c_double), ("card", ctypes.c_double)]


class Alg(ctypes.Structure):
    _fields_ = [("dicg", ctypes.c_int), ("ls", ctypes.c_int), ("secant_iters", ctypes.c_int),
                ("secant_tol", ctypes.c_double), ("secant_denom_eps", ctypes.c_double),
                ("heuristic", ctypes.c_int), ("heuristic_offset", ctypes.c_double)]


class BnbSettings(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("eps", ctypes.c_double), ("eps_decay", ctypes.c_double),
                ("eps_min", ctypes.c_double), ("integer_tol", ctypes.c_double),
                ("abs_gap_tol", ctypes.c_double), ("rel_gap_tol", ctypes.c_double),
                ("branch_k", ctypes.c_int), ("lanes", ctypes.c_int), ("threads", ctypes.c_int),
                ("natural_lower_bound", ctypes.c_double),
                ("time_limit_seconds", ctypes.c_double), ("max_nodes", ctypes.c_longlong),
                ("form", ctypes.c_int), ("refresh", ctypes.c_int), ("num_streams", ctypes.c_int)]


class BnbResult(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_double), ("dual_bound", ctypes.c_double),
                ("nodes_processed", ctypes.c_int), ("reached_tolerance", ctypes.c_int),
                ("reached_time_limit", ctypes.c_int), ("fw_iterations", ctypes.c_longlong),
                ("elapsed_seconds", ctypes.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, res in [("orc_value", ctypes.c_double), ("orc_max_step", ctypes.c_double),
                          ("orc_block_sum", ctypes.c_double),
                          ("orc_quadratic_exact_step", ctypes.c_double),
                          ("orc_certified_fw_lower_bound", ctypes.c_double),
                          ("orc_gap_absolute", ctypes.c_double),
                          ("orc_gap_relative", ctypes.c_double),
                          ("orc_depth_dependent_tolerance", ctypes.c_double),
                          ("orc_box_dicg_maximum_step_cpu", ctypes.c_double),
                          ("orc_unit_simplex_dicg_maximum_step_cpu", ctypes.c_double),
                          ("orc_enumerate_binary_qp", ctypes.c_double),
                          ("orc_secant_search", ctypes.c_double)]:
            getattr(_lib, name).restype = res
        _lib.orc_quadratic_exact_step.argtypes = [ctypes.c_double] * 3
        _lib.orc_certified_fw_lower_bound.argtypes = [ctypes.c_double] * 3
        _lib.orc_gap_absolute.argtypes = [ctypes.c_double] * 2
        _lib.orc_gap_relative.argtypes = [ctypes.c_double] * 2
        _lib.orc_gap_within_tolerance.argtypes = [ctypes.c_double] * 4
        _lib.orc_depth_dependent_tolerance.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int]
    return _lib


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def p(a):
    """Address of a numpy array's data for a ctypes call.  The pointer object
    keeps the array alive, so a temporary array (p(f64(-g))) stays valid for
    the whole call."""
    ptr = ctypes.c_void_p(0 if a is None else a.ctypes.data)
    ptr._keep = a
    return ptr


def dense_objective(Q, c=None, grad_scale=2.0, linear_scale=0.5, constant=0.0, keep=None):
    """GlobalCachedDenseQuadraticObjective: Qt = Q^T column layout."""
    Q = f64(Q)
    n = Q.shape[0]
    Qt = f64(Q.T.copy())  # Qt[k*n+j] = Q[j,k]  ->  row k of Qt holds column k of Q
    # Qt as a flat array with Qt[k*n + j] == Q[j, k] is Q transposed, row-major.
    cc = None if c is None else f64(c)
    o = Objective(OBJ_DENSE, n, p(Qt), p(cc), grad_scale, linear_scale, constant, 0, None, None)
    o._keep = (Qt, cc)
    return o


def qap_objective(A, B):
    A, B = f64(A), f64(B)
    m = A.shape[0]
    o = Objective(OBJ_QAP, m * m, None, None, 1.0, 1.0, 0.0, m, p(A), p(B))
    o._keep = (A, B)
    return o


def box_lmo():
    return Lmo(LMO_BOX, 1.0, 0, None, 0, 0, 0.0, 0.0)


def simplex_lmo(rhs, is_int, equality=False):
    ii = np.ascontiguousarray(is_int, dtype=np.uint8)
    l = Lmo(LMO_SIMPLEX, float(rhs), 1 if equality else 0, p(ii), 0, 0, 0.0, 0.0)
    l._keep = ii
    return l


def birkhoff_lmo(order):
    return Lmo(LMO_BIRKHOFF, 1.0, 0, None, int(order), 0, 0.0, 0.0)


def portfolio_lmo(assets, cap, card, rhs=1.0):
    """Cardinality portfolio: v = (x, z), sum x = rhs, 0 <= x_i <= cap z_i, sum z <= card."""
    return Lmo(LMO_PORTFOLIO, float(rhs), 0, None, 0, int(assets), float(cap), float(card))


def alg(dicg=True, ls=LS_SECANT, secant_iters=4, secant_tol=1e-10, secant_denom_eps=1e-14,
        heuristic=0, heuristic_offset=0.0):
    return Alg(1 if dicg else 0, ls, secant_iters, secant_tol, secant_denom_eps, heuristic,
               heuristic_offset)


def relax_batch(obj, a, lmo, x, lb, ub, eps, iterations, status, flags=None, max_iters=100,
                candidate_tol=1e-6, lanes=64, threads=1):
    """Runs orc_relax_batch on numpy arrays (x, iterations, status updated in place)."""
    lib = load()
    N, n = x.shape
    gap = np.zeros(N)
    value = np.zeros(N)
    bv = np.zeros((N, n))
    bvv = np.full(N, 1e300)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    lib.orc_relax_batch(ctypes.byref(obj), ctypes.byref(a), ctypes.byref(lmo), N, max_iters,
                        p(x), p(lb), p(ub), p(eps), p(iterations), p(status), p(gap), p(value),
                        p(bv), p(bvv), p(fl), ctypes.c_double(candidate_tol), lanes, threads)
    return {"x": x, "gap": gap, "value": value, "iterations": iterations, "status": status,
            "best_vertex": bv, "best_vertex_val": bvv}


def bnb_solve(obj, a, lmo, lb, ub, int_vars, batch_size=1024, max_iters=20000, eps=1e-3,
              eps_decay=0.8, eps_min=1e-6, integer_tol=1e-6, abs_gap_tol=1e-6, rel_gap_tol=1e-2,
              branch_k=3, lanes=64, threads=1, natural_lower_bound=float("-inf"),
              time_limit_seconds=float("inf"), max_nodes=0, form=0, refresh=64, num_streams=8):
    lib = load()
    n = len(lb)
    lb_a, ub_a = f64(lb), f64(ub)
    iv = np.ascontiguousarray(int_vars, dtype=np.int32)
    s = BnbSettings(batch_size, max_iters, eps, eps_decay, eps_min, integer_tol, abs_gap_tol,
                    rel_gap_tol, branch_k, lanes, threads, natural_lower_bound,
                    time_limit_seconds, max_nodes, form, refresh, num_streams)
    r = BnbResult()
    xo = np.zeros(n)
    lib.orc_bnb_solve(ctypes.byref(obj), ctypes.byref(a), ctypes.byref(lmo), n, p(lb_a), p(ub_a),
                      p(iv), len(iv), ctypes.byref(s), ctypes.byref(r), p(xo))
    return {"objective": r.objective, "x": xo, "dual_bound": r.dual_bound,
            "nodes_processed": r.nodes_processed, "reached_tolerance": bool(r.reached_tolerance),
            "reached_time_limit": bool(r.reached_time_limit), "fw_iterations": r.fw_iterations,
            "elapsed_seconds": r.elapsed_seconds}


def enumerate_binary_qp(Q, c, threads=8):
    lib = load()
    Q, c = f64(Q), f64(c)
    n = len(c)
    xo = np.zeros(n)
    f = lib.orc_enumerate_binary_qp(n, p(Q), p(c), threads, p(xo))
    return f, xo


def relax_fast_batch(obj, dicg, x, lb, ub, eps, iterations, status, flags=None, max_iters=100,
                     candidate_tol=1e-6, cutoff=None, refresh=64, threads=1, best_vertex=None,
                     best_vertex_val=None, lmo=None):
    """orc_relax_fast_batch_lmo on numpy arrays (mirror of the batched-form
    kernel; lmo None or a box: the box form, a simplex_lmo: the 0/1 simplex).
    best_vertex / best_vertex_val given: candidates of an earlier phase,
    re-scored first (the FP64 pass after the mixed-precision pass)."""
    lib = load()
    N, n = x.shape
    gap = np.zeros(N)
    value = np.zeros(N)
    rescore = best_vertex_val is not None
    bv = np.zeros((N, n)) if best_vertex is None else np.ascontiguousarray(best_vertex, dtype=np.float64)
    bvv = np.full(N, 1e300) if best_vertex_val is None else np.ascontiguousarray(best_vertex_val, dtype=np.float64)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    cut = None if cutoff is None else f64(cutoff)
    lib.orc_relax_fast_batch_lmo(ctypes.byref(obj), None if lmo is None else ctypes.byref(lmo), int(dicg), N,
                                 max_iters, p(x), p(lb), p(ub), p(eps), p(iterations), p(status), p(gap),
                                 p(value), p(bv), p(bvv), p(fl), ctypes.c_double(candidate_tol), p(cut),
                                 refresh, int(rescore), threads)
    return {"x": x, "gap": gap, "value": value, "iterations": iterations, "status": status,
            "best_vertex": bv, "best_vertex_val": bvv}


def secant_search(obj, a, x, d, gap, max_gamma, lanes=64):
    """orc_secant_search: secant step along d from x (algorithms.cuh:121-258)."""
    lib = load()
    x, d = f64(x), f64(d)
    return lib.orc_secant_search(ctypes.byref(obj), ctypes.byref(a), len(x), lanes, ctypes.c_double(gap),
                                 ctypes.c_double(max_gamma), p(x), p(d))


def least_squares_instance(seed, m, n):
    """Qt, linear, constant of the reference's convergence-matrix instance."""
    lib = load()
    Qt = np.zeros(n * n)
    lin = np.zeros(n)
    k = ctypes.c_double()
    lib.orc_least_squares_instance(ctypes.c_uint(seed), m, n, p(Qt), p(lin), ctypes.byref(k))
    return Qt.reshape(n, n), lin, k.value
