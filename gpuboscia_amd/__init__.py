"""gpuboscia_amd: MI355X (gfx950) Frank-Wolfe / DICG node relaxation.

The product is the C-ABI library ``libgpuboscia_amd.so`` (declared in
``include/gpuboscia_amd.h``) and the C++ drop-in header
``include/gpuboscia/gpuboscia.hpp``.  This module is a thin ctypes binding
used by the tests and the benchmark; it never falls back to any CPU path and
raises if the HIP library is missing.
"""
from __future__ import annotations

import ctypes
import math
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpuboscia_amd.so")

GB_OK, GB_ERR_INVALID_ARGUMENT, GB_ERR_RUNTIME, GB_ERR_DEVICE = 0, 1, 2, 3
ALG_FRANK_WOLFE, ALG_DICG = 0, 1
LS_OPEN_LOOP, LS_SCALED, LS_SAFEGUARDED, LS_SECANT, LS_QUADRATIC_EXACT = 0, 1, 2, 3, 4
OBJ_DENSE, OBJ_QAP = 0, 1
LMO_BOX, LMO_SIMPLEX, LMO_BIRKHOFF, LMO_PORTFOLIO = 0, 1, 2, 3
HEUR_ROUND, HEUR_FLOOR_OFFSET = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
BRANCH_POW2, BRANCH_MOST_FRACTIONAL, BRANCH_LEARNED, BRANCH_CALLBACK, BRANCH_CUSTOM = 0, 1, 2, 3, 4
NODE_INFEASIBLE = 2
NODE_INITIAL_POINT_SUPPLIED = 0x80

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class ObjectiveDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("Qt", ctypes.c_void_p),
                ("linear", ctypes.c_void_p), ("grad_scale", ctypes.c_double),
                ("linear_scale", ctypes.c_double), ("constant", ctypes.c_double),
                ("order", ctypes.c_int), ("A", ctypes.c_void_p), ("B", ctypes.c_void_p)]


class LmoDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("right_side", ctypes.c_double),
                ("equality", ctypes.c_int), ("is_int", ctypes.c_void_p), ("order", ctypes.c_int),
                ("cap", ctypes.c_double), ("cardinality", ctypes.c_double)]


class AlgorithmDesc(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int), ("line_search", ctypes.c_int),
                ("secant_iters", ctypes.c_int), ("secant_tol", ctypes.c_double),
                ("secant_denom_eps", ctypes.c_double), ("heuristic", ctypes.c_int),
                ("heuristic_offset", ctypes.c_double)]


class FastOptions(ctypes.Structure):
    _fields_ = [("refresh_interval", ctypes.c_int), ("waves_per_block", ctypes.c_int),
                ("blocks", ctypes.c_int), ("stats", ctypes.c_void_p), ("gradient", ctypes.c_void_p)]


class MixedOptions(ctypes.Structure):
    _fields_ = [("row_tiles_per_wave", ctypes.c_int), ("columns", ctypes.c_int), ("blocks", ctypes.c_int),
                ("rel_floor", ctypes.c_float),
                ("fp64_max_iters", ctypes.c_int), ("mixed_iterations", ctypes.c_void_p),
                ("fp64_stats", ctypes.c_void_p), ("gradient", ctypes.c_void_p)]


class InstanceBatch(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_void_p), ("lmo", ctypes.c_void_p), ("mixed", ctypes.c_void_p),
                ("N", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in (
                    "x", "lb", "ub", "eps", "iterations", "status", "gap", "value", "best_vertex",
                    "best_vertex_val", "integer_flags", "cutoff")]


class Problem(ctypes.Structure):
    _fields_ = [("n_vars", ctypes.c_int), ("lb", _dp), ("ub", _dp), ("int_vars", _ip),
                ("n_int_vars", ctypes.c_int)]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int)
VALIDATOR_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int,
                                ctypes.c_double)
BRANCH_SCORE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, ctypes.c_int, _ip, ctypes.c_int,
                                   ctypes.c_int, _dp)


class Node(ctypes.Structure):
    """gb_node (node.h:9-33): host arrays of n_vars doubles."""
    _fields_ = [("id", ctypes.c_int), ("parent_id", ctypes.c_int), ("depth", ctypes.c_int),
                ("lb", _dp), ("ub", _dp), ("relaxation_x", _dp), ("relaxation_iterations", ctypes.c_int),
                ("dual_bound", ctypes.c_double)]


BRANCH_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Node), _ip, _dp, ctypes.c_int, _ip,
                             ctypes.c_int, ctypes.c_void_p)


class SchedulingEvent(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("slot_id", ctypes.c_int), ("generation", ctypes.c_ulonglong),
                ("n_nodes", ctypes.c_int)]


SCHEDULING_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(SchedulingEvent))
SCHED_DISPATCHED, SCHED_COMPLETION_DETACHED, SCHED_POSTPROCESS_STARTED = 0, 1, 2


class SolverStats(ctypes.Structure):
    _fields_ = [("incumbent_value", ctypes.c_double), ("dual_bound", ctypes.c_double),
                ("nodes_processed", ctypes.c_int), ("last_gpu_iterations", ctypes.c_int),
                ("last_gpu_gap", ctypes.c_double), ("last_gpu_converged", ctypes.c_int),
                ("reached_tolerance", ctypes.c_int), ("reached_time_limit", ctypes.c_int),
                ("effective_gpu_batch_size", ctypes.c_int), ("effective_num_streams", ctypes.c_int),
                ("paired_gpu_batches", ctypes.c_int), ("fw_iterations", ctypes.c_longlong),
                ("elapsed_seconds", ctypes.c_double), ("relax_seconds", ctypes.c_double)]


class Settings(ctypes.Structure):
    _fields_ = [("gpu_batch_size", ctypes.c_int), ("num_streams", ctypes.c_int),
                ("gpu_block_threads", ctypes.c_int), ("gpu_max_iters", ctypes.c_int),
                ("gpu_eps", ctypes.c_double), ("gpu_eps_decay", ctypes.c_double),
                ("gpu_eps_min", ctypes.c_double), ("integer_tol", ctypes.c_double),
                ("abs_gap_tol", ctypes.c_double), ("rel_gap_tol", ctypes.c_double),
                ("algorithm", AlgorithmDesc), ("branch_rule", ctypes.c_int),
                ("branch_k", ctypes.c_int), ("time_limit_seconds", ctypes.c_double),
                ("natural_lower_bound", ctypes.c_double), ("root_initial_point", _dp),
                ("root_initial_point_size", ctypes.c_int),
                ("root_validator", VALIDATOR_FN), ("root_validator_user", ctypes.c_void_p),
                ("progress", PROGRESS_FN), ("progress_user", ctypes.c_void_p),
                ("branch_score", BRANCH_SCORE_FN), ("branch_score_user", ctypes.c_void_p),
                ("branch", BRANCH_FN), ("branch_user", ctypes.c_void_p),
                ("precision", ctypes.c_int), ("form", ctypes.c_int), ("learned_model", ctypes.c_void_p),
                ("max_nodes", ctypes.c_longlong), ("autotune", ctypes.c_int),
                ("device_frontier", ctypes.c_int), ("device_nodes", ctypes.c_longlong),
                ("blocks_per_node", ctypes.c_int), ("plugin", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_double), ("x", _dp), ("dual_bound", ctypes.c_double),
                ("nodes_processed", ctypes.c_int), ("elapsed_seconds", ctypes.c_double),
                ("reached_tolerance", ctypes.c_int), ("reached_time_limit", ctypes.c_int),
                ("fw_iterations", ctypes.c_longlong), ("relax_seconds", ctypes.c_double)]


_lib = None

#: Every symbol that include/gpuboscia_amd.h declares.
EXPORTED = ["gb_relax_batch", "gb_relax_batch_fast", "gb_default_settings", "gb_solve",
            "gb_lmo_relaxation_feasible_cpu", "gb_last_error", "gb_version", "gb_device_count",
            "gb_brancher_create", "gb_brancher_destroy", "gb_brancher_set_params",
            "gb_brancher_get_params", "gb_brancher_set_normalization", "gb_brancher_predict",
            "gb_brancher_loss_grad", "gb_brancher_train", "gb_branching_dataset",
            "gb_grid_barrier_selftest", "gb_relax_graph_create", "gb_relax_graph_launch",
            "gb_relax_graph_destroy", "gb_mixed_prepare", "gb_mixed_destroy", "gb_mixed_products",
            "gb_relax_batch_mixed", "gb_mixed_autotune", "gb_relax_instances", "gb_shutdown",
            "gb_relax_batch_noalloc", "gb_relax_batch_multiblock", "gb_lmo_compute_v", "gb_mixed_warm_products",
            "gb_children_push", "gb_solver_create", "gb_solver_destroy", "gb_solver_push",
            "gb_solver_set_scheduling_callback", "gb_solver_set_progress_callback", "gb_solver_solve",
            "gb_solver_get_stats", "gb_solver_incumbent_x", "gb_candidate_features",
            "gb_relax_batch_plugin", "gb_qap32_prepare", "gb_qap32_destroy", "gb_relax_batch_qap32",
            "gb_qap32_assignments"]
MLP_FEATURES = 8
MLP_PARAMS = 8961


class GpuBosciaError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class InvalidArgument(GpuBosciaError, ValueError):
    """Mirrors std::invalid_argument in the reference."""


def load():
    """Load the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    # torch bundles its own libamdhip64.so.7 (same soname as /opt/rocm's).
    # Whichever is loaded first serves the whole process; load torch's first
    # so that torch tensors and this library share one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    lib.gb_relax_batch.restype = ctypes.c_int
    lib.gb_relax_batch.argtypes = [ctypes.POINTER(ObjectiveDesc), ctypes.POINTER(AlgorithmDesc),
                                   ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] \
        + [ctypes.c_void_p] * 11 + [ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    lib.gb_relax_batch_noalloc.restype = ctypes.c_int
    lib.gb_relax_batch_noalloc.argtypes = lib.gb_relax_batch.argtypes
    lib.gb_relax_batch_multiblock.restype = ctypes.c_int
    lib.gb_relax_batch_multiblock.argtypes = lib.gb_relax_batch.argtypes[:-1] + [ctypes.c_int, ctypes.c_void_p]
    lib.gb_lmo_compute_v.restype = ctypes.c_int
    lib.gb_lmo_compute_v.argtypes = [ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    lib.gb_mixed_warm_products.restype = ctypes.c_int
    lib.gb_mixed_warm_products.argtypes = [ctypes.POINTER(ObjectiveDesc), ctypes.c_int] + [ctypes.c_void_p] * 5
    lib.gb_qap32_prepare.restype = ctypes.c_int
    lib.gb_qap32_prepare.argtypes = [ctypes.POINTER(ObjectiveDesc), ctypes.c_void_p, ctypes.c_void_p]
    lib.gb_qap32_destroy.restype = ctypes.c_int
    lib.gb_qap32_destroy.argtypes = [ctypes.c_void_p]
    lib.gb_relax_batch_qap32.restype = ctypes.c_int
    lib.gb_relax_batch_qap32.argtypes = [ctypes.c_void_p, ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] \
        + [ctypes.c_void_p] * 11 + [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    lib.gb_qap32_assignments.restype = ctypes.c_int
    lib.gb_qap32_assignments.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4
    lib.gb_children_push.argtypes = [ctypes.c_void_p, ctypes.POINTER(Node)]
    lib.gb_solver_create.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(ObjectiveDesc), ctypes.POINTER(LmoDesc),
                                     ctypes.POINTER(Settings), ctypes.POINTER(ctypes.c_void_p)]
    lib.gb_solver_destroy.argtypes = [ctypes.c_void_p]
    lib.gb_solver_push.argtypes = [ctypes.c_void_p, ctypes.POINTER(Node)]
    lib.gb_solver_set_scheduling_callback.argtypes = [ctypes.c_void_p, SCHEDULING_FN, ctypes.c_void_p]
    lib.gb_solver_set_progress_callback.argtypes = [ctypes.c_void_p, PROGRESS_FN, ctypes.c_void_p]
    lib.gb_solver_solve.argtypes = [ctypes.c_void_p]
    lib.gb_solver_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(SolverStats)]
    lib.gb_solver_incumbent_x.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int)]
    for name in ("gb_children_push", "gb_solver_create", "gb_solver_destroy", "gb_solver_push",
                 "gb_solver_set_scheduling_callback", "gb_solver_set_progress_callback", "gb_solver_solve",
                 "gb_solver_get_stats", "gb_solver_incumbent_x"):
        getattr(lib, name).restype = ctypes.c_int
    lib.gb_candidate_features.restype = ctypes.c_int
    lib.gb_candidate_features.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 10 \
        + [ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    lib.gb_shutdown.restype = ctypes.c_int
    lib.gb_shutdown.argtypes = []
    lib.gb_relax_batch_fast.restype = ctypes.c_int
    lib.gb_relax_batch_fast.argtypes = [ctypes.POINTER(ObjectiveDesc), ctypes.POINTER(AlgorithmDesc),
                                        ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] \
        + [ctypes.c_void_p] * 11 + [ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(FastOptions),
                                     ctypes.c_void_p]
    lib.gb_default_settings.restype = None
    lib.gb_default_settings.argtypes = [ctypes.POINTER(Settings)]
    lib.gb_solve.restype = ctypes.c_int
    lib.gb_solve.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(ObjectiveDesc),
                             ctypes.POINTER(LmoDesc), ctypes.POINTER(Settings),
                             ctypes.POINTER(Result)]
    lib.gb_lmo_relaxation_feasible_cpu.restype = ctypes.c_int
    lib.gb_lmo_relaxation_feasible_cpu.argtypes = [ctypes.POINTER(LmoDesc), ctypes.c_void_p, _dp,
                                                   _dp, _dp, ctypes.c_int, ctypes.c_double]
    vp = ctypes.c_void_p
    lib.gb_brancher_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(vp)]
    lib.gb_brancher_destroy.argtypes = [vp]
    lib.gb_brancher_set_params.argtypes = [vp, vp, ctypes.c_int]
    lib.gb_brancher_get_params.argtypes = [vp, vp, ctypes.c_int]
    lib.gb_brancher_set_normalization.argtypes = [vp, vp, vp]
    lib.gb_brancher_predict.argtypes = [vp, vp, ctypes.c_int, vp, vp]
    lib.gb_brancher_loss_grad.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), vp]
    lib.gb_brancher_train.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.POINTER(ctypes.c_double)]
    lib.gb_branching_dataset.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(ObjectiveDesc),
                                         ctypes.POINTER(LmoDesc), ctypes.POINTER(Settings), ctypes.c_int,
                                         vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    for name in ("gb_brancher_create", "gb_brancher_destroy", "gb_brancher_set_params",
                 "gb_brancher_get_params", "gb_brancher_set_normalization", "gb_brancher_predict",
                 "gb_brancher_loss_grad", "gb_brancher_train", "gb_branching_dataset"):
        getattr(lib, name).restype = ctypes.c_int
    lib.gb_relax_graph_create.argtypes = lib.gb_relax_batch_fast.argtypes[:-1] + [ctypes.POINTER(vp)]
    lib.gb_relax_graph_launch.argtypes = [vp, vp]
    lib.gb_relax_graph_destroy.argtypes = [vp]
    for name in ("gb_relax_graph_create", "gb_relax_graph_launch", "gb_relax_graph_destroy"):
        getattr(lib, name).restype = ctypes.c_int
    lib.gb_mixed_prepare.argtypes = [ctypes.POINTER(ObjectiveDesc), ctypes.c_int, vp, ctypes.POINTER(vp)]
    lib.gb_mixed_destroy.argtypes = [vp]
    lib.gb_mixed_products.argtypes = [vp, vp, ctypes.c_int, vp, vp]
    lib.gb_relax_batch_mixed.argtypes = [vp, ctypes.POINTER(ObjectiveDesc), ctypes.POINTER(AlgorithmDesc),
                                         ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] \
        + [vp] * 11 + [ctypes.c_double, vp, ctypes.POINTER(MixedOptions), vp]
    lib.gb_mixed_autotune.argtypes = [vp, ctypes.POINTER(ObjectiveDesc), ctypes.POINTER(AlgorithmDesc),
                                      ctypes.POINTER(LmoDesc), ctypes.c_int, ctypes.c_int] + [vp] * 7 \
        + [ctypes.c_double, vp, ctypes.POINTER(MixedOptions), ctypes.POINTER(ctypes.c_double), vp]
    lib.gb_relax_instances.argtypes = [ctypes.c_int, vp, ctypes.POINTER(AlgorithmDesc), ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, vp]
    for name in ("gb_mixed_prepare", "gb_mixed_destroy", "gb_mixed_products", "gb_relax_batch_mixed",
                 "gb_mixed_autotune", "gb_relax_instances"):
        getattr(lib, name).restype = ctypes.c_int
    ll = ctypes.POINTER(ctypes.c_longlong)
    lib.gb_grid_barrier_selftest.argtypes = [ctypes.c_int] * 5 + [ctypes.c_longlong, ll, ll,
                                                                  ctypes.POINTER(ctypes.c_double)]
    lib.gb_grid_barrier_selftest.restype = ctypes.c_int
    lib.gb_last_error.restype = ctypes.c_char_p
    lib.gb_version.restype = ctypes.c_char_p
    lib.gb_device_count.restype = ctypes.c_int
    _lib = lib
    return lib


def check(code):
    if code == GB_OK:
        return
    msg = load().gb_last_error().decode()
    if code == GB_ERR_INVALID_ARGUMENT:
        raise InvalidArgument(code, msg)
    raise GpuBosciaError(code, msg)


def default_settings() -> Settings:
    s = Settings()
    load().gb_default_settings(ctypes.byref(s))
    return s


def algorithm(dicg=True, line_search=LS_SECANT, secant_iters=4, secant_tol=1e-10,
              secant_denom_eps=1e-14, heuristic=HEUR_ROUND, heuristic_offset=0.0) -> AlgorithmDesc:
    return AlgorithmDesc(ALG_DICG if dicg else ALG_FRANK_WOLFE, line_search, secant_iters,
                         secant_tol, secant_denom_eps, heuristic, heuristic_offset)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def relax_batch(obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb, ub,
                eps, iterations, status, gap, value, best_vertex, best_vertex_val, flags=None,
                candidate_tol=1e-6, block_threads=0, stream=None):
    """gb_relax_batch on torch device tensors (float64/int32/uint8)."""
    lib = load()
    code = lib.gb_relax_batch(ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                              int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                              _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                              _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                              float(candidate_tol), int(block_threads),
                              ctypes.c_void_p(stream or 0))
    check(code)


def relax_batch_noalloc(obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb, ub,
                        eps, iterations, status, gap, value, best_vertex, best_vertex_val, flags=None,
                        candidate_tol=1e-6, block_threads=0, stream=None):
    """gb_relax_batch_noalloc: gb_relax_batch with its work vectors in the
    library's static device array (no allocation)."""
    lib = load()
    code = lib.gb_relax_batch_noalloc(ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                                      int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                                      _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                                      _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                                      float(candidate_tol), int(block_threads),
                                      ctypes.c_void_p(stream or 0))
    check(code)


def relax_batch_multiblock(obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb, ub,
                           eps, iterations, status, gap, value, best_vertex, best_vertex_val, flags=None,
                           candidate_tol=1e-6, block_threads=0, blocks_per_node=4, stream=None):
    """gb_relax_batch_multiblock: gb_relax_batch with blocks_per_node
    workgroups per node (MultiBlock), bit-identical outputs."""
    lib = load()
    code = lib.gb_relax_batch_multiblock(ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                                         int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                                         _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                                         _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                                         float(candidate_tol), int(block_threads), int(blocks_per_node),
                                         ctypes.c_void_p(stream or 0))
    check(code)


def lmo_compute_v(lmo: LmoDesc, n, N, direction, lb, ub, vertex, stream=None):
    """gb_lmo_compute_v on torch float64 device tensors [N, n]."""
    check(load().gb_lmo_compute_v(ctypes.byref(lmo), int(n), int(N), _ptr(direction), _ptr(lb), _ptr(ub),
                                  _ptr(vertex), ctypes.c_void_p(stream or 0)))


def mixed_warm_products(obj: ObjectiveDesc, N, x, status, iterations, out, stream=None):
    """gb_mixed_warm_products on torch device tensors."""
    check(load().gb_mixed_warm_products(ctypes.byref(obj), int(N), _ptr(x), _ptr(status), _ptr(iterations),
                                        _ptr(out), ctypes.c_void_p(stream or 0)))


def candidate_features(n, N, X, G, LB, UB, depth, gap, value, pc_up, pc_down, Qt, gs, flags, tol, F,
                       stream=None):
    """gb_candidate_features on torch device tensors (F: float32 [N*n, 8])."""
    check(load().gb_candidate_features(int(n), int(N), _ptr(X), _ptr(G), _ptr(LB), _ptr(UB), _ptr(depth), _ptr(gap),
                                       _ptr(value), _ptr(pc_up), _ptr(pc_down), _ptr(Qt), float(gs), _ptr(flags),
                                       float(tol), _ptr(F), ctypes.c_void_p(stream or 0)))


def shutdown():
    """gb_shutdown: release the library's pooled device resources."""
    check(load().gb_shutdown())


def relax_batch_fast(obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb,
                     ub, eps, iterations, status, gap, value, best_vertex, best_vertex_val,
                     flags=None, candidate_tol=1e-6, cutoff=None, refresh=64, waves_per_block=0,
                     blocks=0, stream=None, stats=None):
    """gb_relax_batch_fast on torch device tensors (stats: optional int64 [N,2])."""
    lib = load()
    opt = FastOptions(refresh, waves_per_block, blocks, 0 if stats is None else stats.data_ptr(), 0)
    code = lib.gb_relax_batch_fast(ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                                   int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                                   _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                                   _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                                   float(candidate_tol), _ptr(cutoff), ctypes.byref(opt),
                                   ctypes.c_void_p(stream or 0))
    check(code)


class MixedQ:
    """gb_mixed_*: Q of a dense objective split into bf16 planes for the
    mixed-precision form (planes 1 or 3)."""

    def __init__(self, obj: ObjectiveDesc, planes=3, stream=None):
        lib = load()
        self._obj = obj
        h = ctypes.c_void_p()
        check(lib.gb_mixed_prepare(ctypes.byref(obj), int(planes), ctypes.c_void_p(stream or 0),
                                   ctypes.byref(h)))
        self.handle = h
        self.planes = planes

    def products(self, V, Y, stream=None):
        """Y = Q V through the MFMA fragments (torch float64 [N, n])."""
        check(load().gb_mixed_products(self.handle, _ptr(V), int(V.shape[0]), _ptr(Y),
                                       ctypes.c_void_p(stream or 0)))

    def close(self):
        if self.handle:
            load().gb_mixed_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Qap32:
    """gb_qap32_*: a QAP instance prepared for the FP32 form (FP32 copies of
    A and B, A'A and B B'); A and B stay owned by the caller."""

    def __init__(self, obj: ObjectiveDesc, stream=None):
        h = ctypes.c_void_p()
        check(load().gb_qap32_prepare(ctypes.byref(obj), ctypes.c_void_p(stream or 0), ctypes.byref(h)))
        self.handle = h

    def relax(self, lmo: LmoDesc, N, max_iters, x, lb, ub, eps, iterations, status, gap, value, best_vertex,
              best_vertex_val, flags=None, candidate_tol=1e-6, stop_value=float("-inf"), refresh=0, stream=None):
        """gb_relax_batch_qap32 on torch device tensors (gb_relax_batch's layout)."""
        check(load().gb_relax_batch_qap32(self.handle, ctypes.byref(lmo), int(N), int(max_iters), _ptr(x),
                                          _ptr(lb), _ptr(ub), _ptr(eps), _ptr(iterations), _ptr(status),
                                          _ptr(gap), _ptr(value), _ptr(best_vertex), _ptr(best_vertex_val),
                                          _ptr(flags), float(candidate_tol), float(stop_value), int(refresh),
                                          ctypes.c_void_p(stream or 0)))

    def close(self):
        if self.handle:
            load().gb_qap32_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qap32_assignments(costs, perms, bounds, stream=None):
    """gb_qap32_assignments: costs float32 [T, m, m] (+inf forbidden), perms
    int32 [T, m], bounds float64 [T] (torch device tensors); synchronous."""
    T, m = int(costs.shape[0]), int(costs.shape[1])
    check(load().gb_qap32_assignments(m, T, _ptr(costs), _ptr(perms), _ptr(bounds),
                                      ctypes.c_void_p(stream or 0)))


def relax_batch_mixed(mq: MixedQ, obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x,
                      lb, ub, eps, iterations, status, gap, value, best_vertex, best_vertex_val, flags=None,
                      candidate_tol=1e-6, cutoff=None, rel_floor=1e-6, fp64_max_iters=0, row_tiles_per_wave=0,
                      columns=0, blocks=0, mixed_iterations=None, stream=None):
    """gb_relax_batch_mixed on torch device tensors."""
    opt = MixedOptions(row_tiles_per_wave, columns, blocks, rel_floor, fp64_max_iters,
                       0 if mixed_iterations is None else mixed_iterations.data_ptr(), 0, 0)
    check(load().gb_relax_batch_mixed(mq.handle, ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo),
                                      int(N), int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                                      _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                                      _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                                      float(candidate_tol), _ptr(cutoff), ctypes.byref(opt),
                                      ctypes.c_void_p(stream or 0)))


def mixed_autotune(mq: MixedQ, obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb, ub,
                   eps, iterations, status, flags=None, candidate_tol=1e-6, cutoff=None, stream=None):
    """gb_mixed_autotune: (MixedOptions chosen, seconds of the winning full relaxation)."""
    opt = MixedOptions(0, 0, 0, 1e-6, 0, 0, 0, 0)
    sec = ctypes.c_double()
    check(load().gb_mixed_autotune(mq.handle, ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                                   int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps), _ptr(iterations),
                                   _ptr(status), _ptr(flags), float(candidate_tol), _ptr(cutoff),
                                   ctypes.byref(opt), ctypes.byref(sec), ctypes.c_void_p(stream or 0)))
    return opt, sec.value


def relax_instances(entries, alg: AlgorithmDesc, max_iters, candidate_tol=1e-6, concurrency=4, stream=None):
    """gb_relax_instances.  entries: dicts with obj, lmo, mixed (MixedQ or None) and
    the torch buffers x, lb, ub, eps, iterations, status, gap, value,
    best_vertex, best_vertex_val, flags, cutoff (None allowed for flags/cutoff)."""
    arr = (InstanceBatch * len(entries))()
    keep = []
    for k, e in enumerate(entries):
        keep.append((e["obj"], e["lmo"]))
        b = arr[k]
        b.objective = ctypes.addressof(e["obj"])
        b.lmo = ctypes.addressof(e["lmo"])
        b.mixed = e["mixed"].handle.value if e.get("mixed") is not None else None
        b.N = int(e["x"].shape[0])
        for name, key in (("x", "x"), ("lb", "lb"), ("ub", "ub"), ("eps", "eps"), ("iterations", "iterations"),
                          ("status", "status"), ("gap", "gap"), ("value", "value"),
                          ("best_vertex", "best_vertex"), ("best_vertex_val", "best_vertex_val"),
                          ("integer_flags", "flags"), ("cutoff", "cutoff")):
            t = e.get(key)
            setattr(b, name, None if t is None else t.data_ptr())
    check(load().gb_relax_instances(len(entries), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(alg),
                                    int(max_iters), float(candidate_tol), int(concurrency),
                                    ctypes.c_void_p(stream or 0)))
    del keep


class RelaxGraph:
    """gb_relax_graph_*: a captured batched-form launch over fixed tensors."""

    def __init__(self, obj: ObjectiveDesc, alg: AlgorithmDesc, lmo: LmoDesc, N, max_iters, x, lb, ub,
                 eps, iterations, status, gap, value, best_vertex, best_vertex_val, flags=None,
                 candidate_tol=1e-6, cutoff=None, refresh=64, waves_per_block=0, blocks=0):
        lib = load()
        self._keep = (obj, alg, lmo, x, lb, ub, eps, iterations, status, gap, value, best_vertex,
                      best_vertex_val, flags, cutoff)
        opt = FastOptions(refresh, waves_per_block, blocks, 0, 0)
        h = ctypes.c_void_p()
        check(lib.gb_relax_graph_create(ctypes.byref(obj), ctypes.byref(alg), ctypes.byref(lmo), int(N),
                                        int(max_iters), _ptr(x), _ptr(lb), _ptr(ub), _ptr(eps),
                                        _ptr(iterations), _ptr(status), _ptr(gap), _ptr(value),
                                        _ptr(best_vertex), _ptr(best_vertex_val), _ptr(flags),
                                        float(candidate_tol), _ptr(cutoff), ctypes.byref(opt),
                                        ctypes.byref(h)))
        self.handle = h

    def launch(self, stream=None):
        check(load().gb_relax_graph_launch(self.handle, ctypes.c_void_p(stream or 0)))

    def close(self):
        if self.handle:
            load().gb_relax_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Brancher:
    """The on-GPU MLP branching scorer (gb_brancher_*)."""

    def __init__(self, seed=0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.gb_brancher_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            load().gb_brancher_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        import numpy as np
        out = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(load().gb_brancher_get_params(self.handle, out.ctypes.data, MLP_PARAMS))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(load().gb_brancher_set_params(self.handle, p.ctypes.data, int(p.size)))

    def set_normalization(self, mu, sigma):
        import numpy as np
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sigma = np.ascontiguousarray(sigma, dtype=np.float32)
        check(load().gb_brancher_set_normalization(self.handle, mu.ctypes.data, sigma.ctypes.data))
        self._norm = (mu.copy(), sigma.copy())

    def save(self, path):
        """Writes weights and feature normalization to an .npz file."""
        import numpy as np
        mu, sigma = getattr(self, "_norm", (np.zeros(MLP_FEATURES, np.float32), np.ones(MLP_FEATURES, np.float32)))
        np.savez(path, params=self.params(), mu=mu, sigma=sigma)

    @classmethod
    def from_file(cls, path):
        """A scorer restored from Brancher.save output (plain arrays, no pickles)."""
        import numpy as np
        with np.load(path, allow_pickle=False) as z:
            b = cls(0)
            b.set_params(z["params"])
            b.set_normalization(z["mu"], z["sigma"])
        return b

    def predict(self, X, Y, stream=None):
        """X, Y: torch device float32 tensors [M,8], [M]."""
        check(load().gb_brancher_predict(self.handle, X.data_ptr(), int(X.shape[0]), Y.data_ptr(),
                                         stream or 0))

    def loss_grad(self, X, T):
        import numpy as np
        loss = ctypes.c_double()
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(load().gb_brancher_loss_grad(self.handle, X.data_ptr(), T.data_ptr(), int(X.shape[0]),
                                           ctypes.byref(loss), g.ctypes.data))
        return loss.value, g

    def train(self, X, T, epochs, batch, lr):
        loss = ctypes.c_double()
        check(load().gb_brancher_train(self.handle, X.data_ptr(), T.data_ptr(), int(X.shape[0]),
                                       int(epochs), int(batch), ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value


def grid_barrier_selftest(groups, group_blocks, threads=256, phases_base=2000, phase_step=101,
                          spin_limit=1 << 22):
    """gb_grid_barrier_selftest: (errors, timeouts, seconds per sync)."""
    e, t, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    check(load().gb_grid_barrier_selftest(groups, group_blocks, threads, phases_base, phase_step,
                                          spin_limit, ctypes.byref(e), ctypes.byref(t), ctypes.byref(s)))
    return e.value, t.value, s.value


def branching_dataset(n, lb, ub, int_vars, obj, lmo, settings, max_nodes, capacity):
    """gb_branching_dataset: strong-branching (features, labels) rows."""
    import numpy as np
    lib = load()
    lb_a = (ctypes.c_double * n)(*[float(v) for v in lb])
    ub_a = (ctypes.c_double * n)(*[float(v) for v in ub])
    iv = list(int_vars)
    iv_a = (ctypes.c_int * max(1, len(iv)))(*iv)
    prob = Problem(n, lb_a, ub_a, iv_a, len(iv))
    F = np.zeros((capacity, MLP_FEATURES), dtype=np.float32)
    L = np.zeros(capacity, dtype=np.float32)
    cnt = ctypes.c_int()
    check(lib.gb_branching_dataset(ctypes.byref(prob), ctypes.byref(obj), ctypes.byref(lmo),
                                   ctypes.byref(settings), int(max_nodes), F.ctypes.data, L.ctypes.data,
                                   int(capacity), ctypes.byref(cnt)))
    return F[:cnt.value], L[:cnt.value]


def solve(n, lb, ub, int_vars, obj: ObjectiveDesc, lmo: LmoDesc, settings: Settings | None = None,
          root_initial_point=None):
    """gb_solve with host lists/arrays for the problem; returns a dict."""
    lib = load()
    s = settings if settings is not None else default_settings()
    lb_a = (ctypes.c_double * n)(*[float(v) for v in lb])
    ub_a = (ctypes.c_double * n)(*[float(v) for v in ub])
    iv = list(int_vars)
    iv_a = (ctypes.c_int * max(1, len(iv)))(*iv)
    prob = Problem(n, lb_a, ub_a, iv_a, len(iv))
    keep = []
    if root_initial_point is not None:
        rp = (ctypes.c_double * len(root_initial_point))(*[float(v) for v in root_initial_point])
        keep.append(rp)
        s.root_initial_point = rp
        s.root_initial_point_size = len(root_initial_point)
    xo = (ctypes.c_double * n)()
    res = Result()
    res.x = xo
    code = lib.gb_solve(ctypes.byref(prob), ctypes.byref(obj), ctypes.byref(lmo), ctypes.byref(s),
                        ctypes.byref(res))
    check(code)
    return {"objective": res.objective, "x": [xo[j] for j in range(n)],
            "dual_bound": res.dual_bound, "nodes_processed": res.nodes_processed,
            "elapsed_seconds": res.elapsed_seconds,
            "reached_tolerance": bool(res.reached_tolerance),
            "reached_time_limit": bool(res.reached_time_limit),
            "fw_iterations": res.fw_iterations, "relax_seconds": res.relax_seconds}


class Solver:
    """gb_solver_*: the low-level solver object (DeviceResidentHybridSolver)
    over host lists; `events` collects the scheduling events of each solve
    as (type, slot_id, generation, n_nodes)."""

    def __init__(self, n, lb, ub, int_vars, obj: ObjectiveDesc, lmo: LmoDesc, settings: Settings | None = None,
                 branch=None):
        """branch: optional Python BranchStrategy (parent dict, next_id, x list,
        int_vars) -> (children as dicts with id, parent_id, depth, lb, ub,
        dual_bound; new next_id)."""
        lib = load()
        self.n = n
        self._s = settings if settings is not None else default_settings()
        self._lb = (ctypes.c_double * n)(*[float(v) for v in lb])
        self._ub = (ctypes.c_double * n)(*[float(v) for v in ub])
        iv = list(int_vars)
        self._iv = (ctypes.c_int * max(1, len(iv)))(*iv)
        self._prob = Problem(n, self._lb, self._ub, self._iv, len(iv))
        self._obj, self._lmo = obj, lmo
        self._keep = []
        self.events = []
        if branch is not None:
            def tramp(user, parent, next_id, x, nn, ivp, niv, children):
                p = parent.contents
                view = {"id": p.id, "parent_id": p.parent_id, "depth": p.depth,
                        "lb": [p.lb[j] for j in range(nn)], "ub": [p.ub[j] for j in range(nn)],
                        "relaxation_x": None if not p.relaxation_x else [p.relaxation_x[j] for j in range(nn)],
                        "relaxation_iterations": p.relaxation_iterations, "dual_bound": p.dual_bound}
                kids, nid = branch(view, next_id[0], [x[j] for j in range(nn)], [ivp[t] for t in range(niv)])
                next_id[0] = nid
                for k in kids:
                    clb = (ctypes.c_double * nn)(*k["lb"])
                    cub = (ctypes.c_double * nn)(*k["ub"])
                    node = Node(k["id"], k["parent_id"], k["depth"], clb, cub, None, 0, k["dual_bound"])
                    if lib.gb_children_push(children, ctypes.byref(node)) != GB_OK:
                        return 1
                return 0
            self._branch_cb = BRANCH_FN(tramp)
            self._s.branch_rule = BRANCH_CUSTOM
            self._s.branch = self._branch_cb
        h = ctypes.c_void_p()
        check(lib.gb_solver_create(ctypes.byref(self._prob), ctypes.byref(obj), ctypes.byref(lmo),
                                   ctypes.byref(self._s), ctypes.byref(h)))
        self.handle = h

        def on_event(user, ev):
            e = ev.contents
            self.events.append((e.type, e.slot_id, e.generation, e.n_nodes))
        self._sched_cb = SCHEDULING_FN(on_event)
        check(lib.gb_solver_set_scheduling_callback(self.handle, self._sched_cb, None))

    def push(self, id, parent_id, depth, lb, ub, dual_bound=-1e300, relaxation_x=None, relaxation_iterations=0):
        n = self.n
        lb_a = (ctypes.c_double * len(lb))(*[float(v) for v in lb])
        ub_a = (ctypes.c_double * len(ub))(*[float(v) for v in ub])
        rx = None if relaxation_x is None else (ctypes.c_double * len(relaxation_x))(*relaxation_x)
        if len(lb) != n or len(ub) != n or (rx is not None and len(relaxation_x) != n):
            raise InvalidArgument(GB_ERR_INVALID_ARGUMENT, "invalid node buffer dimensions or resume state")
        self._keep.append((lb_a, ub_a, rx))
        node = Node(id, parent_id, depth, lb_a, ub_a, rx, relaxation_iterations, dual_bound)
        check(load().gb_solver_push(self.handle, ctypes.byref(node)))

    def solve(self):
        self.events.clear()
        check(load().gb_solver_solve(self.handle))
        self._keep.clear()

    def stats(self):
        st = SolverStats()
        check(load().gb_solver_get_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in SolverStats._fields_}

    def incumbent_x(self):
        x = (ctypes.c_double * self.n)()
        has = ctypes.c_int()
        check(load().gb_solver_incumbent_x(self.handle, x, ctypes.byref(has)))
        return [x[j] for j in range(self.n)] if has.value else []

    def close(self):
        if getattr(self, "handle", None):
            load().gb_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def inf():
    return math.inf
