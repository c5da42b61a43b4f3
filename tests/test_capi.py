"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, carries the reference defaults, and rejects invalid input
with the reference's exception classes and message fragments (no GPU work is
started by any call here)."""
import ctypes
import math
import os
import re

import pytest

import gpuboscia_amd as gb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "gpuboscia_amd.h")).read()
    # function declarations: "<type> gb_name(" at the start of a line
    return sorted(set(re.findall(r"^(?:const\s+)?[a-z_]+\s*\*?\s*(gb_[a-z_0-9]+)\s*\(", text,
                                 flags=re.MULTILINE)))


def test_library_exports_header_symbols():
    lib = gb.load()
    syms = header_symbols()
    assert syms, "no declarations parsed"
    for s in syms:
        assert hasattr(lib, s), s
    assert set(gb.EXPORTED) == set(syms)
    assert b"gfx950" in lib.gb_version()


def test_default_settings_match_reference():
    # device_resident_solver.cuh:1006-1017, line_search_options.h:18-23,
    # gpuboscia.cuh:21-32
    s = gb.default_settings()
    assert (s.gpu_batch_size, s.num_streams, s.gpu_block_threads, s.gpu_max_iters) == (1024, 8, 0, 20000)
    assert (s.gpu_eps, s.gpu_eps_decay, s.gpu_eps_min) == (1e-3, 0.8, 1e-6)
    assert (s.integer_tol, s.abs_gap_tol, s.rel_gap_tol) == (1e-6, 1e-6, 1e-2)
    assert s.algorithm.algorithm == gb.ALG_DICG and s.algorithm.line_search == gb.LS_SECANT
    assert (s.algorithm.secant_iters, s.algorithm.secant_tol, s.algorithm.secant_denom_eps) == (4, 1e-10, 1e-14)
    assert s.branch_rule == gb.BRANCH_POW2 and s.branch_k == 3
    assert math.isinf(s.time_limit_seconds) and s.time_limit_seconds > 0
    assert math.isinf(s.natural_lower_bound) and s.natural_lower_bound < 0


def fake_dense(n):
    # A non-null device address that is never dereferenced: every call below
    # fails validation before any device work.
    return gb.ObjectiveDesc(gb.OBJ_DENSE, n, 0x1000, 0, 2.0, 0.5, 0.0, 0, 0, 0)


def box():
    return gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)


@pytest.mark.parametrize("mutate,fragment,exc", [
    (lambda s: setattr(s, "gpu_max_iters", 0), "GPU iterations, batch size and streams must be positive", gb.InvalidArgument),
    (lambda s: setattr(s, "gpu_batch_size", 0), "GPU iterations, batch size and streams must be positive", gb.InvalidArgument),
    (lambda s: setattr(s, "gpu_eps", -1.0), "invalid tolerance configuration", gb.InvalidArgument),
    (lambda s: setattr(s, "gpu_eps_decay", 1.5), "invalid tolerance configuration", gb.InvalidArgument),
    (lambda s: setattr(s, "rel_gap_tol", float("nan")), "invalid tolerance configuration", gb.InvalidArgument),
    (lambda s: setattr(s, "time_limit_seconds", 0.0), "time limit must be positive", gb.InvalidArgument),
    (lambda s: setattr(s, "natural_lower_bound", float("inf")), "natural lower bound must be finite", gb.InvalidArgument),
    (lambda s: setattr(s, "natural_lower_bound", float("nan")), "natural lower bound must be finite", gb.InvalidArgument),
    (lambda s: setattr(s, "gpu_block_threads", 96), "gpu_block_threads must be zero or a positive multiple of 64", gb.InvalidArgument),
    (lambda s: (setattr(s.algorithm, "algorithm", gb.ALG_FRANK_WOLFE),
                setattr(s.algorithm, "line_search", gb.LS_SCALED)), "scaled open-loop is available only for DICG", gb.InvalidArgument),
])
def test_solve_rejects_invalid_settings(mutate, fragment, exc):
    s = gb.default_settings()
    mutate(s)
    with pytest.raises(exc) as e:
        gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(2), box(), s)
    assert fragment in str(e.value)


def test_solve_rejects_bad_problem():
    with pytest.raises(gb.InvalidArgument, match="integer variable index out of range"):
        gb.solve(2, [0, 0], [1, 1], [0, 2], fake_dense(2), box())
    with pytest.raises(gb.GpuBosciaError, match="does not match Problem::n_vars"):
        gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(3), box())


def test_root_initial_point_rules():
    # tests/test_root_initial_point.cpp:22-41
    with pytest.raises(gb.InvalidArgument, match="dimensions"):
        gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(2), box(), root_initial_point=[0.5])
    for bad in ([-0.1, 0.2], [math.nan, 0.0], [0.0, math.inf]):
        with pytest.raises(gb.InvalidArgument, match="nonfinite or violates bounds"):
            gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(2), box(), root_initial_point=bad)
    simplex = gb.LmoDesc(gb.LMO_SIMPLEX, 1.0, 0, 0x1000, 0)
    with pytest.raises(gb.InvalidArgument, match="continuous feasible set"):
        gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(2), simplex, root_initial_point=[0.8, 0.8])
    # a custom validator overrides the LMO's check and is consulted
    called = []

    def validator(user, x, lb, ub, n, tol):
        called.append(True)
        return 1 if x[0] + x[1] <= 1 else 0

    s = gb.default_settings()
    s.root_validator = gb.VALIDATOR_FN(validator)
    with pytest.raises(gb.InvalidArgument, match="continuous feasible set"):
        gb.solve(2, [0, 0], [1, 1], [0, 1], fake_dense(2), box(), s, root_initial_point=[0.8, 0.8])
    assert called


def test_relaxation_feasible_cpu():
    lib = gb.load()

    def feas(lmo, x, tol=1e-9):
        n = len(x)
        xa = (ctypes.c_double * n)(*x)
        z = (ctypes.c_double * n)()
        return lib.gb_lmo_relaxation_feasible_cpu(ctypes.byref(lmo), None, xa, z, z, n, tol)

    assert feas(box(), [0.2, 0.7]) == 1
    assert feas(gb.LmoDesc(gb.LMO_SIMPLEX, 1.0, 0, 0, 0), [0.2, 0.7]) == 1
    assert feas(gb.LmoDesc(gb.LMO_SIMPLEX, 1.0, 0, 0, 0), [0.8, 0.8]) == 0
    assert feas(gb.LmoDesc(gb.LMO_SIMPLEX, 1.0, 1, 0, 0), [0.2, 0.7]) == 0
    assert feas(gb.LmoDesc(gb.LMO_SIMPLEX, 1.0, 1, 0, 0), [0.3, 0.7]) == 1
    assert feas(gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, 2), [0.5, 0.5, 0.5, 0.5]) == 1
    assert feas(gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, 2), [1.0, 0.5, 0.0, 0.5]) == 0


def test_relax_batch_validation_without_device():
    alg = gb.algorithm()
    # N == 0 is a no-op
    gb.relax_batch(fake_dense(4), alg, box(), 0, 10, *([None] * 10))
    with pytest.raises(gb.InvalidArgument, match="multiple of 64"):
        lib = gb.load()
        code = lib.gb_relax_batch(ctypes.byref(fake_dense(4)), ctypes.byref(alg), ctypes.byref(box()),
                                  1, 10, *([ctypes.c_void_p(0x1000)] * 11), 1e-6, 32, None)
        gb.check(code)
    with pytest.raises(gb.InvalidArgument, match="Birkhoff order"):
        lib = gb.load()
        code = lib.gb_relax_batch(ctypes.byref(fake_dense(4)), ctypes.byref(alg),
                                  ctypes.byref(gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, 0)),
                                  1, 10, *([ctypes.c_void_p(0x1000)] * 11), 1e-6, 0, None)
        gb.check(code)
