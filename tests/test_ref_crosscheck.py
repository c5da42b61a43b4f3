"""Cross-checks the oracle's restated host helpers against the reference's
own CPU code compiled into oracle/_ref/ (oracle/ref.mk).  Skipped where the
reference build is absent (it is not shipped to GPU machines)."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
REF_TEST = os.path.join(ROOT, "oracle", "_ref", "test_cpu")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    lib = ctypes.CDLL(REF_LIB)
    for name in ["ref_certified_fw_lower_bound", "ref_gap_absolute", "ref_gap_relative",
                 "ref_depth_dependent_tolerance", "ref_quadratic_exact_step",
                 "ref_box_dicg_maximum_step_cpu", "ref_unit_simplex_dicg_maximum_step_cpu"]:
        getattr(lib, name).restype = ctypes.c_double
    lib.ref_certified_fw_lower_bound.argtypes = [ctypes.c_double] * 3
    lib.ref_gap_absolute.argtypes = [ctypes.c_double] * 2
    lib.ref_gap_relative.argtypes = [ctypes.c_double] * 2
    lib.ref_gap_within_tolerance.argtypes = [ctypes.c_double] * 4
    lib.ref_depth_dependent_tolerance.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int]
    lib.ref_quadratic_exact_step.argtypes = [ctypes.c_double] * 3
    return lib


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def P(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_reference_cpu_kats_pass():
    out = subprocess.run([REF_TEST], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "0 failed" in out.stdout


def test_scalar_helpers_match(ref, orc):
    o = orc.load()
    rng = np.random.default_rng(0)
    specials = [0.0, -0.0, 1.0, -1.0, math.inf, -math.inf, math.nan, 1e-300, 1e300]
    vals = list(rng.standard_normal(200) * 10) + specials
    for _ in range(3000):
        a, b, cc, d = (vals[i] for i in rng.integers(0, len(vals), 4))
        r1, r2 = ref.ref_certified_fw_lower_bound(a, b, cc), o.orc_certified_fw_lower_bound(a, b, cc)
        assert (math.isnan(r1) and math.isnan(r2)) or r1 == r2
        for f in ("gap_absolute", "gap_relative"):
            r1, r2 = getattr(ref, "ref_" + f)(a, b), getattr(o, "orc_" + f)(a, b)
            assert (math.isnan(r1) and math.isnan(r2)) or r1 == r2
        assert ref.ref_gap_within_tolerance(a, b, abs(cc), abs(d)) == \
            o.orc_gap_within_tolerance(a, b, abs(cc), abs(d))
        r1, r2 = ref.ref_quadratic_exact_step(a, b, cc), o.orc_quadratic_exact_step(a, b, cc)
        assert (math.isnan(r1) and math.isnan(r2)) or r1 == r2
    for depth in range(0, 60):
        assert ref.ref_depth_dependent_tolerance(1e-3, 0.8, 1e-6, depth) == \
            o.orc_depth_dependent_tolerance(1e-3, 0.8, 1e-6, depth)


def test_branching_matches(ref, orc):
    o = orc.load()
    rng = np.random.default_rng(1)
    for trial in range(300):
        n = int(rng.integers(1, 60))
        # many exact ties in fractionality exercise std::sort's tie order
        x = c(np.round(rng.random(n) * 4, 1 if trial % 2 else 3))
        iv = np.ascontiguousarray(rng.permutation(n)[: int(rng.integers(1, n + 1))], dtype=np.int32)
        k = int(rng.integers(1, 6))
        out_r = (ctypes.c_int * 64)()
        out_o = (ctypes.c_int * 64)()
        mr = ref.ref_top_k_infeasible(P(x), n, P(iv), len(iv), k, ctypes.c_double(1e-6), out_r)
        mo = o.orc_top_k_infeasible(P(x), P(iv), len(iv), k, ctypes.c_double(1e-6), out_o)
        assert mr == mo and list(out_r[:mr]) == list(out_o[:mo])
        assert ref.ref_most_infeasible(P(x), n, P(iv), len(iv)) == o.orc_most_infeasible(P(x), P(iv), len(iv))


def test_dicg_cpu_helpers_match(ref, orc):
    o = orc.load()
    rng = np.random.default_rng(2)
    for trial in range(500):
        n = int(rng.integers(1, 12))
        lb = c(rng.integers(-2, 2, n))
        ub = c(lb + rng.integers(0, 4, n))
        x = c(np.clip(np.where(rng.random(n) < 0.4, np.round(rng.random(n) * 4 - 2),
                               rng.random(n) * 4 - 2), lb, ub))
        g = c(np.round(rng.standard_normal(n), 1))
        d = c(np.round(rng.standard_normal(n), 2))
        a_r, a_o = np.zeros(n), np.zeros(n)
        ref.ref_box_inface_lmo_cpu(n, P(g), P(x), P(lb), P(ub), P(a_r), ctypes.c_double(1e-6), ctypes.c_double(1e-4))
        o.orc_box_inface_lmo_cpu(n, P(g), P(x), P(lb), P(ub), P(a_o), ctypes.c_double(1e-6), ctypes.c_double(1e-4))
        assert np.array_equal(a_r, a_o)
        assert ref.ref_box_dicg_maximum_step_cpu(n, P(d), P(x), P(lb), P(ub)) == \
            o.orc_box_dicg_maximum_step_cpu(n, P(d), P(x), P(lb), P(ub))
        is_int = np.ascontiguousarray(rng.integers(0, 2, n), dtype=np.uint8)
        lbs = c(np.maximum(lb, 0))
        ubs = c(np.maximum(ub, lbs))
        xs = c(np.clip(np.abs(x), lbs, ubs))
        rhs = float(max(1.0, np.round(xs.sum() + rng.random() * 2, 0)))
        for full in (0, 1):
            a_r, a_o = np.zeros(n), np.zeros(n)
            ref.ref_unit_simplex_inface_lmo_cpu(n, ctypes.c_double(rhs), P(is_int), P(g), P(xs), P(lbs), P(ubs),
                                                P(a_r), ctypes.c_double(1e-6), ctypes.c_double(1e-4), full)
            o.orc_unit_simplex_inface_lmo_cpu(n, ctypes.c_double(rhs), P(is_int), P(g), P(xs), P(lbs), P(ubs),
                                              P(a_o), ctypes.c_double(1e-6), ctypes.c_double(1e-4), full)
            assert np.array_equal(a_r, a_o)
            r1 = ref.ref_unit_simplex_dicg_maximum_step_cpu(n, ctypes.c_double(rhs), P(is_int), P(d), P(xs),
                                                           P(lbs), P(ubs), ctypes.c_double(1e-6), full)
            r2 = o.orc_unit_simplex_dicg_maximum_step_cpu(n, ctypes.c_double(rhs), P(is_int), P(d), P(xs),
                                                         P(lbs), P(ubs), ctypes.c_double(1e-6), full)
            assert (math.isnan(r1) and math.isnan(r2)) or r1 == r2


def test_branch_pow2_children(ref):
    # The product's branching follows the same rule; here the reference's own
    # branch_pow2 pins children bounds and the id sequence.
    lb, ub = c([0, 0, 0]), c([5, 5, 5])
    x = c([1.5, 2.5, 3.9])
    iv = np.array([0, 1, 2], dtype=np.int32)
    olb, oub = np.zeros(8 * 3), np.zeros(8 * 3)
    ids = (ctypes.c_int * 8)()
    nxt = ctypes.c_int()
    cnt = ref.ref_branch_pow2(P(lb), P(ub), 3, P(x), P(iv), 3, 3, ctypes.c_double(1e-6), 10,
                              P(olb), P(oub), ids, ctypes.byref(nxt))
    assert cnt == 8 and list(ids) == list(range(10, 18)) and nxt.value == 18
    olb, oub = olb.reshape(8, 3), oub.reshape(8, 3)
    # mask bit b -> variable vars[b]; vars sorted by distance (ties keep order)
    assert oub[0].tolist() == [1.0, 2.0, 3.0] and olb[7].tolist() == [2.0, 3.0, 4.0]
