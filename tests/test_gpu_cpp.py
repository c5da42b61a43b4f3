"""C++ programs written against the drop-in header (include/gpuboscia/gpuboscia.hpp)
and built by __graft_entry__.build(): the reference's examples/minimal.cu
(x = [0, 1], objective 0.18, plus the std::function hooks and exceptions)
and its solver-level known-answer tests through DeviceResidentHybridSolver
(tests/cpp/reference_kats.cpp), and the device plug-in tests: caller
objective, LMO and heuristic types compiled into the relaxation kernel,
including the sparse Sharpe model's LinkedCardinalityLMO
(tests/cpp/plugin_kats.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "minimal_example")


def test_cpp_dropin_minimal_example():
    assert os.path.exists(EXE), "tests/cpp/minimal_example not built: run __graft_entry__.build()"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "x=[0, 1], objective=0.18000000" in out.stdout
    assert "FAIL" not in out.stdout


def test_cpp_reference_solver_kats():
    exe = os.path.join(ROOT, "tests", "cpp", "reference_kats")
    assert os.path.exists(exe), "tests/cpp/reference_kats not built: run __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout
    assert "0 failure(s)" in out.stdout
    assert "PASS: MultiBlock K=3 search equals the one-workgroup search" in out.stdout


def test_cpp_device_plugins():
    exe = os.path.join(ROOT, "tests", "cpp", "plugin_kats")
    assert os.path.exists(exe), "tests/cpp/plugin_kats not built: run __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout
    assert "0 failure(s)" in out.stdout
    assert "PASS: plug-in objective and LMO: same search" in out.stdout
