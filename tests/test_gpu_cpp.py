"""The reference's examples/minimal.cu, written against the C++ drop-in
header (include/gpuboscia/gpuboscia.hpp) and built by __graft_entry__.build():
x = [0, 1], objective 0.18, plus the std::function hooks and exceptions."""
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
