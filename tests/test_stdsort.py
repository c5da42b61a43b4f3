"""Host-only test of gpuboscia_amd/csrc/gb_stdsort.h, the device sort that
leaves equal keys in std::sort's order (branching candidates of the
device-resident search): built with hipcc as host code and run on the CPU
against std::sort, no GPU involved."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_std_order_sort_matches_std_sort(tmp_path):
    exe = tmp_path / "stdsort_test"
    src = os.path.join(ROOT, "tests", "cpp", "stdsort_test.cpp")
    subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", "-Wall", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout
