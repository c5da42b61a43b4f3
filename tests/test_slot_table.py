"""Host-only test of the rules by which the library lends node counters to
kernel launches (gpuboscia_amd/csrc/gb_slots.h): built with g++ and run on
the CPU, no GPU involved."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_slot_table_rules(tmp_path):
    exe = tmp_path / "slot_table_test"
    src = os.path.join(ROOT, "tests", "cpp", "slot_table_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout
