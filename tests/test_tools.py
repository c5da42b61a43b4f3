"""The example tooling on the CPU (tests/cpp/host_tools, built by
__graft_entry__.build()):

* the MAT v5 reader (include/gpuboscia/mat.hpp) against scipy.io.loadmat on
  the reference's own instance files (copied under tests/golden/mat) and on
  scipy-written variants (tests/golden/make_mat_fixtures.py: plain and zlib
  elements, uint8 / int16 / single / double / logical / sparse, short and
  long names, an empty array), and the graphs against their spectra (rook's
  graph K_k x K_k, triangular graph T(q), 3-regular CHH);
* the logger's layout (logger.h:47-177) line by line;
* the example option parser (examples/common/options.h) on every flag and
  its error messages.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "host_tools")
MAT = os.path.join(ROOT, "tests", "golden", "mat")

sio = pytest.importorskip("scipy.io")
sparse = pytest.importorskip("scipy.sparse")


def run(*args):
    assert os.path.exists(EXE), "tests/cpp/host_tools not built: run __graft_entry__.build()"
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def read_mat(path):
    out = run("mat", path)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    res, i = {}, 0
    while i < len(lines) and lines[i]:
        name, r, c = lines[i].split()
        r, c = int(r), int(c)
        vals = np.array([float(v) for v in lines[i + 1:i + 1 + r * c]], dtype=np.float64)
        res[name] = vals.reshape(r, c)
        i += 1 + r * c
    return res


def scipy_vars(path):
    d = sio.loadmat(path)
    out = {}
    for k, v in d.items():
        if k.startswith("__"):
            continue
        if sparse.issparse(v):
            v = v.toarray()
        out[k] = np.asarray(v, dtype=np.float64)
    return out


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(MAT) if f.endswith(".mat")))
def test_mat_reader_equals_scipy(name):
    path = os.path.join(MAT, name)
    got = read_mat(path)
    want = scipy_vars(path)
    assert set(got) == set(want), (sorted(got), sorted(want))
    for k, v in want.items():
        assert got[k].shape == v.shape, (k, got[k].shape, v.shape)
        assert np.array_equal(got[k], v), k


def _spectrum(A):
    return np.sort(np.round(np.linalg.eigvalsh(A), 6))


@pytest.mark.parametrize("k", [4, 8])
def test_lattice_instances_are_rook_graphs(k):
    A = read_mat(os.path.join(MAT, f"Lattice_{k}_{k * k}.mat"))["A"]
    # K_k x K_k: eigenvalues 2k-2 (once), k-2 (2(k-1) times), -2 ((k-1)^2 times)
    want = np.sort(np.array([2 * k - 2] + [k - 2] * (2 * (k - 1)) + [-2] * ((k - 1) ** 2), dtype=float))
    assert np.array_equal(A, A.T) and np.all(np.diag(A) == 0)
    assert np.allclose(_spectrum(A), want, atol=1e-6)


@pytest.mark.parametrize("q", [7, 10])
def test_triangular_instances_are_triangular_graphs(q):
    A = read_mat(os.path.join(MAT, f"Triangular_{q}_{q * (q - 1) // 2}.mat"))["A"]
    # T(q): eigenvalues 2(q-2) (once), q-4 (q-1 times), -2 (q(q-3)/2 times)
    want = np.sort(np.array([2 * (q - 2)] + [q - 4] * (q - 1) + [-2] * (q * (q - 3) // 2), dtype=float))
    assert np.allclose(_spectrum(A), want, atol=1e-6)


def test_chh_instance_is_cubic():
    A = read_mat(os.path.join(MAT, "CHH_cc_1_1_22_1.mat"))["A"]
    assert A.shape == (22, 22) and np.array_equal(A, A.T) and np.all(A.sum(axis=1) == 3)


def test_price_matrix_orientation():
    d = read_mat(os.path.join(MAT, "FF25EUnew.mat"))["data"]
    assert d.shape == (391, 25)
    assert d[0, 0] == 0.955 and d[0, 1] == 0.9666 and d[1, 0] == 0.9777


def test_mat_reader_errors():
    out = run("mat", os.path.join(MAT, "missing.mat"))
    assert out.returncode == 3 and "cannot open MAT file" in out.stdout
    bad = os.path.join(ROOT, "pytest.ini")
    out = run("mat", bad)
    assert out.returncode == 3 and "little-endian MAT v5" in out.stdout


def test_logger_layout():
    out = run("log")
    assert out.returncode == 0
    t = r"\s+\d\.\d{6}e[+-]\d{2}"
    want = [
        r"Hello:\)",
        r"Parameter settings\.",
        r"  Tree traversal strategy:     best_first",
        r"  Branching strategy:          branch_pow2",
        r"  Algorithm:                   Decomposition-Invariant Frank-Wolfe",
        r"  Line search method:          secant on phi'\(gamma\)",
        r"  Random seed:                 42",
        r"  Termination:                 abs gap <= 1\.000e-06 or rel gap <= 1\.000e-02",
        r"  Time limit \(s\):              60\.000",
        r"  Natural lower bound:         0\.000000e\+00",
        r"  Total number of variables:   4",
        r"  Number of integer variables: 2",
        r"-{93}",
        r"  Iter   Open           Bound       Incumbent       Gap \(abs\)       Gap \(rel\)        Time \(s\)",
        r"-{93}",
        r"     1      1            -Inf             Inf             Inf             Inf" + t,
        r"\*    3      2   -1\.250000e\+00    2\.000000e\+00    3\.250000e\+00    2\.600000e\+00" + t,
        r"     5      1    0\.000000e\+00    2\.000000e\+00    2\.000000e\+00             Inf" + t,
        r"-{93}",
        r"Solution Statistics\.",
        r"  Solution Status:              Optimal",
        r"  Primal Objective:               2\.000000e\+00",
        r"  Dual Bound:                     2\.000000e\+00",
        r"  Dual Gap \(relative\):            0\.000000e\+00",
        r"  Solution dimension:           4",
        r"  Solution nonzeros:             2",
        r"  Solution preview x\[0:3\]:      \[1, 0, 0\.5\] \.\.\.",
        r"Search Statistics\.",
        r"  Total number of nodes processed: 5",
        r"  Total time \(s\):                 5\.000000e-01",
    ]
    lines = out.stdout.rstrip("\n").split("\n")
    assert len(lines) == len(want), out.stdout
    for line, pat in zip(lines, want):
        assert re.fullmatch(pat, line), (line, pat)


def _cli(*args):
    out = run("cli", *args)
    assert out.returncode == 0
    fields = {}
    for line in out.stdout.strip().split("\n"):
        k, _, v = line.partition(" ")
        fields[k] = v
    return fields, out.stderr


def test_cli_defaults_and_flags():
    f, _ = _cli()
    assert f["algorithm"] == "dicg" and f["line_search"] == "3 4 1e-10 1e-14" and f["seed"] == "1 0"
    assert f["eps"] == "0.001 0.80000000000000004 9.9999999999999995e-07"
    assert f["time"] == "3600" and f["natural"] == "-inf" and f["interval"] == "100" and f["criterion"] == "d"
    f, _ = _cli("--algorithm", "fw", "--line-search", "quadratic-exact", "--seed", "77", "--gpu-eps", "0.01",
                "--gpu-eps-decay", "0.5", "--gpu-eps-min", "1e-7", "--abs-gap-tol", "1e-8", "--rel-gap-tol", "0",
                "--time-limit", "12.5", "--natural-lower-bound", "-3", "--log-interval-nodes", "9",
                "--criterion", "A", "--instance", "x.mat", "--secant-iters", "8", "--secant-tol", "1e-9",
                "--secant-denom-eps", "1e-12")
    assert f["algorithm"] == "fw" and f["line_search"] == "4 8 1.0000000000000001e-09 9.9999999999999998e-13"
    assert f["seed"] == "77 1" and f["eps"] == "0.01 0.5 9.9999999999999995e-08"
    assert f["gaps"] == "1e-08 0" and f["time"] == "12.5" and f["natural"] == "-3" and f["interval"] == "9"
    assert f["criterion"] == "a" and f["instance"] == "x.mat"


@pytest.mark.parametrize("args,message", [
    (["--algorithm", "pgd"], "--algorithm must be either 'fw' or 'dicg'."),
    (["--seed", "12x"], "--seed must be an unsigned integer."),
    (["--time-limit", "0"], "--time-limit must be a positive finite number of seconds."),
    (["--log-interval-nodes", "-1"], "--log-interval-nodes must be a positive integer."),
    (["--natural-lower-bound", "inf"], "--natural-lower-bound must be a finite number."),
    (["--line-search", "armijo"], "--line-search must be 'open-loop'"),
    (["--algorithm", "fw", "--line-search", "scaled-open-loop"],
     "scaled-open-loop is currently available only for DICG."),
    (["--secant-iters", "0"], "Secant parameters must be positive finite values."),
    (["--gpu-eps-decay", "1.5"], "Tolerance parameters are invalid."),
    (["--gpu-eps"], "Missing value for argument: --gpu-eps"),
    (["--frobnicate"], "Unknown argument: --frobnicate"),
])
def test_cli_errors(args, message):
    f, err = _cli(*args)
    assert f["exit"] == "1" and message in err, err


def test_cli_help():
    out = run("cli", "--help")
    assert out.stdout.startswith("Usage: cli [--algorithm fw|dicg]") and out.stdout.rstrip().endswith("exit 0")
