"""Pins the CPU MLP reference (oracle/mlp_oracle.py): its analytic gradient
equals central finite differences of its own float64 loss, and its BF16
rounding matches the definition (round to nearest even on the float32 bits)."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5, 0.0],
                 dtype=np.float32)
    r = mo.bf16(x)
    # bf16 keeps 7 stored mantissa bits: near 1 the spacing is 2^-7.
    # 1 + 2^-8: tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 2^-9: below half -> 1.0;
    # 1 + 3*2^-9: above half -> 1 + 2^-7; 1 + 3*2^-8: tie between 1 + 2^-7 and 1 + 2^-6 -> even
    assert list(r) == [1.0, 1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mo.PARAMS) * 0.2
    X = rng.standard_normal((40, mo.F))
    T = rng.random(40)
    mu, sigma = np.zeros(mo.F), np.ones(mo.F)
    loss, g = mo.loss_grad(p, X, T, mu, sigma, rounding=False)
    h = 1e-6
    idx = list(range(0, mo.OFF_B1, 37)) + list(range(mo.OFF_W2, mo.OFF_B2, 211)) + \
        list(range(mo.OFF_W3, mo.OFF_B3, 199)) + list(range(mo.OFF_W4, mo.PARAMS))
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, X, T, mu, sigma, rounding=False)[0]
              - mo.loss_grad(pm, X, T, mu, sigma, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_candidate_features_hand_computed():
    """The feature oracle on one node worked out by hand: n = 4, flags on
    variables 0, 1, 3; x = (0.5, 0.25, 1.0, 2.75); g = (2, -4, 1, 0.5);
    Q diagonal (2, 4, 6, 8), gs = 2; variable 1 has an up pseudocost of 3,
    the others no history; depth 2, gap 0.3, value -0.5, tol 1e-6.
    Candidates: 0 (dist 0.5), 1 (dist 0.25), 3 (dist 0.25); 2 is integral.
    Up gains: 0: (1/2 * 2 * 2 * 0.5) * 0.5 = 0.5; 1: 3 * 0.75 = 2.25;
    3: (1/2 * 2 * 8 * 0.25) * 0.25 = 0.5; mean 13/12.
    Down gains: 0: (2 * 0.5) * 0.5 = 0.5; 1: (4 * 0.25) * 0.25 = 0.25;
    3: (8 * 0.75) * 0.75 = 4.5; mean 5.25 / 3 = 1.75."""
    X = np.array([[0.5, 0.25, 1.0, 2.75]])
    G = np.array([[2.0, -4.0, 1.0, 0.5]])
    LB = np.array([[0.0, 0.0, 1.0, 2.0]])
    UB = np.array([[1.0, 1.0, 1.0, 4.0]])
    F = mo.candidate_features(X, G, LB, UB, np.array([2]), np.array([0.3]), np.array([-0.5]),
                              np.array([-1.0, 3.0, -1.0, -1.0]), np.full(4, -1.0), np.array([2.0, 4, 6, 8]),
                              2.0, np.array([1, 1, 0, 1], np.uint8), 1e-6)
    f = lambda v: np.float32(v)  # noqa: E731
    expect = np.array([
        [0.5, 0.5, 2 / (4 + 1e-12), 0.5 / (13 / 12 + 1e-12), 0.5 / (1.75 + 1e-12), f(2) / f(4), 0.3 / 1.5, 1.0],
        [0.25, 0.25, -4 / (4 + 1e-12), 2.25 / (13 / 12 + 1e-12), 0.25 / (1.75 + 1e-12), f(2) / f(4), 0.2, 1.0],
        [0.0] * 8,
        [2.75, 0.25, 0.5 / (4 + 1e-12), 0.5 / (13 / 12 + 1e-12), 4.5 / (1.75 + 1e-12), f(2) / f(4), 0.2, 2.0],
    ], dtype=np.float32)
    assert F.dtype == np.float32 and F.shape == (4, 8)
    assert np.allclose(F, expect, rtol=1e-6, atol=0), F - expect
