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
