"""CPU reference of the learned-branching MLP (numpy).

TEST INFRASTRUCTURE ONLY.  Mirrors the arithmetic of
gpuboscia_amd/csrc/mlp.hip: BF16-rounded standardized inputs, weights and
hidden activations, FP32 accumulation, ReLU, mean-squared-error loss; the
backward pass uses the same BF16-rounded values.  Pinned by finite
differences (tests/test_mlp_oracle.py).  No reference counterpart.
"""
import numpy as np

F, H = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
PARAMS = OFF_B4 + 1


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _ops(rounding):
    """rounding=True: the kernels' arithmetic (float32, BF16-rounded operands);
    rounding=False: exact float64 network (for finite-difference checks)."""
    if rounding:
        return bf16, np.float32
    return (lambda a: np.asarray(a, dtype=np.float64)), np.float64


def unpack_as(p, dt):
    return tuple(np.asarray(t, dtype=dt) for t in unpack(np.asarray(p, dtype=np.float64)))


def forward(p, X, mu, sigma, rounding=True):
    r, dt = _ops(rounding)
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack_as(p, dt) if not rounding else unpack(p)
    xn = r((np.asarray(X, dt) - np.asarray(mu, dt)) / np.asarray(sigma, dt))
    h1 = r(np.maximum(xn @ r(W1) + b1, 0))
    h2 = r(np.maximum(h1 @ r(W2) + b2, 0))
    h3 = r(np.maximum(h2 @ r(W3) + b3, 0))
    y = h3 @ r(w4) + b4
    return y.astype(dt), (xn, h1, h2, h3)


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(H, H), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4],
            p[OFF_B4])


def loss_grad(p, X, T, mu, sigma, rounding=True):
    r, dt = _ops(rounding)
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack_as(p, dt) if not rounding else unpack(np.asarray(p, np.float32))
    y, (xn, h1, h2, h3) = forward(p, X, mu, sigma, rounding)
    M = len(T)
    d = y - np.asarray(T, dt)
    loss = float(np.mean(d.astype(np.float64) ** 2))
    dy = (2.0 / M) * d
    dz3 = np.where(h3 > 0, np.outer(dy, r(w4)), 0).astype(dt)
    dz2 = np.where(h2 > 0, dz3 @ r(W3).T, 0).astype(dt)
    dz1 = np.where(h1 > 0, dz2 @ r(W2).T, 0).astype(dt)
    g = np.zeros(PARAMS, dt)
    g[OFF_W1:OFF_B1] = (xn.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = (h2.T @ dz3).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(0)
    g[OFF_W4:OFF_B4] = h3.T @ dy
    g[OFF_B4] = dy.sum()
    return loss, g
