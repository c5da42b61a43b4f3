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


def candidate_features(X, G, LB, UB, depth, gap, value, pc_up, pc_down, qdiag, gs, flags, tol):
    """Feature rows of the learned rule, restated from their definitions (the
    order documented next to GB_MLP_FEATURES): for node i and fractional
    candidate j (flag set, min(frac, 1 - frac) > tol), in FP64 then rounded
    to float32:
      0 x_j, 1 min(frac, 1 - frac), 2 g_j / (max_k |g_k| + 1e-12),
      3 / 4 up / down gain over its mean across the node's candidates (+1e-12):
        gain = (pseudocost mean if >= 0, else 1/2 gs Q_jj d) * d, d the
        distance to ceil (up) or floor (down),
      5 depth / n (float32 division), 6 gap / (1 + |value|), 7 ub_j - lb_j;
    other rows are zero.  Returns float32 [N * n, 8]."""
    N, n = X.shape
    F = np.zeros((N * n, 8), dtype=np.float32)
    for i in range(N):
        x, g = X[i], G[i]
        frac = x - np.floor(x)
        dist = np.minimum(frac, 1.0 - frac)
        cand = (flags != 0) & (dist > tol)
        gmax = float(np.max(np.abs(g)))
        du = np.ceil(x) - x
        dd = x - np.floor(x)
        curv = 0.5 * gs * qdiag
        gu = np.where(pc_up >= 0.0, pc_up, curv * du) * du
        gd = np.where(pc_down >= 0.0, pc_down, curv * dd) * dd
        cnt = int(cand.sum())
        mu = float(gu[cand].sum()) / (cnt if cnt else 1.0)
        md = float(gd[cand].sum()) / (cnt if cnt else 1.0)
        rgap = gap[i] / (1.0 + abs(value[i]))
        for j in np.nonzero(cand)[0]:
            r = i * n + j
            F[r, 0] = x[j]
            F[r, 1] = dist[j]
            F[r, 2] = g[j] / (gmax + 1e-12)
            F[r, 3] = gu[j] / (mu + 1e-12)
            F[r, 4] = gd[j] / (md + 1e-12)
            F[r, 5] = np.float32(depth[i]) / np.float32(n)
            F[r, 6] = rgap
            F[r, 7] = UB[i, j] - LB[i, j]
    return F
