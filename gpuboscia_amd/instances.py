"""Seeded synthetic instances for the benchmark configurations.

Random numbers come from a repo-owned SplitMix64 stream (Steele, Lea and
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) with
Box-Muller normals, so an instance is bit-identical on every machine and
numpy version.  Each instance reports a SHA-256 of its data.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, count: int) -> np.ndarray:
    """count outputs of SplitMix64 started at state seed ^ (stream << 32)."""
    with np.errstate(over="ignore"):
        base = np.uint64((int(seed) ^ (int(stream) << 32)) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, count, lo=0.0, hi=1.0):
    """U[lo, hi) with 53-bit resolution."""
    u = (splitmix64(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed, stream, count):
    """N(0,1) by Box-Muller on pairs of U(0,1] draws."""
    m = (count + 1) // 2
    z = splitmix64(seed, stream, 2 * m)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * math.pi * u2)
    out[1::2] = r * np.sin(2.0 * math.pi * u2)
    return out[:count]


def sha256(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def binary_qp(n: int, seed: int, rows: int | None = None):
    """min x'Qx + c'x over x in {0,1}^n with Q = A'A/m + 0.1 I, A ~ N(0,1)
    (m x n, m = rows or round(1.5 n)), c ~ U(-1,1).  Config C0 uses n=20,
    rows=30; the n=200 set uses rows=300.  Returns (Q, c, sha256)."""
    m = rows if rows is not None else int(round(1.5 * n))
    A = normal(seed, 1, m * n).reshape(m, n)
    Q = A.T @ A / m + 0.1 * np.eye(n)
    Q = 0.5 * (Q + Q.T)
    c = uniform(seed, 2, n, -1.0, 1.0)
    return Q, c, sha256(Q, c)


def qap_instance(m: int, seed: int, p: float = 0.3):
    """Symmetric Erdos-Renyi(p) 0/1 adjacency A, B = Pi A Pi^T for a seeded
    permutation Pi; the optimum of ||A P - P B||_F^2 over permutations is 0
    at P = Pi^T.  Returns (A, B, perm, sha256)."""
    u = uniform(seed, 3, m * m).reshape(m, m)
    A = (np.triu(u, 1) < p).astype(np.float64)
    A = np.triu(A, 1)
    A = A + A.T
    keys = uniform(seed, 4, m)
    perm = np.argsort(keys, kind="stable")
    Pi = np.zeros((m, m))
    Pi[np.arange(m), perm] = 1.0
    B = Pi @ A @ Pi.T
    return A, B, perm, sha256(A, B)


def student_t(seed, stream, count, df=5):
    """Student-t(df) draws: N(0,1) / sqrt(chi2_df / df), chi2 from df squared normals."""
    z = normal(seed, stream, count)
    g = normal(seed, stream + 1, count * df).reshape(count, df)
    return z / np.sqrt((g * g).sum(axis=1) / df)


def portfolio(assets: int = 100, seed: int = 1, periods: int = 600, factors: int = 3, lam: float = 1.0,
              cap: float = 0.2, card: int = 10):
    """Cardinality-constrained mean-variance portfolio (the cardinality
    portfolio configuration).  Monthly returns r_t = B f_t + e_t: factor
    returns f ~ N(0, 0.04^2), loadings B ~ N(1, 0.3), idiosyncratic e =
    0.05 t_5 / sqrt(5/3) (Student-t with 5 degrees of freedom scaled to a 5%
    standard deviation); Sigma and mu are the sample covariance and mean of T
    periods.  Objective over v = (x, z), n = 2 assets:
        f(v) = x' Sigma x - lam mu' x = 1/2 v'Qv + c'v,  Q = blockdiag(2 Sigma, 0),
        c = (-lam mu, 0),
    LMO: sum x = 1, 0 <= x_i <= cap z_i, z in {0,1}, sum z <= card.
    Returns (Q, c, Sigma, mu, sha256)."""
    m, T, k = assets, periods, factors
    F = 0.04 * normal(seed, 11, T * k).reshape(T, k)
    B = 1.0 + 0.3 * normal(seed, 12, m * k).reshape(m, k)
    E = 0.05 * student_t(seed, 13, T * m).reshape(T, m) / np.sqrt(5.0 / 3.0)
    R = F @ B.T + E
    mu = R.mean(axis=0)
    D = R - mu
    Sigma = D.T @ D / (T - 1)
    Sigma = 0.5 * (Sigma + Sigma.T)
    n = 2 * m
    Q = np.zeros((n, n))
    Q[:m, :m] = 2.0 * Sigma
    c = np.zeros(n)
    c[:m] = -lam * mu
    return Q, c, Sigma, mu, sha256(Q, c)


def batched_sizes(count: int = 256, seed: int = 1, lo: int = 32, hi: int = 512) -> np.ndarray:
    """Sizes of the heterogeneous-batch configuration: n_i = round(exp U(ln lo,
    ln hi)), i < count; instance i is binary_qp(n_i, seed * 100000 + i)."""
    u = uniform(seed, 21, count, math.log(lo), math.log(hi))
    return np.rint(np.exp(u)).astype(np.int64)


def batched_lmos(count: int = 256, seed: int = 1, sizes=None):
    """LMO of each heterogeneous-batch instance, chosen by seed: ("box", 0)
    or ("simplex", K), the bounded simplex {x in [0,1]^n, sum x <= K} with
    K = ceil(n / 4) (about half of the instances each)."""
    sizes = batched_sizes(count, seed) if sizes is None else sizes
    pick = uniform(seed, 22, count)
    out = []
    for i, n in enumerate(sizes):
        if pick[i] < 0.5:
            out.append(("box", 0))
        else:
            out.append(("simplex", int(math.ceil(int(n) / 4))))
    return out


def open_nodes(n: int, N: int, seed: int, card: int = 0):
    """N open B&B nodes of a binary problem: node 0 is the root, the others
    carry random fixings of up to 30 % of the variables (lower bound 1 or
    upper bound 0; depth = number of fixings).  With a cardinality limit
    card > 0 (bounded simplex sum x <= card) at most card - 1 variables of a
    node are fixed at 1, so most nodes are feasible; the rest are fixed at 0.
    Returns LB, UB [N][n] and the per-node FW tolerance
    max(1e-6, 1e-3 * 0.8^depth)."""
    u = uniform(seed, 23, 3 * N * n).reshape(3, N, n)
    frac = uniform(seed, 24, N) * 0.3
    fix = u[0] < frac[:, None]
    one = u[1] < 0.5
    fix[0] = False
    LB = np.where(fix & one, 1.0, 0.0)
    UB = np.where(fix & ~one, 0.0, 1.0)
    if card > 0:
        # keep the first card - 1 ones of each node (in index order), fix the others at 0
        ones = np.cumsum(LB, axis=1)
        drop = (LB == 1.0) & (ones > card - 1)
        LB[drop] = 0.0
        UB[drop] = 0.0
    depth = fix.sum(axis=1)
    eps = np.maximum(1e-6, 1e-3 * 0.8 ** depth)
    return LB, UB, eps
