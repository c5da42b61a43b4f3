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
