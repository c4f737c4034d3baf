"""Synthetic inputs of the measured workloads (host numpy, seeded, identical on every machine).

plain_gaussian     512 x 256 i.i.d. N(0,1), seed 0 (the plain FP64 CPU-runnable workload)
tall_geometric     4096 x 1024 A = Q1 diag(sigma) Q2^T, Q1 = qr(N(0,1) 4096x1024, seed 1),
                   Q2 = qr(N(0,1) 1024x1024, seed 2), sigma_i = 10^(-8 i/(n-1))
graded             16384 x 4096 G diag(10^(-12 j/(n-1))), G ~ N(0,1), seed 5
batch_mixed        20,000 matrices, m ~ U{8..256}, n ~ U{4..min(m,128)}, entries N(0,1) * 10^U(-3,3)
                   per matrix; even index requests FP32 (tol 1e-6), odd FP64 (tol 1e-14); seed 3
"""
from __future__ import annotations

import numpy as np


def plain_gaussian(m: int = 512, n: int = 256, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((m, n))


def tall_geometric(m: int = 4096, n: int = 1024, decades: float = 8.0, seeds=(1, 2)):
    """Returns (A row-major, exact sigma descending)."""
    q1, _ = np.linalg.qr(np.random.default_rng(seeds[0]).standard_normal((m, n)))
    q2, _ = np.linalg.qr(np.random.default_rng(seeds[1]).standard_normal((n, n)))
    sigma = 10.0 ** (-decades * np.arange(n) / (n - 1))
    a = (q1 * sigma) @ q2.T
    return np.ascontiguousarray(a), sigma


def graded(m: int = 16384, n: int = 4096, decades: float = 12.0, seed: int = 5) -> np.ndarray:
    g = np.random.default_rng(seed).standard_normal((m, n))
    g *= 10.0 ** (-decades * np.arange(n) / (n - 1))
    return g


def batch_mixed(count: int = 20000, seed: int = 3, m_range=(8, 256), n_max: int = 128):
    """Returns a list of (A, requested precision 'f32'|'f64', tol)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(rng.integers(m_range[0], m_range[1] + 1))
        n = int(rng.integers(4, min(m, n_max) + 1))
        scale = 10.0 ** rng.uniform(-3, 3)
        a = rng.standard_normal((m, n)) * scale
        if i % 2 == 0:
            out.append((a, "f32", 1e-6))
        else:
            out.append((a, "f64", 1e-14))
    return out
