"""Numeric helpers shared by the tests (invariants of tex/sections/03-gpu-parallel-reduction.tex:29-43)."""
import numpy as np


def rho_a(a, u, s, v):
    """||A - U S V^T||_F / max(1, ||A||_F)."""
    return float(np.linalg.norm(a - (u * s) @ v.T) / max(1.0, np.linalg.norm(a)))


def rho_orth(x, cols=None):
    """||X^T X - I||_F over the columns `cols` (all by default)."""
    if cols is not None:
        x = x[:, cols]
    return float(np.linalg.norm(x.T @ x - np.eye(x.shape[1])))


def gaussian(m, n, seed):
    return np.random.default_rng(seed).standard_normal((m, n))


def sorted_desc(s):
    return np.sort(np.asarray(s))[::-1]


def match_columns(x, y):
    """Largest |cos| distance between matched columns of x and y after sign alignment."""
    signs = np.sign(np.sum(x * y, axis=0))
    signs[signs == 0] = 1.0
    return float(np.max(np.abs(x - y * signs)))
