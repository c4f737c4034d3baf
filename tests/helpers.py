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


def circle_at(slot, rnd, n):
    """Column at position `slot` of round `rnd` of the round-robin circle of n (even) columns."""
    return 0 if slot == 0 else 1 + ((slot - 1 - rnd) % (n - 1))


def block_rotation(app, aqq, apq, eps):
    """Rotation of the block paths' inner sweep: the reference skip test, then t = sign(d) g / (|d| + hypot(d, g))
    and c = 1 / sqrt(1 + t^2) with d = aqq - app, g = 2 apq (the reference's tau form, rearranged)."""
    rhs = eps * np.sqrt(max(app, 0.0) * max(aqq, 0.0))
    if not (abs(apq) > rhs and abs(apq) > 1e-300):
        return 1.0, 0.0, False
    d, g = aqq - app, 2.0 * apq
    den = abs(d) + np.hypot(d, g)
    t = (g if d >= 0 else -g) / den
    c = 1.0 / np.sqrt(1.0 + t * t)
    return c, t * c, True


def gram_space_sweep(g, eps, cross=False):
    """One cyclic Jacobi sweep on the symmetric n x n Gram matrix g in round-robin order (pair u of a round =
    circle positions u and n-1-u), or with cross=True only over the (n/2)^2 pairs across the two halves (round r
    pairs column u with column n/2 + (u + r) mod n/2, n/2 rounds): g <- J^T g J per rotation, Q accumulates J.
    Returns (Q, rotations, g)."""
    g = np.array(g, dtype=np.float64)
    n = g.shape[0]
    h = n // 2
    q = np.eye(n)
    rotations = 0
    for rnd in range(h if cross else n - 1):
        params = []
        for u in range(n // 2):
            if cross:
                p, r = u, h + (u + rnd) % h
            else:
                p, r = circle_at(u, rnd, n), circle_at(n - 1 - u, rnd, n)
            c, s, rot = block_rotation(g[p, p], g[r, r], g[p, r], eps)
            rotations += rot
            params.append((p, r, c, s))
        j = np.eye(n)
        for p, r, c, s in params:
            # new column p = c col_p - s col_r, new column r = s col_p + c col_r
            j[p, p], j[r, p], j[p, r], j[r, r] = c, -s, s, c
        g = j.T @ g @ j
        q = q @ j
    return q, rotations, g
