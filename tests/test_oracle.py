"""Pins the CPU oracle (oracle/jacobi_oracle.c) before results are checked against it.

The reference ships no tests or golden vectors, so the oracle is pinned by:
  * hand-derived known answers (round-robin schedules n=4, n=5; 2x2 SVD) in tests/golden/closed_form.json,
  * closed-form properties of the rotation (src/kernels.cu:485-521),
  * LAPACK dgesvj fixtures (tests/golden/dgesvj_small.npz), an independent FP64 Jacobi,
  * invariants rho_A, rho_U, rho_V (tex/sections/03-gpu-parallel-reduction.tex:29-43),
  * special inputs whose SVD is known exactly (diagonal, orthogonal columns, zero columns).
Each check is chosen so that a plausible slip (a sign flip of s, a transposed V, a dropped term of
tau, a wrong pair orientation) fails it; test_planted_* plant such slips on the CPU to show that.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from helpers import gaussian, match_columns, rho_a, rho_orth, sorted_desc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def closed():
    with open(os.path.join(GOLDEN, "closed_form.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def fixtures():
    return np.load(os.path.join(GOLDEN, "dgesvj_small.npz"))


# ---------------------------------------------------------------------------------------------- schedule

def test_schedule_known_answers(closed):
    assert [list(map(list, r)) for r in oracle.round_robin(4)] == closed["schedule_n4"]
    assert [list(map(list, r)) for r in oracle.round_robin(5)] == closed["schedule_n5"]


@pytest.mark.parametrize("n", list(range(2, 40)) + [64, 65, 127, 256])
def test_schedule_properties(n):
    rounds = oracle.round_robin(n)
    assert len(rounds) == (n - 1 if n % 2 == 0 else n)
    seen = set()
    for r in rounds:
        cols = [c for pq in r for c in pq]
        assert len(cols) == len(set(cols)), "pairs of a round must be disjoint"
        assert len(r) == n // 2
        for p, q in r:
            key = (min(p, q), max(p, q))
            assert key not in seen
            seen.add(key)
    assert len(seen) == n * (n - 1) // 2, "every unordered pair once per sweep"


def test_schedule_orientation_mixed():
    # about half of the pairs have p > q; the rotation's roles of c and s depend on it
    flipped = sum(p > q for r in oracle.round_robin(64) for p, q in r)
    assert 0 < flipped < 64 * 63 // 2


# ---------------------------------------------------------------------------------------------- rotation

def _rotated_offdiag(app, aqq, apq, c, s):
    # Gram of [c a_p - s a_q, s a_p + c a_q]: off-diagonal = (c^2 - s^2) apq + c s (app - aqq)
    return (c * c - s * s) * apq + c * s * (app - aqq)


def _check_rotation(app, aqq, apq, rotates, c, s):
    assert rotates == 1
    assert abs(c * c + s * s - 1.0) < 1.2e-15
    t = s / c
    assert abs(t) <= 1.0 + 1e-15
    scale = max(abs(app), abs(aqq), abs(apq))
    assert abs(_rotated_offdiag(app, aqq, apq, c, s)) <= 2e-15 * scale


def test_rotation_closed_forms():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = rng.standard_normal(6)
        y = rng.standard_normal(6) * 10.0 ** rng.uniform(-4, 4)
        app, aqq, apq = float(x @ x), float(y @ y), float(x @ y)
        r, c, s = oracle.rotation(app, aqq, apq, 1e-15)
        _check_rotation(app, aqq, apq, r, c, s)


def test_rotation_special_cases():
    r, c, s = oracle.rotation(2.0, 2.0, 0.5, 1e-15)   # alpha == beta -> tau = 0 -> t = 1
    assert r == 1 and c == 1.0 / math.sqrt(2.0) and abs(s - 1.0 / math.sqrt(2.0)) < 1e-16
    assert oracle.rotation(1.0, 3.0, 0.0, 1e-15) == (0, 1.0, 0.0)       # gamma == 0: no rotation
    assert oracle.rotation(1e-320, 1e-320, 1e-301, 1e-15) == (0, 1.0, 0.0)  # |gamma| <= 1e-300
    # exactly at the threshold |gamma| == eps*sqrt(alpha*beta): skip (strict inequality)
    assert oracle.rotation(4.0, 9.0, 0.5 * 6.0, 0.5) == (0, 1.0, 0.0)
    assert oracle.rotation(4.0, 9.0, 0.5 * 6.0 * (1 + 1e-15), 0.5)[0] == 1
    # negative tau branch: t = -1 / (-tau + sqrt(1 + tau^2))
    r, c, s = oracle.rotation(5.0, 1.0, 1.0, 1e-15)
    tau = (1.0 - 5.0) / 2.0
    assert r == 1 and abs(s / c - (-1.0 / (-tau + math.sqrt(1 + tau * tau)))) < 1e-15


def test_planted_rotation_sign_flip_is_caught():
    app, aqq, apq = 3.0, 1.0, 0.7
    r, c, s = oracle.rotation(app, aqq, apq, 1e-15)
    with pytest.raises(AssertionError):
        _check_rotation(app, aqq, apq, r, c, -s)   # a flipped sine leaves a large off-diagonal


# ---------------------------------------------------------------------------------------------- full solve

def test_svd_2x2_closed_form(closed):
    a = np.array(closed["svd2x2_a"])
    r = oracle.svd(a, eps=1e-15, max_sweeps=30)
    assert np.allclose(sorted_desc(r.sigma), closed["svd2x2_sigma_desc"], rtol=1e-15, atol=0)
    assert rho_a(a, r.u, r.sigma, r.v) < 1e-15


@pytest.mark.parametrize("case", ["g4x3", "g8x8", "g16x7", "g33x17", "g64x64", "g97x40", "graded48x24"])
def test_svd_matches_dgesvj_fixture(fixtures, case):
    a = fixtures[f"{case}_a"]
    s_ref = fixtures[f"{case}_sigma"]
    r = oracle.svd(a, eps=1e-15, max_sweeps=60)
    order = np.argsort(-r.sigma, kind="stable")
    s = r.sigma[order]
    assert np.max(np.abs(s - s_ref) / s_ref) < 2e-13
    # singular vectors: matched, sign-aligned columns (sigma are well separated in these fixtures)
    gaps = np.min(np.abs(np.diff(s_ref)) / s_ref[:-1])
    assert gaps > 1e-6
    assert match_columns(r.v[:, order], fixtures[f"{case}_v"]) < 1e-10
    assert match_columns(r.u[:, order], fixtures[f"{case}_u"]) < 1e-10
    assert rho_a(a, r.u, r.sigma, r.v) < 5e-14
    assert rho_orth(r.v) < 1e-13
    assert rho_orth(r.u) < 1e-13


def test_planted_transposed_v_is_caught(fixtures):
    a = fixtures["g33x17_a"]
    r = oracle.svd(a, eps=1e-15, max_sweeps=60)
    assert rho_a(a, r.u, r.sigma, r.v) < 1e-14
    assert rho_a(a, r.u, r.sigma, r.v.T) > 1e-2   # returning V^T instead of V fails rho_A


def test_svd_diagonal_input_one_sweep():
    d = np.array([3.0, -1.0, 7.0, 0.5])
    a = np.zeros((6, 4))
    a[np.arange(4), np.arange(4)] = d
    r = oracle.svd(a, eps=1e-15, max_sweeps=30)
    assert r.sweeps == 1 and list(r.rot_per_sweep) == [0]
    assert np.array_equal(r.sigma, np.abs(d))            # Jacobi column order, not sorted
    assert np.array_equal(r.v, np.eye(4))
    assert np.array_equal(np.abs(r.u[:4]), np.eye(4))


def test_svd_orthogonal_columns_exact():
    q, _ = np.linalg.qr(gaussian(40, 12, 3))
    sig = np.linspace(1, 12, 12)
    a = q * sig
    r = oracle.svd(a, eps=1e-9, max_sweeps=30)
    assert np.max(np.abs(r.sigma - sig) / sig) < 1e-14


def test_svd_zero_column_cutoff():
    a = gaussian(10, 4, 5)
    a[:, 2] = 0.0
    r = oracle.svd(a, eps=1e-12, max_sweeps=30)
    k = int(np.argmin(r.sigma))
    assert r.sigma[k] == 0.0 and np.all(r.u[:, k] == 0.0)   # U column zeroed when sigma <= eps


def test_svd_nonconvergence_is_not_error():
    a = gaussian(64, 32, 6)
    r = oracle.svd(a, eps=1e-15, max_sweeps=2)
    assert r.sweeps == 2 and np.all(r.rot_per_sweep > 0)


def test_svd_single_column_and_invalid():
    a = gaussian(7, 1, 7)
    r = oracle.svd(a, eps=1e-9, max_sweeps=5)
    assert r.sweeps == 1 and abs(r.sigma[0] - np.linalg.norm(a)) < 1e-14
    with pytest.raises(ValueError):
        oracle.svd(gaussian(3, 4, 0))
    with pytest.raises(ValueError):
        oracle.svd(gaussian(4, 3, 0), eps=0.0)
    with pytest.raises(ValueError):
        oracle.svd(gaussian(4, 3, 0), max_sweeps=0)


def test_svd_config1_shape_against_lapack():
    """512 x 256 N(0,1) seed 0 at eps 1e-15 (the plain-FP64 CPU-runnable workload)."""
    a = gaussian(512, 256, 0)
    r = oracle.svd(a, eps=1e-15, max_sweeps=30, threads=4)
    _, s_ref, _, _ = oracle.dgesvj(a)
    assert np.max(np.abs(sorted_desc(r.sigma) - s_ref) / s_ref) < 1e-12
    assert 9 <= r.sweeps <= 14
    assert rho_a(a, r.u, r.sigma, r.v) < 1e-13
    assert rho_orth(r.u) < 1e-12 and rho_orth(r.v) < 1e-11
    # the parallel (OpenMP) solve equals the serial one bit for bit: pairs of a round are independent
    r1 = oracle.svd(a, eps=1e-15, max_sweeps=30, threads=1)
    assert np.array_equal(r.sigma, r1.sigma) and np.array_equal(r.v, r1.v) and r.sweeps == r1.sweeps


def test_u_sigma_closed_form():
    a = np.diag([2.0, 0.0, 5.0])
    u, s = oracle.u_sigma(a, 1e-9)
    assert np.array_equal(s, [2.0, 0.0, 5.0])
    assert np.array_equal(u, np.diag([1.0, 0.0, 1.0]))


def test_pair_stats_brute_force():
    a = gaussian(50, 6, 8)
    for p, q in [(0, 5), (3, 1), (2, 2)]:
        x, y, z = oracle.pair_stats(a, p, q)
        assert abs(x - a[:, p] @ a[:, p]) < 1e-12 and abs(y - a[:, q] @ a[:, q]) < 1e-12
        assert abs(z - a[:, p] @ a[:, q]) < 1e-12


def test_gram_space_sweep_reference():
    """The numpy Gram-space sweep that the block inner step is checked against: each sweep is an orthogonal
    similarity (Q^T G Q equals the swept G, Q^T Q = I), a few sweeps diagonalise G to its eigenvalues
    (numpy.linalg.eigvalsh), and a planted sign flip in the rotation breaks the annihilation."""
    from helpers import gram_space_sweep, block_rotation
    rng = np.random.default_rng(3)
    x = rng.standard_normal((120, 16)) * 10.0 ** (-2 * np.arange(16) / 15)
    g = x.T @ x
    q, rot, g1 = gram_space_sweep(g, 1e-15)
    assert rot == 16 * 15 // 2
    assert np.linalg.norm(q.T @ q - np.eye(16)) < 1e-14
    assert np.max(np.abs(q.T @ g @ q - g1)) < 1e-14 * np.max(np.abs(g))
    for _ in range(6):
        _, rot, g1 = gram_space_sweep(g1, 1e-15)
    assert rot == 0
    ev = np.linalg.eigvalsh(g)
    assert np.max(np.abs(np.sort(np.diag(g1)) - ev) / ev) < 1e-12
    # the rotation annihilates its 2x2 off-diagonal entry; a flipped sine does not
    app, aqq, apq = 3.0, 1.0, 0.7
    c, s, _ = block_rotation(app, aqq, apq, 1e-15)
    assert abs(apq * (c * c - s * s) + (app - aqq) * c * s) < 1e-15
    assert abs(apq * (c * c - s * s) + (app - aqq) * c * (-s)) > 0.1


def test_gram_space_cross_sweep_reference():
    """The cross ordering of the numpy Gram-space sweep: every pair (i, h + j) across the two halves is rotated
    exactly once per sweep (the first sweep of a full-rank G rotates all h^2 of them), each sweep is an orthogonal
    similarity, and alternating it with full sweeps of each half (the intra-block step) diagonalises G to its
    eigenvalues; pairs inside a half are never touched by the cross sweep."""
    from helpers import gram_space_sweep
    rng = np.random.default_rng(4)
    n, h = 16, 8
    x = rng.standard_normal((120, n)) * 10.0 ** (-2 * np.arange(n) / (n - 1))
    g = x.T @ x
    q, rot, g1 = gram_space_sweep(g, 1e-15, cross=True)
    assert rot == h * h
    assert np.linalg.norm(q.T @ q - np.eye(n)) < 1e-14
    assert np.max(np.abs(q.T @ g @ q - g1)) < 1e-14 * np.max(np.abs(g))
    gi = g1
    for _ in range(8):
        for lo in (0, h):
            qi, _, _ = gram_space_sweep(gi[lo:lo + h, lo:lo + h], 1e-15)
            full = np.eye(n)
            full[lo:lo + h, lo:lo + h] = qi
            gi = full.T @ gi @ full
        _, rot, gi = gram_space_sweep(gi, 1e-15, cross=True)
    assert rot == 0
    ev = np.linalg.eigvalsh(g)
    assert np.max(np.abs(np.sort(np.diag(gi)) - ev) / ev) < 1e-11
