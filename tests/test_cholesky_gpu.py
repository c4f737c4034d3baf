"""Tiled FP64 Cholesky (jsvd_cholesky_f64, the factorisation inside the mixed path's Cholesky QR) on cuda:0.

Checks: R^T R reproduces A to a relative Frobenius residual <= 1e-15 n; R equals LAPACK's upper factor
(numpy.linalg.cholesky, transposed) to 1e-12 relative on well-conditioned inputs; the strictly lower triangle
and nothing beyond the padding is touched; a matrix that is not positive definite reports the failing row.
"""
import numpy as np
import pytest

import jsvd

pytestmark = pytest.mark.gpu


def _spd(n, seed, shift=1.0):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n + 7, n))
    return b.T @ b / n + shift * np.eye(n)


def _run(torch, a_sym, lda_extra=0):
    n = a_sym.shape[0]
    n64 = (n + 63) // 64 * 64
    lda = n64 + lda_extra
    store = np.full((n64, lda), 7.0)          # store[c, r] = A[r][c]: column-major lda x n64
    store[:n, :n] = a_sym.T
    d = torch.from_numpy(store).cuda()
    info = torch.full((1,), -5, dtype=torch.int32, device="cuda")
    jsvd.cholesky(d, n, info)
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    return out, int(info.cpu().item()), store


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 1024])
def test_cholesky_matches_lapack(torch_cuda, n):
    a = _spd(n, n)
    out, info, before = _run(torch_cuda, a, lda_extra=3)
    assert info == 0
    r = np.triu(out[:n, :n].T)
    assert np.linalg.norm(r.T @ r - a) / np.linalg.norm(a) < 1e-15 * max(n, 8)
    r_ref = np.linalg.cholesky(a).T
    assert np.max(np.abs(r - r_ref)) / np.max(np.abs(r_ref)) < 1e-12
    # strictly lower triangle of the n x n block untouched, rows beyond the padding untouched
    low = np.tril(np.ones((n, n), dtype=bool), -1)
    assert np.array_equal(out[:n, :n].T[low], before[:n, :n].T[low])
    n64 = (n + 63) // 64 * 64
    assert np.array_equal(out[:, n64:], before[:, n64:])
    # padding set to the identity
    if n64 > n:
        pad = out[n:n64, n:n64]
        assert np.array_equal(pad, np.eye(n64 - n))


def test_cholesky_ill_conditioned(torch_cuda):
    """kappa(A) = 1e12 (the shifted Cholesky QR's first pass sees such Grams): R^T R still reproduces A."""
    rng = np.random.default_rng(3)
    n = 300
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * np.logspace(0, -12, n)) @ q.T
    a = (a + a.T) / 2 + 1e-14 * np.eye(n)
    out, info, _ = _run(torch_cuda, a)
    assert info == 0
    r = np.triu(out[:n, :n].T)
    assert np.linalg.norm(r.T @ r - a) / np.linalg.norm(a) < 1e-14


def test_cholesky_reports_indefinite(torch_cuda):
    n = 130
    a = _spd(n, 4)
    a[100, 100] = -1.0  # leading minors up to order 100 stay positive definite: row 100 is the first bad pivot
    _, info, _ = _run(torch_cuda, a)
    assert info == 101


def test_cholesky_bad_arguments(torch_cuda):
    import torch
    d = torch.zeros((64, 64), dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        jsvd.cholesky(d, 65, info)   # needs 128 columns and lda >= 128
    with pytest.raises(ValueError):
        jsvd.cholesky(d[:, :32].contiguous(), 10, info)   # lda 32 < 64
