"""The FP64 dense kernels of the mixed path on cuda:0 (jsvd_gemm_f64, jsvd_trsm_right_upper_f64) against torch
float64 references, at the shapes the mixed path launches them with (4096 x 1024 B^T B with k split over several
workgroups, 1024^3 products, the 4096 x 1024 triangular solve) and ragged ones.  Tolerances are normwise, a few
hundred units of FP64 rounding: |D - D_ref| / (|alpha| |op(A)| |op(B)| + |beta| |C|) <= 1e-13 for the products,
|X - X_ref| / |X_ref| <= 1e-12 for the solve (R condition number about 10)."""
import numpy as np
import pytest

import jsvd

pytestmark = pytest.mark.gpu


def _cm(torch, x, ld):
    """Column-major flat storage of x (rows x cols) with leading dimension ld, padding NaN (never read)."""
    rows, cols = x.shape
    buf = torch.full((cols, ld), float("nan"), dtype=torch.float64, device="cuda")
    buf[:, :rows] = x.T
    return buf


def _from_cm(buf, rows):
    return buf[:, :rows].T


@pytest.mark.parametrize("ta,tb,m,n,k", [(0, 0, 300, 200, 100), (1, 0, 129, 257, 65), (0, 1, 64, 64, 64),
                                         (1, 1, 190, 70, 333)])
def test_gemm_transposes(torch_cuda, ta, tb, m, n, k):
    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(m + n + k)
    opa = torch.randn(m, k, dtype=torch.float64, device="cuda", generator=g)
    opb = torch.randn(k, n, dtype=torch.float64, device="cuda", generator=g)
    c = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g)
    a_store = opa.T.contiguous() if ta else opa
    b_store = opb.T.contiguous() if tb else opb
    lda, ldb = a_store.shape[0] + 3, b_store.shape[0] + 1
    a, b, cc = _cm(torch, a_store, lda), _cm(torch, b_store, ldb), _cm(torch, c, m + 2)
    d = torch.zeros(n, m + 5, dtype=torch.float64, device="cuda")
    jsvd.gemm(ta, tb, m, n, k, 0.75, a, lda, b, ldb, -1.25, cc, m + 2, d, m + 5)
    torch.cuda.synchronize()
    ref = 0.75 * opa @ opb - 1.25 * c
    scale = 0.75 * torch.linalg.norm(opa) * torch.linalg.norm(opb) + 1.25 * torch.linalg.norm(c)
    assert float(torch.linalg.norm(_from_cm(d, m) - ref) / scale) < 1e-13


def test_gram_upper_split_k_deterministic(torch_cuda):
    """B^T B of a 4096 x 1024 B on the upper tiles (k split across workgroups): matches, lower tiles untouched,
    bit-identical on every repetition (partials summed in split order), scratch reused across launches."""
    torch = torch_cuda
    m, n = 4096, 1024
    b = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    bb = _cm(torch, b, m)
    outs = []
    for rep in range(3):
        d = torch.full((n, n), 7.0, dtype=torch.float64, device="cuda")
        jsvd.gemm(True, False, n, n, m, 1.0, bb, m, bb, m, 0.0, None, n, d, n, jsvd.GEMM_UPPER_OUT)
        # another product in between reuses the split scratch
        e = torch.zeros(256, 256, dtype=torch.float64, device="cuda")
        jsvd.gemm(True, False, 256, 256, 4096, 1.0, bb, m, bb, m, 0.0, None, 256, e, 256, jsvd.GEMM_UPPER_OUT)
        outs.append(d.clone())
    torch.cuda.synchronize()
    got = _from_cm(outs[0], n)
    ref = b.T @ b
    tiles = torch.arange(n, device="cuda") // 64
    upper = tiles[:, None] <= tiles[None, :]
    err = torch.linalg.norm(torch.where(upper, got - ref, 0.0)) / torch.linalg.norm(ref)
    assert float(err) < 1e-13
    assert bool(torch.all(got[~upper] == 7.0))
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_mirror_symmetric_and_newton_schulz(torch_cuda):
    """V^T V in full (upper tiles mirrored), then 1.5 V - 0.5 V G with C read from V and D a separate buffer."""
    torch = torch_cuda
    n = 1024
    q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda",
                                       generator=torch.Generator(device="cuda").manual_seed(4)))
    v = q + 1e-6 * torch.randn(n, n, dtype=torch.float64, device="cuda")
    vv = _cm(torch, v, n)
    g = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    jsvd.gemm(True, False, n, n, n, 1.0, vv, n, vv, n, 0.0, None, n, g, n, jsvd.GEMM_UPPER_OUT | jsvd.GEMM_MIRROR)
    nxt = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    jsvd.gemm(False, False, n, n, n, -0.5, vv, n, g, n, 1.5, vv, n, nxt, n)
    torch.cuda.synchronize()
    gm = _from_cm(g, n)
    assert torch.equal(gm, gm.T)
    assert float(torch.linalg.norm(gm - v.T @ v) / n ** 0.5) < 1e-13
    ref = 1.5 * v - 0.5 * v @ (v.T @ v)
    assert float(torch.linalg.norm(_from_cm(nxt, n) - ref) / n ** 0.5) < 1e-13
    # one Newton-Schulz step squares the orthogonality error (here about 1e-3 -> 1e-7)
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    orth = _from_cm(nxt, n)
    e0 = float(torch.linalg.norm(v.T @ v - eye))
    e1 = float(torch.linalg.norm(orth.T @ orth - eye))
    assert e1 < e0 ** 2


def test_triangular_product_and_row_major(torch_cuda):
    torch = torch_cuda
    n, m = 1000, 3000
    gen = torch.Generator(device="cuda").manual_seed(5)
    r1 = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    r2 = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    lower = torch.tril(torch.ones(n, n, dtype=torch.bool, device="cuda"), -1)
    a1, a2 = _cm(torch, r1.masked_fill(lower, float("nan")), n), _cm(torch, r2.masked_fill(lower, float("nan")), n)
    d = torch.full((n, n), 3.0, dtype=torch.float64, device="cuda")
    jsvd.gemm(False, False, n, n, n, 1.0, a1, n, a2, n, 0.0, None, n, d, n,
              jsvd.GEMM_TRI_A | jsvd.GEMM_TRI_B | jsvd.GEMM_ZERO_LOWER)
    q = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=gen)
    u = torch.zeros(m, n, dtype=torch.float64, device="cuda")
    jsvd.gemm(False, False, m, n, n, 1.0, _cm(torch, q, m), m, _cm(torch, r2, n), n, 0.0, None, n, u, n,
              jsvd.GEMM_ROW_MAJOR_CD)
    torch.cuda.synchronize()
    got = _from_cm(d, n)
    ref = torch.triu(r1) @ torch.triu(r2)
    assert float(torch.linalg.norm(got - ref) / (torch.linalg.norm(torch.triu(r1)) * torch.linalg.norm(torch.triu(r2)))) < 1e-13
    assert bool(torch.all(got[lower] == 0.0))
    assert float(torch.linalg.norm(u - q @ r2) / (torch.linalg.norm(q) * torch.linalg.norm(r2))) < 1e-13


@pytest.mark.parametrize("rows,n", [(4096, 1024), (1000, 1100), (77, 50)])
def test_trsm_right_upper(torch_cuda, rows, n):
    torch = torch_cuda
    gen = torch.Generator(device="cuda").manual_seed(rows + n)
    r = 0.1 * torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen) / n ** 0.5
    r = torch.triu(r) + torch.diag(1.0 + torch.rand(n, dtype=torch.float64, device="cuda", generator=gen))
    stored = r + torch.tril(torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda"), -1)
    b = torch.randn(rows, n, dtype=torch.float64, device="cuda", generator=gen)
    bb = _cm(torch, b, rows + 3)
    jsvd.trsm_right_upper(bb, rows + 3, rows, n, _cm(torch, stored, n + 1), n + 1)
    torch.cuda.synchronize()
    x = _from_cm(bb, rows)
    ref = torch.linalg.solve_triangular(r, b, upper=True, left=False)
    assert float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref)) < 1e-12
    assert bool(torch.all(torch.isnan(bb[:, rows:])))  # padding rows untouched
