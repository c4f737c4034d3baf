"""The FP64 dense kernels of the mixed path (csrc/dense_kernels.hpp) on the CPU model of a workgroup, no GPU
needed: the product kernel under every flag it is launched with (upper-triangle tiles with and without the mirror,
triangular operands, zeroed lower triangle, row-major output, all four transpose combinations, ragged sizes,
scalar and 16-byte operand loads, k split over several workgroups and reduced by the second launch) and the
triangular solve (n not a
multiple of 16, and past one 1024-column range, so the range update runs), against numpy.  The model checks
every barrier, wavefront operation and buffer access (AddressSanitizer + UBSan, exact-size buffers)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "tests", "simt_model")
BUILD = os.path.join(MODEL, "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")

UPPER_OUT, MIRROR, TRI_A, TRI_B, ZERO_LOWER, ROW_MAJOR_CD = 1, 2, 4, 8, 16, 32


@pytest.fixture(scope="module")
def model():
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "dense_model")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-Wno-unknown-pragmas", "-Wno-psabi", "-I", MODEL, "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "jacobi-svd-cuda_amd", "csrc"),
                    os.path.join(MODEL, "dense_model.cpp"), "-o", out], check=True)
    return out


def _run(binary, payload, out_count, out_dtype=np.float64, extra_i64=0):
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            for part in payload:
                f.write(part.tobytes())
        r = subprocess.run([binary, fi, fo], capture_output=True, text=True, env=ENV, timeout=600)
        assert r.returncode == 0, r.stderr[-4000:]
        raw = np.fromfile(fo, dtype=np.uint8)
    vals = raw[:out_count * 8].view(np.float64)
    rest = raw[out_count * 8:].view(np.int64)
    return vals, rest, r.stderr


def _gemm(binary, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, d0, ldd, flags, cus=16,
          work_elems=1 << 20):
    hdr = np.array([0, ta, tb, m, n, k, lda, ldb, ldc, ldd, flags, cus, work_elems,
                    a.size, b.size, c.size, d0.size], dtype=np.int64)
    d, _, log = _run(binary, [hdr, np.array([alpha, beta]), a, b, c, d0], d0.size)
    return d, log


def _colmajor(x, ld):
    """x (rows x cols) into a column-major buffer with leading dimension ld."""
    buf = np.full(ld * x.shape[1], np.nan)
    for j in range(x.shape[1]):
        buf[j * ld:j * ld + x.shape[0]] = x[:, j]
    return buf


def _from_colmajor(buf, rows, cols, ld):
    return np.stack([buf[j * ld:j * ld + rows] for j in range(cols)], axis=1)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("even", [False, True])  # odd leading dimensions: scalar loads; even: 16-byte loads
def test_gemm_general(model, ta, tb, even):
    rng = np.random.default_rng(1 + 2 * ta + tb)
    m, n, k = 101, 70, 45
    opa, opb = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    a_store, b_store = (opa.T if ta else opa), (opb.T if tb else opb)

    def ld(rows):  # at least rows + 1, of the chosen parity
        return rows + 1 + ((rows + 1) % 2 != 0) if even else rows + 1 + ((rows + 1) % 2 == 0)

    lda, ldb = ld(a_store.shape[0]), ld(b_store.shape[0])
    c = rng.standard_normal((m, n))
    d, log = _gemm(model, ta, tb, m, n, k, 1.3, _colmajor(a_store, lda), lda, _colmajor(b_store, ldb), ldb, -0.7,
                   _colmajor(c, m + 2), m + 2, np.zeros((m + 1) * n), m + 1, 0)
    assert f"vec {int(lda % 2 == 0 and ldb % 2 == 0)}" in log, log
    got = _from_colmajor(d, m, n, m + 1)
    np.testing.assert_allclose(got, 1.3 * opa @ opb - 0.7 * c, rtol=0, atol=1e-12)


def test_gemm_upper_symmetric_split_k(model):
    """B^T B with only the upper tiles written, k split seven ways; untouched tiles keep their contents."""
    rng = np.random.default_rng(5)
    m, n = 1000, 130
    b = rng.standard_normal((m, n))
    d0 = np.full(n * n, 7.0)
    d, log = _gemm(model, 1, 0, n, n, m, 1.0, _colmajor(b, m), m, _colmajor(b, m), m, 0.0, np.zeros(1), 1,
                   d0, n, UPPER_OUT, cus=64)
    assert "splits 7" in log, log
    got = _from_colmajor(d, n, n, n)
    want = b.T @ b
    for ti in range(3):
        for tj in range(3):
            blk = (slice(64 * ti, 64 * ti + 64), slice(64 * tj, 64 * tj + 64))
            if ti <= tj:
                np.testing.assert_allclose(got[blk], want[blk], rtol=0, atol=1e-10)
            else:
                assert np.all(got[blk] == 7.0)


def test_gemm_mirror_gives_full_symmetric(model):
    rng = np.random.default_rng(6)
    n = 200
    v = rng.standard_normal((n, n))
    d, log = _gemm(model, 1, 0, n, n, n, 1.0, _colmajor(v, n), n, _colmajor(v, n), n, 0.0, np.zeros(1), 1,
                   np.zeros(n * n), n, UPPER_OUT | MIRROR, cus=16)
    got = _from_colmajor(d, n, n, n)
    np.testing.assert_allclose(got, v.T @ v, rtol=0, atol=1e-11)
    assert np.array_equal(got, got.T)


def test_gemm_triangular_operands_zero_lower(model):
    """R_pass R_acc: both operands upper triangular with garbage below the diagonal, result upper."""
    rng = np.random.default_rng(7)
    n = 150
    r1 = rng.standard_normal((n, n))
    r2 = rng.standard_normal((n, n))
    lower = np.tril_indices(n, -1)
    r1[lower] = np.nan  # never read
    r2[lower] = np.nan
    d, log = _gemm(model, 0, 0, n, n, n, 1.0, _colmajor(r1, n), n, _colmajor(r2, n), n, 0.0, np.zeros(1), 1,
                   np.full(n * n, 3.0), n, TRI_A | TRI_B | ZERO_LOWER)
    assert "splits 1" in log
    got = _from_colmajor(d, n, n, n)
    np.testing.assert_allclose(got, np.triu(np.nan_to_num(r1)) @ np.triu(np.nan_to_num(r2)), rtol=0, atol=1e-12)
    assert np.all(np.tril(got, -1) == 0.0)


def test_gemm_row_major_output(model):
    """U = Q V_M written row-major (the final product of the mixed path)."""
    rng = np.random.default_rng(8)
    m, n = 190, 70
    q, vm = rng.standard_normal((m, n)), rng.standard_normal((n, n))
    d, _ = _gemm(model, 0, 0, m, n, n, 1.0, _colmajor(q, m + 6), m + 6, _colmajor(vm, n), n, 0.0, np.zeros(1), 1,
                 np.zeros(m * n), n, ROW_MAJOR_CD)
    np.testing.assert_allclose(d.reshape(m, n), q @ vm, rtol=0, atol=1e-12)


def test_gemm_newton_schulz_form_in_place(model):
    """nxt = 1.5 cur - 0.5 cur G with C aliasing D's role (separate buffers here, C read, D written)."""
    rng = np.random.default_rng(9)
    n = 96
    cur = np.linalg.qr(rng.standard_normal((n, n)))[0] + 1e-6 * rng.standard_normal((n, n))
    g = cur.T @ cur
    d, _ = _gemm(model, 0, 0, n, n, n, -0.5, _colmajor(cur, n), n, _colmajor(g, n), n, 1.5, _colmajor(cur, n), n,
                 np.zeros(n * n), n, 0)
    np.testing.assert_allclose(_from_colmajor(d, n, n, n), 1.5 * cur - 0.5 * cur @ g, rtol=0, atol=1e-13)


def _trsm(binary, b, ldb, r, ldr, cus=16, padded=False):
    rows, n = b.shape
    r_cols = r.shape[1]
    hdr = np.array([1, rows, n, ldb, ldr, ldb * n, ldr * r_cols, cus, 1 << 20, int(padded)], dtype=np.int64)
    out, _, _ = _run(binary, [hdr, _colmajor(b, ldb), _colmajor(r, ldr)], ldb * n)
    return _from_colmajor(out, rows, n, ldb)


@pytest.mark.parametrize("rows,n", [(70, 50), (20, 1030)])
def test_trsm_right_upper(model, rows, n):
    rng = np.random.default_rng(rows + n)
    r = rng.standard_normal((n, n)) * 0.1
    r[np.diag_indices(n)] = 1.0 + rng.random(n)   # well conditioned upper part
    garbage_lower = r.copy()
    garbage_lower[np.tril_indices(n, -1)] = np.nan  # the strictly lower triangle must not be read
    b = rng.standard_normal((rows, n))
    x = _trsm(model, b, rows + 3, garbage_lower, n + 1)
    want = np.linalg.solve(np.triu(r).T, b.T).T
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-11 * np.max(np.abs(want)))


def test_trsm_right_upper_prepadded(model):
    """R passed in place, padded as the Cholesky factorisation leaves it (identity up to the next multiple of 64,
    strictly lower triangle garbage), as the mixed path calls the solve."""
    rng = np.random.default_rng(17)
    rows, n = 40, 100
    np64 = 128
    r = rng.standard_normal((n, n)) * 0.1
    r[np.diag_indices(n)] = 1.0 + rng.random(n)
    stored = np.eye(np64 + 3, np64)
    stored[:n, :n] = r
    stored[np.tril_indices(np64, -1)] = np.nan
    b = rng.standard_normal((rows, n))
    x = _trsm(model, b, rows, stored, np64 + 3, padded=True)
    want = np.linalg.solve(np.triu(r).T, b.T).T
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-11 * np.max(np.abs(want)))
