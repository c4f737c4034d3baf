"""The blocked Cholesky kernels of the mixed path's QR (csrc/cholesky_kernels.hpp) on the CPU model of a workgroup,
no GPU needed: R^T R = A on the upper triangle for n not a multiple of 64 (identity padding) and over several tile
steps, the strictly lower triangle left untouched, and the first non-positive pivot reported exactly (info = 1 +
its row).  The model checks every barrier, wavefront operation (readlane, DPP) and buffer access."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "tests", "simt_model")
BUILD = os.path.join(MODEL, "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")


@pytest.fixture(scope="module")
def model():
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "cholesky_model")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-Wno-unknown-pragmas", "-Wno-psabi", "-I", MODEL, "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "jacobi-svd-cuda_amd", "csrc"),
                    os.path.join(MODEL, "cholesky_model.cpp"), "-o", out], check=True)
    return out


def _run(binary, a, ld):
    n = a.shape[0]
    n64 = (n + 63) // 64 * 64
    buf = np.full((n64, ld), -3.25)  # column-major: buf[j, i] = A(i, j); below-diagonal sentinel
    for j in range(n):
        buf[j, :j + 1] = a[:j + 1, j]
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([n, ld], dtype=np.int64).tofile(f)
            buf.tofile(f)
        r = subprocess.run([binary, fi, fo], capture_output=True, text=True, env=ENV, timeout=600)
        assert r.returncode == 0, r.stderr[-4000:]
        raw = np.fromfile(fo, dtype=np.uint8)
    out = raw[:buf.size * 8].view(np.float64).reshape(n64, ld)
    info = int(raw[buf.size * 8:].view(np.int64)[0])
    return out, info


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n))
    return x @ x.T / n + np.eye(n)


@pytest.mark.parametrize("n", [64, 150])
def test_factor_matches(model, n):
    a = _spd(n, n)
    out, info = _run(model, a, n + 64)
    assert info == 0
    r = np.triu(out[:n, :n].T)
    assert np.linalg.norm(r.T @ r - a) / np.linalg.norm(a) < 1e-14
    np.testing.assert_allclose(r, np.linalg.cholesky(a).T, rtol=0, atol=1e-12)
    lower = np.tril_indices(n, -1)
    assert np.all(out[:n, :n].T[lower] == -3.25)  # strictly lower triangle untouched


def test_first_bad_pivot(model):
    n = 130
    a = _spd(n, 4)
    a[100, 100] = -1.0  # leading minors up to order 100 stay positive definite: row 100 is the first bad pivot
    _, info = _run(model, a, n + 62)
    assert info == 101
