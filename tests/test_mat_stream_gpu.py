"""Streamed *.mat solve (jsvd_mat_stream_solve): mmap -> pinned double buffer -> HBM -> device decode -> solve.

The streamed results must equal, bit for bit, the same path's solve of the same matrix through the host entry
point (the kernels are deterministic and the device decode is an exact byte swap); output records come in
input order as U (m x n), sigma (1 x n), V (n x n), the reference pipeline's packets (src/pipeline.cpp:711-739).
"""
import os

import numpy as np
import pytest

import jsvd
from helpers import gaussian

pytestmark = pytest.mark.gpu


def _solve_direct(a, path, cfg):
    res, _ = jsvd.solve(a, cfg, path=path)
    return res


@pytest.mark.parametrize("path", [jsvd.PATH_PLAIN_F64, jsvd.PATH_MIXED])
def test_stream_equals_direct_solves(torch_cuda, tmp_path, path):
    shapes = [(300, 100), (64, 64), (500, 40), (20, 5), (300, 100)]
    mats = [gaussian(m, n, 10 + k) for k, (m, n) in enumerate(shapes)]
    src, dst = str(tmp_path / "in.mat"), str(tmp_path / "out.mat")
    jsvd.write_mat(src, mats)
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    rep = jsvd.mat_stream_solve(src, dst, cfg, path=path)
    assert rep.records == len(mats)
    assert rep.payload_bytes == sum(m * n * 8 for m, n in shapes)
    out = jsvd.read_mat(dst)
    assert len(out) == 3 * len(mats)
    for k, a in enumerate(mats):
        m, n = a.shape
        u, s, v = out[3 * k], out[3 * k + 1], out[3 * k + 2]
        assert u.shape == (m, n) and s.shape == (1, n) and v.shape == (n, n)
        # records narrower than 33 columns take the plain path
        ref = _solve_direct(a, path if n >= 33 else jsvd.PATH_PLAIN_F64, cfg)
        assert np.array_equal(u, ref.u), k
        assert np.array_equal(s[0], ref.sigma), k
        assert np.array_equal(v, ref.v), k


def test_stream_without_output_and_empty_file(torch_cuda, tmp_path):
    src = str(tmp_path / "in.mat")
    jsvd.write_mat(src, [gaussian(128, 64, 1)])
    rep = jsvd.mat_stream_solve(src, None, jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60))
    assert rep.records == 1 and rep.sweeps_fp64_last > 0
    empty = str(tmp_path / "empty.mat")
    open(empty, "wb").close()
    dst = str(tmp_path / "out.mat")
    rep = jsvd.mat_stream_solve(empty, dst)
    assert rep.records == 0 and os.path.getsize(dst) == 0


def test_stream_rejects_bad_files(torch_cuda, tmp_path):
    good = str(tmp_path / "good.mat")
    jsvd.write_mat(good, [gaussian(16, 8, 2), gaussian(16, 8, 3)])
    data = open(good, "rb").read()
    trunc = str(tmp_path / "trunc.mat")
    open(trunc, "wb").write(data[:-8])
    with pytest.raises(RuntimeError, match="Truncated \\*.mat payload"):
        jsvd.mat_stream_solve(trunc, None)
    open(trunc, "wb").write(data[:16 + 16 * 8 * 8 + 5])
    with pytest.raises(RuntimeError, match="Truncated \\*.mat header"):
        jsvd.mat_stream_solve(trunc, None)
    wide = str(tmp_path / "wide.mat")
    jsvd.write_mat(wide, [gaussian(4, 9, 4)])
    with pytest.raises(ValueError):
        jsvd.mat_stream_solve(wide, None)
