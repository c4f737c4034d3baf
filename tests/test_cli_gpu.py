"""The CLI end to end on cuda:0: a two-record *.mat file through the default (plain FP64, reference semantics)
pipeline and through --solver mixed; outputs equal the library's direct solves bit for bit, the JSON report
carries the reference's fields."""
import json
import os
import subprocess

import numpy as np
import pytest

import jsvd
from helpers import gaussian, sorted_desc

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(__file__), "..", "jacobi-svd-cuda_amd", "lib", "jacobi-svd")


def _run(*args):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout, p.stderr


def test_cli_plain_and_mixed(torch_cuda, tmp_path):
    mats = [gaussian(64, 40, 1), gaussian(100, 50, 2)]
    src = tmp_path / "in.mat"
    jsvd.write_mat(str(src), mats)
    rc, out, err = _run(src, "--epsilon", "1e-14", "--max-sweeps", "60", "--json-report")
    assert rc == 0, err
    assert "Pipeline completed." in out
    report = json.loads(out[out.index("{"):])
    assert report["testcase_count"] == 2 and report["emitted_matrix_count"] == 6
    assert report["total_sweeps"] > 0 and report["layout_transpose_mode"] == "auto"
    got = jsvd.read_mat(str(tmp_path / "in.svd.mat"))
    assert len(got) == 6
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    for k, a in enumerate(mats):
        ref = jsvd.one_sided_jacobi_svd(a, cfg)
        assert np.array_equal(got[3 * k], ref.u)
        assert np.array_equal(got[3 * k + 1].ravel(), ref.sigma)
        assert np.array_equal(got[3 * k + 2], ref.v)
    # the mixed solver through the streamed *.mat path: same bytes as the library's streamed solve
    rc, out, err = _run(src, tmp_path / "mixed.mat", "--solver", "mixed", "--epsilon", "1e-14", "--max-sweeps",
                        "60", "-q")
    assert rc == 0, err
    assert out == ""
    jsvd.mat_stream_solve(str(src), str(tmp_path / "lib.mat"), cfg, path=jsvd.PATH_MIXED)
    assert (tmp_path / "mixed.mat").read_bytes() == (tmp_path / "lib.mat").read_bytes()
    mixed = jsvd.read_mat(str(tmp_path / "mixed.mat"))
    for k, a in enumerate(mats):
        s_ref = np.linalg.svd(a, compute_uv=False)
        assert np.max(np.abs(sorted_desc(mixed[3 * k + 1].ravel()) - s_ref)) / s_ref[0] < 1e-12
    # an existing output needs --force
    rc, _, err = _run(src, tmp_path / "mixed.mat", "--solver", "mixed")
    assert rc == 2 and "--force" in err
