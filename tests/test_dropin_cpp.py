"""Runs the C++ drop-in test driver (tests/cpp/test_dropin.cpp, built by build()) in its host-only mode
here and in its GPU mode on cuda:0: the reference's C++ entry points (kernels.hpp, io.hpp, pipeline.hpp)
called as a user of those headers would."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "jacobi-svd-cuda_amd", "lib", "test_dropin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _run(mode, tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("test_dropin is not built; run __graft_entry__.build()")
    res = subprocess.run([BIN, mode, str(tmp_path), GOLDEN], capture_output=True, text=True, timeout=300)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "passed" in res.stdout


def test_dropin_host_only(tmp_path):
    _run("cpu", tmp_path)


@pytest.mark.gpu
def test_dropin_gpu(tmp_path, torch_cuda):
    _run("gpu", tmp_path)
