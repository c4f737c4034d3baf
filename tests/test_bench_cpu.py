"""bench.py's host-side contract, on the CPU: the multi-GPU refusal, the committed traffic figures it reads,
the workload generators' stated shapes and distributions, and the flop count behind the eigen-kernel roofline."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from jsvd import workloads  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_traffic_figures_come_from_the_committed_summary():
    # the figures bench.py reports are the ones in the summary's table rows for the same kernels
    rows = {}
    for line in open(os.path.join(ROOT, "profiles", "traffic.md")):
        if line.startswith("| `"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            rows[cells[0].strip("`")] = float(cells[2].replace(",", ""))
    assert bench.profiled_traffic("round_kernel_tall") == rows["jsvd::round_kernel<256, 8, 2>"]
    assert bench.profiled_traffic("block_eigen_f32") == rows["jsvd::block_eigen_kernel<float, 32, true>"]
    assert 0.98 < bench.profiled_traffic("round_kernel_tall") / 83886080 < 1.02
    assert bench.profiled_traffic("no_such_kernel") is None


def test_eigen_flop_count():
    # 32 inner rounds of the cross ordering; per round 528 2x2 blocks of the 32 x 32 pair grid's upper triangle
    # (24 flops each, two-sided) and 64 x 32 (row, pair) entries of Q (6 flops each)
    assert 32 * 33 // 2 == 528
    assert bench.EIGEN_FLOPS_PER_ROUND == 528 * 24 + 64 * 32 * 6
    assert bench.EIGEN_FLOPS_PER_PAIR == 32 * bench.EIGEN_FLOPS_PER_ROUND


def test_tall_geometric_workload():
    a, sigma, eps, max_sweeps, name = bench.make_workload("tall")
    assert a.shape == (4096, 1024) and name == "tall_geometric_4096x1024_fp64"
    assert eps == 1e-14 and max_sweeps == 60
    assert sigma[0] == 1.0 and abs(sigma[-1] - 1e-8) < 1e-20
    assert np.allclose(np.diff(np.log10(sigma)), -8.0 / 1023)
    # the construction's singular values are those of A (spot check with the Gram of a column subset bound)
    assert np.linalg.norm(a, 2) <= 1.0 + 1e-12


def test_batch_mix_generator():
    items = workloads.batch_mixed(400)
    ms = np.array([a.shape[0] for a, _, _ in items])
    ns = np.array([a.shape[1] for a, _, _ in items])
    assert ms.min() >= 8 and ms.max() <= 256
    assert np.all(ns >= 4) and np.all(ns <= np.minimum(ms, 128))
    assert [p for _, p, _ in items[:4]] == ["f32", "f64", "f32", "f64"]
    assert [t for _, _, t in items[:2]] == [1e-6, 1e-14]
