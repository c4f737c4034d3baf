"""Request loop (continuous batching) on cuda:0.

Requests of mixed shapes and precisions are submitted from a thread while earlier ones run; each result
must equal, bit for bit, the result of the same request served alone by a fresh one-lane server (the
kernels compute every request independently of its batch), and meet its accuracy: sorted sigma normwise
error vs numpy (LAPACK gesdd) <= 1e-5 for relerr-1e-6 requests (FP32) and <= 1e-12 for relerr-1e-12
requests (FP64).
"""
import threading
import time

import numpy as np
import pytest

import jsvd
from helpers import sorted_desc

pytestmark = pytest.mark.gpu


def _requests(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(np.exp(rng.uniform(np.log(16), np.log(600))))
        n = int(rng.integers(4, min(m, 160) + 1))
        kappa = 10.0 ** rng.uniform(1, 6)
        q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
        q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        s = kappa ** (-np.arange(n) / max(n - 1, 1))
        out.append(((q1 * s) @ q2.T, 1e-6 if i % 2 == 0 else 1e-12))
    return out


def test_server_mix_equals_alone_and_accuracy(torch_cuda):
    reqs = _requests(48, 5)
    srv = jsvd.Server(lanes=4, small_max_rows=256, small_max_columns=128)
    ids = [None] * len(reqs)

    def producer():
        for i, (a, rel) in enumerate(reqs):
            ids[i] = srv.submit(a, rel)
            time.sleep(0.002)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    results = [srv.wait(i) for i in ids]
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == len(reqs) and stats["batches"] > 0 and stats["plans"] > 0
    assert stats["max_in_flight"] > 1, "requests overlapped on the device"

    alone = jsvd.Server(lanes=1)
    for (a, rel), (u, s, v, info) in zip(reqs, results):
        assert info.status == 0
        rid = alone.submit(a, rel)
        u1, s1, v1, info1 = alone.wait(rid)
        assert info1.batch_size == 1
        assert np.array_equal(s, s1) and np.array_equal(u, u1) and np.array_equal(v, v1)
        ref = sorted_desc(np.linalg.svd(a, compute_uv=False))
        err = np.max(np.abs(sorted_desc(s) - ref)) / ref[0]
        assert err < (1e-5 if rel >= 1e-7 else 1e-12), (a.shape, rel, err)
    alone.close()


def test_server_budget_admission(torch_cuda):
    reqs = _requests(12, 9)
    biggest = max(8 * (3 * a.shape[0] * a.shape[1] + 2 * a.shape[1] ** 2 + a.shape[1]) for a, _ in reqs)
    budget = 2 * biggest
    srv = jsvd.Server(lanes=4, hbm_budget_bytes=budget)
    ids = [srv.submit(a, rel) for a, rel in reqs]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert all(r[3].status == 0 for r in res)
    assert st["budget_peak"] <= budget and st["budget_in_use"] == 0


def test_poisson_serving_short_run(torch_cuda):
    """A short open-loop run of the serving harness (Poisson arrivals, serving-workload shapes up to 512 x 128):
    every request completes, latency percentiles are reported, and sigma meets the requested accuracy
    (normwise vs the construction's exact singular values: 1e-5 for relerr 1e-6, 1e-12 for relerr 1e-12)."""
    from jsvd import serving
    pool = serving.request_pool(24, seed=9, max_m=512, max_n=128)
    srv = jsvd.Server(lanes=4, large_path=jsvd.PATH_MIXED)
    try:
        rep = serving.run_poisson(srv, pool, rate=200.0, duration=1.0, check_every=1)
    finally:
        srv.close()
    assert rep["failed"] == 0 and rep["completed"] == rep["submitted"] > 50
    assert rep["p50_ms"] <= rep["p99_ms"]
    assert rep["max_sigma_err_fp32_requests"] < 1e-5
    assert rep["max_sigma_err_fp64_requests"] < 1e-12
    assert rep["server"]["max_in_flight"] > 1


def test_poisson_serving_workload_at_500_per_s(torch_cuda):
    """The serving workload at its configured shapes (m = exp U(ln 16, ln 2048), n ~ U{4..min(m, 512)},
    kappa = 10^U(1, 10), relerr 1e-6 / 1e-12 alternating) offered at 500 requests/s: every request completes,
    the server keeps up (completed rate within 10 % of the offered rate), each precision class meets its
    accuracy (normwise sigma error vs the construction's exact values: 1e-5 for relerr 1e-6, 1e-12 for relerr
    1e-12, checked on every request), and the first results equal, bit for bit, the same matrices served alone."""
    from jsvd import serving
    pool = serving.request_pool(32, seed=11, max_m=2048, max_n=512)
    assert max(a.shape[0] for a, _, _ in pool) > 1024 and max(a.shape[1] for a, _, _ in pool) > 256
    srv = jsvd.Server(lanes=4, large_path=jsvd.PATH_MIXED)
    try:
        print("[c4] 500/s for 3 s over 32 matrices up to 2048 x 512", flush=True)
        rep = serving.run_poisson(srv, pool, rate=500.0, duration=3.0, check_every=1, keep=16)
    finally:
        srv.close()
    assert rep["failed"] == 0 and rep["completed"] == rep["submitted"] > 1200
    assert rep["solves_per_s"] > 0.9 * rep["offered_rate"], rep
    assert rep["max_sigma_err_fp32_requests"] < 1e-5
    assert rep["max_sigma_err_fp64_requests"] < 1e-12
    assert rep["server"]["batches"] > 0 and rep["server"]["plans"] > 0 and rep["server"]["max_in_flight"] > 1
    alone = jsvd.Server(lanes=1, large_path=jsvd.PATH_MIXED)
    try:
        for p, u, s, v in rep["kept"]:
            a, rel, _ = pool[p]
            u1, s1, v1, info1 = alone.wait(alone.submit(a, rel))
            assert info1.status == 0
            assert np.array_equal(s, s1) and np.array_equal(u, u1) and np.array_equal(v, v1), (a.shape, rel)
    finally:
        alone.close()
