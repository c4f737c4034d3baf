"""Open-loop serving harness (jsvd.serving.run_poisson) against a stand-in server on the CPU: results are
collected while arrivals continue, arrivals past the outstanding-request bound are shed rather than queued,
and the accuracy check and kept results refer to the right pool entries."""
import threading
import time
from types import SimpleNamespace

import numpy as np

from jsvd import serving


class _SlowServer:
    """Completes each request `service_s` after the previous one (one worker), returning the exact sigma of a
    diagonal test matrix."""

    def __init__(self, service_s):
        self.service_s = service_s
        self.cv = threading.Condition()
        self.next_id = 0
        self.done = {}
        self.pending = []
        self.collected = 0
        self.max_outstanding_seen = 0
        self.worker = threading.Thread(target=self._run, daemon=True)
        self.stop = False
        self.worker.start()

    def _run(self):
        while True:
            with self.cv:
                while not self.pending and not self.stop:
                    self.cv.wait()
                if self.stop and not self.pending:
                    return
                rid, a, t = self.pending.pop(0)
            time.sleep(self.service_s)
            s = np.sort(np.abs(np.diag(a)))[::-1].copy()
            with self.cv:
                self.done[rid] = (np.eye(*a.shape), s, np.eye(a.shape[1]),
                                  SimpleNamespace(status=0, latency_s=time.perf_counter() - t, queued_s=0.0))
                self.cv.notify_all()

    def submit(self, a, rel):
        with self.cv:
            rid = self.next_id
            self.next_id += 1
            self.pending.append((rid, a, time.perf_counter()))
            self.max_outstanding_seen = max(self.max_outstanding_seen, rid + 1 - self.collected)
            self.cv.notify_all()
            return rid

    def wait(self, rid):
        with self.cv:
            while rid not in self.done:
                self.cv.wait()
            self.collected += 1
            return self.done.pop(rid)

    def stats(self):
        return {}

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        self.worker.join()


def _pool(count):
    out = []
    for i in range(count):
        sigma = np.array([1.0, 0.5, 0.25, 0.125]) * (i + 1)
        out.append((np.diag(sigma), 1e-6 if i % 2 == 0 else 1e-12, sigma))
    return out


def test_collects_during_arrivals_and_checks_accuracy():
    srv = _SlowServer(0.0005)
    try:
        rep = serving.run_poisson(srv, _pool(6), rate=400.0, duration=0.5, check_every=1, keep=5)
    finally:
        srv.close()
    assert rep["shed"] == 0 and rep["failed"] == 0
    assert rep["completed"] == rep["submitted"] == rep["arrivals"] > 100
    assert rep["max_sigma_err_fp32_requests"] == 0.0 and rep["max_sigma_err_fp64_requests"] == 0.0
    assert [p for p, *_ in rep["kept"]] == [0, 1, 2, 3, 4]
    for p, u, s, v in rep["kept"]:
        assert np.array_equal(s, _pool(6)[p][2])


def test_overload_sheds_past_the_outstanding_bound():
    srv = _SlowServer(0.01)  # capacity about 100 requests/s, offered 1000/s
    try:
        rep = serving.run_poisson(srv, _pool(4), rate=1000.0, duration=0.5, max_outstanding=8)
    finally:
        srv.close()
    assert rep["shed"] > 0 and rep["submitted"] + rep["shed"] == rep["arrivals"]
    assert rep["completed"] == rep["submitted"]
    assert srv.max_outstanding_seen <= 8
