"""Open-loop serving harness for the request loop (jsvd.Server): Poisson arrivals over a pool of requests.

Request shapes and spectra follow the serving workload: m = floor(exp U(ln 16, ln 2048)), n ~ U{4..min(m, 512)},
A = Q1 diag(sigma) Q2^T with sigma geometric from 1 to 1/kappa, kappa = 10^U(1, 10), requested relative error
alternating 1e-6 and 1e-12.  Matrices are generated once (seeded) and cycled; every submission is a new request
to the server.  Latency is measured by the server from submit to completion, so the order in which results are
collected does not affect it.
"""
from __future__ import annotations

import time

import numpy as np


def request_pool(count: int, seed: int = 4, min_m: int = 16, max_m: int = 2048, max_n: int = 512):
    """`count` (matrix, requested relative error) pairs of the serving workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(np.exp(rng.uniform(np.log(min_m), np.log(max_m))))
        m = max(min_m, min(m, max_m))
        n = int(rng.integers(4, min(m, max_n) + 1))
        kappa = 10.0 ** rng.uniform(1.0, 10.0)
        q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
        q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        sigma = kappa ** (-np.arange(n) / max(n - 1, 1))
        out.append((np.ascontiguousarray((q1 * sigma) @ q2.T), 1e-6 if i % 2 == 0 else 1e-12, sigma))
    return out


def run_poisson(server, pool, rate: float, duration: float, seed: int = 0, check_every: int = 0, keep: int = 0,
                max_outstanding: int = 20000):
    """Submits requests from `pool` (cycled) at Poisson arrival times of `rate` per second for `duration`
    seconds while a collector thread takes the results in submission order.  Returns throughput, latency
    percentiles and server statistics.

    An arrival that finds `max_outstanding` requests submitted but not yet collected is shed (counted, not
    submitted): past the server's capacity the backlog, and the host memory the queued matrices hold, stay
    bounded.  With check_every > 0, every check_every-th result's sorted sigma is compared with the pool's
    exact singular values (normwise error).  With keep > 0, the results (U, sigma, V) of the first `keep`
    requests are returned under "kept" as (pool index, u, sigma, v) for comparison with solves of the same
    matrices elsewhere."""
    import queue
    import threading

    rng = np.random.default_rng(seed)
    lat, queued, errs = [], [], {1e-6: 0.0, 1e-12: 0.0}
    kept = []
    bad = [0]
    collected = [0]
    todo = queue.Queue()
    failure = []

    def collect():
        j = 0
        while True:
            item = todo.get()
            if item is None:
                return
            rid, p = item
            try:
                u, s, v, info = server.wait(rid)
            except Exception as e:  # reported after the run; the loop keeps draining
                failure.append(e)
                bad[0] += 1
                collected[0] += 1
                j += 1
                continue
            if info.status != 0:
                bad[0] += 1
            else:
                lat.append(info.latency_s)
                queued.append(info.queued_s)
                if j < keep:
                    kept.append((p, u, s, v))
                if check_every and j % check_every < 2:  # consecutive pairs: both requested accuracies
                    _, rel, exact = pool[p]
                    err = float(np.max(np.abs(np.sort(s)[::-1] - exact)) / exact[0])
                    errs[rel] = max(errs[rel], err)
            collected[0] += 1
            j += 1

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    submitted = shed = 0
    t0 = time.perf_counter()
    next_t = rng.exponential(1.0 / rate)
    k = 0
    try:
        while next_t < duration:
            now = time.perf_counter() - t0
            if now < next_t:
                time.sleep(next_t - now)
            if submitted - collected[0] >= max_outstanding:
                shed += 1
            else:
                a, rel, _ = pool[k % len(pool)]
                todo.put((server.submit(a, rel), k % len(pool)))
                submitted += 1
            k += 1
            next_t += rng.exponential(1.0 / rate)
        submit_end = time.perf_counter() - t0
    finally:
        todo.put(None)
        th.join()
    if failure:
        raise failure[0]
    makespan = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    stats = server.stats()
    rep = {
        "rate_target": rate,
        "arrivals": k,
        "submitted": submitted,
        "shed": shed,
        "offered_rate": k / submit_end if submit_end > 0 else 0.0,
        "completed": len(lat),
        "failed": bad[0],
        "solves_per_s": len(lat) / makespan if makespan > 0 else 0.0,
        "p50_ms": float(np.percentile(lat_ms, 50)) if lat else None,
        "p99_ms": float(np.percentile(lat_ms, 99)) if lat else None,
        "mean_queued_ms": float(np.mean(queued) * 1e3) if queued else None,
        "max_sigma_err_fp32_requests": errs[1e-6] if check_every else None,
        "max_sigma_err_fp64_requests": errs[1e-12] if check_every else None,
        "server": stats,
    }
    if keep:
        rep["kept"] = kept
    return rep


def main(argv=None) -> None:
    """python -m jsvd.serving --rates 500,1000,2000,5000 --duration 20: one JSON line per arrival rate."""
    import argparse
    import json
    import sys

    import jsvd

    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="500,1000,2000,5000")
    ap.add_argument("--duration", type=float, default=20.0)
    ap.add_argument("--pool", type=int, default=128)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--large-path", default="mixed", choices=["plain", "mixed"])
    ap.add_argument("--check-every", type=int, default=25)
    ap.add_argument("--max-outstanding", type=int, default=20000,
                    help="arrivals beyond this many uncollected requests are shed")
    ap.add_argument("--small-rows", type=int, default=256, help="largest m served by the batched kernel")
    ap.add_argument("--small-cols", type=int, default=128, help="largest n served by the batched kernel")
    args = ap.parse_args(argv)
    print(f"[serving] generating {args.pool} requests", file=sys.stderr, flush=True)
    pool = request_pool(args.pool)
    large = jsvd.PATH_MIXED if args.large_path == "mixed" else jsvd.PATH_PLAIN_F64
    for rate in (float(r) for r in args.rates.split(",")):
        srv = jsvd.Server(lanes=args.lanes, large_path=large, small_max_rows=args.small_rows,
                          small_max_columns=args.small_cols)
        try:
            rep = run_poisson(srv, pool, rate, args.duration, check_every=args.check_every,
                              max_outstanding=args.max_outstanding)
        finally:
            srv.close()
        rep["lanes"] = args.lanes
        rep["large_path"] = args.large_path
        rep["batched_up_to"] = [args.small_rows, args.small_cols]
        rep["duration_s"] = args.duration
        print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
