"""The plain round kernels on the CPU model of a workgroup (tests/simt_model), no GPU needed.

tests/simt_model/plain_model.cpp compiles csrc/plain_kernels.hpp unchanged with g++, AddressSanitizer and UBSan and
runs the same solve twice: one round_kernel launch per round, and round_span_kernel launches over spans of rounds
(each workgroup computing its own slots plus the halo slots that feed them, the launches alternating between two
buffers, the second one starting as garbage).  The span solve must equal the per-round solve bit for bit -- A, V,
and every sweep's rotation count and largest pair cosine -- on partitions with one workgroup and with many, odd and
even m and n (the dummy column of an odd n), every reduction order the span kernel emulates (64-, 128- and
256-thread round shapes) and rounds per launch from 2 to the whole sweep; the model also fails the run on any
barrier or wavefront operation the whole workgroup / wavefront does not reach and on any access outside LDS or the
exact-size buffers.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "tests", "simt_model")
BUILD = os.path.join(MODEL, "_build")
FLAGS = ["-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-Wno-unknown-pragmas", "-Wno-psabi", "-I", MODEL, "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "jacobi-svd-cuda_amd", "csrc")]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")


@pytest.fixture(scope="module")
def plain_model():
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "plain_model")
    subprocess.run(["g++", *FLAGS, os.path.join(MODEL, "plain_model.cpp"), "-o", out], check=True)
    return out


def _run(binary, m, n, threads, ea, nv, waves, max_rounds, sweeps, eps=1e-15, seed=0):
    lda, ldv = m + (m & 1), n + (n & 1)
    a = np.zeros((n, lda))
    a[:, :m] = np.random.default_rng(seed).standard_normal((n, m))
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([m, n, lda, ldv, threads, ea, nv, waves, max_rounds, sweeps], np.int64).tofile(f)
            np.array([eps]).tofile(f)
            a.tofile(f)
        r = subprocess.run([binary, fi, fo], capture_output=True, text=True, env=ENV)
        assert r.returncode == 0, r.stderr[-4000:]
        buf = open(fo, "rb").read()
    pos = 0

    def take(count, dtype):
        nonlocal pos
        x = np.frombuffer(buf, dtype, count, pos)
        pos += x.nbytes
        return x

    runs = []
    for _ in range(2):
        av = take(lda * n, np.uint64)
        vv = take(ldv * n, np.uint64)
        s = int(take(1, np.int64)[0])
        stats = [(int(take(1, np.int64)[0]), int(take(1, np.uint64)[0])) for _ in range(s)]
        runs.append((av, vv, stats))
    launches = int(take(1, np.int64)[0])
    part = take(3, np.int64)
    return runs, launches, part


@pytest.mark.parametrize("m,n,threads,ea,nv,waves,max_rounds,sweeps,groups", [
    (40, 24, 64, 1, 1, 8, 4, 3, 3),     # three workgroups: both turns in the end groups, one interior
    (41, 23, 64, 1, 1, 8, 4, 3, 3),     # odd m (padding row), odd n (dummy column)
    (40, 24, 64, 1, 1, 8, 2, 3, 2),
    (100, 30, 64, 2, 1, 8, 3, 2, 3),    # two chunks per virtual thread
    (60, 30, 128, 1, 1, 16, 8, 2, 1),   # two virtual wavefronts; one workgroup holds the circle: 2 launches a sweep
    (20, 3, 64, 1, 1, 8, 4, 3, 1),      # the smallest span circle (4 positions, 3 rounds)
    (130, 60, 64, 2, 1, 8, 2, 2, 5),
])
def test_span_equals_rounds_on_model(plain_model, m, n, threads, ea, nv, waves, max_rounds, sweeps, groups):
    (one, span), launches, part = _run(plain_model, m, n, threads, ea, nv, waves, max_rounds, sweeps)
    assert part[0] == groups
    assert np.array_equal(one[0], span[0]) and np.array_equal(one[1], span[1])
    assert one[2] == span[2] and len(one[2]) >= 1
    assert one[2][0][0] > 0  # the first sweep rotates
    N = n + (n & 1)
    per_sweep = launches // len(span[2])
    assert per_sweep % 2 == 0 and per_sweep * len(span[2]) == launches
    assert per_sweep <= N - 1
