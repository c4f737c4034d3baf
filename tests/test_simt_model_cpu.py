"""Kernels of the library on the CPU model of a workgroup (tests/simt_model), no GPU needed.

tests/simt_model compiles the library's kernel sources (csrc/batch_kernel.hpp, csrc/block_kernels.hpp,
unchanged) with g++, AddressSanitizer and UBSan and runs every GPU thread as a fiber: a barrier or wavefront
operation (DPP move, ballot, MFMA) that the whole workgroup / wavefront does not reach, or any access outside LDS,
scratch slices or the buffers, stops the run with an error.

* Batched kernel: the inputs of the batch test held off the GPU (1x1, 5x1, 2x2, 64x64, the at-limit 256x128 and
  255x127 items, and a batch with an item beyond the limits), under every launch shape of the kernel.
* Block kernels: the sweep loop of the block and mixed paths (column ordering, intra-block step, Gram / inner
  sweep / apply rounds) at width 32 in both inner orderings, FP64 and FP32, against LAPACK's singular values.
* Column ordering: the bitonic network gives the permutation of numpy's stable sort by descending norm, ties by
  column index.

The model's own checks are pinned by tests/simt_model/selftest.cpp (planted faults).
"""
import concurrent.futures as cf
import os
import subprocess
import tempfile

import numpy as np
import pytest

import jsvd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "tests", "simt_model")
BUILD = os.path.join(MODEL, "_build")
FLAGS = ["-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-Wno-unknown-pragmas", "-I", MODEL, "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "jacobi-svd-cuda_amd", "csrc")]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")


def _build(name):
    out = os.path.join(BUILD, name)
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run(["g++", *FLAGS, os.path.join(MODEL, name + ".cpp"), "-o", out], check=True)
    return out


@pytest.fixture(scope="module")
def model_bins():
    with cf.ThreadPoolExecutor(3) as ex:
        batch, selftest, block, order = ex.map(_build, ["batch_model", "selftest", "block_model", "order_model"])
    return batch, selftest, block, order


def _run_model(binary, mats, prec, tol, threads, lds, max_sweeps, max_rows=256, max_cols=128):
    """Packs like jsvd.Batch.solve (cost order), runs the model, returns (sigma list, sweeps, model stats)."""
    pk = jsvd.BatchPacked(mats, prec, tol, max_sweeps)
    order = pk.cost_order()
    items = np.ascontiguousarray(pk.items[order])
    lens = [max(x, 1) for x in pk.lens]
    a = np.zeros(lens[0])
    a[:pk.lens[0]] = pk.a
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([threads, lds, max_rows, max_cols, len(mats), len(mats), *lens], dtype=np.int64).tofile(f)
            items.tofile(f)
            a.tofile(f)
        r = subprocess.run([binary, fi, fo], capture_output=True, text=True, env=ENV)
        assert r.returncode == 0, r.stderr[-4000:]
        out = np.fromfile(fo, dtype=np.uint8)
    sizes = [lens[1] * 8, lens[2] * 8, lens[3] * 8, len(mats) * 4, 16]
    parts = np.split(out, np.cumsum(sizes)[:-1])
    s = parts[1].view(np.float64)
    sw_q = parts[3].view(np.int32)
    sweeps = np.zeros(len(mats), np.int32)
    sweeps[order] = sw_q
    sig = [s[int(it["sigma_offset"]):int(it["sigma_offset"]) + int(it["columns"])] for it in pk.items]
    return sig, sweeps, parts[4].view(np.int64)


def _edge_inputs():
    rng = np.random.default_rng(11)
    mats = [rng.standard_normal(s) for s in [(1, 1), (5, 1), (2, 2), (256, 128), (255, 127), (64, 64)]]
    big = rng.standard_normal((300, 10))
    return mats, big


def test_model_catches_planted_faults(model_bins):
    selftest = model_bins[1]
    r = subprocess.run([selftest, "sum"], capture_output=True, text=True, env=ENV)
    assert r.returncode == 0 and r.stdout.split() == ["2016.0"] * 4, r.stdout + r.stderr
    for case, needles in (("barrier", ["left the kernel"]), ("divergent", ["lanes at a wave op"]),
                          ("twosites", ["different barriers"]),
                          ("lds", ["global-buffer-overflow", "insufficient space for an object"])):
        r = subprocess.run([selftest, case], capture_output=True, text=True, env=ENV)
        assert r.returncode != 0 and any(n in r.stderr for n in needles), (case, r.stdout, r.stderr[-2000:])


def test_batch_edge_items_on_model_every_shape(model_bins):
    """1x1, 5x1, 2x2, 64x64 (max 60 sweeps) and the rejection batch, under all six launch shapes."""
    binary = model_bins[0]
    mats, big = _edge_inputs()
    small = [mats[0], mats[1], mats[2], mats[5]]
    refs = [oracle.svd(a, eps=1e-14, max_sweeps=60) for a in small]
    jobs = []
    with cf.ThreadPoolExecutor(8) as ex:
        for threads, lds in jsvd.BATCH_SHAPES:
            jobs.append((threads, lds, "edge",
                         ex.submit(_run_model, binary, small, [0] * 4, [1e-14] * 4, threads, lds, 60)))
            jobs.append((threads, lds, "reject",
                         ex.submit(_run_model, binary, [mats[2], big, mats[5]], [0] * 3, [1e-14] * 3, threads, lds,
                                   60)))
        for threads, lds, kind, fut in jobs:
            sig, sweeps, _ = fut.result()
            if kind == "edge":
                for i, ref in enumerate(refs):
                    assert np.max(np.abs(sig[i] - ref.sigma) / ref.sigma) < 1e-12, (threads, lds, i)
                    assert sweeps[i] == ref.sweeps, (threads, lds, i, sweeps[i], ref.sweeps)
            else:
                assert sweeps[1] == -1 and sweeps[0] == refs[2].sweeps and sweeps[2] == refs[3].sweeps


def test_batch_at_limit_items_on_model(model_bins):
    """256x128 and 255x127 (A and V in the scratch slice at the 38/76 KiB shapes, V in LDS at 152 KiB): two
    sweeps cover every index the later sweeps use.  Two sweeps leave sigma unconverged, where the kernel's and the
    oracle's summation orders give slightly different rotations: column norms agree to 1e-9 relative."""
    binary = model_bins[0]
    mats, _ = _edge_inputs()
    big = [mats[3], mats[4]]
    refs = [oracle.svd(a, eps=1e-14, max_sweeps=2) for a in big]
    with cf.ThreadPoolExecutor(4) as ex:
        futs = {(t, l): ex.submit(_run_model, binary, big, [0, 0], [1e-14] * 2, t, l, 2)
                for t, l in (jsvd.BATCH_SHAPES[0], jsvd.BATCH_SHAPES[1], jsvd.BATCH_SHAPES[5])}
        for key, fut in futs.items():
            sig, sweeps, _ = fut.result()
            for i, ref in enumerate(refs):
                assert sweeps[i] == 2, key
                assert np.max(np.abs(sig[i] - ref.sigma) / ref.sigma) < 1e-9, key


def _run_block_model(binary, a, prec, order, kb, max_sweeps, eps):
    """The block sweep loop on the model; returns (sweeps, rotations per sweep, A (m x n), V (n x n))."""
    m, n = a.shape
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([m, n], np.int64).tofile(f)
            np.ascontiguousarray(a.T).tofile(f)  # column-major
        r = subprocess.run([binary, prec, order, str(kb), str(max_sweeps), repr(eps), fi, fo], capture_output=True,
                           text=True, env=ENV)
        assert r.returncode == 0, r.stderr[-4000:]
        raw = open(fo, "rb").read()
    sweeps, lda, ldv, npad = (int(x) for x in np.frombuffer(raw[:32], np.int64))
    o = 48
    rots = np.frombuffer(raw[o:o + 4 * sweeps], np.int32)
    o += 12 * sweeps
    a_out = np.frombuffer(raw[o:o + 8 * lda * npad], np.float64).reshape(npad, lda).T
    o += 8 * lda * npad
    v_out = np.frombuffer(raw[o:o + 8 * ldv * npad], np.float64).reshape(npad, ldv).T
    return sweeps, rots, a_out[:m, :n], v_out[:n, :n]


def test_block_sweeps_on_model(model_bins):
    """Width-32 block sweeps of a column-graded 160 x 128 matrix (4 blocks) to convergence: the cross ordering
    (intra-block step + cross pairs) visits every column pair once per sweep; both orderings end on a rotation-free
    sweep with LAPACK's singular values (FP64 1e-13, FP32 1e-5 of sigma_max), V orthogonal and A V = X."""
    block = model_bins[2]
    rng = np.random.default_rng(5)
    n = 128
    a = rng.standard_normal((n + 32, n)) * 10.0 ** (-3 * np.arange(n) / n)
    ref = np.linalg.svd(a, compute_uv=False)
    cases = [("f64", "cross", 1e-14), ("f64", "full", 1e-14), ("f32", "cross", 1e-6)]
    with cf.ThreadPoolExecutor(3) as ex:
        futs = [ex.submit(_run_block_model, block, a, p, o, 32, 30, e) for p, o, e in cases]
        for (prec, order, _), fut in zip(cases, futs):
            sweeps, rots, x, v = fut.result()
            assert rots[-1] == 0 and sweeps < 30, (prec, order, list(rots))
            if order == "cross":
                # each column pair is visited once per sweep (and in FP64 every pair rotates in the first sweep)
                assert np.all(rots <= n * (n - 1) // 2), (prec, list(rots))
                assert prec != "f64" or rots[0] == n * (n - 1) // 2, list(rots)
            s = np.sort(np.linalg.norm(x, axis=0))[::-1]
            if prec == "f64":
                assert np.max(np.abs(s - ref) / ref) < 1e-13, order
                assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-12, order
                assert np.linalg.norm(a @ v - x) / np.linalg.norm(a) < 1e-14, order
            else:
                assert np.max(np.abs(s - ref)) / ref[0] < 1e-5, order
                assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-4, order


@pytest.mark.parametrize("n_pad", [1, 96, 1000, 1024, 2080, 4096])
def test_column_order_on_model(model_bins, n_pad):
    """column_order_kernel's permutation (bitonic network over the next power of two, padded with keys below every
    norm) equals a stable sort by descending squared norm with ties by column index, with repeated keys, zero norms
    (padding columns) and signed zeros among the keys, for powers of two and between them."""
    order = model_bins[3]
    rng = np.random.default_rng(n_pad)
    keys = rng.choice([0.0, -0.0, 1.0, 2.5, 1e-300], size=n_pad) * (rng.random(n_pad) < 0.3)
    keys = np.where(rng.random(n_pad) < 0.5, rng.random(n_pad) * 10.0, keys)
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([n_pad], np.int64).tofile(f)
            keys.astype(np.float64).tofile(f)
        r = subprocess.run([order, fi, fo], capture_output=True, text=True, env=ENV)
        assert r.returncode == 0, r.stderr[-4000:]
        perm = np.fromfile(fo, np.int32)
    expected = np.lexsort((np.arange(n_pad), -keys))
    assert np.array_equal(perm, expected)
