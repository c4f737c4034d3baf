"""Generates the committed fixtures under tests/golden/ (run from the repository root).

The reference ships no tests, golden vectors or fixtures, so these fixtures come from sources that do
not share code with the oracle:

* dgesvj_small.npz -- seeded inputs and LAPACK dgesvj (SciPy-bundled OpenBLAS) singular values and
  vectors, an independent FP64 one-sided Jacobi.  Keys: <case>_a, <case>_sigma (descending),
  <case>_u, <case>_v.
* mat_1x2.bin -- the hand-computed .mat record of the 1x2 matrix [1.0, -2.0]: big-endian u64 rows,
  u64 cols, then big-endian IEEE doubles (the record layout of the reference's MatFilePolicy,
  src/io.cu:1180-1286).
* closed_form.json -- hand-derived known answers: the round-robin schedules for n=4 and n=5 and the
  2x2 SVD of [[3, 0], [4, 5]] (sigma = sqrt(45), sqrt(5)).

Usage: python tests/golden/make_golden.py
"""
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402  (dgesvj binding only)

CASES = {
    "g4x3": (4, 3, 11),
    "g8x8": (8, 8, 12),
    "g16x7": (16, 7, 13),
    "g33x17": (33, 17, 14),
    "g64x64": (64, 64, 15),
    "g97x40": (97, 40, 16),
}


def main():
    arrays = {}
    for name, (m, n, seed) in CASES.items():
        a = np.random.default_rng(seed).standard_normal((m, n))
        u, s, v, _ = oracle.dgesvj(a)
        arrays[f"{name}_a"] = a
        arrays[f"{name}_sigma"] = s
        arrays[f"{name}_u"] = u
        arrays[f"{name}_v"] = v
    # column-graded case: G * diag(10^(-6 j/(n-1)))
    g = np.random.default_rng(17).standard_normal((48, 24)) * (10.0 ** (-6.0 * np.arange(24) / 23.0))
    u, s, v, _ = oracle.dgesvj(g)
    arrays.update({"graded48x24_a": g, "graded48x24_sigma": s, "graded48x24_u": u, "graded48x24_v": v})
    np.savez(os.path.join(HERE, "dgesvj_small.npz"), **arrays)

    rec = struct.pack(">QQ", 1, 2) + struct.pack(">dd", 1.0, -2.0)
    with open(os.path.join(HERE, "mat_1x2.bin"), "wb") as f:
        f.write(rec)

    closed = {
        "schedule_n4": [[[0, 3], [1, 2]], [[0, 2], [3, 1]], [[0, 1], [2, 3]]],
        "schedule_n5": [[[1, 4], [2, 3]], [[0, 4], [1, 2]], [[0, 3], [4, 2]], [[0, 2], [3, 1]], [[0, 1], [3, 4]]],
        "svd2x2_a": [[3.0, 0.0], [4.0, 5.0]],
        "svd2x2_sigma_desc": [45.0 ** 0.5, 5.0 ** 0.5],
        "mat_1x2_hex": "0000000000000001" "0000000000000002" "3ff0000000000000" "c000000000000000",
    }
    with open(os.path.join(HERE, "closed_form.json"), "w") as f:
        json.dump(closed, f, indent=1)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
