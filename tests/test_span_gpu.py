"""Round spans of the plain FP64 path (several rounds of the round-robin circle per launch) on cuda:0.

A span launch runs up to max_rounds rounds: every workgroup computes the slots it owns plus the halo slots that
feed them, so no workgroup waits for another, and the launches alternate between two buffers.  The claim is
bit-identity with one launch per round (the reference's host loop of one round after another,
src/kernels.cu:889-981): U, sigma, V, the sweep count and every sweep's rotation count and largest pair cosine
must be equal bit for bit, for every (wavefronts, rounds per launch) setting, on odd and even shapes, on shapes
held by one workgroup and by many, and after a reshape.  The per-round results themselves are checked against the
oracle by test_plain_gpu.py.
"""
import numpy as np
import pytest

import jsvd
from helpers import gaussian

pytestmark = pytest.mark.gpu


def _solve(plan, a):
    r = plan.solve_host(a)
    rot, off = plan.sweep_stats()
    return r, rot, off


def _same(x, y):
    (r1, rot1, off1), (r2, rot2, off2) = x, y
    assert r1.sweeps == r2.sweeps
    assert np.array_equal(rot1, rot2)
    assert np.array_equal(off1.view(np.uint64), off2.view(np.uint64))
    for p, q in ((r1.u, r2.u), (r1.sigma, r2.sigma), (r1.v, r2.v)):
        assert np.array_equal(np.asarray(p).view(np.uint64), np.asarray(q).view(np.uint64))


SHAPES = [
    # (m, n, layout mode, [(waves, max_rounds), ...]); mode 1 = FORCE_ENABLE (256-thread round shape)
    (512, 256, 0, [(16, 8), (16, 2), (16, 5), (8, 4), (8, 2)]),   # the reference's own case
    (512, 256, 1, [(16, 8), (8, 3)]),
    (201, 97, 0, [(16, 8), (8, 4), (16, 3)]),                      # odd m, odd n (dummy column)
    (96, 40, 0, [(16, 31), (16, 8), (8, 4)]),
    (33, 32, 0, [(16, 31), (16, 2), (8, 4)]),                      # 16 slots: one workgroup at 16 wavefronts
    (20, 3, 0, [(16, 3), (8, 2)]),                                 # the smallest span shape
    (64, 64, 0, [(16, 8), (8, 4)]),
    (300, 255, 0, [(16, 8), (8, 4)]),
    (1000, 500, 0, [(8, 4), (8, 2)]),                              # 8 A + 4 V chunks per lane: 8 wavefronts
]


@pytest.mark.parametrize("m,n,mode,settings", SHAPES)
def test_span_bit_identical_to_rounds(torch_cuda, m, n, mode, settings):
    a = gaussian(m, n, 1000 * m + n)
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30, layout_transpose_mode=mode)
    print(f"span: {m}x{n} mode={mode} settings={settings}", flush=True)
    plan = jsvd.Plan(m, n, cfg)
    plan.set_round_span(0)
    ref = _solve(plan, a)
    for waves, rounds in settings:
        plan.set_round_span(waves, rounds)
        got = _solve(plan, a)
        _same(ref, got)
    plan.close()


def test_reference_case_uses_spans_by_default(torch_cuda):
    plan = jsvd.Plan(512, 256, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30))
    sp = plan.round_span()
    assert sp["waves"] in (8, 16) and sp["max_rounds"] >= 2
    assert sp["launches_per_sweep"] % 2 == 0 and sp["launches_per_sweep"] < 255


def test_span_after_reshape(torch_cuda):
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30)
    a = gaussian(150, 70, 7)
    plan = jsvd.Plan(512, 256, cfg)
    plan.reshape(150, 70, 1e-15)
    assert plan.round_span()["waves"] > 0
    got = _solve(plan, a)
    fresh = jsvd.Plan(150, 70, cfg)
    fresh.set_round_span(0)
    _same(_solve(fresh, a), got)


def test_span_refusals(torch_cuda):
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30)
    big = jsvd.Plan(2048, 600, cfg)   # beyond the span kernels' register-resident columns
    assert big.round_span()["waves"] == 0
    with pytest.raises(ValueError):
        big.set_round_span(16, 8)
    mixed = jsvd.Plan(256, 64, cfg, path=jsvd.PATH_MIXED)
    with pytest.raises(ValueError):
        mixed.set_round_span(16, 8)
    small = jsvd.Plan(64, 2, cfg)     # one round per sweep: no spans
    with pytest.raises(ValueError):
        small.set_round_span(16, 8)
    plan = jsvd.Plan(512, 256, cfg)
    with pytest.raises(ValueError):
        plan.set_round_span(12, 8)    # 8 or 16 wavefronts only
