"""Autotuner (jsvd_autotune) on cuda:0: the report, the cache that plans consult, and solves through the
tuned launch parameters.

Tolerances: a tuned plan's sorted sigma against LAPACK's, normwise max|d sigma| / sigma_max <= 1e-12, and
rho_A <= 1e-12 -- the same bar as the untuned plans, since the tuner only changes launch parameters.
"""
import numpy as np
import pytest

import jsvd
from helpers import gaussian, rho_a, sorted_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_cache():
    jsvd.autotune_clear()
    yield
    jsvd.autotune_clear()


def _plain_cfg():
    return jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=40)


def test_report_and_plain_plan_use_choice(torch_cuda):
    m, n = 4096, 1024
    before = jsvd.Plan(m, n, _plain_cfg()).launch_params()
    rep = jsvd.autotune(m, n, repetitions=1)
    assert rep.candidates_tried >= 5
    assert rep.round_threads in (64, 256, 1024)
    assert rep.round_ea in (0, 1, 2, 4, 8) and rep.round_ev in (0, 1, 2, 4, 8)
    assert rep.apply_workgroups in (256, 512, 1024)
    assert rep.block_width in (16, 32, 64)
    assert 0.0 < rep.round_us_best <= rep.round_us_default
    assert 0.0 < rep.block_round_us_best <= rep.block_round_us_default
    sweeps = list(rep.block_sweep_us)
    assert all(x > 0.0 for x in sweeps)
    assert sweeps[(16, 32, 64).index(rep.block_width)] == min(sweeps)
    after = jsvd.Plan(m, n, _plain_cfg()).launch_params()
    assert (after["round_threads"], after["round_ea"], after["round_ev"]) == (
        rep.round_threads, rep.round_ea, rep.round_ev)
    # the mixed path sweeps the n x n R^T, the block FP64 path sweeps A itself
    mixed = jsvd.Plan(m, n, _plain_cfg(), path=jsvd.PATH_MIXED).launch_params()
    assert mixed["round_threads"] == 0
    assert (mixed["apply_workgroups"], mixed["block_width"]) == (rep.rt_apply_workgroups, rep.rt_block_width)
    blk = jsvd.Plan(m, n, _plain_cfg(), path=jsvd.PATH_BLOCK_F64).launch_params()
    assert (blk["apply_workgroups"], blk["block_width"]) == (rep.apply_workgroups, rep.block_width)
    assert rep.rt_block_width in (16, 32, 64) and rep.rt_apply_workgroups in (256, 512, 1024)
    # an explicit width wins over the tuned one
    other = 16 if rep.block_width != 16 else 32
    forced_w = jsvd.Plan(m, n, _plain_cfg(), path=jsvd.PATH_MIXED,
                         mixed=jsvd.mixed_config(jsvd.PATH_MIXED, block_width=other)).launch_params()
    assert forced_w["block_width"] == other
    # a forced layout mode keeps the built-in rule
    forced = jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=1e-15, layout_transpose_mode=2)).launch_params()
    assert forced == before
    jsvd.autotune_clear()
    assert jsvd.Plan(m, n, _plain_cfg()).launch_params() == before


@pytest.mark.parametrize("m,n", [(301, 77), (1000, 999), (2049, 300)])
def test_tuned_plain_plan_solves(torch_cuda, m, n):
    a = gaussian(m, n, 11)
    rep = jsvd.autotune(m, n, repetitions=1)
    plan = jsvd.Plan(m, n, _plain_cfg())
    p = plan.launch_params()
    assert (p["round_threads"], p["round_ea"], p["round_ev"]) == (rep.round_threads, rep.round_ea, rep.round_ev)
    got = plan.solve_host(a)
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(sorted_desc(got.sigma) - s_ref)) / s_ref[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-12


def test_tuned_mixed_plan_and_reshape(torch_cuda):
    m, n = 1024, 256
    rep = jsvd.autotune(m, n, repetitions=1)
    plan = jsvd.Plan(2048, 512, _plain_cfg(), path=jsvd.PATH_MIXED)
    p0 = plan.launch_params()
    assert (p0["apply_workgroups"], p0["block_width"]) == (512, 32)  # no tuning for 2048 x 512: built-in
    plan.reshape(m, n, 1e-15)
    p1 = plan.launch_params()
    assert (p1["apply_workgroups"], p1["block_width"]) == (rep.rt_apply_workgroups, rep.rt_block_width)
    a = gaussian(m, n, 5)
    got = plan.solve_host(a)
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(sorted_desc(got.sigma) - s_ref)) / s_ref[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-12


def test_small_shapes_and_bad_arguments(torch_cuda):
    rep = jsvd.autotune(5, 1, repetitions=1)  # one column: nothing to time
    assert rep.candidates_tried == 0
    rep = jsvd.autotune(8, 8, repetitions=2)  # square, narrower than two blocks: plain candidates only
    assert rep.candidates_tried >= 1 and rep.block_round_us_best == 0.0 and rep.block_width == 0
    with pytest.raises(ValueError):
        jsvd.autotune(4, 8)
    with pytest.raises(ValueError):
        jsvd.autotune(8, 4, repetitions=0)
