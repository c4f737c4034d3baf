"""Host-side checks of libjsvd.so that need no GPU: it loads, exports every symbol include/jsvd.h
declares, carries the reference defaults, validates inputs like the reference (src/kernels.cu:807-833,
338-356) before touching the device, and its closed-form schedule equals the oracle's schedule."""
import os

import numpy as np
import pytest

import jsvd
import oracle


def test_library_exports_every_declared_symbol():
    lib = jsvd.lib()
    names = jsvd.declared_symbols()
    assert "jsvd_plain_f64" in names and "jsvd_round_f64" in names
    for name in names:
        assert hasattr(lib, name), name


def test_default_config_matches_reference():
    c = jsvd.default_config()
    assert c.epsilon == 1e-9 and c.max_sweeps == 128 and c.threads_per_block == 256
    assert c.layout_transpose_mode == jsvd.LAYOUT_AUTO
    assert c.layout_transpose_min_columns == 16 and c.layout_transpose_min_elements == 4096
    assert c.layout_transpose_auto_tune == 0
    assert c.layout_transpose_benchmark_repetitions == 2 and c.layout_transpose_benchmark_sweeps == 8


@pytest.mark.parametrize("shape,cfg,msg", [
    ((3, 4), {}, "requires rows >= columns"),
    ((0, 0), {}, "shape must be non-zero"),
    ((4, 3), {"epsilon": 0.0}, "epsilon must be positive"),
    ((4, 3), {"epsilon": -1.0}, "epsilon must be positive"),
    ((4, 3), {"max_sweeps": 0}, "max_sweeps must be positive"),
    ((4, 3), {"layout_transpose_min_columns": 0}, "layout_transpose_min_columns must be positive"),
    ((4, 3), {"layout_transpose_min_elements": 0}, "layout_transpose_min_elements must be positive"),
    ((4, 3), {"layout_transpose_benchmark_repetitions": 0}, "repetitions must be positive"),
    ((4, 3), {"layout_transpose_benchmark_sweeps": 0}, "benchmark_sweeps must be positive"),
])
def test_invalid_arguments_raise_before_device(shape, cfg, msg):
    a = np.zeros(shape)
    with pytest.raises(ValueError, match=msg):
        jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(**cfg))


@pytest.mark.parametrize("n", list(range(1, 34)) + [64, 99, 128, 255, 256, 1023, 1024])
def test_closed_form_schedule_equals_oracle(n):
    ref = oracle.round_robin(n)
    if n < 2:
        assert ref == []
        assert jsvd.round_robin_pairs(n, 0) == []
        return
    for r, pairs in enumerate(ref):
        assert jsvd.round_robin_pairs(n, r) == pairs


def test_schedule_round_out_of_range():
    with pytest.raises(ValueError):
        jsvd.round_robin_pairs(8, 7)


def test_mat_helpers_match_hand_computed_record(tmp_path):
    """jsvd.write_mat / read_mat against the hand-computed bytes of a 1 x 2 record [1.0, -2.0]."""
    import jsvd
    golden = open(os.path.join(os.path.dirname(__file__), "golden", "mat_1x2.bin"), "rb").read()
    p = str(tmp_path / "x.mat")
    jsvd.write_mat(p, [np.array([[1.0, -2.0]])])
    assert open(p, "rb").read() == golden
    back = jsvd.read_mat(p)
    assert len(back) == 1 and back[0].shape == (1, 2) and np.array_equal(back[0], [[1.0, -2.0]])
    open(p, "wb").write(golden[:-3])
    with pytest.raises(RuntimeError, match="payload"):
        jsvd.read_mat(p)


def test_config_validate_without_a_matrix():
    """jsvd_config_validate: the reference's validation messages, no matrix and no device needed."""
    jsvd.validate_config(jsvd.JacobiSvdConfig())
    with pytest.raises(ValueError, match="epsilon"):
        jsvd.validate_config(jsvd.JacobiSvdConfig(epsilon=0.0))
    with pytest.raises(ValueError, match="max_sweeps"):
        jsvd.validate_config(jsvd.JacobiSvdConfig(max_sweeps=0))
