"""The library's streaming copy kernel (jsvd_hbm_copy, the measured practical HBM roof of bench.py): a bit-exact
copy for sizes that fill whole chunks, partial chunks and a single vector, in both store variants; bad sizes and
alignments are refused."""
import numpy as np
import pytest

import jsvd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("count", [2, 1024 * 64, 1024 * 64 * 17 + 6, 3 * 1024 * 1024 + 10])
def test_hbm_copy_bitwise(torch_cuda, variant, count):
    torch = torch_cuda
    src = torch.randn(count, dtype=torch.float64, device="cuda")
    dst = torch.zeros_like(src)
    jsvd.hbm_copy(dst, src, variant)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)


def test_hbm_copy_refuses_bad_arguments(torch_cuda):
    torch = torch_cuda
    src = torch.zeros(16, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        jsvd.hbm_copy(torch.zeros(16, dtype=torch.float32, device="cuda"), src, 2)
    with pytest.raises(ValueError):
        jsvd.hbm_copy(torch.zeros(3, dtype=torch.float32, device="cuda"), src[:3], 0)   # 12 bytes
    with pytest.raises(ValueError):
        jsvd.hbm_copy(torch.zeros(15, dtype=torch.float32, device="cuda")[:12], src[1:13], 0)   # misaligned
