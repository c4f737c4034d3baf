"""jsvd_batch_solve_noalloc (the batched solve with the library's fixed device scratch) against
jsvd_batch_solve_device (batch object with its own allocations): bit-identical U, sigma, V and sweep counts on two
calls in a row, the second with a smaller batch; no device memory allocated by the call; oversized limits refused."""
import numpy as np
import pytest

import jsvd

pytestmark = pytest.mark.gpu


def _batch(seed, count):
    rng = np.random.default_rng(seed)
    mats, prec, tol = [], [], []
    shapes = [(256, 128), (200, 100), (255, 127)]  # A and V in the scratch slices at the default launch shape
    for i in range(count):
        if i < len(shapes):
            m, n = shapes[i]
        else:
            m = int(rng.integers(8, 257))
            n = int(rng.integers(jsvd.BATCH_MIN_COLUMNS, min(m, 128) + 1))
        mats.append(rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-3, 3))
        prec.append(jsvd.PRECISION_F32 if i % 2 else jsvd.PRECISION_F64)
        tol.append(1e-6 if i % 2 else 1e-14)
    return mats, prec, tol


def _device(torch, mats, prec, tol):
    pk = jsvd.BatchPacked(mats, prec, tol, 60)
    pk.check_limits(256, 128)
    d_items = torch.from_numpy(pk.items[pk.cost_order()].view(np.uint8).copy()).to("cuda")
    d = [torch.from_numpy(pk.a).to("cuda")]
    d += [torch.zeros(max(n, 1), dtype=torch.float64, device="cuda") for n in pk.lens[1:]]
    d.append(torch.zeros(len(mats), dtype=torch.int32, device="cuda"))
    return d_items, d


def _reference(torch, d_items, d, count):
    b = jsvd.Batch(256, 128)
    b.solve_device(d_items, count, *d)
    out = [x.cpu().numpy().copy() for x in d[1:]]
    b.close()
    for x in d[1:]:
        x.fill_(-7)
    torch.cuda.synchronize()
    return out


def test_noalloc_equals_batch_object_twice(torch_cuda):
    torch = torch_cuda
    for seed, count in ((21, 48), (22, 9)):  # the second call is smaller
        mats, prec, tol = _batch(seed, count)
        d_items, d = _device(torch, mats, prec, tol)
        ref = _reference(torch, d_items, d, count)
        free_before = torch.cuda.mem_get_info()[0]
        jsvd.batch_solve_noalloc(d_items, count, *d, max_rows=256, max_cols=128)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free_before
        got = [x.cpu().numpy() for x in d[1:]]
        assert np.all(got[-1] >= 1), (seed, got[-1])
        for name, x, y in zip(("u", "sigma", "v", "sweeps"), got, ref):
            assert np.array_equal(x, y), (seed, name)


def test_noalloc_refuses_oversized_limits(torch_cuda):
    torch = torch_cuda
    mats, prec, tol = _batch(23, 4)
    d_items, d = _device(torch, mats, prec, tol)
    with pytest.raises(ValueError, match="scratch slice"):
        jsvd.batch_solve_noalloc(d_items, 4, *d, max_rows=4096, max_cols=4096)
    with pytest.raises(ValueError, match="limits"):
        jsvd.batch_solve_noalloc(d_items, 4, *d, max_rows=10, max_cols=20)
