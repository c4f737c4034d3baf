"""Full-size workloads on cuda:0, checked through size-independent properties.

* Tall geometric 4096 x 1024 (A = Q1 diag(10^(-8 i / 1023)) Q2^T, the bench workload): both paths against the
  exact singular values of the construction, normwise max |d sigma| / sigma_max <= 1e-12 and per-sigma relative
  <= 1e-12 where sigma_i >= 1e-3 sigma_max; ||A - U S V^T||_F / ||A||_F <= 1e-12; ||U^T U - I||_F and
  ||V^T V - I||_F <= 1e-10; the two paths agree normwise to 1e-12; both converge (a sweep without rotations).
* Column-graded 16384 x 4096 (G diag(10^(-12 j / 4095)), G ~ N(0, 1)): the mixed path against the plain FP64
  path (the reference algorithm) per singular value, relative <= 1e-10 over all 4096 of them (column scales
  span 12 orders; one-sided Jacobi is relatively accurate on column-graded inputs), plus the residual and
  orthogonality bounds above (1e-11 for U, V at this size).
* The 20,000-matrix uneven batch: no item rejected, every item converged within its cap; a sample of FP64 items
  against the oracle column by column (sigma relative 1e-12, U and V entries 1e-9, sweeps +-1), a sample of
  FP32 items against LAPACK (sorted sigma, 2e-5 of sigma_max); items re-solved alone equal their results in the
  full batch bit for bit.
"""
import numpy as np
import pytest

import jsvd
import oracle
from helpers import sorted_desc

pytestmark = pytest.mark.gpu


def _device_solve(torch, d_a, path, eps, max_sweeps):
    m, n = d_a.shape
    d_u = torch.empty((m, n), dtype=torch.float64, device="cuda")
    d_s = torch.empty(n, dtype=torch.float64, device="cuda")
    d_v = torch.empty((n, n), dtype=torch.float64, device="cuda")
    plan = jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=eps, max_sweeps=max_sweeps), path=path)
    plan.solve_device(d_a, d_u, d_s, d_v)
    torch.cuda.synchronize()
    info = plan.last_info()
    plan.close()
    return d_u, d_s, d_v, info


def _invariants(torch, d_a, d_u, d_s, d_v):
    n = d_s.numel()
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    rho_u = float(torch.linalg.norm(d_u.T @ d_u - eye))
    rho_v = float(torch.linalg.norm(d_v.T @ d_v - eye))
    rho_a = float(torch.linalg.norm(d_a - (d_u * d_s) @ d_v.T) / torch.linalg.norm(d_a))
    return rho_a, rho_u, rho_v


def test_tall_geometric_4096x1024_both_paths(torch_cuda):
    torch = torch_cuda
    from jsvd import workloads
    a, s_true = workloads.tall_geometric(4096, 1024)
    d_a = torch.from_numpy(a).cuda()
    sig = {}
    for path in (jsvd.PATH_MIXED, jsvd.PATH_PLAIN_F64):
        print(f"tall 4096x1024 path {jsvd.PATH_NAMES[path]}", flush=True)
        d_u, d_s, d_v, info = _device_solve(torch, d_a, path, 1e-14, 60)
        assert info.converged == 1, path
        s = sorted_desc(d_s.cpu().numpy())
        assert np.max(np.abs(s - s_true)) / s_true[0] < 1e-12, path
        big = s_true >= 1e-3 * s_true[0]
        assert np.max(np.abs(s - s_true)[big] / s_true[big]) < 1e-12, path
        rho_a, rho_u, rho_v = _invariants(torch, d_a, d_u, d_s, d_v)
        assert rho_a < 1e-12 and rho_u < 1e-10 and rho_v < 1e-10, (path, rho_a, rho_u, rho_v)
        sig[path] = s
        del d_u, d_v
    assert np.max(np.abs(sig[jsvd.PATH_MIXED] - sig[jsvd.PATH_PLAIN_F64])) / s_true[0] < 1e-12


def test_column_graded_16384x4096_mixed_against_plain(torch_cuda):
    torch = torch_cuda
    from jsvd import graded
    a = graded.graded_matrix(16384, 4096, seed=5)
    d_a = torch.from_numpy(a).cuda()
    del a
    sig = {}
    for path in (jsvd.PATH_MIXED, jsvd.PATH_PLAIN_F64):
        print(f"graded 16384x4096 path {jsvd.PATH_NAMES[path]}", flush=True)
        d_u, d_s, d_v, info = _device_solve(torch, d_a, path, 2e-14, 80)
        assert info.converged == 1, path
        rho_a, rho_u, rho_v = _invariants(torch, d_a, d_u, d_s, d_v)
        assert rho_a < 1e-12 and rho_u < 1e-10 and rho_v < 1e-10, (path, rho_a, rho_u, rho_v)
        sig[path] = sorted_desc(d_s.cpu().numpy())
        del d_u, d_v
        torch.cuda.empty_cache()
    ref = sig[jsvd.PATH_PLAIN_F64]
    assert ref[-1] > 0 and ref[0] / ref[-1] > 1e11   # the spectrum spans the column scales
    assert np.max(np.abs(sig[jsvd.PATH_MIXED] - ref) / ref) < 1e-10


def test_uneven_batch_20000(torch_cuda):
    torch = torch_cuda
    from jsvd import workloads
    items = workloads.batch_mixed(20000)
    mats = [a for a, _, _ in items]
    prec = [jsvd.PRECISION_F32 if p == "f32" else jsvd.PRECISION_F64 for _, p, _ in items]
    tol = [t for _, _, t in items]
    b = jsvd.Batch(256, 128)
    print("batch 20000 solve", flush=True)
    us, ss, vs, sweeps = b.solve(mats, prec, tol, max_sweeps=60)
    assert np.all(sweeps >= 1) and np.all(sweeps < 60)
    rng = np.random.default_rng(0)
    f64 = [i for i in range(len(mats)) if prec[i] == jsvd.PRECISION_F64]
    f32 = [i for i in range(len(mats)) if prec[i] == jsvd.PRECISION_F32]
    for i in rng.choice(f64, 60, replace=False):
        ref = oracle.svd(mats[i], eps=1e-14, max_sweeps=60, threads=4)
        assert abs(int(sweeps[i]) - ref.sweeps) <= 1, (i, sweeps[i], ref.sweeps)
        assert np.max(np.abs(ss[i] - ref.sigma) / ref.sigma) < 1e-12, i
        assert np.max(np.abs(vs[i] - ref.v)) < 1e-9 and np.max(np.abs(us[i] - ref.u)) < 1e-9, i
    for i in rng.choice(f32, 60, replace=False):
        s_ref = sorted_desc(np.linalg.svd(mats[i], compute_uv=False))
        assert np.max(np.abs(sorted_desc(ss[i]) - s_ref)) / s_ref[0] < 2e-5, i
    # the same items alone: bit-identical to their results inside the full batch
    picks = [int(x) for x in rng.choice(len(mats), 6, replace=False)]
    u1, s1, v1, sw1 = b.solve([mats[k] for k in picks], [prec[k] for k in picks], [tol[k] for k in picks],
                              max_sweeps=60)
    for j, k in enumerate(picks):
        assert np.array_equal(s1[j], ss[k]) and np.array_equal(u1[j], us[k]) and np.array_equal(v1[j], vs[k])
        assert sw1[j] == sweeps[k]
    b.close()
