"""Block FP64 and mixed FP32/FP64 paths on cuda:0 against the CPU oracle and LAPACK dgesvj.

These paths use a different (blocked) rotation order than the reference, so columns come out in a
different Jacobi order: sigma is compared sorted, U and V through invariants and matched columns.
Tolerances:
  * sorted sigma vs the oracle / dgesvj: normwise max|d sigma| / sigma_max <= 1e-12 (the metric's
    "sigma relerr <= 1e-12"), and per-sigma relative <= 1e-12 where sigma_i >= 1e-3 sigma_max;
  * rho_A <= 1e-13, ||U^T U - I||_F and ||V^T V - I||_F <= 1e-11 (n up to 256);
  * column-graded inputs (relative accuracy of Jacobi): per-sigma relative error vs dgesvj <= 1e-10.
"""
import numpy as np
import pytest

import jsvd
import oracle
from helpers import gaussian, match_columns, rho_a, rho_orth, sorted_desc

pytestmark = pytest.mark.gpu

PATHS = [jsvd.PATH_BLOCK_F64, jsvd.PATH_MIXED]


def _solve(a, path, eps=1e-14, max_sweeps=60):
    return jsvd.solve(a, jsvd.JacobiSvdConfig(epsilon=eps, max_sweeps=max_sweeps), path=path)


@pytest.mark.parametrize("path", PATHS)
def test_gaussian_512x256_against_oracle(torch_cuda, path):
    a = gaussian(512, 256, 0)
    got, info = _solve(a, path, eps=1e-15, max_sweeps=40)
    ref = oracle.svd(a, eps=1e-15, max_sweeps=30, threads=4)
    s_ref = sorted_desc(ref.sigma)
    s = sorted_desc(got.sigma)
    assert info.converged == 1 and info.path == path
    if path == jsvd.PATH_MIXED:
        assert info.sweeps_fp32 > 0
    assert np.max(np.abs(s - s_ref)) / s_ref[0] < 1e-12
    assert np.max(np.abs(s - s_ref) / s_ref) < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
    assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11
    # singular vectors equal the oracle's up to sign after sorting (well separated sigma)
    o1 = np.argsort(-got.sigma, kind="stable")
    o2 = np.argsort(-ref.sigma, kind="stable")
    assert match_columns(got.v[:, o1], ref.v[:, o2]) < 1e-8


@pytest.mark.parametrize("path", PATHS)
def test_geometric_spectrum_1024x256(torch_cuda, path):
    """A = Q1 diag(1 .. 1e-8) Q2^T: normwise sigma error vs the exact construction and vs dgesvj."""
    from jsvd import workloads
    a, s_true = workloads.tall_geometric(1024, 256)
    got, info = _solve(a, path)
    s = sorted_desc(got.sigma)
    _, s_lapack, _, _ = oracle.dgesvj(a)
    assert info.converged == 1
    assert np.max(np.abs(s - s_true)) / s_true[0] < 1e-12
    assert np.max(np.abs(s - s_lapack)) / s_lapack[0] < 1e-12
    big = s_true >= 1e-3
    assert np.max(np.abs(s - s_true)[big] / s_true[big]) < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-12
    assert rho_orth(got.v) < 1e-11


@pytest.mark.parametrize("path", PATHS)
def test_column_graded_relative_accuracy(torch_cuda, path):
    rng = np.random.default_rng(5)
    m, n = 600, 160
    a = rng.standard_normal((m, n)) * 10.0 ** (-12.0 * np.arange(n) / (n - 1))
    got, info = _solve(a, path)
    _, s_lapack, _, _ = oracle.dgesvj(a)
    s = sorted_desc(got.sigma)
    assert info.converged == 1
    assert np.max(np.abs(s - s_lapack) / s_lapack) < 1e-10


@pytest.mark.parametrize("m,n", [(70, 33), (129, 65), (300, 100), (257, 257)])
@pytest.mark.parametrize("path", PATHS)
def test_padded_shapes(torch_cuda, path, m, n):
    """Column counts that are not multiples of the 32-column block and rows not multiples of the slab."""
    a = gaussian(m, n, m * 7 + n)
    got, info = _solve(a, path)
    _, s_lapack, _, _ = oracle.dgesvj(a)
    s = sorted_desc(got.sigma)
    assert info.converged == 1
    assert np.max(np.abs(s - s_lapack)) / s_lapack[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
    assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11


def test_mixed_plan_reuse_deterministic(torch_cuda):
    torch = torch_cuda
    m, n = 400, 96
    a = gaussian(m, n, 3)
    plan = jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60), path=jsvd.PATH_MIXED)
    d_a = torch.from_numpy(a).to("cuda")
    outs = []
    for _ in range(2):
        u = torch.zeros((m, n), dtype=torch.float64, device="cuda")
        s = torch.zeros(n, dtype=torch.float64, device="cuda")
        v = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        plan.solve_device(d_a, u, s, v)
        outs.append((u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy(), plan.last_info().as_dict()))
    for x, y in zip(outs[0][:3], outs[1][:3]):
        assert np.array_equal(x, y)
    assert outs[0][3] == outs[1][3]
    plan.close()


def test_block_paths_reject_small_n():
    with pytest.raises(ValueError, match="two 32-column blocks"):
        jsvd.solve(gaussian(40, 32, 0), path=jsvd.PATH_MIXED)


@pytest.mark.parametrize("kw", [32, 64, 128])
@pytest.mark.parametrize("dtype,tol", [("float64", 1e-12), ("float32", 2e-4)])
def test_block_eigen_matches_gram_space_sweep(torch_cuda, dtype, tol, kw):
    """The inner step of a block round (jsvd_block_eigen) at every block width (kw = 2 x 16, 32, 64 columns)
    against the numpy restatement of one Gram-space sweep (tests/helpers.py): same rotation order and formula,
    so Q agrees to rounding (FP64) or to FP32 accumulation (FP32); Q is orthogonal and Q^T G Q equals the
    swept G."""
    import torch
    from helpers import gram_space_sweep
    rng = np.random.default_rng(11 + kw)
    pairs = 5
    grams = []
    for p in range(pairs):
        x = rng.standard_normal((300, kw)) * 10.0 ** (-(p + 1) * np.arange(kw) / (kw - 1.0))
        x[:, kw - 1 - p] = 0.0  # a zero column never rotates
        grams.append(x.T @ x)
    g = np.stack(grams)
    tdt = getattr(torch, dtype)
    d_g = torch.from_numpy(g).to(device="cuda", dtype=tdt).contiguous()
    d_q = torch.empty_like(d_g)
    d_rot = torch.zeros(pairs, dtype=torch.int32, device="cuda")
    eps = 1e-15 if dtype == "float64" else 1e-7
    jsvd.block_eigen(d_g, eps, d_q, d_rot)
    torch.cuda.synchronize()
    q = d_q.cpu().numpy().astype(np.float64)
    rot = d_rot.cpu().numpy()
    for p in range(pairs):
        g_in = d_g[p].cpu().numpy().astype(np.float64)
        q_ref, rot_ref, g_ref = gram_space_sweep(g_in, eps)
        assert np.max(np.abs(q[p] - q_ref)) < tol, p
        if dtype == "float64":
            assert rot[p] == rot_ref
        assert rho_orth(q[p]) < (1e-13 if dtype == "float64" else 1e-5)
        swept = q[p].T @ g_in @ q[p]
        assert np.max(np.abs(swept - g_ref)) / np.max(np.abs(g_in)) < (1e-13 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-12), ("float32", 2e-4)])
def test_block_eigen_cross_matches_gram_space_sweep(torch_cuda, dtype, tol):
    """The cross ordering of the block rounds (jsvd_block_eigen_ordered, kw = 64) against the numpy restatement of
    the same ordering (tests/helpers.py, cross=True): Q agrees to rounding, rotation counts agree (FP64), Q is
    orthogonal and Q^T G Q equals the swept G."""
    import torch
    from helpers import gram_space_sweep
    rng = np.random.default_rng(77)
    pairs, kw = 5, 64
    grams = []
    for p in range(pairs):
        x = rng.standard_normal((300, kw)) * 10.0 ** (-(p + 1) * np.arange(kw) / (kw - 1.0))
        x[:, kw - 1 - p] = 0.0
        grams.append(x.T @ x)
    tdt = getattr(torch, dtype)
    d_g = torch.from_numpy(np.stack(grams)).to(device="cuda", dtype=tdt).contiguous()
    d_q = torch.empty_like(d_g)
    d_rot = torch.zeros(pairs, dtype=torch.int32, device="cuda")
    eps = 1e-15 if dtype == "float64" else 1e-7
    jsvd.block_eigen(d_g, eps, d_q, d_rot, inner_order=jsvd.INNER_CROSS)
    torch.cuda.synchronize()
    q = d_q.cpu().numpy().astype(np.float64)
    rot = d_rot.cpu().numpy()
    for p in range(pairs):
        g_in = d_g[p].cpu().numpy().astype(np.float64)
        q_ref, rot_ref, g_ref = gram_space_sweep(g_in, eps, cross=True)
        assert np.max(np.abs(q[p] - q_ref)) < tol, p
        if dtype == "float64":
            assert rot[p] == rot_ref
        assert rho_orth(q[p]) < (1e-13 if dtype == "float64" else 1e-5)
        swept = q[p].T @ g_in @ q[p]
        assert np.max(np.abs(swept - g_ref)) / np.max(np.abs(g_in)) < (1e-13 if dtype == "float64" else 1e-5)
    with pytest.raises(ValueError, match="inner_order"):
        g16 = torch.zeros((2, 32, 32), dtype=tdt, device="cuda")
        jsvd.block_eigen(g16, eps, torch.empty_like(g16), torch.zeros(2, dtype=torch.int32, device="cuda"),
                         inner_order=jsvd.INNER_CROSS)


def test_block_eigen_rejects_bad_arguments(torch_cuda):
    import torch
    g = torch.zeros((2, 64, 64), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        jsvd.block_eigen(g[:, :32, :], 1e-15, torch.empty_like(g), torch.zeros(2, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        jsvd.block_eigen(g, -1.0, torch.empty_like(g), torch.zeros(2, dtype=torch.int32, device="cuda"))
    g48 = torch.zeros((2, 48, 48), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        jsvd.block_eigen(g48, 1e-15, torch.empty_like(g48), torch.zeros(2, dtype=torch.int32, device="cuda"))


@pytest.mark.parametrize("block_width", [16, 32, 64])
@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("m,n", [(512, 256), (300, 97)])
def test_block_widths_against_oracle(torch_cuda, path, block_width, m, n):
    """Every instantiated block width (16, 32, 64 columns; 300 x 97 leaves an odd block count at width 16 and
    a padded last block at every width) solves to the same bar as the default width."""
    a = gaussian(m, n, 3 + block_width)
    mx = jsvd.mixed_config(path, block_width=block_width)
    plan = jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=40), path=path, mixed=mx)
    got = plan.solve_host(a)
    info = plan.last_info()
    s_ref = np.linalg.svd(a, compute_uv=False)
    s = sorted_desc(got.sigma)
    assert info.converged == 1
    assert np.max(np.abs(s - s_ref)) / s_ref[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
    assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11


def test_block_width_argument_checks(torch_cuda):
    with pytest.raises(ValueError):
        jsvd.Plan(100, 60, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, block_width=48))
    # a width that would leave one block is halved: 40 columns at width 64 run at width 32
    plan = jsvd.Plan(100, 40, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=40), path=jsvd.PATH_BLOCK_F64,
                     mixed=jsvd.mixed_config(jsvd.PATH_BLOCK_F64, block_width=64))
    assert plan.launch_params()["block_width"] == 32
    a = gaussian(100, 40, 9)
    got = plan.solve_host(a)
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(sorted_desc(got.sigma) - s_ref)) / s_ref[0] < 1e-12


@pytest.mark.parametrize("path,qr", [(jsvd.PATH_PLAIN_F64, 1), (jsvd.PATH_BLOCK_F64, 1), (jsvd.PATH_MIXED, 1),
                                     (jsvd.PATH_MIXED, 2)])
def test_plan_reshape_equals_fresh_plans(torch_cuda, path, qr):
    """One plan re-targeted at several smaller shapes and epsilons (jsvd_plan_reshape) gives, bit for bit, the
    results of fresh plans created for each shape (mixed: with one QR and with the second QR); shapes beyond the
    capacity are rejected."""
    import torch
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    mx = jsvd.mixed_config(path, qr_precondition=qr)
    big = jsvd.Plan(400, 160, cfg, path=path, mixed=mx)
    shapes = [(400, 160), (257, 97), (300, 40), (101, 101), (399, 159)]
    if path == jsvd.PATH_PLAIN_F64:
        shapes.append((50, 7))
    for k, (m, n) in enumerate(shapes):
        eps = 1e-14 if k % 2 == 0 else 1e-12
        a = gaussian(m, n, 40 + k)
        big.reshape(m, n, eps)
        d_a = torch.from_numpy(a).cuda()
        outs = []
        for plan in (big, jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=eps, max_sweeps=60), path=path, mixed=mx)):
            d_u = torch.empty((m, n), dtype=torch.float64, device="cuda")
            d_s = torch.empty(n, dtype=torch.float64, device="cuda")
            d_v = torch.empty((n, n), dtype=torch.float64, device="cuda")
            plan.solve_device(d_a, d_u, d_s, d_v)
            outs.append((d_u.cpu().numpy(), d_s.cpu().numpy(), d_v.cpu().numpy()))
        for x, y in zip(outs[0], outs[1]):
            assert np.array_equal(x, y), (m, n)
    with pytest.raises(ValueError):
        big.reshape(401, 160, 1e-14)
    with pytest.raises(ValueError):
        big.reshape(300, 161, 1e-14)
    big.close()


@pytest.mark.parametrize("m,n", [(1024, 256), (300, 97), (257, 257)])
def test_mixed_qr_first_agrees_with_direct_form(torch_cuda, m, n):
    """The mixed path's QR-first form (A = Q R, sweeps on R^T) and its direct form (sweeps on A) solve the same
    problem to the same bar; on a geometric spectrum the QR-first form needs fewer FP32 and FP64 sweeps."""
    from jsvd import workloads
    a, s_true = workloads.tall_geometric(m, n) if m > n else (gaussian(m, n, 12), None)
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    res = {}
    for qr in (True, False):
        plan = jsvd.Plan(m, n, cfg, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, qr_precondition=qr))
        got = plan.solve_host(a)
        res[qr] = (got, plan.last_info())
        assert res[qr][1].converged == 1
        assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
        assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11
    s1, s0 = sorted_desc(res[True][0].sigma), sorted_desc(res[False][0].sigma)
    assert np.max(np.abs(s1 - s0)) / s0[0] < 1e-12
    if s_true is not None:
        assert np.max(np.abs(s1 - s_true)) / s_true[0] < 1e-12
        assert res[True][1].sweeps_fp32 < res[False][1].sweeps_fp32
        assert res[True][1].sweeps_fp64 < res[False][1].sweeps_fp64


def test_mixed_rank_deficient_and_zero_inputs(torch_cuda):
    """Exactly rank-deficient inputs: a zero matrix and a matrix with a zero column (the shifted Cholesky QR
    breaks down and the solve falls back to the direct form), and a rank-60 matrix of 80 columns (sigma below
    epsilon: those U columns are zero as in the reference; V stays orthogonal)."""
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    zero = np.zeros((100, 40))
    got = jsvd.Plan(100, 40, cfg, path=jsvd.PATH_MIXED).solve_host(zero)
    assert np.all(got.sigma == 0.0) and np.all(got.u == 0.0)
    assert rho_orth(got.v) < 1e-12
    a = gaussian(120, 50, 21)
    a[:, 17] = 0.0
    got = jsvd.Plan(120, 50, cfg, path=jsvd.PATH_MIXED).solve_host(a)
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(sorted_desc(got.sigma) - s_ref)) / s_ref[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13 and rho_orth(got.v) < 1e-11
    rng = np.random.default_rng(22)
    b = rng.standard_normal((200, 60)) @ rng.standard_normal((60, 80))
    b /= np.linalg.norm(b, 2)  # sigma_max = 1: the 20 null singular values land at rounding level, below 1e-14
    got = jsvd.Plan(200, 80, cfg, path=jsvd.PATH_MIXED).solve_host(b)
    s = got.sigma
    big = s > 1e-14
    assert big.sum() == 60
    assert np.all(got.u[:, ~big] == 0.0)
    assert rho_orth(got.v) < 1e-11
    assert rho_orth(got.u[:, big]) < 1e-11
    assert rho_a(b, got.u, got.sigma, got.v) < 1e-13


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("shape", [(700, 224), (300, 200)])
def test_inner_orders_agree(torch_cuda, path, shape):
    """Block rounds at width 32 in the cross ordering (intra-block step + cross pairs, the default) and in the full
    ordering give the same sorted sigma (1e-12 relative) and both meet the path's accuracy bounds; 200 columns
    leave an odd block count (a dummy block in the circle) and a padded last block."""
    m, n = shape
    a = gaussian(m, n, 17) * 10.0 ** (-4 * np.arange(n) / n)
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    out = {}
    for order in (jsvd.INNER_CROSS, jsvd.INNER_FULL):
        got, info = jsvd.solve(a, cfg, path=path, mixed=jsvd.mixed_config(path, block_width=32, inner_order=order))
        assert info.converged == 1, order
        assert rho_a(a, got.u, got.sigma, got.v) < 1e-13, order
        assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11, order
        out[order] = sorted_desc(got.sigma)
    s_ref = sorted_desc(np.linalg.svd(a, compute_uv=False))
    for order, s in out.items():
        assert np.max(np.abs(s - s_ref) / s_ref) < 1e-12, order
    assert np.max(np.abs(out[jsvd.INNER_CROSS] - out[jsvd.INNER_FULL]) / out[jsvd.INNER_FULL]) < 1e-12


def test_inner_order_validated(torch_cuda):
    a = gaussian(80, 64, 1)
    with pytest.raises(ValueError, match="inner_order"):
        jsvd.solve(a, jsvd.JacobiSvdConfig(epsilon=1e-14), path=jsvd.PATH_MIXED,
                   mixed=jsvd.mixed_config(jsvd.PATH_MIXED, inner_order=7))


@pytest.mark.parametrize("fraction", [0.0, 0.05, 0.3])
def test_fp32_phase_stop_fraction(torch_cuda, fraction):
    """The FP32 phase ends on the first sweep that rotates at most fraction x pairs (0: a rotation-free sweep);
    every earlier FP32 sweep rotated more; the FP64 phase still ends on a rotation-free sweep; accuracy holds."""
    m, n = 1024, 256
    rng = np.random.default_rng(21)
    q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    a = (q1 * sig) @ q2.T
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    plan = jsvd.Plan(m, n, cfg, path=jsvd.PATH_MIXED,
                     mixed=jsvd.mixed_config(jsvd.PATH_MIXED, fp32_stop_fraction=fraction))
    got = plan.solve_host(a)
    info = plan.last_info()
    rot, _ = plan.sweep_stats()
    k = info.sweeps_fp32
    stop = int(fraction * n * (n - 1) / 2)
    assert k >= 1 and len(rot) == k + info.sweeps_fp64
    assert rot[k - 1] <= stop and all(r > stop for r in rot[:k - 1])
    assert rot[-1] == 0 and info.converged == 1
    s = sorted_desc(got.sigma)
    assert np.max(np.abs(s - sig)) / sig[0] < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
    assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11
    with pytest.raises(ValueError):
        jsvd.Plan(m, n, cfg, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, fp32_stop_fraction=1.0))


@pytest.mark.parametrize("m,n,kind", [(1024, 256, "geometric"), (600, 200, "gaussian"), (256, 256, "geometric")])
def test_mixed_second_qr(torch_cuda, m, n, kind):
    """qr_precondition 2 (A = Q R, then R^T = Q2 R2, sweeps on R2^T, U = Q Y, V = Q2 V_N) against one QR: the same
    sorted sigma (1e-12 normwise; 1e-12 of the exact sigma of the geometric construction), U and V orthogonal,
    A = U S V^T, converged; on the kappa = 1e8 construction the FP32 phase needs no more sweeps than with one QR."""
    rng = np.random.default_rng(13)
    if kind == "geometric":
        q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
        q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        exact = 10.0 ** (-8.0 * np.arange(n) / (n - 1))
        a = (q1 * exact) @ q2.T
    else:
        exact = None
        a = rng.standard_normal((m, n))
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    out = {}
    for qr in (1, 2):
        print(f"second qr: {m}x{n} {kind} qr_precondition={qr}", flush=True)
        got, info = jsvd.solve(a, cfg, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, qr_precondition=qr))
        assert info.converged == 1, qr
        assert rho_a(a, got.u, got.sigma, got.v) < 1e-13, qr
        assert rho_orth(got.u) < 1e-11 and rho_orth(got.v) < 1e-11, qr
        out[qr] = (sorted_desc(got.sigma), info)
    s1, i1 = out[1]
    s2, i2 = out[2]
    assert np.max(np.abs(s2 - s1)) / s1[0] < 1e-12
    if exact is not None:
        assert np.max(np.abs(s2 - exact)) / exact[0] < 1e-12
        assert i2.sweeps_fp32 <= i1.sweeps_fp32, (i1.as_dict(), i2.as_dict())
    with pytest.raises(ValueError, match="qr_precondition"):
        jsvd.Plan(m, n, cfg, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, qr_precondition=4))


@pytest.mark.parametrize("m,n,kind,expected", [(2048, 512, "geometric", 2), (1024, 256, "geometric", 1),
                                               (1200, 600, "gaussian", 1)])
def test_mixed_second_qr_automatic(torch_cuda, m, n, kind, expected):
    """qr_precondition 3 (the default) takes the second QR only for n >= 512 and an ill-conditioned R (diagonal
    spread >= 100): its results equal, bit for bit, those of the form it chose -- two QRs on a kappa = 1e8 matrix of
    512 columns, one QR on the same construction at 256 columns and on a well-conditioned Gaussian of 600."""
    rng = np.random.default_rng(17)
    if kind == "geometric":
        q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
        q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        a = (q1 * 10.0 ** (-8.0 * np.arange(n) / (n - 1))) @ q2.T
    else:
        a = rng.standard_normal((m, n))
    cfg = jsvd.JacobiSvdConfig(epsilon=1e-14, max_sweeps=60)
    assert jsvd.mixed_config(jsvd.PATH_MIXED).qr_precondition == 3
    print(f"auto second qr: {m}x{n} {kind}", flush=True)
    got = {}
    for qr in (3, expected):
        r, info = jsvd.solve(a, cfg, path=jsvd.PATH_MIXED, mixed=jsvd.mixed_config(jsvd.PATH_MIXED, qr_precondition=qr))
        got[qr] = (r, info.as_dict())
    (ra, ia), (re, ie) = got[3], got[expected]
    assert ia == ie
    for x, y in ((ra.u, re.u), (ra.sigma, re.sigma), (ra.v, re.v)):
        assert np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))
