"""Parity of the plain FP64 HIP path (libjsvd.so on cuda:0) with the CPU oracle.

Kernel by kernel on seeded inputs, then end to end on the plain-FP64 workload (512 x 256 N(0,1),
seed 0, eps 1e-15, max 30 sweeps) and edge shapes.  Tolerances:
  * transpose, schedule, rotation parameters: bit-exact (rotation math is evaluated without FMA
    contraction on both sides);
  * Gram / apply / sigma: |diff| <= 64 * eps_mach * (product of the operand norms) -- only the
    summation order differs from the oracle's row-order sums;
  * full solve vs oracle, same column order and signs: sigma relative 1e-12, U and V entries 1e-9
    (after ~10 sweeps of rounding-order differences), sweep count equal (+-1 for borderline ties);
  * full solve vs LAPACK dgesvj: sorted sigma relative 1e-12.
"""
import numpy as np
import pytest

import jsvd
import oracle
from helpers import gaussian, rho_a, rho_orth, sorted_desc

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps


def _dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _col_padded(a, lda):
    m, n = a.shape
    out = np.zeros((n, lda))
    out[:, :m] = a.T
    return out  # row-major n x lda == column-major m x n with leading dimension lda


@pytest.mark.parametrize("rows,cols", [(1, 1), (5, 3), (64, 64), (65, 130), (200, 7), (513, 257)])
def test_transpose_bitwise(torch_cuda, rows, cols):
    torch = torch_cuda
    a = gaussian(rows, cols, rows * 1000 + cols)
    src = _dev(torch, a)
    dst = torch.full((cols, rows + 3), np.nan, dtype=torch.float64, device="cuda")
    jsvd.transpose(src, dst, rows, cols, cols, rows + 3)
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    assert np.array_equal(out[:, :rows], a.T)
    assert np.all(np.isnan(out[:, rows:]))   # nothing written outside the matrix


def test_rotation_params_bitwise(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(2)
    k = 4096
    x = rng.standard_normal((k, 5))
    y = rng.standard_normal((k, 5)) * 10.0 ** rng.uniform(-6, 6, size=(k, 1))
    al = np.sum(x * x, 1)
    be = np.sum(y * y, 1)
    ga = np.sum(x * y, 1)
    # special cases: equal norms, zero gamma, denormal gamma, threshold ties
    al[:4] = [2.0, 1.0, 1e-320, 4.0]
    be[:4] = [2.0, 3.0, 1e-320, 9.0]
    ga[:4] = [0.5, 0.0, 1e-301, 3.0]
    eps = 1e-15
    epsv = np.full(k, eps)
    epsv[3] = 0.5
    c = torch.zeros(k, dtype=torch.float64, device="cuda")
    s = torch.zeros_like(c)
    rot = torch.zeros(k, dtype=torch.int32, device="cuda")
    d_al, d_be, d_ga = _dev(torch, al), _dev(torch, be), _dev(torch, ga)
    jsvd.rotation_params(d_al, d_be, d_ga, eps, c, s, rot)
    c3 = torch.zeros(1, dtype=torch.float64, device="cuda")
    s3 = torch.zeros_like(c3)
    r3 = torch.zeros(1, dtype=torch.int32, device="cuda")
    jsvd.rotation_params(d_al[3:4], d_be[3:4], d_ga[3:4], 0.5, c3, s3, r3)
    torch.cuda.synchronize()
    c, s, rot = c.cpu().numpy(), s.cpu().numpy(), rot.cpu().numpy()
    for i in range(k):
        if i == 3:
            continue
        r, ci, si = oracle.rotation(al[i], be[i], ga[i], eps)
        assert rot[i] == r and c[i] == ci and s[i] == si, i
    assert int(r3.item()) == 0 and c3.item() == 1.0 and s3.item() == 0.0


@pytest.mark.parametrize("m", [1, 7, 512, 4097])
def test_pair_gram_matches_oracle(torch_cuda, m):
    torch = torch_cuda
    n = 6
    a = gaussian(m, n, m)
    lda = m + (m % 2)
    d_a = _dev(torch, _col_padded(a, lda))
    pairs = np.array([[0, 5], [3, 1], [2, 4]], dtype=np.int32)
    d_p = _dev(torch, pairs.reshape(-1))
    al = torch.zeros(3, dtype=torch.float64, device="cuda")
    be, ga = torch.zeros_like(al), torch.zeros_like(al)
    jsvd.pair_gram(d_a, lda, m, d_p, al, be, ga)
    torch.cuda.synchronize()
    for k, (p, q) in enumerate(pairs):
        x, y, z = oracle.pair_stats(a, p, q)
        tol = 64 * EPS64 * max(m, 1) ** 0.5
        assert abs(al[k].item() - x) <= tol * x
        assert abs(be[k].item() - y) <= tol * y
        assert abs(ga[k].item() - z) <= tol * np.sqrt(x * y)


@pytest.mark.parametrize("m,n", [(512, 256), (33, 9), (3000, 64), (9000, 40), (20000, 6)])
def test_one_round_matches_oracle(torch_cuda, m, n):
    """One fused round (register-resident and streaming variants) vs oracle stats + rotation + apply."""
    torch = torch_cuda
    a = gaussian(m, n, m + n)
    lda, ldv = m + (m % 2), n + (n % 2)
    d_a = _dev(torch, _col_padded(a, lda))
    v0 = np.eye(n)
    d_v = _dev(torch, _col_padded(v0, ldv))
    rot = torch.zeros(1, dtype=torch.int32, device="cuda")
    off = torch.zeros(1, dtype=torch.int64, device="cuda")
    rnd = 1 % max(1, n - 1 + (n % 2))
    jsvd.round_(d_a, lda, m, d_v, ldv, n, rnd, 1e-15, rot, off)
    torch.cuda.synchronize()
    ga = d_a.cpu().numpy()[:, :m].T
    gv = d_v.cpu().numpy()[:, :n].T
    ref_a = a.copy()
    ref_v = v0.copy()
    pairs = oracle.round_robin(n)[rnd]
    nrot = 0
    offmax = 0.0
    for p, q in pairs:
        x, y, z = oracle.pair_stats(a, p, q)
        offmax = max(offmax, abs(z) / np.sqrt(x * y))
        r, c, s = oracle.rotation(x, y, z, 1e-15)
        nrot += r
        ap, aq = ref_a[:, p].copy(), ref_a[:, q].copy()
        ref_a[:, p], ref_a[:, q] = c * ap - s * aq, s * ap + c * aq
        vp, vq = ref_v[:, p].copy(), ref_v[:, q].copy()
        ref_v[:, p], ref_v[:, q] = c * vp - s * vq, s * vp + c * vq
    assert int(rot.item()) == nrot
    got_off = np.array([off.item()], dtype=np.int64).view(np.float64)[0]
    assert abs(got_off - offmax) <= 1e-12 * offmax
    scale = np.max(np.linalg.norm(a, axis=0))
    assert np.max(np.abs(ga - ref_a)) <= 1e-12 * scale
    assert np.max(np.abs(gv - ref_v)) <= 1e-12
    # padding row of an odd m stays zero
    if m % 2:
        assert np.all(d_a.cpu().numpy()[:, m:] == 0.0)


@pytest.mark.parametrize("m,n,first,count", [(300, 40, 5, 60), (97, 9, 0, 9)])
def test_rounds_equal_single_rounds(torch_cuda, m, n, first, count):
    """jsvd_rounds_f64 (consecutive rounds, wrapping into the next sweep) equals the same rounds launched one
    by one, bit for bit, rotation counts included."""
    torch = torch_cuda
    a = gaussian(m, n, 3)
    lda, ldv = m + (m % 2), n + (n % 2)
    per_round = n - 1 + (n % 2)
    results = []
    for mode in ("single", "batched", "slots"):
        d_a = _dev(torch, _col_padded(a, lda))
        d_v = _dev(torch, _col_padded(np.eye(n), ldv))
        rot = torch.zeros(8, dtype=torch.int32, device="cuda")
        off = torch.zeros(8, dtype=torch.int64, device="cuda")
        if mode == "batched":
            jsvd.rounds(d_a, lda, m, d_v, ldv, n, first, count, 1e-15, rot_count=rot, offmax=off)
        elif mode == "slots":  # counters spread over 8 slots: same totals
            jsvd.rounds(d_a, lda, m, d_v, ldv, n, first, count, 1e-15, rot_count=rot, offmax=off, count_slots=8)
        else:
            for k in range(count):
                jsvd.round_(d_a, lda, m, d_v, ldv, n, (first + k) % per_round, 1e-15, rot_count=rot, offmax=off)
        torch.cuda.synchronize()
        results.append((d_a.cpu().numpy(), d_v.cpu().numpy(), int(rot.sum().item()), int(off.max().item())))
    for r in results[1:]:
        assert np.array_equal(results[0][0], r[0])
        assert np.array_equal(results[0][1], r[1])
        assert results[0][2] == r[2] > 0
        assert results[0][3] == r[3] > 0
    with pytest.raises(ValueError):
        jsvd.rounds(d_a, lda, m, d_v, ldv, n, per_round, 1, 1e-15)


def test_round_skip_flag(torch_cuda):
    torch = torch_cuda
    a = gaussian(64, 8, 1)
    d_a = _dev(torch, _col_padded(a, 64))
    d_v = _dev(torch, _col_padded(np.eye(8), 8))
    zero = torch.zeros(1, dtype=torch.int32, device="cuda")
    jsvd.round_(d_a, 64, 64, d_v, 8, 8, 0, 1e-15, skip_if_zero=zero)
    torch.cuda.synchronize()
    assert np.array_equal(d_a.cpu().numpy()[:, :64].T, a)


@pytest.mark.parametrize("m,n", [(10, 3), (700, 300)])
def test_u_sigma_matches_oracle(torch_cuda, m, n):
    torch = torch_cuda
    a = gaussian(m, n, 9)
    a[:, 1] = 0.0
    lda = m + (m % 2)
    d_a = _dev(torch, _col_padded(a, lda))
    u = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    s = torch.zeros(n, dtype=torch.float64, device="cuda")
    jsvd.u_sigma(d_a, lda, m, n, 1e-9, u, s)
    torch.cuda.synchronize()
    u_ref, s_ref = oracle.u_sigma(a, 1e-9)
    assert np.max(np.abs(s.cpu().numpy() - s_ref) / np.maximum(s_ref, 1e-300)) < 1e-14
    assert s.cpu().numpy()[1] == 0.0 and np.all(u.cpu().numpy()[:, 1] == 0.0)
    assert np.max(np.abs(u.cpu().numpy() - u_ref)) < 1e-14


def test_offnorm_two_stage(torch_cuda):
    torch = torch_cuda
    m, n = 300, 70
    a = gaussian(m, n, 10)
    lda = m
    d_a = _dev(torch, _col_padded(a, lda))
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    jsvd.offnorm(d_a, lda, m, n, out)
    torch.cuda.synchronize()
    g = a.T @ a
    d = np.sqrt(np.diag(g))
    cos = np.abs(g) / np.outer(d, d)
    np.fill_diagonal(cos, 0.0)
    assert abs(out.item() - cos.max()) < 1e-13


# ------------------------------------------------------------------------------------------ end to end

def _compare_to_oracle(a, eps, max_sweeps, sweep_slack=1):
    got = jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(epsilon=eps, max_sweeps=max_sweeps))
    ref = oracle.svd(a, eps=eps, max_sweeps=max_sweeps, threads=4)
    assert abs(got.sweeps - ref.sweeps) <= sweep_slack, (got.sweeps, ref.sweeps)
    # same Jacobi column order and the same signs as the reference algorithm
    assert np.max(np.abs(got.sigma - ref.sigma) / np.maximum(ref.sigma, 1e-300)) < 1e-12
    assert np.max(np.abs(got.v - ref.v)) < 1e-9
    assert np.max(np.abs(got.u - ref.u)) < 1e-9
    return got, ref


def test_plain_config1_parity(torch_cuda):
    """512 x 256 N(0,1) seed 0, eps 1e-15, max 30 sweeps."""
    a = gaussian(512, 256, 0)
    got, ref = _compare_to_oracle(a, 1e-15, 30)
    _, s_lapack, _, _ = oracle.dgesvj(a)
    assert np.max(np.abs(sorted_desc(got.sigma) - s_lapack) / s_lapack) < 1e-12
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13
    assert rho_orth(got.u) < 1e-12
    assert rho_orth(got.v) < 1e-11


@pytest.mark.parametrize("m,n,seed", [(4, 3, 1), (8, 8, 2), (9, 5, 3), (31, 31, 4), (100, 1, 5), (257, 129, 6),
                                      (1030, 77, 7)])
def test_plain_shapes_parity(torch_cuda, m, n, seed):
    a = gaussian(m, n, seed)
    got, _ = _compare_to_oracle(a, 1e-12, 40)
    assert rho_a(a, got.u, got.sigma, got.v) < 1e-13


def test_plain_golden_fixtures(torch_cuda):
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dgesvj_small.npz"))
    for case in ["g4x3", "g8x8", "g16x7", "g33x17", "g64x64", "g97x40", "graded48x24"]:
        a = fx[f"{case}_a"]
        got = jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=60))
        s_ref = fx[f"{case}_sigma"]
        assert np.max(np.abs(sorted_desc(got.sigma) - s_ref) / s_ref) < 2e-13, case
        assert rho_a(a, got.u, got.sigma, got.v) < 5e-14, case


def test_plain_edge_cases(torch_cuda):
    # diagonal input: one rotation-free sweep, sigma = |d| in column order, V = I
    d = np.array([3.0, -1.0, 7.0, 0.5])
    a = np.zeros((6, 4))
    a[np.arange(4), np.arange(4)] = d
    r = jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30))
    assert r.sweeps == 1 and np.array_equal(r.sigma, np.abs(d)) and np.array_equal(r.v, np.eye(4))
    # 2 x 2 closed form
    r = jsvd.one_sided_jacobi_svd(np.array([[3.0, 0.0], [4.0, 5.0]]), jsvd.JacobiSvdConfig(epsilon=1e-15))
    assert np.allclose(sorted_desc(r.sigma), [45 ** 0.5, 5 ** 0.5], rtol=1e-15, atol=0)
    # zero column: sigma 0 and a zero U column (absolute-eps cutoff)
    a = gaussian(10, 4, 5)
    a[:, 2] = 0.0
    r = jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(epsilon=1e-12))
    k = int(np.argmin(r.sigma))
    assert r.sigma[k] == 0.0 and np.all(r.u[:, k] == 0.0)
    # non-convergence at max_sweeps is not an error
    r = jsvd.one_sided_jacobi_svd(gaussian(64, 32, 6), jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=2))
    assert r.sweeps == 2
    # rank deficient: duplicated columns
    a = gaussian(20, 6, 8)
    a[:, 5] = a[:, 0]
    r = jsvd.one_sided_jacobi_svd(a, jsvd.JacobiSvdConfig(epsilon=1e-14))
    assert rho_a(a, r.u, r.sigma, r.v) < 1e-14 and np.min(r.sigma) < 1e-13


def test_plan_device_solve_deterministic(torch_cuda):
    torch = torch_cuda
    m, n = 384, 200
    a = gaussian(m, n, 21)
    plan = jsvd.Plan(m, n, jsvd.JacobiSvdConfig(epsilon=1e-15, max_sweeps=30))
    d_a = _dev(torch, a)
    outs = []
    for _ in range(2):
        u = torch.zeros((m, n), dtype=torch.float64, device="cuda")
        s = torch.zeros(n, dtype=torch.float64, device="cuda")
        v = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        sw = plan.solve_device(d_a, u, s, v)
        outs.append((u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy(), sw))
    assert outs[0][3] == outs[1][3]
    for x, y in zip(outs[0][:3], outs[1][:3]):
        assert np.array_equal(x, y)
    host = plan.solve_host(a)
    assert np.array_equal(host.sigma, outs[0][1]) and np.array_equal(host.v, outs[0][2])
    rot, off = plan.sweep_stats()
    assert len(rot) == host.sweeps and rot[-1] == 0 and np.all(rot[:-1] > 0)
    assert off[-1] <= 1e-15
    assert np.array_equal(d_a.cpu().numpy(), a)   # the input is not modified
    plan.close()
