// mixed.hip -- driver of the block FP64 and mixed FP32/FP64 one-sided Jacobi solvers.
//
// Mixed path, all on the device, no host round trip per sweep:
//   A (row-major) -> A64 column-major, padded (LDS-tiled transpose)
//   QR: d_j = |a_j|, B = A D^-1; three Cholesky QR passes, the first with a diagonal shift:
//       G = B^T B (+ s I), R_k = chol(G), B <- B R_k^-1; Q = B, R = R_3 R_2 R_1 D   (A = Q R)
//   M = R^T (n x n); with qr_precondition 2 a second QR, M = Q2 R2 by the same column-scaled shifted Cholesky
//       QR, and the sweeps run on N = R2^T instead of M (an LQ step of R: N N^T = R2^T R2 = M^T M is the same
//       Gram matrix, and N's columns are closer to orthogonal; otherwise N below stands for M)
//   N32 = fp32(N), V32 = I; FP32 block sweeps to the FP32 tolerance (MFMA, FP32 accumulation)
//   V64 = fp64(V32); Newton-Schulz V <- V (3I - V^T V) / 2 (FP64 GEMMs) until V is orthogonal in FP64
//   X = N V64 (FP64 GEMM); FP64 block sweeps on (X, V64) to epsilon
//   sigma_j = |x_j|, Y = X / sigma; with the second QR U_A = Q Y and V_A = Q2 V64, with one QR V_A = Y and
//   U_A = Q V64 (row-major, one GEMM each); U columns with sigma_j <= epsilon zeroed as the reference zeroes them
// The products (B^T B and V^T V on their upper triangle only, R_k R_acc on triangular operands) and the triangular
// solves are this library's FP64 MFMA kernels (dense.hip), as are the tiled Cholesky factorisations
// (cholesky.hip).  The shifted Cholesky QR follows Fukaya, Kannan, Nakatsukasa,
// Yamamoto and Yanagisawa, "Shifted Cholesky QR for computing the QR factorization of ill-conditioned
// matrices" (SIAM J. Sci. Comput. 42, 2020): shift s = 11 (m n + n (n + 1)) u |B|^2 with |B|_F^2 <= n for
// unit columns; the column scaling keeps the factorisation's backward error columnwise, so column-graded
// matrices keep their relative accuracy.
#include "mixed.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "autotune.hpp"
#include "cholesky.hpp"
#include "dense.hpp"
#include "plain_f64.hpp"

namespace jsvd
{
    namespace
    {
        double default_fp32_tolerance(int64_t rows)
        {
            return std::max(1e-6, 4.0 * std::sqrt(static_cast<double>(rows)) * std::ldexp(1.0, -24));
        }

        // Geometry of an m x n sweep problem: block width `requested` (16, 32 or 64), or for 0 the autotuner's
        // choice when it measured this shape, else the default; the apply grid size measured with that width.
        BlockGeometry solver_geometry(int64_t m, int64_t n, int requested)
        {
            Tuning t;
            const bool tuned = lookup_tuning(m, n, t) && t.has_block;
            const int kb = effective_block_width((requested == 0 && tuned) ? t.block_width : requested, n);
            BlockGeometry g = BlockGeometry::make(m, n, kb);
            if (tuned && t.block_width == kb)
            {
                g.apply_wgs = t.apply_wgs;
            }
            return g;
        }

        // Column j (one workgroup each, j < n_pad): dcol[j] = |a_j| (0 for padding columns), a_j <- a_j / |a_j|.
        __global__ __launch_bounds__(256) void column_scale_kernel(double *__restrict__ a, int64_t lda, int64_t m,
                                                                   int64_t n, double *__restrict__ dcol)
        {
            __shared__ double red[4];
            const int64_t j = blockIdx.x;
            double *col = a + j * lda;
            double acc = 0.0;
            if (j < n)
            {
                for (int64_t i = threadIdx.x; i < m; i += 256)
                {
                    acc += col[i] * col[i];
                }
            }
            acc = dev::wave_sum(acc);
            if ((threadIdx.x % 64) == 0)
            {
                red[threadIdx.x / 64] = acc;
            }
            __syncthreads();
            const double d = sqrt((red[0] + red[1]) + (red[2] + red[3]));
            if (threadIdx.x == 0)
            {
                dcol[j] = d;
            }
            if (j < n && d > 0.0)
            {
                const double inv = 1.0 / d;
                for (int64_t i = threadIdx.x; i < m; i += 256)
                {
                    col[i] *= inv;
                }
            }
        }

        __global__ void add_diagonal_kernel(double *__restrict__ g, int64_t ld, int64_t n, double s)
        {
            const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i < n)
            {
                g[i * ld + i] += s;
            }
        }

        // dst (n x n, ldd) <- upper triangle of src (lds), zeros below the diagonal.
        __global__ void copy_upper_kernel(const double *__restrict__ src, int64_t lds, double *__restrict__ dst,
                                          int64_t ldd, int64_t n)
        {
            const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (e >= n * n)
            {
                return;
            }
            const int64_t j = e / n;
            const int64_t i = e % n;
            dst[j * ldd + i] = (i <= j) ? src[j * lds + i] : 0.0;
        }

        // M (ldm x n_pad, column-major) <- (R D)^T for the upper triangular R (n x n, ldr) and column scales d:
        // M[i][j] = R[j][i] d_i for j <= i < n, zero elsewhere (padding rows and columns included).
        __global__ void transpose_scaled_kernel(const double *__restrict__ r, int64_t ldr,
                                                const double *__restrict__ d, int64_t n, double *__restrict__ mo,
                                                int64_t ldm, int64_t n_pad)
        {
            const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (e >= ldm * n_pad)
            {
                return;
            }
            const int64_t j = e / ldm;
            const int64_t i = e % ldm;
            mo[e] = (i < n && j < n && j <= i) ? r[i * ldr + j] * d[i] : 0.0;
        }

        // *spread = max_i |r_ii| / min_i |r_ii| over the n diagonal entries of r (leading dimension ldr): a cheap
        // estimate of the condition number of the column-scaled matrix r factors (+inf where an entry is zero).
        // One workgroup of 256 threads.
        __global__ __launch_bounds__(256) void diag_spread_kernel(const double *__restrict__ r, int64_t ldr, int64_t n,
                                                                  double *__restrict__ spread)
        {
            __shared__ double hi_w[4], lo_w[4];
            double hi = 0.0, lo = HUGE_VAL;
            for (int64_t i = threadIdx.x; i < n; i += 256)
            {
                const double d = fabs(r[i * ldr + i]);
                hi = fmax(hi, d);
                lo = fmin(lo, d);
            }
            hi = dev::wave_max(hi);
            lo = -dev::wave_max(-lo);
            if ((threadIdx.x % 64) == 0)
            {
                hi_w[threadIdx.x / 64] = hi;
                lo_w[threadIdx.x / 64] = lo;
            }
            __syncthreads();
            if (threadIdx.x == 0)
            {
                hi = fmax(fmax(hi_w[0], hi_w[1]), fmax(hi_w[2], hi_w[3]));
                lo = fmin(fmin(lo_w[0], lo_w[1]), fmin(lo_w[2], lo_w[3]));
                *spread = (lo > 0.0) ? hi / lo : HUGE_VAL;
            }
        }

        // Row-major u (m x n): zero column j where sigma_j <= eps (the reference's U cutoff).
        __global__ void mask_columns_kernel(double *__restrict__ u, int64_t m, int64_t n,
                                            const double *__restrict__ sigma, double eps)
        {
            const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (e < m * n && !(sigma[e % n] > eps))
            {
                u[e] = 0.0;
            }
        }

        unsigned grid_of(int64_t count) { return static_cast<unsigned>((count + 255) / 256); }
    } // namespace

    void BlockSolver::set_geometry(int64_t m, int64_t n, const BlockGeometry &sweep)
    {
        g_ = sweep;
        ga_ = BlockGeometry::make(m, n, sweep.kb);
        ga_.apply_wgs = sweep.apply_wgs;
        g_.cross = ga_.cross = g_.kb == 32 && mixed_.inner_order == JSVD_INNER_CROSS;
    }

    BlockSolver::BlockSolver(int64_t m, int64_t n, const jsvd_config &cfg, const jsvd_mixed_config &mixed,
                             hipStream_t stream)
        : cfg_(cfg), mixed_(mixed), stream_(stream)
    {
        set_geometry(m, n, qr_first() ? solver_geometry(n, n, mixed_.block_width)
                                      : solver_geometry(m, n, mixed_.block_width));
        if (ga_.nb < 2)
        {
            throw_invalid("block and mixed paths need at least two 32-column blocks (n >= 33).");
        }
        fp32_tol_requested_ = mixed_.fp32_tolerance;
        cap_ = std::max(cfg_.max_sweeps, mixed_.fp32_max_sweeps);
        // buffers sized for the A-side geometry, which also covers the n x n sweep geometry and direct solves
        const size_t a_elems = static_cast<size_t>(ga_.lda) * ga_.n_pad;
        const size_t v_elems = static_cast<size_t>(ga_.ldv) * ga_.n_pad;
        const size_t part_elems = ga_.partial_elems();
        cap_a_elems_ = a_elems;
        cap_v_elems_ = v_elems;
        cap_part_elems_ = part_elems;
        cap_pairs_ = ga_.pairs();
        cap_npad_ = ga_.n_pad;
        cap_q_elems_ = ga_.q_elems();
        JSVD_HIP_CHECK(hipMalloc(&a64_, sizeof(double) * a_elems));
        JSVD_HIP_CHECK(hipMalloc(&v64_, sizeof(double) * v_elems));
        JSVD_HIP_CHECK(hipMalloc(&partial_, sizeof(double) * part_elems));
        JSVD_HIP_CHECK(hipMalloc(&q_, sizeof(double) * ga_.q_elems()));
        JSVD_HIP_CHECK(hipMalloc(&flags_, sizeof(int) * ga_.flag_elems()));
        JSVD_HIP_CHECK(hipMalloc(&perm_, sizeof(int) * ga_.n_pad));
        JSVD_HIP_CHECK(hipMalloc(&norm2_, sizeof(double) * ga_.n_pad));
        JSVD_HIP_CHECK(hipMalloc(&rot_, sizeof(int) * 2 * cap_));
        JSVD_HIP_CHECK(hipMalloc(&cont_, sizeof(int) * 2 * cap_));
        JSVD_HIP_CHECK(hipMalloc(&off_, sizeof(unsigned long long) * 2 * cap_));
        JSVD_HIP_CHECK(hipHostMalloc(&h_rot_, sizeof(int) * 2 * cap_, hipHostMallocDefault));
        for (hipEvent_t &e : done_)
        {
            JSVD_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        // Padding rows and columns stay zero for the plan's lifetime (zero rows rotate into zero rows, zero
        // columns never rotate).
        JSVD_HIP_CHECK(hipMemsetAsync(a64_, 0, sizeof(double) * a_elems, stream_));
        if (mixed_.path == JSVD_PATH_MIXED)
        {
            JSVD_HIP_CHECK(hipMalloc(&a64b_, sizeof(double) * a_elems));
            JSVD_HIP_CHECK(hipMalloc(&w64_, sizeof(double) * v_elems));
            JSVD_HIP_CHECK(hipMalloc(&spread_, sizeof(double)));
            if (qr_first() && (mixed_.qr_precondition == 2 ||
                               (mixed_.qr_precondition == 3 && ga_.n >= kSecondQrMinColumns)))
            {
                // the second QR's Q2 and N = R2^T, in the sweep geometry (ldv x n_pad covers it: ldv >= lda of an
                // n x n sweep, both multiples of the slab height)
                JSVD_HIP_CHECK(hipMalloc(&q2_, sizeof(double) * v_elems));
                JSVD_HIP_CHECK(hipMalloc(&n2_, sizeof(double) * v_elems));
                JSVD_HIP_CHECK(hipMemsetAsync(q2_, 0, sizeof(double) * v_elems, stream_));
                JSVD_HIP_CHECK(hipMemsetAsync(n2_, 0, sizeof(double) * v_elems, stream_));
            }
            // n_pad x n_pad products and the Cholesky factors, padded to the factorisation's 64-row tiles
            const int64_t n64 = cholesky_padded_size(ga_.n_pad);
            JSVD_HIP_CHECK(hipMalloc(&gram64_, sizeof(double) * n64 * n64));
            JSVD_HIP_CHECK(hipMalloc(&racc_, sizeof(double) * n64 * n64));
            JSVD_HIP_CHECK(hipMalloc(&dcol_, sizeof(double) * ga_.n_pad));
            JSVD_HIP_CHECK(hipMalloc(&chol_info_, sizeof(int) * 3));
            JSVD_HIP_CHECK(hipHostMalloc(&h_sigma_, sizeof(double) * ga_.n_pad, hipHostMallocDefault));
            JSVD_HIP_CHECK(hipMalloc(&a32_, sizeof(float) * a_elems));
            JSVD_HIP_CHECK(hipMalloc(&v32_, sizeof(float) * v_elems));
            JSVD_HIP_CHECK(hipMemsetAsync(a64b_, 0, sizeof(double) * a_elems, stream_));
            JSVD_HIP_CHECK(hipMemsetAsync(w64_, 0, sizeof(double) * v_elems, stream_));
            // scratch of the dense kernels: inverted diagonal blocks of the triangular solve and split-k partial tiles
            JSVD_HIP_CHECK(hipMalloc(&dinv_, sizeof(double) * trsm_dinv_elems(ga_.n_pad)));
            dws_.work_elems = dense_work_elems();
            JSVD_HIP_CHECK(hipMalloc(&dws_.work, sizeof(double) * dws_.work_elems));
        }
    }

    BlockSolver::~BlockSolver()
    {
        for (void *p : {static_cast<void *>(dinv_), static_cast<void *>(dws_.work), static_cast<void *>(q2_), static_cast<void *>(n2_), static_cast<void *>(spread_), static_cast<void *>(a64_), static_cast<void *>(a64b_), static_cast<void *>(v64_),
                        static_cast<void *>(w64_), static_cast<void *>(gram64_), static_cast<void *>(a32_),
                        static_cast<void *>(v32_), static_cast<void *>(partial_), static_cast<void *>(q_),
                        static_cast<void *>(flags_), static_cast<void *>(rot_), static_cast<void *>(cont_), static_cast<void *>(off_),
                        static_cast<void *>(perm_), static_cast<void *>(norm2_), static_cast<void *>(racc_),
                        static_cast<void *>(dcol_), static_cast<void *>(chol_info_)})
        {
            (void)hipFree(p);
        }
        (void)hipHostFree(h_rot_);
        (void)hipHostFree(h_sigma_);
        for (hipEvent_t e : done_)
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
        }
    }

    // Block sweeps with the same device-side early exit and lagged host polling as the plain path.  The phase
    // ends after the first sweep with at most stop_rotations rotations (0: the first rotation-free sweep, the
    // reference's rule); returns the executed sweep count, that sweep included.
    template <typename T>
    int BlockSolver::run_sweeps(const BlockGeometry &g, T *a, T *v, double eps, int max_sweeps, int *rot,
                                unsigned long long *off, double &last_off, int stop_rotations)
    {
        int *cont = cont_ + (rot - rot_);
        JSVD_HIP_CHECK(hipMemsetAsync(rot, 0, sizeof(int) * max_sweeps, stream_));
        JSVD_HIP_CHECK(hipMemsetAsync(off, 0, sizeof(unsigned long long) * max_sweeps, stream_));
        int *h = h_rot_ + (rot - rot_);
        std::fill(h, h + max_sweeps, -1);
        T *partial = reinterpret_cast<T *>(partial_);
        T *q = reinterpret_cast<T *>(q_);
        int enqueued = 0;
        // Sweep 0 runs in the incoming column order; every later sweep first orders the columns by descending
        // norm (the norms left by the previous sweep), which groups columns of similar size into the same
        // blocks and cuts the sweep count.  A and V columns are addressed through perm_, nothing moves.
        launch_iota(perm_, g.n_pad, stream_);
        for (int s = 0; s < max_sweeps; ++s)
        {
            const int *skip = (s > 0) ? cont + (s - 1) : nullptr;
            if (s > 0)
            {
                launch_column_order(norm2_, g.n_pad, perm_, skip, stream_);
            }
            // cross ordering: the pairs inside every block first, then the outer rounds' cross pairs
            launch_block_intra<T>(g, a, v, eps, partial, q, flags_, rot + s, off + s, perm_, norm2_, skip, stream_);
            for (int r = 0; r < g.rounds(); ++r)
            {
                launch_block_round<T>(g, a, v, r, eps, partial, q, flags_, rot + s, off + s, perm_, norm2_, skip,
                                      stream_);
            }
            launch_sweep_continue(rot + s, stop_rotations, cont + s, stream_);
            JSVD_HIP_CHECK(hipMemcpyAsync(h + s, rot + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
            JSVD_HIP_CHECK(hipEventRecord(done_[s & 1], stream_));
            enqueued = s + 1;
            if (s >= 1)
            {
                JSVD_HIP_CHECK(hipEventSynchronize(done_[(s - 1) & 1]));
                if (h[s - 1] <= stop_rotations)
                {
                    break;
                }
            }
        }
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
        int executed = enqueued;
        for (int s = 0; s < enqueued; ++s)
        {
            if (h[s] <= stop_rotations)
            {
                executed = s + 1;
                break;
            }
        }
        std::vector<unsigned long long> bits(executed);
        JSVD_HIP_CHECK(hipMemcpy(bits.data(), off, sizeof(unsigned long long) * executed, hipMemcpyDeviceToHost));
        for (int s = 0; s < executed; ++s)
        {
            double x = 0.0;
            std::memcpy(&x, &bits[s], sizeof(double));
            stat_rot_.push_back(h[s]);
            stat_off_.push_back(x);
        }
        last_off = stat_off_.back();
        return executed;
    }

    // V <- V (3I - V^T V) / 2, written as W = 1.5 V - 0.5 V (V^T V); converges quadratically from an FP32
    // orthogonality error of ~1e-6 to FP64 rounding in two steps.
    void BlockSolver::orthogonalize_v(double *v, double *w)
    {
        const int64_t np = g_.n_pad;
        const int64_t ldv = g_.ldv;
        double *cur = v;
        double *nxt = w;
        for (int step = 0; step < mixed_.orthogonalize_steps; ++step)
        {
            // G = V^T V (upper tiles, mirrored), then nxt = 1.5 cur - 0.5 cur G over all ldv rows (padding rows of
            // cur are zero and stay zero)
            launch_gemm_f64(true, false, np, np, np, 1.0, cur, ldv, cur, ldv, 0.0, nullptr, 0, gram64_, np,
                            dense::kUpperOut | dense::kMirror, dws_, stream_);
            launch_gemm_f64(false, false, ldv, np, np, -0.5, cur, ldv, gram64_, np, 1.5, cur, ldv, nxt, ldv, 0, dws_,
                            stream_);
            std::swap(cur, nxt);
        }
        if (cur != v)
        {
            JSVD_HIP_CHECK(hipMemcpyAsync(v, cur, sizeof(double) * g_.ldv * g_.n_pad, hipMemcpyDeviceToDevice, stream_));
        }
    }

    bool BlockSolver::reshape(int64_t m, int64_t n)
    {
        auto fits = [&](const BlockGeometry &x) {
            return x.nb >= 2 && static_cast<size_t>(x.lda) * x.n_pad <= cap_a_elems_ &&
                   static_cast<size_t>(x.ldv) * x.n_pad <= cap_v_elems_ && x.partial_elems() <= cap_part_elems_ &&
                   x.pairs() <= cap_pairs_ && x.n_pad <= cap_npad_ && x.q_elems() <= cap_q_elems_;
        };
        BlockGeometry sweep = qr_first() ? solver_geometry(n, n, mixed_.block_width)
                                         : solver_geometry(m, n, mixed_.block_width);
        if (!fits(BlockGeometry::make(m, n, sweep.kb)) || !fits(sweep))
        {
            // the preferred width needs more rotation or partial space than allocated: keep the current width
            const int kb = effective_block_width(g_.kb, n);
            sweep = BlockGeometry::make(qr_first() ? n : m, n, kb);
            if (!fits(BlockGeometry::make(m, n, kb)) || !fits(sweep))
            {
                return false;
            }
        }
        const BlockGeometry a_side = BlockGeometry::make(m, n, sweep.kb);
        if (a_side.m != ga_.m || a_side.n != ga_.n || a_side.kb != ga_.kb)
        {
            // padding rows and columns of the new extent must read as zero
            const size_t a_elems = static_cast<size_t>(a_side.lda) * a_side.n_pad;
            JSVD_HIP_CHECK(hipMemsetAsync(a64_, 0, sizeof(double) * a_elems, stream_));
            if (a64b_)
            {
                JSVD_HIP_CHECK(hipMemsetAsync(a64b_, 0, sizeof(double) * a_elems, stream_));
            }
        }
        set_geometry(m, n, sweep);
        return true;
    }

    void BlockSolver::fp32_phase(const BlockGeometry &g, const double *x64, double *y64, double tol,
                                 jsvd_solve_info &info)
    {
        const int64_t x_elems = g.lda * g.n_pad;
        const int64_t v_elems = g.ldv * g.n_pad;
        launch_convert_f64_to_f32(x64, a32_, x_elems, stream_);
        launch_identity_f32(v32_, g.ldv, g.n_pad, stream_);
        const int fp32_cap = std::max(1, mixed_.fp32_max_sweeps);
        // the FP32 phase preconditions the FP64 one: it may end on a sweep that still rotates a few pairs
        const double pairs = 0.5 * static_cast<double>(g.n) * static_cast<double>(g.n - 1);
        const int stop = static_cast<int>(mixed_.fp32_stop_fraction * pairs);
        info.sweeps_fp32 = run_sweeps<float>(g, a32_, v32_, tol, fp32_cap, rot_, off_, info.offmax_fp32, stop);
        launch_convert_f32_to_f64(v32_, v64_, v_elems, stream_);
        orthogonalize_v(v64_, w64_);
        launch_gemm_f64(false, false, g.lda, g.n_pad, g.n_pad, 1.0, x64, g.lda, v64_, g.ldv, 0.0, nullptr, 0, y64,
                        g.lda, 0, dws_, stream_);
    }

    bool BlockSolver::factor_qr(double *x, int64_t m, int64_t n, int64_t lda, double *m_out, double *spread)
    {
        const int64_t ldr = cholesky_padded_size(g_.n_pad);
        hipLaunchKernelGGL(column_scale_kernel, dim3(static_cast<unsigned>(g_.n_pad)), dim3(256), 0, stream_, x,
                           lda, m, n, dcol_);
        JSVD_HIP_CHECK(hipGetLastError());
        // |B|_2^2 <= |B|_F^2 <= n for unit columns
        const double u = std::ldexp(1.0, -53);
        const double shift = 11.0 * (static_cast<double>(m) * n + static_cast<double>(n) * (n + 1)) * u * n;
        for (int pass = 0; pass < 3; ++pass)
        {
            // G = B^T B on the upper triangle (the only part the factorisation reads)
            launch_gemm_f64(true, false, n, n, m, 1.0, x, lda, x, lda, 0.0, nullptr, 0, gram64_, ldr,
                            dense::kUpperOut, dws_, stream_);
            if (pass == 0)
            {
                hipLaunchKernelGGL(add_diagonal_kernel, dim3(grid_of(n)), dim3(256), 0, stream_, gram64_,
                                   static_cast<int64_t>(ldr), static_cast<int64_t>(n), shift);
                JSVD_HIP_CHECK(hipGetLastError());
            }
            launch_cholesky_upper(gram64_, ldr, n, chol_info_ + pass, stream_);
            // B <- B R^-1 (the factor's identity padding up to ldr is the padding the solve needs)
            launch_trsm_right_upper_f64(x, lda, m, n, gram64_, ldr, true, nullptr, dinv_, dws_, stream_);
            if (pass == 0)
            {
                hipLaunchKernelGGL(copy_upper_kernel, dim3(grid_of(static_cast<int64_t>(n) * n)), dim3(256), 0,
                                   stream_, gram64_, static_cast<int64_t>(ldr), racc_, static_cast<int64_t>(ldr),
                                   static_cast<int64_t>(n));
                JSVD_HIP_CHECK(hipGetLastError());
            }
            else
            {
                // racc <- R_pass racc (upper times upper stays upper), through w64_ (free until the sweeps)
                launch_gemm_f64(false, false, n, n, n, 1.0, gram64_, ldr, racc_, ldr, 0.0, nullptr, 0, w64_, ldr,
                                dense::kTriA | dense::kTriB | dense::kZeroLower, dws_, stream_);
                JSVD_HIP_CHECK(hipMemcpyAsync(racc_, w64_, sizeof(double) * ldr * n, hipMemcpyDeviceToDevice,
                                              stream_));
            }
        }
        if (spread != nullptr)
        {
            hipLaunchKernelGGL(diag_spread_kernel, dim3(1), dim3(256), 0, stream_, static_cast<const double *>(racc_),
                               static_cast<int64_t>(ldr), n, spread_);
            JSVD_HIP_CHECK(hipGetLastError());
            JSVD_HIP_CHECK(hipMemcpyAsync(spread, spread_, sizeof(double), hipMemcpyDeviceToHost, stream_));
        }
        int status[3] = {0, 0, 0};
        JSVD_HIP_CHECK(hipMemcpyAsync(status, chol_info_, sizeof(status), hipMemcpyDeviceToHost, stream_));
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
        if (status[0] != 0 || status[1] != 0 || status[2] != 0)
        {
            return false; // a factor was not positive definite
        }
        // (R D)^T into m_out with the sweep geometry's padding
        hipLaunchKernelGGL(transpose_scaled_kernel, dim3(grid_of(g_.lda * g_.n_pad)), dim3(256), 0, stream_, racc_,
                           static_cast<int64_t>(ldr), dcol_, static_cast<int64_t>(n), m_out, g_.lda, g_.n_pad);
        JSVD_HIP_CHECK(hipGetLastError());
        return true;
    }

    bool BlockSolver::solve_qr_first(const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                     jsvd_solve_info &info)
    {
        launch_transpose(d_a, ga_.n, a64_, ga_.lda, ga_.m, ga_.n, stream_);
        // A = Q R: Q in a64_, M = R^T in a64b_; the spread of R's diagonal estimates the conditioning
        double spread = 0.0;
        if (!factor_qr(a64_, ga_.m, ga_.n, ga_.lda, a64b_, &spread))
        {
            return false; // the caller solves A directly
        }
        // M = Q2 R2 on a copy of M: Q2 in q2_, N = R2^T in n2_; the sweeps run on N, or on M when a factor of the
        // second QR breaks down.  Automatic mode: the second QR where its three Cholesky factorisations cost less
        // than the sweeps it saves -- n >= 512 (a sweep has n / kb - 1 outer rounds) and an ill-conditioned R
        // (diagonal spread >= 100; measured at 4096 x 1024: 0.3-0.9 ms faster from kappa 1e4 to 1e8, 0.9-1.2 ms
        // slower on well-conditioned Gaussian inputs; at n = 256 slower at every kappa).
        const double *sweep_in = a64b_;
        bool second = false;
        const bool want_second = mixed_.qr_precondition == 2 ||
                                 (mixed_.qr_precondition == 3 && ga_.n >= kSecondQrMinColumns &&
                                  spread >= kSecondQrMinSpread);
        if (want_second && q2_ != nullptr)
        {
            JSVD_HIP_CHECK(hipMemcpyAsync(q2_, a64b_, sizeof(double) * g_.lda * g_.n_pad, hipMemcpyDeviceToDevice,
                                          stream_));
            second = factor_qr(q2_, g_.n, g_.n, g_.lda, n2_, nullptr);
            sweep_in = second ? n2_ : a64b_;
        }
        // FP32 phase, then X = N V (or M V), FP64 phase
        const double tol = (fp32_tol_requested_ > 0.0) ? fp32_tol_requested_ : default_fp32_tolerance(g_.m);
        fp32_phase(g_, sweep_in, w64_, tol, info);
        info.sweeps_fp64 = run_sweeps<double>(g_, w64_, v64_, cfg_.epsilon, cfg_.max_sweeps, rot_ + cap_, off_ + cap_,
                                              info.offmax_fp64);
        info.converged = (h_rot_[cap_ + info.sweeps_fp64 - 1] == 0) ? 1 : 0;
        // sigma_j = |x_j| and Y = X / sigma (row-major into d_v), no cutoff: a column is normalised whenever
        // sigma_j > 0
        launch_u_sigma(w64_, g_.lda, g_.n, g_.n, 0.0, d_v, d_sigma, stream_);
        if (second)
        {
            // A = Q M^T = Q N Q2^T = Q Y Sigma (Q2 V)^T: U_A = Q Y (Y row-major in d_v is Y^T column-major with
            // leading dimension n), then V_A = Q2 V over it, both produced row-major
            launch_gemm_f64(false, true, ga_.m, ga_.n, ga_.n, 1.0, a64_, ga_.lda, d_v, ga_.n, 0.0, nullptr, 0, d_u,
                            ga_.n, dense::kRowMajorCD, dws_, stream_);
            launch_gemm_f64(false, false, ga_.n, ga_.n, ga_.n, 1.0, q2_, g_.lda, v64_, g_.ldv, 0.0, nullptr, 0, d_v,
                            ga_.n, dense::kRowMajorCD, dws_, stream_);
        }
        else
        {
            // V_A = Y; U_A = Q V_M, produced row-major directly: U^T = V_M^T Q^T is column-major n x m with leading
            // dimension n
            launch_gemm_f64(false, false, ga_.m, ga_.n, ga_.n, 1.0, a64_, ga_.lda, v64_, g_.ldv, 0.0, nullptr, 0, d_u,
                            ga_.n, dense::kRowMajorCD, dws_, stream_);
        }
        hipLaunchKernelGGL(mask_columns_kernel, dim3(grid_of(ga_.m * ga_.n)), dim3(256), 0, stream_, d_u, ga_.m,
                           ga_.n, static_cast<const double *>(d_sigma), cfg_.epsilon);
        JSVD_HIP_CHECK(hipGetLastError());
        JSVD_HIP_CHECK(hipMemcpyAsync(h_sigma_, d_sigma, sizeof(double) * ga_.n, hipMemcpyDeviceToHost, stream_));
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
        for (int64_t j = 0; j < ga_.n; ++j)
        {
            if (!(h_sigma_[j] > 0.0))
            {
                return false; // a zero column of X leaves V_A without a unit column there
            }
        }
        return true;
    }

    void BlockSolver::solve_direct(const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                   jsvd_solve_info &info)
    {
        // row-major m x n -> column-major lda x n (padding rows/columns keep their zeros)
        launch_transpose(d_a, ga_.n, a64_, ga_.lda, ga_.m, ga_.n, stream_);
        double *a_final = a64_;
        if (mixed_.path == JSVD_PATH_MIXED)
        {
            const double tol = (fp32_tol_requested_ > 0.0) ? fp32_tol_requested_ : default_fp32_tolerance(ga_.m);
            fp32_phase(ga_, a64_, a64b_, tol, info);
            a_final = a64b_;
        }
        else
        {
            launch_pad_copy_f64(nullptr, 0, 0, 0, v64_, ga_.ldv, ga_.n_pad, true, stream_);
        }
        info.sweeps_fp64 = run_sweeps<double>(ga_, a_final, v64_, cfg_.epsilon, cfg_.max_sweeps, rot_ + cap_,
                                              off_ + cap_, info.offmax_fp64);
        info.converged = (h_rot_[cap_ + info.sweeps_fp64 - 1] == 0) ? 1 : 0;
        launch_u_sigma(a_final, ga_.lda, ga_.m, ga_.n, cfg_.epsilon, d_u, d_sigma, stream_);
        launch_transpose(v64_, ga_.ldv, d_v, ga_.n, ga_.n, ga_.n, stream_);
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void BlockSolver::solve_device(const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                   jsvd_solve_info &info)
    {
        std::memset(&info, 0, sizeof(info));
        stat_rot_.clear();
        stat_off_.clear();
        info.path = mixed_.path;
        if (qr_first())
        {
            if (solve_qr_first(d_a, d_u, d_sigma, d_v, info))
            {
                return;
            }
            // the direct form on A (zero sigma or a Cholesky breakdown); the n x n problem's buffers are reused
            std::memset(&info, 0, sizeof(info));
            stat_rot_.clear();
            stat_off_.clear();
            info.path = mixed_.path;
        }
        solve_direct(d_a, d_u, d_sigma, d_v, info);
    }
} // namespace jsvd
