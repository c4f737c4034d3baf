// mixed.hip -- driver of the block FP64 and mixed FP32/FP64 one-sided Jacobi solvers.
//
// Mixed path, all on the device, no host round trip per sweep:
//   A (row-major) -> A64 column-major, padded (LDS-tiled transpose)
//   A32 = fp32(A64), V32 = I; FP32 block sweeps to the FP32 tolerance (MFMA, FP32 accumulation)
//   V64 = fp64(V32); Newton-Schulz V <- V (3I - V^T V) / 2 (FP64 GEMMs) until V is orthogonal in FP64
//   A1 = A64 V64 (FP64 GEMM); FP64 block sweeps on (A1, V64) to epsilon
//   sigma / U from A1 (column norms, fused transpose + divide), V64 -> row-major V
// The GEMMs of the re-orthogonalisation and of A V are plain library GEMMs (rocBLAS dgemm).
#include "mixed.hpp"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "autotune.hpp"
#include "plain_f64.hpp"

namespace jsvd
{
    namespace
    {
        void blas_check(rocblas_status st, const char *what)
        {
            if (st != rocblas_status_success)
            {
                throw Error(JSVD_ERR_DEVICE, std::string("rocBLAS call failed: ") + what + ", status=" +
                                                 rocblas_status_to_string(st));
            }
        }

        double default_fp32_tolerance(int64_t m)
        {
            return std::max(1e-6, 4.0 * std::sqrt(static_cast<double>(m)) * std::ldexp(1.0, -24));
        }
        // Geometry of an m x n solve: block width `requested` (16, 32 or 64), or for 0 the autotuner's choice
        // when it measured this shape, else the default; the apply grid size measured with that width.
        BlockGeometry solver_geometry(int64_t m, int64_t n, int requested)
        {
            Tuning t;
            const bool tuned = lookup_tuning(m, n, t);
            const int kb = effective_block_width((requested == 0 && tuned) ? t.block_width : requested, n);
            BlockGeometry g = BlockGeometry::make(m, n, kb);
            if (tuned && t.block_width == kb)
            {
                g.apply_wgs = t.apply_wgs;
            }
            return g;
        }
    } // namespace

    BlockSolver::BlockSolver(int64_t m, int64_t n, const jsvd_config &cfg, const jsvd_mixed_config &mixed,
                             hipStream_t stream)
        : g_(solver_geometry(m, n, mixed.block_width)), cfg_(cfg), mixed_(mixed), stream_(stream)
    {
        if (g_.nb < 2)
        {
            throw_invalid("block and mixed paths need at least two 32-column blocks (n >= 33).");
        }
        fp32_tol_requested_ = mixed_.fp32_tolerance;
        if (mixed_.fp32_tolerance <= 0.0)
        {
            mixed_.fp32_tolerance = default_fp32_tolerance(m);
        }
        cap_ = std::max(cfg_.max_sweeps, mixed_.fp32_max_sweeps);
        const size_t a_elems = static_cast<size_t>(g_.lda) * g_.n_pad;
        const size_t v_elems = static_cast<size_t>(g_.ldv) * g_.n_pad;
        const size_t part_elems = g_.partial_elems();
        cap_a_elems_ = a_elems;
        cap_v_elems_ = v_elems;
        cap_part_elems_ = part_elems;
        cap_pairs_ = g_.pairs();
        cap_npad_ = g_.n_pad;
        cap_q_elems_ = g_.q_elems();
        JSVD_HIP_CHECK(hipMalloc(&a64_, sizeof(double) * a_elems));
        JSVD_HIP_CHECK(hipMalloc(&v64_, sizeof(double) * v_elems));
        JSVD_HIP_CHECK(hipMalloc(&partial_, sizeof(double) * part_elems));
        JSVD_HIP_CHECK(hipMalloc(&q_, sizeof(double) * g_.q_elems()));
        JSVD_HIP_CHECK(hipMalloc(&flags_, sizeof(int) * g_.pairs()));
        JSVD_HIP_CHECK(hipMalloc(&perm_, sizeof(int) * g_.n_pad));
        JSVD_HIP_CHECK(hipMalloc(&norm2_, sizeof(double) * g_.n_pad));
        JSVD_HIP_CHECK(hipMalloc(&rot_, sizeof(int) * 2 * cap_));
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
            JSVD_HIP_CHECK(hipMalloc(&gram64_, sizeof(double) * g_.n_pad * g_.n_pad));
            JSVD_HIP_CHECK(hipMalloc(&a32_, sizeof(float) * a_elems));
            JSVD_HIP_CHECK(hipMalloc(&v32_, sizeof(float) * v_elems));
            JSVD_HIP_CHECK(hipMemsetAsync(a64b_, 0, sizeof(double) * a_elems, stream_));
            rocblas_handle h = nullptr;
            blas_check(rocblas_create_handle(&h), "rocblas_create_handle");
            blas_ = h;
            blas_check(rocblas_set_stream(h, stream_), "rocblas_set_stream");
        }
    }

    BlockSolver::~BlockSolver()
    {
        if (blas_)
        {
            (void)rocblas_destroy_handle(static_cast<rocblas_handle>(blas_));
        }
        for (void *p : {static_cast<void *>(a64_), static_cast<void *>(a64b_), static_cast<void *>(v64_),
                        static_cast<void *>(w64_), static_cast<void *>(gram64_), static_cast<void *>(a32_),
                        static_cast<void *>(v32_), static_cast<void *>(partial_), static_cast<void *>(q_),
                        static_cast<void *>(flags_), static_cast<void *>(rot_), static_cast<void *>(off_),
                        static_cast<void *>(perm_), static_cast<void *>(norm2_)})
        {
            (void)hipFree(p);
        }
        (void)hipHostFree(h_rot_);
        for (hipEvent_t e : done_)
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
        }
    }

    // Block sweeps with the same device-side early exit and lagged host polling as the plain path.
    // Returns the executed sweep count (first rotation-free sweep included).
    template <typename T>
    int BlockSolver::run_sweeps(T *a, T *v, double eps, int max_sweeps, int *rot, unsigned long long *off,
                                double &last_off)
    {
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
        launch_iota(perm_, g_.n_pad, stream_);
        for (int s = 0; s < max_sweeps; ++s)
        {
            const int *skip = (s > 0) ? rot + (s - 1) : nullptr;
            if (s > 0)
            {
                launch_column_order(norm2_, g_.n_pad, perm_, skip, stream_);
            }
            for (int r = 0; r < g_.rounds(); ++r)
            {
                launch_block_round<T>(g_, a, v, r, eps, partial, q, flags_, rot + s, off + s, perm_, norm2_, skip,
                                      stream_);
            }
            JSVD_HIP_CHECK(hipMemcpyAsync(h + s, rot + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
            JSVD_HIP_CHECK(hipEventRecord(done_[s & 1], stream_));
            enqueued = s + 1;
            if (s >= 1)
            {
                JSVD_HIP_CHECK(hipEventSynchronize(done_[(s - 1) & 1]));
                if (h[s - 1] == 0)
                {
                    break;
                }
            }
        }
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
        int executed = enqueued;
        for (int s = 0; s < enqueued; ++s)
        {
            if (h[s] == 0)
            {
                executed = s + 1;
                break;
            }
        }
        unsigned long long bits = 0;
        JSVD_HIP_CHECK(hipMemcpy(&bits, off + (executed - 1), sizeof(bits), hipMemcpyDeviceToHost));
        std::memcpy(&last_off, &bits, sizeof(double));
        return executed;
    }

    // V <- V (3I - V^T V) / 2, written as W = 1.5 V - 0.5 V (V^T V); converges quadratically from an FP32
    // orthogonality error of ~1e-6 to FP64 rounding in two steps.
    void BlockSolver::orthogonalize_v(double *v, double *w)
    {
        rocblas_handle h = static_cast<rocblas_handle>(blas_);
        const rocblas_int np = static_cast<rocblas_int>(g_.n_pad);
        const rocblas_int ldv = static_cast<rocblas_int>(g_.ldv);
        double *cur = v;
        double *nxt = w;
        for (int step = 0; step < mixed_.orthogonalize_steps; ++step)
        {
            const double one = 1.0, zero = 0.0, minus_half = -0.5, three_halves = 1.5;
            blas_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, np, np, np, &one, cur, ldv,
                                     cur, ldv, &zero, gram64_, np),
                       "dgemm V^T V");
            JSVD_HIP_CHECK(hipMemcpyAsync(nxt, cur, sizeof(double) * g_.ldv * g_.n_pad, hipMemcpyDeviceToDevice, stream_));
            blas_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, np, np, np, &minus_half, cur,
                                     ldv, gram64_, np, &three_halves, nxt, ldv),
                       "dgemm V (V^T V)");
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
        BlockGeometry g = solver_geometry(m, n, mixed_.block_width);
        if (!fits(g))
        {
            // the preferred width needs more rotation or partial space than allocated: keep the current width
            g = BlockGeometry::make(m, n, effective_block_width(g_.kb, n));
            if (!fits(g))
            {
                return false;
            }
        }
        const size_t a_elems = static_cast<size_t>(g.lda) * g.n_pad;
        if (g.m != g_.m || g.n != g_.n || g.kb != g_.kb)
        {
            // padding rows and columns of the new extent must read as zero
            JSVD_HIP_CHECK(hipMemsetAsync(a64_, 0, sizeof(double) * a_elems, stream_));
            if (a64b_)
            {
                JSVD_HIP_CHECK(hipMemsetAsync(a64b_, 0, sizeof(double) * a_elems, stream_));
            }
        }
        g_ = g;
        mixed_.fp32_tolerance = (fp32_tol_requested_ > 0.0) ? fp32_tol_requested_ : default_fp32_tolerance(m);
        return true;
    }

    void BlockSolver::solve_device(const double *d_a, double *d_u, double *d_sigma, double *d_v,
                                   jsvd_solve_info &info)
    {
        std::memset(&info, 0, sizeof(info));
        info.path = mixed_.path;
        // row-major m x n -> column-major lda x n (padding rows/columns keep their zeros)
        launch_transpose(d_a, g_.n, a64_, g_.lda, g_.m, g_.n, stream_);
        double *a_final = a64_;
        if (mixed_.path == JSVD_PATH_MIXED)
        {
            const int64_t a_elems = g_.lda * g_.n_pad;
            const int64_t v_elems = g_.ldv * g_.n_pad;
            launch_convert_f64_to_f32(a64_, a32_, a_elems, stream_);
            launch_identity_f32(v32_, g_.ldv, g_.n_pad, stream_);
            const int fp32_cap = std::max(1, mixed_.fp32_max_sweeps);
            info.sweeps_fp32 = run_sweeps<float>(a32_, v32_, mixed_.fp32_tolerance, fp32_cap, rot_, off_,
                                                 info.offmax_fp32);
            launch_convert_f32_to_f64(v32_, v64_, v_elems, stream_);
            orthogonalize_v(v64_, w64_);
            rocblas_handle h = static_cast<rocblas_handle>(blas_);
            const double one = 1.0, zero = 0.0;
            blas_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none,
                                     static_cast<rocblas_int>(g_.lda), static_cast<rocblas_int>(g_.n_pad),
                                     static_cast<rocblas_int>(g_.n_pad), &one, a64_, static_cast<rocblas_int>(g_.lda),
                                     v64_, static_cast<rocblas_int>(g_.ldv), &zero, a64b_,
                                     static_cast<rocblas_int>(g_.lda)),
                       "dgemm A V");
            a_final = a64b_;
        }
        else
        {
            launch_pad_copy_f64(nullptr, 0, 0, 0, v64_, g_.ldv, g_.n_pad, true, stream_);
        }
        info.sweeps_fp64 = run_sweeps<double>(a_final, v64_, cfg_.epsilon, cfg_.max_sweeps, rot_ + cap_, off_ + cap_,
                                              info.offmax_fp64);
        info.converged = (h_rot_[cap_ + info.sweeps_fp64 - 1] == 0) ? 1 : 0;
        launch_u_sigma(a_final, g_.lda, g_.m, g_.n, cfg_.epsilon, d_u, d_sigma, stream_);
        launch_transpose(v64_, g_.ldv, d_v, g_.n, g_.n, g_.n, stream_);
        JSVD_HIP_CHECK(hipStreamSynchronize(stream_));
    }
} // namespace jsvd
