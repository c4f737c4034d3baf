// mixed.hpp -- block and mixed-precision one-sided Jacobi solvers (mixed.hip).
#pragma once

#include <memory>
#include <vector>

#include "block.hpp"

namespace jsvd
{
    // Device-resident block solver for one shape.  Path JSVD_PATH_BLOCK_F64 runs FP64 block sweeps only;
    // JSVD_PATH_MIXED runs FP32 block sweeps (MFMA Gram and rotate, FP32 accumulation) to the FP32 tolerance,
    // re-orthogonalises the FP32 V in FP64 (Newton-Schulz), forms A V in FP64 and finishes with FP64 block
    // sweeps to the caller's epsilon.
    class BlockSolver
    {
    public:
        BlockSolver(int64_t m, int64_t n, const jsvd_config &cfg, const jsvd_mixed_config &mixed, hipStream_t stream);
        ~BlockSolver();
        BlockSolver(const BlockSolver &) = delete;
        BlockSolver &operator=(const BlockSolver &) = delete;

        // d_a row-major m x n (not modified); outputs row-major U (m x n), sigma (n), V (n x n).
        void solve_device(const double *d_a, double *d_u, double *d_sigma, double *d_v, jsvd_solve_info &info);

        // Re-targets the solver at an m x n problem inside the buffers allocated at construction (the capacity
        // shape); false (nothing changed) when it does not fit or n < 33.
        bool reshape(int64_t m, int64_t n);
        void set_epsilon(double eps) { cfg_.epsilon = eps; }
        [[nodiscard]] const BlockGeometry &geometry() const { return g_; }

    private:
        template <typename T>
        int run_sweeps(T *a, T *v, double eps, int max_sweeps, int *rot, unsigned long long *off, double &last_off);
        void orthogonalize_v(double *v, double *w);

        BlockGeometry g_;
        jsvd_config cfg_;
        jsvd_mixed_config mixed_;
        hipStream_t stream_;
        void *blas_ = nullptr; // rocblas_handle
        double *a64_ = nullptr;
        double *a64b_ = nullptr;
        double *v64_ = nullptr;
        double *w64_ = nullptr;
        double *gram64_ = nullptr;
        float *a32_ = nullptr;
        float *v32_ = nullptr;
        double *partial_ = nullptr; // reused as float for the FP32 phase
        double *q_ = nullptr;
        int *flags_ = nullptr;
        int *perm_ = nullptr;      // [n_pad] logical -> physical column of the current sweep
        double *norm2_ = nullptr;  // [n_pad] squared column norms after each column's last rotation
        int *rot_ = nullptr;                 // [2 * cap] rotation counts (FP32 sweeps, then FP64 sweeps)
        unsigned long long *off_ = nullptr;  // [2 * cap]
        int *h_rot_ = nullptr;               // pinned mirror
        int cap_ = 0;
        // capacity of the geometry-sized buffers
        size_t cap_a_elems_ = 0;
        size_t cap_v_elems_ = 0;
        size_t cap_part_elems_ = 0;
        size_t cap_q_elems_ = 0;
        int cap_pairs_ = 0;
        int64_t cap_npad_ = 0;
        double fp32_tol_requested_ = 0.0;
        hipEvent_t done_[2] = {nullptr, nullptr};
    };
} // namespace jsvd
