// mixed.hpp -- block and mixed-precision one-sided Jacobi solvers (mixed.hip).
#pragma once

#include <memory>
#include <vector>

#include "block.hpp"
#include "dense.hpp"

namespace jsvd
{
    // Device-resident block solver for one shape.  Path JSVD_PATH_BLOCK_F64 runs FP64 block sweeps only.
    // Path JSVD_PATH_MIXED first factors A = Q R in FP64 (column-scaled shifted Cholesky QR, three passes) and
    // solves the n x n problem M = R^T: FP32 block sweeps (MFMA Gram and rotate, FP32 accumulation) to the FP32
    // tolerance, the FP32 V re-orthogonalised in FP64 (Newton-Schulz), M V formed in FP64, FP64 block sweeps to
    // the caller's epsilon; then sigma_j = |x_j|, V_A = X / sigma (X = the converged M V) and U_A = Q V_M.  Sweeping
    // R^T instead of A diagonalises R R^T, one QR-algorithm step further along than A^T A, which cuts the sweep
    // count of both phases; the sweeps also run on n instead of m rows.  qr_precondition = 2 adds a second QR
    // (3, the default, adds it where it pays: n >= 512 and an ill-conditioned R),
    // M = Q2 R2 (the same column-scaled shifted Cholesky QR), and sweeps N = R2^T instead (one step further again,
    // the LQ step of preconditioned Jacobi); then Y = X / sigma, U_A = Q Y and V_A = Q2 V.  When a factor of the
    // second QR breaks down the solve goes on with one QR; when one of the first breaks down or a sigma comes out
    // exactly zero (V_A cannot be normalised) the solve is redone on A directly (FP32 sweeps, Newton-Schulz, A V,
    // FP64 sweeps) with the reference's output conventions.  qr_precondition = 0 selects that direct form outright.
    // Automatic second QR (qr_precondition 3): only for n >= this many columns and an R whose diagonal spread
    // max |r_ii| / min |r_ii| is at least kSecondQrMinSpread.
    constexpr int64_t kSecondQrMinColumns = 512;
    constexpr double kSecondQrMinSpread = 100.0;

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
        // Per-sweep rotated-pair counts and largest pair cosines of the last solve: the FP32 sweeps (mixed path),
        // then the FP64 sweeps.
        [[nodiscard]] const std::vector<int> &sweep_rotations() const { return stat_rot_; }
        [[nodiscard]] const std::vector<double> &sweep_offmax() const { return stat_off_; }

    private:
        template <typename T>
        int run_sweeps(const BlockGeometry &g, T *a, T *v, double eps, int max_sweeps, int *rot,
                       unsigned long long *off, double &last_off, int stop_rotations = 0);
        void orthogonalize_v(double *v, double *w);
        [[nodiscard]] bool qr_first() const { return mixed_.path == JSVD_PATH_MIXED && mixed_.qr_precondition != 0; }
        void set_geometry(int64_t m, int64_t n, const BlockGeometry &sweep);
        // FP32 sweeps on (x32 = fp32(x64), V32 = I) with geometry g, V64 <- orthogonalised fp64(V32), y = x64 V64.
        void fp32_phase(const BlockGeometry &g, const double *x64, double *y64, double tol, jsvd_solve_info &info);
        // Column-scaled shifted Cholesky QR of x (m x n, leading dimension lda): x <- Q, (R D)^T into m_out (sweep
        // geometry).  False when a Cholesky factor breaks down.
        // With spread, also max |r_ii| / min |r_ii| of R (read back with the factorisation's status).
        bool factor_qr(double *x, int64_t m, int64_t n, int64_t lda, double *m_out, double *spread);
        void solve_direct(const double *d_a, double *d_u, double *d_sigma, double *d_v, jsvd_solve_info &info);
        bool solve_qr_first(const double *d_a, double *d_u, double *d_sigma, double *d_v, jsvd_solve_info &info);

        BlockGeometry ga_; // A-side geometry (m x n): input transpose, QR, direct solves
        BlockGeometry g_;  // sweep geometry: n x n for the QR-first mixed path, else m x n
        jsvd_config cfg_;
        jsvd_mixed_config mixed_;
        hipStream_t stream_;
        DenseWorkspace dws_;         // split-k scratch of the dense products
        double *dinv_ = nullptr;     // inverted diagonal blocks of the triangular solve
        double *a64_ = nullptr;
        double *a64b_ = nullptr;
        double *q2_ = nullptr;       // second QR of the mixed path: Q2 (n x n, sweep geometry), or null
        double *n2_ = nullptr;       // second QR: N = (R2 D2)^T, the matrix the sweeps run on
        double *spread_ = nullptr;   // [1] diagonal spread of the first QR's R
        double *v64_ = nullptr;
        double *w64_ = nullptr;
        double *gram64_ = nullptr;
        float *a32_ = nullptr;
        float *v32_ = nullptr;
        double *partial_ = nullptr; // reused as float for the FP32 phase
        double *racc_ = nullptr;    // [n_pad x n_pad] accumulated Cholesky QR factor
        double *dcol_ = nullptr;    // [n_pad] column norms of A (QR column scaling)
        int *chol_info_ = nullptr;  // [3] Cholesky status of the three QR passes
        double *h_sigma_ = nullptr; // [n_pad] pinned copy of sigma (zero check)
        double *q_ = nullptr;
        int *flags_ = nullptr;
        int *perm_ = nullptr;      // [n_pad] logical -> physical column of the current sweep
        double *norm2_ = nullptr;  // [n_pad] squared column norms after each column's last rotation
        int *rot_ = nullptr;                 // [2 * cap] rotation counts (FP32 sweeps, then FP64 sweeps)
        int *cont_ = nullptr;                // [2 * cap] 1 while the phase goes on after sweep s
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
        std::vector<int> stat_rot_;
        std::vector<double> stat_off_;
    };
} // namespace jsvd
