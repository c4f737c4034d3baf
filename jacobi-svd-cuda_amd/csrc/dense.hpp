// dense.hpp -- host launchers of the FP64 dense kernels (dense_kernels.hpp) used by the mixed path.
#pragma once

#include "common.hpp"

namespace jsvd
{
    namespace dense
    {
        // Structure flags of a product (the JSVD_GEMM_* values of jsvd.h).
        enum GemmFlags : int
        {
            kUpperOut = 1,   // only output tiles with tile row <= tile column
            kMirror = 2,     // with kUpperOut: an off-diagonal tile is also written transposed below the diagonal
            kTriA = 4,       // op(A) upper triangular: entries (i, k) with i > k read as zero
            kTriB = 8,       // op(B) upper triangular: entries (k, j) with k > j read as zero
            kZeroLower = 16, // result entries (i, j) with i > j are written as zero
            kRowMajorCD = 32 // C and D row-major: entry (i, j) at i * ld + j
        };
        static_assert(kUpperOut == JSVD_GEMM_UPPER_OUT && kMirror == JSVD_GEMM_MIRROR && kTriA == JSVD_GEMM_TRI_A &&
                          kTriB == JSVD_GEMM_TRI_B && kZeroLower == JSVD_GEMM_ZERO_LOWER &&
                          kRowMajorCD == JSVD_GEMM_ROW_MAJOR_CD,
                      "flag values of the C ABI");
    } // namespace dense

    // Scratch of the dense products: split-k partial tiles.
    struct DenseWorkspace
    {
        double *work = nullptr;
        int64_t work_elems = 0;
    };

    // Size of a workspace that lets every product split k as far as the launcher would (any size works; a smaller
    // one, or none, only limits the splitting).
    int64_t dense_work_elems();

    // D = alpha op(A) op(B) + beta C with op(X) = X or X^T (trans_a / trans_b); column-major A (lda), B (ldb); C and
    // D column-major unless flags has dense::kRowMajorCD.  flags: dense::GemmFlags.  C may alias D.
    void launch_gemm_f64(bool trans_a, bool trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                         int64_t lda, const double *b, int64_t ldb, double beta, const double *c, int64_t ldc,
                         double *d, int64_t ldd, int flags, const DenseWorkspace &ws, hipStream_t stream);

    // The triangular solve pads R to a multiple of 64 (identity outside n x n).  Doubles of that padded copy and of
    // the inverted 64 x 64 diagonal blocks that launch_trsm_right_upper_f64 needs for an n x n R.
    int64_t trsm_padded_size(int64_t n);
    int64_t trsm_rp_elems(int64_t n);
    int64_t trsm_dinv_elems(int64_t n);

    // B (rows x n, column-major, ldb) <- B R^-1 for the upper triangle of R (n x n, column-major, ldr; the strictly
    // lower triangle is not read).  r_padded: R is already padded as the solve needs it -- ldr >= trsm_padded_size(n)
    // and R(k, j) = identity for max(k, j) >= n within that size (the Cholesky factorisation's padding) -- and is used
    // in place; else rp (trsm_rp_elems(n) doubles) receives a padded copy.  dinv: trsm_dinv_elems(n) doubles.
    void launch_trsm_right_upper_f64(double *b, int64_t ldb, int64_t rows, int64_t n, const double *r, int64_t ldr,
                                     bool r_padded, double *rp, double *dinv, const DenseWorkspace &ws,
                                     hipStream_t stream);
} // namespace jsvd
