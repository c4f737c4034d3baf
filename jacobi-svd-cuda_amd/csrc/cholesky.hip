// cholesky.hip -- blocked Cholesky factorisation A = R^T R (R upper triangular) of a symmetric positive definite
// FP64 matrix, used by the mixed path's Cholesky QR.
//
// Right-looking over 64 x 64 tiles, three launches per tile step k:
//   cholesky_diag_kernel    four wavefronts: factor the diagonal tile A_kk in LDS, four block rows of 16 (one
//                           wavefront's lane-broadcast pivot steps, then an MFMA rank-16 trailing update)
//   cholesky_panel_kernel   four wavefronts per tile column j > k: solve R_kk^T R_kj = A_kj by forward
//                           substitution, each column held by a quad of lanes (16 rows each)
//   cholesky_update_kernel  one workgroup per trailing tile (i, j), k < i <= j: A_ij -= R_ki^T R_kj on FP64 MFMA
//                           (v_mfma_f64_16x16x4f64), both operand tiles staged in LDS
// Only the upper triangle is read and written; the strictly lower triangle keeps whatever it held.
#include "cholesky.hpp"
#include "cholesky_kernels.hpp"

namespace jsvd
{
    using namespace chol;

    int64_t cholesky_padded_size(int64_t n) { return round_up(n, kT); }

    void launch_cholesky_upper(double *a, int64_t ld, int64_t n, int *info, hipStream_t stream)
    {
        const int64_t n64 = cholesky_padded_size(n);
        if (ld < n64)
        {
            throw_invalid("cholesky: leading dimension below the padded size.");
        }
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, stream, info);
        JSVD_HIP_CHECK(hipGetLastError());
        if (n64 != n)
        {
            hipLaunchKernelGGL(pad_identity_kernel, dim3(static_cast<unsigned>((n64 * n64 + 255) / 256)), dim3(256),
                               0, stream, a, ld, n, n64);
            JSVD_HIP_CHECK(hipGetLastError());
        }
        const int nt = static_cast<int>(n64 / kT);
        for (int k = 0; k < nt; ++k)
        {
            hipLaunchKernelGGL(cholesky_diag_kernel, dim3(1), dim3(kT * kDiagWaves), 0, stream, a, ld, k, info);
            JSVD_HIP_CHECK(hipGetLastError());
            const int rest = nt - k - 1;
            if (rest > 0)
            {
                hipLaunchKernelGGL(cholesky_panel_kernel, dim3(static_cast<unsigned>(rest)), dim3(kT * kPanelWaves),
                                   0, stream, a, ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
                hipLaunchKernelGGL(cholesky_update_kernel, dim3(static_cast<unsigned>(rest), static_cast<unsigned>(rest)),
                                   dim3(256), 0, stream, a, ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
            }
        }
    }
} // namespace jsvd
