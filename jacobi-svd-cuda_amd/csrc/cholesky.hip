// cholesky.hip -- blocked Cholesky factorisation A = R^T R (R upper triangular) of a symmetric positive definite
// FP64 matrix, used by the mixed path's Cholesky QR.
//
// Right-looking over 64 x 64 tiles, three launches per tile step k:
//   cholesky_diag_kernel    one workgroup: factors the diagonal tile A_kk in LDS (64 steps of a rank-1 update of
//                           the trailing rows with the unscaled pivot row, then the row scaled) and writes R_kk
//   cholesky_panel_kernel   one workgroup per tile column j > k: solves R_kk^T R_kj = A_kj by forward
//                           substitution in LDS
//   cholesky_update_kernel  one workgroup per trailing tile (i, j), k < i <= j: A_ij -= R_ki^T R_kj on FP64 MFMA
//                           (v_mfma_f64_16x16x4f64), both operand tiles staged in LDS
// Only the upper triangle is read and written; the strictly lower triangle keeps whatever it held.
#include "cholesky.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kT = 64; // tile size

        typedef double f64x4c __attribute__((ext_vector_type(4)));

        __global__ __launch_bounds__(256) void cholesky_diag_kernel(double *__restrict__ a, int64_t ld, int k,
                                                                    int *__restrict__ info)
        {
            __shared__ double d[kT][kT + 1]; // diagonal tile, d[col][row]
            __shared__ int bad;
            const int tid = threadIdx.x;
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            double *akk = a + r0 * ld + r0;
            for (int e = tid; e < kT * kT; e += 256)
            {
                const int c = e / kT, r = e % kT;
                d[c][r] = akk[static_cast<int64_t>(c) * ld + r];
            }
            if (tid == 0)
            {
                bad = 0;
            }
            __syncthreads();
            // Step c updates rows c+1.. with the unscaled row c (A[i][l] -= A[c][i] A[c][l] / A[c][c]), then
            // scales row c by 1 / sqrt(A[c][c]).
            for (int c = 0; c < kT; ++c)
            {
                double piv = d[c][c];
                if (!(piv > 0.0))
                {
                    if (tid == 0)
                    {
                        bad = 1;
                    }
                    piv = 1.0; // keep the arithmetic finite; the factorisation is reported as failed
                }
                const int w = kT - 1 - c; // trailing size
                for (int e = tid; e < w * w; e += 256)
                {
                    const int i = c + 1 + e / w;
                    const int l = c + 1 + e % w;
                    if (i <= l)
                    {
                        d[l][i] -= d[i][c] * d[l][c] / piv;
                    }
                }
                __syncthreads();
                if (tid >= c && tid < kT)
                {
                    const double r = sqrt(piv);
                    d[tid][c] = (tid == c) ? r : d[tid][c] / r;
                }
                __syncthreads();
            }
            for (int e = tid; e < kT * kT; e += 256)
            {
                const int c = e / kT, r = e % kT;
                if (r <= c)
                {
                    akk[static_cast<int64_t>(c) * ld + r] = d[c][r];
                }
            }
            if (tid == 0 && bad)
            {
                atomicCAS(info, 0, k * kT + 1);
            }
        }

        // Tile column j = k + 1 + blockIdx.x of block row k: R_kk^T R_kj = A_kj, row by row (Y[r][:] = X[r][:] /
        // R[r][r], then X[p][:] -= R[r][p] Y[r][:] for p > r).
        __global__ __launch_bounds__(256) void cholesky_panel_kernel(double *__restrict__ a, int64_t ld, int k)
        {
            __shared__ double d[kT][kT + 1]; // R_kk, d[col][row] (upper part read)
            __shared__ double x[kT][kT + 1]; // the panel tile, x[col][row]
            const int j = k + 1 + static_cast<int>(blockIdx.x);
            const int tid = threadIdx.x;
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            const double *akk = a + r0 * ld + r0;
            double *akj = a + static_cast<int64_t>(j) * kT * ld + r0;
            for (int e = tid; e < kT * kT; e += 256)
            {
                const int c = e / kT, r = e % kT;
                d[c][r] = akk[static_cast<int64_t>(c) * ld + r];
                x[c][r] = akj[static_cast<int64_t>(c) * ld + r];
            }
            __syncthreads();
            for (int r = 0; r < kT; ++r)
            {
                if (tid < kT)
                {
                    x[tid][r] /= d[r][r];
                }
                __syncthreads();
                const int w = kT - 1 - r;
                for (int e = tid; e < w * kT; e += 256)
                {
                    const int p = r + 1 + e / kT;
                    const int col = e % kT;
                    x[col][p] -= d[p][r] * x[col][r];
                }
                __syncthreads();
            }
            for (int e = tid; e < kT * kT; e += 256)
            {
                const int c = e / kT, r = e % kT;
                akj[static_cast<int64_t>(c) * ld + r] = x[c][r];
            }
        }

        // Tile (i, j) = (k + 1 + blockIdx.y, k + 1 + blockIdx.x), skipped below the diagonal.
        __global__ __launch_bounds__(256) void cholesky_update_kernel(double *__restrict__ a, int64_t ld, int k)
        {
            __shared__ double ri[kT][kT + 1]; // R_ki, [col][row]
            __shared__ double rj[kT][kT + 1]; // R_kj, [col][row]
            const int i = k + 1 + static_cast<int>(blockIdx.y);
            const int j = k + 1 + static_cast<int>(blockIdx.x);
            if (i > j)
            {
                return;
            }
            const int tid = threadIdx.x;
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            const double *pi = a + static_cast<int64_t>(i) * kT * ld + r0;
            const double *pj = a + static_cast<int64_t>(j) * kT * ld + r0;
            for (int e = tid; e < kT * kT; e += 256)
            {
                const int c = e / kT, r = e % kT;
                ri[c][r] = pi[static_cast<int64_t>(c) * ld + r];
                rj[c][r] = pj[static_cast<int64_t>(c) * ld + r];
            }
            __syncthreads();
            const int lane = tid % 64;
            const int wave = tid / 64;
            // C = R_ki^T R_kj: wave w computes rows 16w..16w+15 of C, all four 16-column tiles.
            f64x4c acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                acc[t] = f64x4c{0.0, 0.0, 0.0, 0.0};
            }
            for (int kk = 0; kk < kT / 4; ++kk)
            {
                const int p = 4 * kk + lane / 16;
                const double av = ri[16 * wave + lane % 16][p]; // A[row][p] = R_ki[p][row]
#pragma unroll
                for (int t = 0; t < 4; ++t)
                {
                    const double bv = rj[16 * t + lane % 16][p]; // B[p][col] = R_kj[p][col]
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
                }
            }
            double *out = a + static_cast<int64_t>(j) * kT * ld + static_cast<int64_t>(i) * kT;
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                const int col = 16 * t + lane % 16;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int row = 16 * wave + lane / 16 + 4 * v;
                    if (i < j || row <= col) // a diagonal tile keeps its strictly lower triangle
                    {
                        out[static_cast<int64_t>(col) * ld + row] -= acc[t][v];
                    }
                }
            }
        }

        // The padding of an n x n matrix to n64 x n64: identity rows and columns n..n64-1.
        __global__ void pad_identity_kernel(double *__restrict__ a, int64_t ld, int64_t n, int64_t n64)
        {
            const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (e >= n64 * n64)
            {
                return;
            }
            const int64_t c = e / n64;
            const int64_t r = e % n64;
            if (r >= n || c >= n)
            {
                a[c * ld + r] = (r == c) ? 1.0 : 0.0;
            }
        }

        __global__ void zero_int_kernel(int *__restrict__ p)
        {
            *p = 0;
        }
    } // namespace

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
            hipLaunchKernelGGL(cholesky_diag_kernel, dim3(1), dim3(256), 0, stream, a, ld, k, info);
            JSVD_HIP_CHECK(hipGetLastError());
            const int rest = nt - k - 1;
            if (rest > 0)
            {
                hipLaunchKernelGGL(cholesky_panel_kernel, dim3(static_cast<unsigned>(rest)), dim3(256), 0, stream, a,
                                   ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
                hipLaunchKernelGGL(cholesky_update_kernel, dim3(static_cast<unsigned>(rest), static_cast<unsigned>(rest)),
                                   dim3(256), 0, stream, a, ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
            }
        }
    }
} // namespace jsvd
