// cholesky.hip -- blocked Cholesky factorisation A = R^T R (R upper triangular) of a symmetric positive definite
// FP64 matrix, used by the mixed path's Cholesky QR.
//
// Right-looking over 64 x 64 tiles, three launches per tile step k:
//   cholesky_diag_kernel    four wavefronts: factor the diagonal tile A_kk with its columns in registers (64
//                           steps: pivot, scaled row through LDS, rank-1 update) and write R_kk
//   cholesky_panel_kernel   one wavefront per tile column j > k: solves R_kk^T R_kj = A_kj by forward
//                           substitution, each lane one column in registers
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

        // v of lane `lane` (a compile-time constant at every call) read into scalar registers.
        __device__ inline double lane_value(double v, int lane)
        {
            const long long bits = __double_as_longlong(v);
            const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
            const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
            return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
        }

        constexpr int kDiagWaves = 4;               // wavefronts of the diagonal-tile kernel
        constexpr int kDiagRows = kT / kDiagWaves;  // tile rows each of them holds

        // Four wavefronts: thread (w, l) holds rows 16 w .. 16 w + 15 of column l of the diagonal tile in
        // registers.  Step c: the wavefront owning row c reads the pivot from lane c (v_readlane), forms the
        // scaled row R[c][l] = A[c][l] / sqrt(pivot) and publishes it in LDS (two buffers, alternating, so one
        // barrier per step suffices); every wavefront then applies the rank-1 update to its rows below c,
        // reading R[c][i] as LDS broadcasts.  A lane's entries below its column's diagonal take part in the
        // arithmetic but are never written back.
        __global__ __launch_bounds__(kT * kDiagWaves) void cholesky_diag_kernel(double *__restrict__ a, int64_t ld,
                                                                               int k, int *__restrict__ info)
        {
            __shared__ __align__(16) double row[2][kT];
            const int l = threadIdx.x % kT;
            const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kT)); // wave-uniform
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            double *col = a + (r0 + l) * ld + r0 + w * kDiagRows;
            double x[kDiagRows];
#pragma unroll
            for (int i = 0; i < kDiagRows; ++i)
            {
                x[i] = col[i];
            }
            bool bad = false;
#pragma unroll
            for (int c = 0; c < kT; ++c)
            {
                const int b = c & 1;
                const int ow = c / kDiagRows; // owner wavefront and its register of row c (compile-time)
                const int oi = c % kDiagRows;
                if (w == ow)
                {
                    double piv = lane_value(x[oi], c);
                    if (!(piv > 0.0))
                    {
                        bad = true;
                        piv = 1.0; // keep the arithmetic finite; the factorisation is reported as failed
                    }
                    const double r = sqrt(piv);
                    const double u = x[oi] / r; // R[c][l] for l > c
                    x[oi] = (l == c) ? r : u;
                    row[b][l] = u;
                }
                __syncthreads();
                const double ul = row[b][l];
                if (w > ow)
                {
                    const double *rc = &row[b][w * kDiagRows];
#pragma unroll
                    for (int i = 0; i < kDiagRows; ++i)
                    {
                        x[i] = fma(-rc[i], ul, x[i]);
                    }
                }
                else if (w == ow)
                {
#pragma unroll
                    for (int i = oi + 1; i < kDiagRows; ++i)
                    {
                        x[i] = fma(-row[b][ow * kDiagRows + i], ul, x[i]);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < kDiagRows; ++i)
            {
                if (w * kDiagRows + i <= l)
                {
                    col[i] = x[i];
                }
            }
            if (l == 0 && bad)
            {
                atomicCAS(info, 0, k * kT + 1);
            }
        }

        // Tile column j = k + 1 + blockIdx.x of block row k, one wavefront: lane c holds column c of the tile in
        // registers and solves R_kk^T y = x by forward substitution (y[r] = x[r] / R[r][r], then x[p] -= R[r][p]
        // y[r] for p > r), R_kk staged in LDS and read as broadcasts.
        __global__ __launch_bounds__(64) void cholesky_panel_kernel(double *__restrict__ a, int64_t ld, int k)
        {
            __shared__ double rt[kT][kT]; // rt[p][r] = R_kk[r][p] (column p of R_kk), upper part read
            __shared__ double inv[kT];    // 1 / R_kk[r][r]
            const int j = k + 1 + static_cast<int>(blockIdx.x);
            const int lane = threadIdx.x;
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            const double *akk = a + r0 * ld + r0;
            for (int p = 0; p < kT; ++p)
            {
                rt[p][lane] = akk[static_cast<int64_t>(p) * ld + lane];
            }
            double *col = a + (static_cast<int64_t>(j) * kT + lane) * ld + r0;
            double x[kT];
#pragma unroll
            for (int r = 0; r < kT; ++r)
            {
                x[r] = col[r];
            }
            __syncthreads();
            inv[lane] = 1.0 / rt[lane][lane];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kT; ++r)
            {
                x[r] *= inv[r];
#pragma unroll
                for (int p = r + 1; p < kT; ++p)
                {
                    x[p] = fma(-rt[p][r], x[r], x[p]);
                }
            }
#pragma unroll
            for (int r = 0; r < kT; ++r)
            {
                col[r] = x[r];
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
            hipLaunchKernelGGL(cholesky_diag_kernel, dim3(1), dim3(kT * kDiagWaves), 0, stream, a, ld, k, info);
            JSVD_HIP_CHECK(hipGetLastError());
            const int rest = nt - k - 1;
            if (rest > 0)
            {
                hipLaunchKernelGGL(cholesky_panel_kernel, dim3(static_cast<unsigned>(rest)), dim3(64), 0, stream, a,
                                   ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
                hipLaunchKernelGGL(cholesky_update_kernel, dim3(static_cast<unsigned>(rest), static_cast<unsigned>(rest)),
                                   dim3(256), 0, stream, a, ld, k);
                JSVD_HIP_CHECK(hipGetLastError());
            }
        }
    }
} // namespace jsvd
