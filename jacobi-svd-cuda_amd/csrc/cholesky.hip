// cholesky.hip -- blocked Cholesky factorisation A = R^T R (R upper triangular) of a symmetric positive definite
// FP64 matrix, used by the mixed path's Cholesky QR.
//
// Right-looking over 64 x 64 tiles, three launches per tile step k:
//   cholesky_diag_kernel    four wavefronts: factor the diagonal tile A_kk with its columns in registers (64
//                           steps: pivot, scaled row through LDS, rank-1 update) and write R_kk
//   cholesky_panel_kernel   four wavefronts per tile column j > k: solve R_kk^T R_kj = A_kj by forward
//                           substitution, each column held by a quad of lanes (16 rows each)
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
            __shared__ int tile_first_bad; // first step of this tile with a non-positive pivot (kT: none)
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
            if (threadIdx.x == 0)
            {
                tile_first_bad = kT; // ordered before every update by the first step's barrier
            }
            int first_bad = kT;
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
                        first_bad = (first_bad == kT) ? c : first_bad;
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
            if (l == 0 && first_bad < kT)
            {
                atomicMin(&tile_first_bad, first_bad);
            }
            __syncthreads();
            // tiles are factored in order, so the first tile to fail sets info: 1 + its first bad row
            if (threadIdx.x == 0 && tile_first_bad < kT)
            {
                atomicCAS(info, 0, k * kT + tile_first_bad + 1);
            }
        }

        constexpr int kPanelWaves = 4;             // wavefronts of the panel kernel
        constexpr int kPanelRows = kT / 4;         // rows of a column each lane of a quad holds
        constexpr int kRzStride = kT + 2;          // LDS row stride of R_kk (16-byte aligned pairs)

        // Broadcast of v from lane `src` of every quad (DPP quad_perm), under full exec.
        template <int SRC>
        __device__ inline double quad_broadcast(double v)
        {
            constexpr int kCtrl = SRC | (SRC << 2) | (SRC << 4) | (SRC << 6);
            const long long bits = __double_as_longlong(v);
            int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(bits), kCtrl, 0xF, 0xF, false);
            int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(bits >> 32), kCtrl, 0xF, 0xF, false);
            asm volatile("" : "+v"(lo), "+v"(hi));
            return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
        }

        // Tile column j = k + 1 + blockIdx.x of block row k, four wavefronts: each quad of lanes holds one
        // column of the tile (lane q of the quad rows 16 q .. 16 q + 15) and solves R_kk^T y = x by forward
        // substitution: step r, the lane holding row r forms y[r] = x[r] / R[r][r] (times the reciprocal), the
        // quad shares it by a DPP broadcast and every lane applies x[p] -= R[r][p] y[r] to its rows.  R_kk sits in
        // LDS with its entries on and below the diagonal replaced by zeros, so rows p <= r take a zero update and
        // no lane needs a per-row test; the arithmetic of every entry above is that of the one-lane-per-column
        // substitution.
        __global__ __launch_bounds__(kT * kPanelWaves) void cholesky_panel_kernel(double *__restrict__ a, int64_t ld,
                                                                                 int k)
        {
            __shared__ __align__(16) double rz[kT][kRzStride]; // rz[r][p] = R_kk[r][p] for p > r, else 0
            __shared__ double inv[kT];                          // 1 / R_kk[r][r]
            const int tid = threadIdx.x;
            const int j = k + 1 + static_cast<int>(blockIdx.x);
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            const double *akk = a + r0 * ld + r0;
            // all loads of R_kk and of this lane's column in flight at once, then the LDS stores
            constexpr int kLoads = kT * kT / (kT * kPanelWaves);
            double v[kLoads];
#pragma unroll
            for (int t = 0; t < kLoads; ++t)
            {
                const int e = tid + t * kT * kPanelWaves;
                v[t] = akk[static_cast<int64_t>(e / kT) * ld + e % kT]; // column e / kT, row e % kT
            }
            const int lane = tid % kT;
            const int q = lane & 3;                             // row quarter of this lane
            const int c = (tid / kT) * (kT / 4) + (lane >> 2);  // tile column of this quad
            double *col = a + (static_cast<int64_t>(j) * kT + c) * ld + r0 + q * kPanelRows;
            double x[kPanelRows];
#pragma unroll
            for (int i = 0; i < kPanelRows; ++i)
            {
                x[i] = col[i];
            }
#pragma unroll
            for (int t = 0; t < kLoads; ++t)
            {
                const int e = tid + t * kT * kPanelWaves;
                const int p = e / kT;
                const int r = e % kT;
                rz[r][p] = (p > r) ? v[t] : 0.0;
                if (p == r)
                {
                    inv[r] = 1.0 / v[t];
                }
            }
            __syncthreads();
            const double *rq = &rz[0][q * kPanelRows];
#pragma unroll
            for (int r = 0; r < kT; ++r)
            {
                const int owner = r / kPanelRows; // compile-time after unrolling
                const int ir = r % kPanelRows;
                const double y = x[ir] * inv[r];
                x[ir] = (q == owner) ? y : x[ir];
                const double yb = (owner == 0)   ? quad_broadcast<0>(y)
                                  : (owner == 1) ? quad_broadcast<1>(y)
                                  : (owner == 2) ? quad_broadcast<2>(y)
                                                 : quad_broadcast<3>(y);
#pragma unroll
                for (int i = 0; i < kPanelRows; ++i)
                {
                    x[i] = fma(-rq[r * kRzStride + i], yb, x[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < kPanelRows; ++i)
            {
                col[i] = x[i];
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
            // all loads of both operand tiles and of the output tile in flight at once
            constexpr int kLoads = kT * kT / 256;
            double vi[kLoads], vj[kLoads];
#pragma unroll
            for (int t = 0; t < kLoads; ++t)
            {
                const int e = tid + t * 256;
                const int c = e / kT, r = e % kT;
                vi[t] = pi[static_cast<int64_t>(c) * ld + r];
                vj[t] = pj[static_cast<int64_t>(c) * ld + r];
            }
            const int lane = tid % 64;
            const int wave = tid / 64;
            double *out = a + static_cast<int64_t>(j) * kT * ld + static_cast<int64_t>(i) * kT;
            double prev[4][4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    prev[t][v] = out[static_cast<int64_t>(16 * t + lane % 16) * ld + 16 * wave + lane / 16 + 4 * v];
                }
            }
#pragma unroll
            for (int t = 0; t < kLoads; ++t)
            {
                const int e = tid + t * 256;
                const int c = e / kT, r = e % kT;
                ri[c][r] = vi[t];
                rj[c][r] = vj[t];
            }
            __syncthreads();
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
                        out[static_cast<int64_t>(col) * ld + row] = prev[t][v] - acc[t][v];
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
