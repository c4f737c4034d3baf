// cholesky_kernels.hpp -- device code of the blocked Cholesky factorisation (cholesky.hip launches it): the
// diagonal-tile, panel and trailing-update kernels, the identity padding and the status reset.
#pragma once

#include "common.hpp"

namespace jsvd
{
    namespace chol
    {
        constexpr int kT = 64; // tile size

        typedef dev::vec_t<double, 4> f64x4c;

        // v of lane `lane` (a compile-time constant at every call) read into scalar registers.
        __device__ inline double lane_value(double v, int lane)
        {
            const long long bits = __double_as_longlong(v);
            const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
            const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
            return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
        }

        constexpr int kDiagWaves = 4; // wavefronts of the diagonal-tile kernel
        constexpr int kSub = 16;      // rows of a block row of the diagonal tile

        // Blocked factorisation of the diagonal tile A_kk (64 x 64, upper triangle) in LDS, four block rows of 16:
        //   block row b, wavefront 0: lane l holds rows 16 b .. 16 b + 15 of tile column l in registers and runs the
        //     16 pivot steps of those rows across the whole tile -- pivot from its lane (v_readlane), square root,
        //     scaled row R[c][l] = A[c][l] / R[c][c] for l > c, then x[i] -= R[c][16 b + i] R[c][l] for the rows
        //     below c, R[c][16 b + i] read from lane 16 b + i -- with no barrier inside;
        //   all four wavefronts: the rank-16 update of the trailing 16 x 16 sub-blocks on or above the diagonal,
        //     A_ij -= R_bi^T R_bj for b < i <= j (FP64 MFMA, operands from LDS), two barriers per block row.
        // Entries below a column's diagonal take part in the arithmetic but are never written back; the strictly lower
        // triangle of the tile in memory keeps whatever it held.
        __global__ __launch_bounds__(kT * kDiagWaves) void cholesky_diag_kernel(double *__restrict__ a, int64_t ld,
                                                                               int k, int *__restrict__ info)
        {
            __shared__ double t[kT][kT + 1]; // t[row][col]
            __shared__ int tile_first_bad;   // first step of this tile with a non-positive pivot (kT: none)
            const int tid = threadIdx.x;
            const int lane = tid % kT;
            const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(tid / kT)); // wave-uniform
            const int64_t r0 = static_cast<int64_t>(k) * kT;
            double *tile = a + r0 * ld + r0;
            // the whole tile (both triangles; the lower one is never written back) in flight at once
            constexpr int kLoads = kT * kT / (kT * kDiagWaves);
            double v[kLoads];
#pragma unroll
            for (int q = 0; q < kLoads; ++q)
            {
                const int e = tid + q * kT * kDiagWaves; // column e / kT, row e % kT
                v[q] = tile[static_cast<int64_t>(e / kT) * ld + e % kT];
            }
#pragma unroll
            for (int q = 0; q < kLoads; ++q)
            {
                const int e = tid + q * kT * kDiagWaves;
                t[e % kT][e / kT] = v[q];
            }
            if (tid == 0)
            {
                tile_first_bad = kT;
            }
            __syncthreads();
            int first_bad = kT;
#pragma unroll
            for (int b = 0; b < kT / kSub; ++b)
            {
                if (w == 0)
                {
                    double x[kSub];
#pragma unroll
                    for (int i = 0; i < kSub; ++i)
                    {
                        x[i] = t[kSub * b + i][lane];
                    }
#pragma unroll
                    for (int c = 0; c < kSub; ++c)
                    {
                        const int pc = kSub * b + c; // tile row of this pivot step (compile-time)
                        double piv = lane_value(x[c], pc);
                        if (!(piv > 0.0))
                        {
                            first_bad = (first_bad == kT) ? pc : first_bad;
                            piv = 1.0; // keep the arithmetic finite; the factorisation is reported as failed
                        }
                        // 1 / sqrt(piv): the hardware estimate and two Newton steps y (3 - piv y^2) / 2 (full FP64
                        // precision), then r = piv / sqrt(piv) and the pivot row scaled by a product: a shorter
                        // dependent chain per pivot than a square root and a division
                        double inv = __builtin_amdgcn_rsq(piv);
                        inv = inv * fma(-0.5 * piv * inv, inv, 1.5);
                        inv = inv * fma(-0.5 * piv * inv, inv, 1.5);
                        const double r = piv * inv;
                        const double u = x[c] * inv; // R[pc][lane] for lane > pc
                        x[c] = (lane == pc) ? r : u;
#pragma unroll
                        for (int i = c + 1; i < kSub; ++i)
                        {
                            x[i] = fma(-lane_value(x[c], kSub * b + i), x[c], x[i]);
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kSub; ++i)
                    {
                        t[kSub * b + i][lane] = x[i];
                    }
                }
                __syncthreads();
                // trailing sub-blocks (i, j), b < i <= j, in a fixed order, wavefront p % 4 takes pair p
                int p = 0;
#pragma unroll
                for (int i = b + 1; i < kT / kSub; ++i)
                {
#pragma unroll
                    for (int j = i; j < kT / kSub; ++j, ++p)
                    {
                        if (w == p % kDiagWaves)
                        {
                            f64x4c acc = f64x4c{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                            for (int s = 0; s < kSub / 4; ++s)
                            {
                                const int kr = kSub * b + 4 * s + lane / 16; // a row of R_b
                                const double av = t[kr][kSub * i + lane % 16];   // R_bi^T (row lane % 16, k)
                                const double bv = t[kr][kSub * j + lane % 16];   // R_bj (k, col lane % 16)
                                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                            }
#pragma unroll
                            for (int q = 0; q < 4; ++q)
                            {
                                t[kSub * i + lane / 16 + 4 * q][kSub * j + lane % 16] -= acc[q];
                            }
                        }
                    }
                }
                __syncthreads();
            }
            // write back the upper triangle (row <= column)
#pragma unroll
            for (int q = 0; q < kLoads; ++q)
            {
                const int e = tid + q * kT * kDiagWaves;
                const int row = e % kT, col = e / kT;
                if (row <= col)
                {
                    tile[static_cast<int64_t>(col) * ld + row] = t[row][col];
                }
            }
            if (lane == 0 && first_bad < kT)
            {
                atomicMin(&tile_first_bad, first_bad);
            }
            __syncthreads();
            // tiles are factored in order, so the first tile to fail sets info: 1 + its first bad row
            if (tid == 0 && tile_first_bad < kT)
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
    } // namespace chol
} // namespace jsvd
