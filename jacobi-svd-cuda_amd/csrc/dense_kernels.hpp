// dense_kernels.hpp -- FP64 dense kernels of the mixed path (device code; dense.hip launches them): the products
// and the triangular solve of the QR-first preconditioning (column-scaled shifted Cholesky QR: B^T B, B R^-1,
// R_pass R_acc), the Newton-Schulz re-orthogonalisation of V (V^T V, 1.5 V - 0.5 V G), X = M V and U = Q V_M.
//
// gemm_kernel: D = alpha op(A) op(B) + beta C on v_mfma_f64_16x16x4f64.  Column-major operands (C and D optionally
// row-major).  A workgroup of 256 threads computes a 64 x 64 output tile; op(A) and op(B) are staged through LDS
// in k-chunks of 16 (two buffers: the next chunk's global loads are in flight while the current one feeds the
// MFMAs, one barrier per chunk); wavefront w owns the 32 x 32 quarter (w / 2, w % 2) as 2 x 2 MFMA tiles.  Flags
// restrict the work to what the structure needs: only tiles on or above the diagonal (symmetric products, whose
// lower triangle is either not read or mirrored from the upper one), upper-triangular operands (k-ranges of a
// tile shrink to where both are non-zero, entries below the diagonal read as zero, whatever the buffer holds
// there), a zero lower triangle of the result.  Operands that are 16-byte aligned with even leading dimensions are
// read as 16-byte vectors, out-of-range elements zeroed by selection rather than by branches.  Small tile grids
// split k over several workgroups, each writing its partial tile to scratch; gemm_reduce_kernel then sums the
// partials of a tile in split order (no fence or counter inside the product kernel: a device-scope fence per
// workgroup writes the L2 back and costs more than the second launch).
//
// trsm_panel_kernel: B <- B R^-1 for upper-triangular R by block substitution over 64-column blocks, one workgroup
// per 16-row panel of B held in registers for the whole solve (wavefront w holds the 16-column slice w of every
// block).  Step J: X_J = B_J inv(R_JJ) with the pre-inverted 64 x 64 diagonal block (the slices exchanged through
// LDS), stored and published in LDS; every wavefront then subtracts X_J R_J,q from its slices of the blocks q > J.
// The products are formed transposed (B^T <- inv(R)^T B^T), which puts a panel row on the lanes of an MFMA output
// column (128-byte row segments of the column-major B), and the MFMA k index is mapped so that the 16 values a
// lane takes from a column of R are contiguous.
#pragma once

#include "common.hpp"
#include "dense.hpp"

namespace jsvd
{
    namespace dense
    {
        constexpr int kTile = 64;      // output tile edge of gemm_kernel
        constexpr int kChunk = 16;     // k-chunk staged in LDS
        constexpr int kLdsStride = 65; // LDS row stride (doubles) of a staged chunk
        constexpr int kRing = 4;       // k-chunks whose global loads are in flight (register sets; even)

        struct GemmParams
        {
            int64_t M, N, K;
            const double *A;
            int64_t lda;
            const double *B;
            int64_t ldb;
            const double *C; // read only when beta != 0
            int64_t ldc;
            double *D;
            int64_t ldd;
            double alpha, beta;
            int flags;
            int tiles_m, tiles_n; // tile grid of the output
            int tiles;            // tiles launched (upper-triangle tiles with kUpperOut)
            int splits;           // k-splits (gridDim.y)
            double *work;         // splits * tiles * 4096 partials when splits > 1
            bool vec;             // A and B 16-byte aligned with even leading dimensions: 16-byte loads
        };

        typedef dev::vec_t<double, 4> f64x4;

        constexpr int64_t kMaxSplits = 8;

        // Launch parameters of one product (host side): tile grid, and the k-split -- about two workgroups per CU
        // when the tile grid alone leaves the device idle, at least 8 chunks per split, within the workspace.  No
        // split for triangular operands or a zeroed lower triangle.
        inline GemmParams plan_gemm(int64_t m, int64_t n, int64_t k, double alpha, const double *a, int64_t lda,
                                    const double *b, int64_t ldb, double beta, const double *c, int64_t ldc, double *d,
                                    int64_t ldd, int flags, int cus, double *work, int64_t work_elems)
        {
            GemmParams p{};
            p.M = m;
            p.N = n;
            p.K = k > 0 ? k : 0;
            p.A = a;
            p.lda = lda;
            p.B = b;
            p.ldb = ldb;
            p.C = c;
            p.ldc = ldc;
            p.D = d;
            p.ldd = ldd;
            p.alpha = alpha;
            p.beta = beta;
            p.flags = flags;
            p.tiles_m = static_cast<int>((m + kTile - 1) / kTile);
            p.tiles_n = static_cast<int>((n + kTile - 1) / kTile);
            p.tiles = (flags & kUpperOut) ? p.tiles_n * (p.tiles_n + 1) / 2 : p.tiles_m * p.tiles_n;
            int64_t splits = 1;
            if (!(flags & (kZeroLower | kTriA | kTriB)) && work)
            {
                const int64_t chunks = (p.K + kChunk - 1) / kChunk;
                splits = (2 * static_cast<int64_t>(cus) + p.tiles - 1) / p.tiles;
                splits = splits < chunks / 8 ? splits : (chunks / 8 > 1 ? chunks / 8 : 1);
                splits = splits < kMaxSplits ? splits : kMaxSplits;
                while (splits > 1 && splits * p.tiles * 16 * 256 > work_elems)
                {
                    --splits;
                }
            }
            p.splits = static_cast<int>(splits);
            p.work = work;
            p.vec = (reinterpret_cast<uintptr_t>(a) % 16) == 0 && (reinterpret_cast<uintptr_t>(b) % 16) == 0 &&
                    (lda % 2) == 0 && (ldb % 2) == 0;
            return p;
        }


        // Output tile (ti, tj) of launched tile t.
        __device__ inline void tile_of(const GemmParams &p, int t, int &ti, int &tj)
        {
            if (p.flags & kUpperOut)
            {
                int row = 0;
                while (t >= p.tiles_n - row)
                {
                    t -= p.tiles_n - row;
                    ++row;
                }
                ti = row;
                tj = row + t;
            }
            else
            {
                ti = t / p.tiles_n;
                tj = t % p.tiles_n;
            }
        }

        template <bool TA, bool TB>
        __device__ inline void load_chunk(const GemmParams &p, int64_t m0, int64_t n0, int64_t kc, int64_t k_hi, int tid,
                                          double (&ra)[4], double (&rb)[4])
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                // op(A)(m, k)
                const int64_t am = TA ? m0 + tid / 16 + 16 * q : m0 + tid % 64;
                const int64_t ak = TA ? kc + tid % 16 : kc + tid / 64 + 4 * q;
                bool ok = am < p.M && ak < k_hi;
                if (p.flags & kTriA)
                {
                    ok = ok && am <= ak;
                }
                ra[q] = ok ? (TA ? p.A[ak + am * p.lda] : p.A[am + ak * p.lda]) : 0.0;
                // op(B)(k, n)
                const int64_t bk = TB ? kc + tid / 64 + 4 * q : kc + tid % 16;
                const int64_t bn = TB ? n0 + tid % 64 : n0 + tid / 16 + 16 * q;
                ok = bn < p.N && bk < k_hi;
                if (p.flags & kTriB)
                {
                    ok = ok && bk <= bn;
                }
                rb[q] = ok ? (TB ? p.B[bn + bk * p.ldb] : p.B[bk + bn * p.ldb]) : 0.0;
            }
        }

        template <bool TA, bool TB>
        __device__ inline void store_chunk(double *as, double *bs, int tid, const double (&ra)[4],
                                           const double (&rb)[4])
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int am = TA ? tid / 16 + 16 * q : tid % 64;
                const int ak = TA ? tid % 16 : tid / 64 + 4 * q;
                as[ak * kLdsStride + am] = ra[q];
                const int bk = TB ? tid / 64 + 4 * q : tid % 16;
                const int bn = TB ? tid % 64 : tid / 16 + 16 * q;
                bs[bk * kLdsStride + bn] = rb[q];
            }
        }

        typedef dev::vec_t<double, 2> f64x2;

        // One operand's k-chunk with 16-byte loads: rows r0 .. r0 + 63 (m for A, n for B) by k = kc .. kc + 15.
        // RCONTIG: the stored matrix has r as its fast index (x[r + k ld]), else k (x[k + r ld]).  TRI: 1 = entries
        // with r > k are zero (an upper-triangular op(A)), 2 = entries with k > r are zero (an upper-triangular
        // op(B)).  Requires x 16-byte aligned and ld even; an element outside rows < R, k < k_hi reads as zero and
        // its load address is clamped into the matrix (a 16-byte pair never leaves its column: r or k even, ld even).
        template <bool RCONTIG>
        __device__ inline void load_vec(const double *x, int64_t ld, int64_t r0, int64_t rows, int64_t kc, int64_t k_hi,
                                        int tri, int tid, double (&out)[4])
        {
#pragma unroll
            for (int q = 0; q < 2; ++q)
            {
                const int64_t r = RCONTIG ? r0 + 2 * (tid % 32) : r0 + tid / 8 + 32 * q;
                const int64_t k = RCONTIG ? kc + tid / 32 + 8 * q : kc + 2 * (tid % 8);
                // second element of the pair: (r + 1, k) or (r, k + 1)
                const int64_t r2 = RCONTIG ? r + 1 : r;
                const int64_t k2 = RCONTIG ? k : k + 1;
                bool ok1 = r < rows && k < k_hi;
                bool ok2 = r2 < rows && k2 < k_hi;
                if (tri == 1)
                {
                    ok1 = ok1 && r <= k;
                    ok2 = ok2 && r2 <= k2;
                }
                else if (tri == 2)
                {
                    ok1 = ok1 && k <= r;
                    ok2 = ok2 && k2 <= r2;
                }
                const int64_t rc = (r < rows) ? r : (r0 < rows ? r0 : 0);
                const int64_t kk = (k < k_hi) ? k : kc;
                const f64x2 v = *reinterpret_cast<const f64x2 *>(RCONTIG ? x + rc + kk * ld : x + kk + rc * ld);
                out[2 * q] = ok1 ? v[0] : 0.0;
                out[2 * q + 1] = ok2 ? v[1] : 0.0;
            }
        }

        template <bool RCONTIG>
        __device__ inline void store_vec(double *lds, int tid, const double (&in)[4])
        {
#pragma unroll
            for (int q = 0; q < 2; ++q)
            {
                const int r = RCONTIG ? 2 * (tid % 32) : tid / 8 + 32 * q;
                const int k = RCONTIG ? tid / 32 + 8 * q : 2 * (tid % 8);
                lds[k * kLdsStride + r] = in[2 * q];
                if (RCONTIG)
                {
                    lds[k * kLdsStride + r + 1] = in[2 * q + 1];
                }
                else
                {
                    lds[(k + 1) * kLdsStride + r] = in[2 * q + 1];
                }
            }
        }

        template <bool TA, bool TB>
        __device__ inline void load_chunk_any(const GemmParams &p, int64_t m0, int64_t n0, int64_t kc, int64_t k_hi,
                                              int tid, double (&ra)[4], double (&rb)[4])
        {
            if (p.vec)
            {
                load_vec<!TA>(p.A, p.lda, m0, p.M, kc, k_hi, (p.flags & kTriA) ? 1 : 0, tid, ra);
                load_vec<TB>(p.B, p.ldb, n0, p.N, kc, k_hi, (p.flags & kTriB) ? 2 : 0, tid, rb);
            }
            else
            {
                load_chunk<TA, TB>(p, m0, n0, kc, k_hi, tid, ra, rb);
            }
        }

        template <bool TA, bool TB>
        __device__ inline void store_chunk_any(const GemmParams &p, double *as, double *bs, int tid,
                                               const double (&ra)[4], const double (&rb)[4])
        {
            if (p.vec)
            {
                store_vec<!TA>(as, tid, ra);
                store_vec<TB>(bs, tid, rb);
            }
            else
            {
                store_chunk<TA, TB>(as, bs, tid, ra, rb);
            }
        }

        __device__ inline void write_out(const GemmParams &p, int64_t i, int64_t j, double acc, bool mirror)
        {
            if (i >= p.M || j >= p.N)
            {
                return;
            }
            const bool rm = (p.flags & kRowMajorCD) != 0;
            double val = p.alpha * acc;
            if (p.beta != 0.0)
            {
                val = fma(p.beta, rm ? p.C[i * p.ldc + j] : p.C[i + j * p.ldc], val);
            }
            if ((p.flags & kZeroLower) && i > j)
            {
                val = 0.0;
            }
            if (rm)
            {
                p.D[i * p.ldd + j] = val;
            }
            else
            {
                p.D[i + j * p.ldd] = val;
                if (mirror)
                {
                    p.D[j + i * p.ldd] = val;
                }
            }
        }

        template <bool TA, bool TB>
        __global__ __launch_bounds__(256) void gemm_kernel(GemmParams p)
        {
            __shared__ double as[2][kChunk * kLdsStride];
            __shared__ double bs[2][kChunk * kLdsStride];
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            int ti = 0, tj = 0;
            tile_of(p, static_cast<int>(blockIdx.x), ti, tj);
            const int64_t m0 = static_cast<int64_t>(ti) * kTile;
            const int64_t n0 = static_cast<int64_t>(tj) * kTile;
            const int wr = 32 * (wave / 2);
            const int wc = 32 * (wave % 2);
            if ((p.flags & kZeroLower) && ti > tj)
            {
                // strictly below the diagonal: zeros, nothing to multiply
                for (int e = tid; e < kTile * kTile; e += 256)
                {
                    write_out(p, m0 + e % kTile, n0 + e / kTile, 0.0, false);
                }
                return;
            }
            // k-range of this tile where both operands can be non-zero, then this workgroup's split of it
            int64_t k_lo = 0, k_hi = p.K;
            if (p.flags & kTriA)
            {
                k_lo = m0;
            }
            if (p.flags & kTriB)
            {
                k_hi = min(k_hi, n0 + kTile);
            }
            const int64_t chunks = (k_hi > k_lo) ? (k_hi - k_lo + kChunk - 1) / kChunk : 0;
            const int split = static_cast<int>(blockIdx.y);
            const int64_t c_begin = chunks * split / p.splits;
            const int64_t c_end = chunks * (split + 1) / p.splits;

            f64x4 acc[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
            {
#pragma unroll
                for (int b = 0; b < 2; ++b)
                {
                    acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
                }
            }
            // A ring of kRing register sets keeps the global loads of chunks c + 1 .. c + kRing - 1 in flight while
            // chunk c feeds the MFMAs: at step c, set (c + 1) % kRing is written to the other LDS buffer (its loads
            // were issued kRing - 1 steps earlier) and set c % kRing, stored to LDS one step ago, takes the loads of
            // chunk c + kRing.  The ring slot is a compile-time index (the step loop is unrolled by kRing).
            double ra[kRing][4], rb[kRing][4];
#pragma unroll
            for (int u = 0; u < kRing; ++u)
            {
                if (c_begin + u < c_end)
                {
                    load_chunk_any<TA, TB>(p, m0, n0, k_lo + (c_begin + u) * kChunk, k_hi, tid, ra[u], rb[u]);
                }
            }
            if (c_begin < c_end)
            {
                store_chunk_any<TA, TB>(p, as[0], bs[0], tid, ra[0], rb[0]);
            }
            __syncthreads();
            for (int64_t c0 = c_begin; c0 < c_end; c0 += kRing)
            {
#pragma unroll
                for (int u = 0; u < kRing; ++u)
                {
                    const int64_t c = c0 + u;
                    if (c < c_end)
                    {
                        const int buf = u & 1; // (c - c_begin) & 1, kRing being even
                        const double *a_s = as[buf];
                        const double *b_s = bs[buf];
#pragma unroll
                        for (int s = 0; s < kChunk / 4; ++s)
                        {
                            const int k = 4 * s + lane / 16;
                            const double a0 = a_s[k * kLdsStride + wr + lane % 16];
                            const double a1 = a_s[k * kLdsStride + wr + 16 + lane % 16];
                            const double b0 = b_s[k * kLdsStride + wc + lane % 16];
                            const double b1 = b_s[k * kLdsStride + wc + 16 + lane % 16];
                            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
                            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
                            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
                            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
                        }
                        if (c + 1 < c_end)
                        {
                            const int next = (u + 1) % kRing; // a constant once the loop is unrolled
                            store_chunk_any<TA, TB>(p, as[buf ^ 1], bs[buf ^ 1], tid, ra[next], rb[next]);
                        }
                        if (c + kRing < c_end)
                        {
                            load_chunk_any<TA, TB>(p, m0, n0, k_lo + (c + kRing) * kChunk, k_hi, tid, ra[u], rb[u]);
                        }
                        __syncthreads();
                    }
                }
            }

            if (p.splits > 1)
            {
                // this split's partial tile, one coalesced row of 256 per value; gemm_reduce_kernel finishes the tile
                const int64_t base = (static_cast<int64_t>(split) * p.tiles + blockIdx.x) * 16;
#pragma unroll
                for (int a = 0; a < 2; ++a)
                {
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                    {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            p.work[(base + 8 * a + 4 * b + v) * 256 + tid] = acc[a][b][v];
                        }
                    }
                }
                return;
            }

            const bool mirror = (p.flags & kMirror) && ti != tj;
#pragma unroll
            for (int a = 0; a < 2; ++a)
            {
#pragma unroll
                for (int b = 0; b < 2; ++b)
                {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                    {
                        const int64_t i = m0 + wr + 16 * a + lane / 16 + 4 * v;
                        const int64_t j = n0 + wc + 16 * b + lane % 16;
                        write_out(p, i, j, acc[a][b][v], mirror);
                    }
                }
            }
        }

        // Sums the split partials of each tile in split order and writes the tile as gemm_kernel's epilogue does
        // (same thread-to-entry map).  Grid: one workgroup of 256 per tile.
        __global__ __launch_bounds__(256) void gemm_reduce_kernel(GemmParams p)
        {
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            int ti = 0, tj = 0;
            tile_of(p, static_cast<int>(blockIdx.x), ti, tj);
            const int64_t m0 = static_cast<int64_t>(ti) * kTile;
            const int64_t n0 = static_cast<int64_t>(tj) * kTile;
            const int wr = 32 * (wave / 2);
            const int wc = 32 * (wave % 2);
            const bool mirror = (p.flags & kMirror) && ti != tj;
            // split by split, the 16 partial values of a thread loaded together (in flight at once), summed in split
            // order
            double sum[16];
            for (int s = 0; s < p.splits; ++s)
            {
                const double *w = p.work + ((static_cast<int64_t>(s) * p.tiles + blockIdx.x) * 16) * 256 + tid;
                double x[16];
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    x[r] = w[r * 256];
                }
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    sum[r] = (s == 0) ? x[r] : sum[r] + x[r];
                }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
            {
#pragma unroll
                for (int b = 0; b < 2; ++b)
                {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                    {
                        const int64_t i = m0 + wr + 16 * a + lane / 16 + 4 * v;
                        const int64_t j = n0 + wc + 16 * b + lane % 16;
                        write_out(p, i, j, sum[8 * a + 4 * b + v], mirror);
                    }
                }
            }
        }

        constexpr int kSolveBlock = 64;  // columns per substitution step of the triangular solve
        constexpr int kPanelRows = 16;   // rows of B per workgroup
        constexpr int kPanelSteps = 16;  // 64-column blocks per panel launch (1024 columns)

        // rp = R padded for the triangular solve: rp(k, j) = R(k, j) for k <= j < n (R column-major, ldr; its
        // strictly lower triangle is not read), the identity outside n x n up to np (a multiple of 64), zero below
        // the diagonal; column-major with leading dimension np.
        __global__ __launch_bounds__(256) void pad_upper_kernel(const double *__restrict__ r, int64_t ldr, int64_t n,
                                                                int64_t np, double *__restrict__ rp)
        {
            const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
            if (e >= np * np)
            {
                return;
            }
            const int64_t k = e % np, j = e / np;
            double x = (k == j) ? 1.0 : 0.0;
            if (k < n && j < n)
            {
                x = (k <= j) ? r[k + j * ldr] : 0.0;
            }
            rp[e] = x;
        }

        // inv(R_bb) of every 64 x 64 diagonal block b of the padded upper-triangular R (column-major, ld), stored
        // transposed: dinvt[4096 b + 64 j + k] = inv(R_bb)(k, j).  Thread j solves R_bb x = e_j by back substitution
        // with x in registers; the rows of R_bb are LDS broadcasts.
        __global__ __launch_bounds__(64) void diag_inverse_kernel(const double *__restrict__ rp, int64_t ld,
                                                                  double *__restrict__ dinvt)
        {
            __shared__ double rb[64][65]; // rb[k][j] = R_bb(k, j)
            const int j = threadIdx.x;
            const int64_t d0 = static_cast<int64_t>(blockIdx.x) * 64;
            double x[64];
            // column j of the block: thread j reads row k of it (R_bb(k, j)); all 64 loads in flight before the
            // first wait (the padded R holds the whole 64 x 64 block, so the loads below the diagonal stay inside it)
#pragma unroll
            for (int k = 0; k < 64; ++k)
            {
                x[k] = rp[(d0 + k) + (d0 + j) * ld];
            }
#pragma unroll
            for (int k = 0; k < 64; ++k)
            {
                rb[k][j] = (k <= j) ? x[k] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int i = 63; i >= 0; --i)
            {
                // the dot product over k > i as four interleaved partial sums (four short dependency chains)
                double s4[4] = {(i == j) ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k = i + 1; k < 64; ++k)
                {
                    s4[k & 3] = fma(-rb[i][k], x[k], s4[k & 3]);
                }
                x[i] = (i <= j) ? ((s4[0] + s4[1]) + (s4[2] + s4[3])) / rb[i][i] : 0.0;
            }
            double *out = dinvt + static_cast<int64_t>(blockIdx.x) * 4096 + static_cast<int64_t>(j) * 64;
#pragma unroll
            for (int k = 0; k < 64; ++k)
            {
                out[k] = x[k];
            }
        }

        // Columns c0 .. c0 + 64 nblocks - 1 of B (column-major, ldb; `rows` x `cols` valid, columns at or past cols
        // read as zero and are not written) <- B R^-1 restricted to those columns, given that B's columns before c0
        // are final and their contribution is already subtracted.  rp: the padded R (column-major, ldp); dinvt: the
        // transposed inverses of its 64 x 64 diagonal blocks.  c0 is a multiple of 64, nblocks <= kPanelSteps.
        //
        // One workgroup per 16-row panel; wavefront w holds the 16-column slice w of every 64-column block,
        // transposed in MFMA output layout: acc[q][v] = B(r0 + lane % 16, c0 + 64 q + 16 w + lane / 16 + 4 v).
        // MFMA k map: in step s, the lanes of group g = lane / 16 carry k = 16 g + s of the 64-wide block, so the 16
        // operand values a lane takes from one column of R (or of inv(R_JJ)^T) are contiguous in memory.
        __global__ __launch_bounds__(256) void trsm_panel_kernel(double *__restrict__ b, int64_t ldb, int64_t rows,
                                                                 int64_t cols, int64_t c0, int nblocks,
                                                                 const double *__restrict__ rp, int64_t ldp,
                                                                 const double *__restrict__ dinvt)
        {
            __shared__ double bt[kSolveBlock][kPanelRows + 1]; // B~_J as [k][row]
            __shared__ double xt[kSolveBlock][kPanelRows + 1]; // X_J as [k][row]
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            const int g = lane / 16;
            const int i = lane % 16;
            const int64_t row = static_cast<int64_t>(blockIdx.x) * kPanelRows + i;
            const bool row_ok = row < rows;
            f64x4 acc[kPanelSteps];
#pragma unroll
            for (int q = 0; q < kPanelSteps; ++q)
            {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int64_t col = c0 + kSolveBlock * q + 16 * wave + g + 4 * v;
                    acc[q][v] = (q < nblocks && row_ok && col < cols) ? b[row + col * ldb] : 0.0;
                }
            }
            for (int J = 0; J < nblocks; ++J)
            {
                // publish this wavefront's slice of B~_J
#pragma unroll
                for (int q = 0; q < kPanelSteps; ++q)
                {
                    if (q == J)
                    {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            bt[16 * wave + g + 4 * v][i] = acc[q][v];
                        }
                    }
                }
                __syncthreads();
                // X_J^T slice w: A operand (col' = lane % 16, k) = inv(R_JJ)(k, 16 w + col'), B operand (k, row) =
                // B~_J(row, k), k = 16 g + s
                const double *dj = dinvt + (c0 / kSolveBlock + J) * 4096 + (16 * wave + i) * 64 + 16 * g;
                f64x4 x = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s < 16; ++s)
                {
                    x = __builtin_amdgcn_mfma_f64_16x16x4f64(dj[s], bt[16 * g + s][i], x, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    xt[16 * wave + g + 4 * v][i] = x[v];
                    const int64_t col = c0 + kSolveBlock * J + 16 * wave + g + 4 * v;
                    if (row_ok && col < cols)
                    {
                        b[row + col * ldb] = x[v];
                    }
                }
                __syncthreads();
                // B_q^T slice w -= R_Jq^T X_J^T for the blocks q > J: A operand (col', k) = -R(c0 + 64 J + k, c0 + 64 q +
                // 16 w + col'), B operand (k, row) = X_J(row, k)
                double xk[16];
#pragma unroll
                for (int s = 0; s < 16; ++s)
                {
                    xk[s] = xt[16 * g + s][i];
                }
                const double *rj = rp + (c0 + kSolveBlock * J + 16 * g) + (c0 + 16 * wave + i) * ldp;
#pragma unroll
                for (int q = 0; q < kPanelSteps; ++q)
                {
                    if (q > J && q < nblocks)
                    {
                        const double *rq = rj + static_cast<int64_t>(kSolveBlock) * q * ldp;
#pragma unroll
                        for (int s = 0; s < 16; ++s)
                        {
                            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-rq[s], xk[s], acc[q], 0, 0, 0);
                        }
                    }
                }
            }
        }

        // Host-side schedule of the triangular solve over n columns (padded to 64): panel(c0, nblocks) solves a
        // column range of up to 1024 with trsm_panel_kernel, update(c0, width, rest) subtracts the solved range's
        // contribution from the `rest` columns after it (a product with R's rows c0 .. c0 + width - 1).
        template <typename Panel, typename Update>
        inline void trsm_schedule(int64_t n, Panel &&panel, Update &&update)
        {
            const int64_t np = (n + kSolveBlock - 1) / kSolveBlock * kSolveBlock;
            constexpr int64_t kRange = kSolveBlock * kPanelSteps;
            for (int64_t c0 = 0; c0 < np; c0 += kRange)
            {
                const int64_t width = (np - c0 < kRange) ? np - c0 : kRange;
                panel(c0, static_cast<int>(width / kSolveBlock));
                const int64_t rest = n - (c0 + width);
                if (rest > 0)
                {
                    update(c0, width, rest);
                }
            }
        }
    } // namespace dense
} // namespace jsvd
