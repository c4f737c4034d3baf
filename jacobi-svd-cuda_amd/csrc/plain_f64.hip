// plain_f64.hip -- the plain FP64 one-sided Jacobi kernels for gfx950 (CDNA4, wave64).
//
//   transpose_tile_kernel     LDS-tiled out-of-place transpose (row-major <-> column-major copies of A,
//                             final V), optionally dividing by sigma on the fly (U extraction)
//   column_norms_kernel       sigma_j = sqrt(sum_i a_ij^2), one wavefront per column
//   round_kernel              one fused round: column-pair Gram (alpha, beta, gamma) with wavefront
//                             shuffles, FP64 rotation parameters with the reference skip test, Givens
//                             apply to the A and V column pairs, rotation count and off-orthogonality
//   pair_gram_kernel          stand-alone column-pair Gram (kernel-level entry point)
//   offnorm_*                 two-stage max |a_i.a_j|/(|a_i||a_j|)
//
// Layout: A is column-major (m rows, leading dimension lda, lda even so every column starts on a
// 16-byte boundary and rows move as double2), V is column-major (n rows, ldv even).  Rows m..lda-1 are
// workspace padding and are read as zero.
#include "plain_f64.hpp"
#include "plain_kernels.hpp"

#include <algorithm>

namespace jsvd
{
    namespace
    {
        using plain::block_sum3;
        using plain::round_kernel;

        constexpr int kTile = 64;      // transpose tile edge (doubles)
        constexpr int kTileRows = 4;   // block is kTile x kTileRows threads

        template <bool Scale>
        __global__ __launch_bounds__(kTile *kTileRows) void transpose_tile_kernel(
            const double *__restrict__ src, int64_t lds, double *__restrict__ dst, int64_t ldd, int64_t rows,
            int64_t cols, int64_t tiles_x, const double *__restrict__ row_scale, double eps)
        {
            __shared__ double tile[kTile][kTile + 1];
            const int64_t tid_x = threadIdx.x;
            const int64_t tid_y = threadIdx.y;
            const int64_t bx = static_cast<int64_t>(blockIdx.x) % tiles_x;
            const int64_t by = static_cast<int64_t>(blockIdx.x) / tiles_x;
            const int64_t c = bx * kTile + tid_x;
            // every load of the tile in flight before the first wait: unconditional loads (element 0 outside the
            // matrix), the bounds applied to the LDS stores
            constexpr int kPer = kTile / kTileRows;
            double x[kPer];
            double sr[Scale ? kPer : 1];
#pragma unroll
            for (int j = 0; j < kPer; ++j)
            {
                const int64_t r = by * kTile + tid_y + j * kTileRows;
                const bool in = r < rows && c < cols;
                x[j] = src[in ? r * lds + c : 0];
                if constexpr (Scale)
                {
                    sr[j] = row_scale[in ? r : 0];
                }
            }
#pragma unroll
            for (int j = 0; j < kPer; ++j)
            {
                const int64_t r = by * kTile + tid_y + j * kTileRows;
                if (r < rows && c < cols)
                {
                    double v = x[j];
                    if constexpr (Scale)
                    {
                        v = (sr[j] > eps) ? (v / sr[j]) : 0.0;
                    }
                    tile[tid_y + j * kTileRows][tid_x] = v;
                }
            }
            __syncthreads();
            const int64_t oc = by * kTile + tid_x;
#pragma unroll
            for (int j = 0; j < kTile; j += kTileRows)
            {
                const int64_t orow = bx * kTile + tid_y + j;
                if (orow < cols && oc < rows)
                {
                    dst[orow * ldd + oc] = tile[tid_x][tid_y + j];
                }
            }
        }

        // One wavefront per column: sigma_j = sqrt(max(sum a_ij^2, 0)).  Optionally also writes the squared
        // norm (used by the off-norm reduction).
        __global__ __launch_bounds__(256) void column_norms_kernel(const double *__restrict__ a, int64_t lda,
                                                                   int64_t m, int64_t n, double *__restrict__ sigma,
                                                                   double *__restrict__ norm2)
        {
            const int64_t col = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
            const int lane = threadIdx.x % 64;
            if (col >= n)
            {
                return;
            }
            const double *cj = a + col * lda;
            double acc = 0.0;
            for (int64_t i = lane; i < m; i += 64)
            {
                const double x = cj[i];
                acc += x * x;
            }
            acc = dev::wave_sum(acc);
            if (lane == 0)
            {
                if (sigma)
                {
                    sigma[col] = sqrt(fmax(acc, 0.0));
                }
                if (norm2)
                {
                    norm2[col] = acc;
                }
            }
        }

        __global__ void identity_kernel(double *__restrict__ v, int64_t ldv, int64_t n)
        {
            const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            const int64_t total = ldv * n;
            if (idx >= total)
            {
                return;
            }
            const int64_t col = idx / ldv;
            const int64_t row = idx % ldv;
            v[idx] = (row == col) ? 1.0 : 0.0;
        }

        // Folds the counter slots of one sweep: one 64-lane wave, slots <= 64 x 64.
        __global__ __launch_bounds__(64) void sweep_totals_kernel(const int *__restrict__ rot_slots,
                                                                  const unsigned long long *__restrict__ off_slots,
                                                                  int slots, int *__restrict__ rot_total,
                                                                  unsigned long long *__restrict__ off_total)
        {
            int count = 0;
            unsigned long long best = 0; // non-negative doubles order like their bit patterns
            for (int k = threadIdx.x; k < slots; k += 64)
            {
                count += rot_slots[k];
                best = max(best, off_slots[k]);
            }
#pragma unroll
            for (int offset = 32; offset > 0; offset >>= 1)
            {
                count += __shfl_xor(count, offset, 64);
                best = max(best, static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(best), offset, 64)));
            }
            if (threadIdx.x == 0)
            {
                *rot_total = count;
                *off_total = best;
            }
        }

        __global__ __launch_bounds__(256) void pair_gram_kernel(const double *__restrict__ a, int64_t lda,
                                                                int64_t m, const int *__restrict__ pairs,
                                                                double *__restrict__ alpha,
                                                                double *__restrict__ beta,
                                                                double *__restrict__ gamma)
        {
            __shared__ double scratch[4][3];
            const int64_t k = blockIdx.x;
            const int p = pairs[2 * k];
            const int q = pairs[2 * k + 1];
            const double *cp = a + static_cast<int64_t>(p) * lda;
            const double *cq = a + static_cast<int64_t>(q) * lda;
            double spp = 0.0, sqq = 0.0, spq = 0.0;
            for (int64_t i = threadIdx.x; i < m; i += 256)
            {
                const double x = cp[i];
                const double y = cq[i];
                spp += x * x;
                sqq += y * y;
                spq += x * y;
            }
            block_sum3<256>(spp, sqq, spq, scratch);
            if (threadIdx.x == 0)
            {
                alpha[k] = spp;
                beta[k] = sqq;
                gamma[k] = spq;
            }
        }

        __global__ __launch_bounds__(256) void rotation_params_kernel(const double *__restrict__ alpha,
                                                                      const double *__restrict__ beta,
                                                                      const double *__restrict__ gamma,
                                                                      int64_t count, double eps,
                                                                      double *__restrict__ c, double *__restrict__ s,
                                                                      int *__restrict__ rotate)
        {
            const int64_t k = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
            if (k >= count)
            {
                return;
            }
            double ck, sk;
            rotate[k] = dev::rotation_params(alpha[k], beta[k], gamma[k], eps, ck, sk) ? 1 : 0;
            c[k] = ck;
            s[k] = sk;
        }

        // Off-norm stage 1: one 32 x 32 column tile (I, J) of the Gram matrix per workgroup, rows streamed
        // through LDS in chunks of 32; each workgroup writes the max cosine of its strictly-upper entries.
        constexpr int kOffTile = 32;
        __global__ __launch_bounds__(256) void offnorm_stage1_kernel(const double *__restrict__ a, int64_t lda,
                                                                     int64_t m, int64_t n,
                                                                     const double *__restrict__ norm2,
                                                                     int64_t tiles, double *__restrict__ partial)
        {
            __shared__ double ti[kOffTile][kOffTile + 1];
            __shared__ double tj[kOffTile][kOffTile + 1];
            __shared__ double red[4];
            // Map the linear block id onto the upper triangle (I <= J) of the tile grid.
            int64_t b = blockIdx.x;
            int64_t I = 0;
            while (b >= tiles - I)
            {
                b -= tiles - I;
                ++I;
            }
            const int64_t J = I + b;
            const int tid = threadIdx.x;
            const int li = tid % kOffTile;       // local column in tile I
            const int lj0 = (tid / kOffTile) * 4; // four local columns in tile J
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int64_t r0 = 0; r0 < m; r0 += kOffTile)
            {
                for (int e = tid; e < kOffTile * kOffTile; e += 256)
                {
                    const int rr = e % kOffTile;
                    const int cc = e / kOffTile;
                    const int64_t row = r0 + rr;
                    const int64_t ci = I * kOffTile + cc;
                    const int64_t cj = J * kOffTile + cc;
                    ti[cc][rr] = (row < m && ci < n) ? a[ci * lda + row] : 0.0;
                    tj[cc][rr] = (row < m && cj < n) ? a[cj * lda + row] : 0.0;
                }
                __syncthreads();
#pragma unroll 4
                for (int rr = 0; rr < kOffTile; ++rr)
                {
                    const double x = ti[li][rr];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                    {
                        acc[t] += x * tj[lj0 + t][rr];
                    }
                }
                __syncthreads();
            }
            double best = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                const int64_t gi = I * kOffTile + li;
                const int64_t gj = J * kOffTile + lj0 + t;
                if (gi < n && gj < n && gi < gj)
                {
                    best = fmax(best, dev::pair_cosine(norm2[gi], norm2[gj], acc[t]));
                }
            }
            best = dev::wave_max(best);
            if ((tid % 64) == 0)
            {
                red[tid / 64] = best;
            }
            __syncthreads();
            if (tid == 0)
            {
                partial[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            }
        }

        // Off-norm stage 2: one workgroup reduces the per-tile partial maxima.
        __global__ __launch_bounds__(256) void offnorm_stage2_kernel(const double *__restrict__ partial,
                                                                     int64_t count, double *__restrict__ out)
        {
            __shared__ double red[4];
            double best = 0.0;
            for (int64_t i = threadIdx.x; i < count; i += 256)
            {
                best = fmax(best, partial[i]);
            }
            best = dev::wave_max(best);
            if ((threadIdx.x % 64) == 0)
            {
                red[threadIdx.x / 64] = best;
            }
            __syncthreads();
            if (threadIdx.x == 0)
            {
                out[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            }
        }

        int pow2_ceil(int64_t x)
        {
            int k = 1;
            while (k < x)
            {
                k *= 2;
            }
            return k;
        }

        using RoundFn = void (*)(double *, int64_t, int64_t, double *, int64_t, int64_t, int, int, double, int *,
                                 unsigned long long *, const int *, int);

        // Instantiated (EA, EV) pairs; EV <= EA because n <= m.  EV == 0 streams V.
        template <int T>
        RoundFn pick_round(int ea, int ev)
        {
            switch (ea)
            {
            case 0: return round_kernel<T, 0, 0>;
            case 1: return round_kernel<T, 1, 1>;
            case 2: return ev <= 1 ? round_kernel<T, 2, 1> : round_kernel<T, 2, 2>;
            case 4:
                if (ev == 0) return round_kernel<T, 4, 0>;
                if (ev == 1) return round_kernel<T, 4, 1>;
                if (ev == 2) return round_kernel<T, 4, 2>;
                return round_kernel<T, 4, 4>;
            default:
                if (ev == 0) return round_kernel<T, 8, 0>;
                if (ev == 1) return round_kernel<T, 8, 1>;
                if (ev == 2) return round_kernel<T, 8, 2>;
                if constexpr (T < 1024)
                {
                    if (ev == 4) return round_kernel<T, 8, 4>;
                    return round_kernel<T, 8, 8>;
                }
                return round_kernel<T, 8, 0>; // 128-VGPR budget of 1024-thread groups: stream V
            }
        }
    } // namespace

    RoundShape choose_round_shape(int64_t m, int64_t n, bool wide)
    {
        RoundShape shape{};
        const int64_t m2 = (m + 1) / 2;
        const int64_t n2 = (n + 1) / 2;
        if (m2 <= 64 * 8 && !wide)
        {
            shape.threads = 64;
        }
        else if (m2 <= 256 * 8)
        {
            shape.threads = 256;
        }
        else
        {
            shape.threads = 1024;
        }
        const int64_t per_a = (m2 + shape.threads - 1) / shape.threads;
        if (per_a > 8)
        {
            shape.ea = 0;
            shape.ev = 0;
            return shape;
        }
        shape.ea = pow2_ceil(per_a);
        const int64_t per_v = (n2 + shape.threads - 1) / shape.threads;
        shape.ev = std::min(shape.ea, pow2_ceil(std::max<int64_t>(per_v, 1)));
        // 1024-thread workgroups get 128 VGPRs per lane: with 8 A chunks only 2 V chunks still fit.
        if (shape.threads == 1024 && shape.ea == 8 && shape.ev > 2)
        {
            shape.ev = 0;
        }
        return shape;
    }

    void launch_round(const RoundShape &shape, double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n,
                      int round, double eps, int *rot_count, unsigned long long *offmax, const int *skip_if_zero,
                      hipStream_t stream, int count_slots)
    {
        if (n < 2)
        {
            return;
        }
        if (count_slots < 1 || (count_slots & (count_slots - 1)) != 0)
        {
            throw_invalid("count_slots must be a power of two.");
        }
        const int N = static_cast<int>(n + (n & 1));
        RoundFn fn = nullptr;
        switch (shape.threads)
        {
        case 64: fn = pick_round<64>(shape.ea, shape.ev); break;
        case 256: fn = pick_round<256>(shape.ea, shape.ev); break;
        default: fn = pick_round<1024>(shape.ea, shape.ev); break;
        }
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(N / 2)), dim3(static_cast<unsigned>(shape.threads)), 0,
                           stream, a, lda, m, v, ldv, n, N, round, eps, rot_count, offmax, skip_if_zero, count_slots);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    namespace
    {
        using SpanFn = void (*)(const double *, double *, int64_t, int64_t, const double *, double *, int64_t,
                                int64_t, int, int, int, int, int, double, int *, unsigned long long *, const int *,
                                int);

        // The span kernel for (TT, EA, NV, W), or null where the build has none: V chunks beyond the A chunks
        // (n <= m), or a stage of W columns that would not fit in LDS.
        template <int TT, int EA, int NV, int W>
        SpanFn span_fn()
        {
            constexpr int CA = (TT / 64) * EA;
            // (16 wavefronts get 128 registers per lane: at most 8 chunks of each of the two columns there)
            if constexpr (NV <= CA && W * (CA + NV) * 1024 <= 160 * 1024 && (W < 16 || CA + NV <= 8))
            {
                return plain::round_span_kernel<TT, EA, NV, W>;
            }
            return nullptr;
        }

        template <int TT, int EA, int W>
        SpanFn pick_span_nv(int nv)
        {
            switch (nv)
            {
            case 1: return span_fn<TT, EA, 1, W>();
            case 2: return span_fn<TT, EA, 2, W>();
            case 4: return span_fn<TT, EA, 4, W>();
            default: return nullptr;
            }
        }

        template <int W>
        SpanFn pick_span_w(int threads, int ea, int nv)
        {
            if (threads == 64)
            {
                switch (ea)
                {
                case 1: return pick_span_nv<64, 1, W>(nv);
                case 2: return pick_span_nv<64, 2, W>(nv);
                case 4: return pick_span_nv<64, 4, W>(nv);
                case 8: return pick_span_nv<64, 8, W>(nv);
                default: return nullptr;
                }
            }
            if (threads == 256)
            {
                switch (ea)
                {
                case 1: return pick_span_nv<256, 1, W>(nv);
                case 2: return pick_span_nv<256, 2, W>(nv);
                default: return nullptr;
                }
            }
            return nullptr;
        }

        // V chunks per lane of a column: a power of two covering ceil(n / 2) double2 chunks in 64 lanes.
        int span_nv(int64_t n)
        {
            const int64_t per = ((n + 1) / 2 + 63) / 64;
            return per <= 1 ? 1 : (per <= 2 ? 2 : (per <= 4 ? 4 : 0));
        }

        SpanFn pick_span(const RoundShape &shape, int waves, int64_t n)
        {
            const int nv = span_nv(n);
            if (nv == 0 || shape.ea < 1)
            {
                return nullptr;
            }
            const int ca = (shape.threads / 64) * shape.ea;
            if (ca > 8 || nv > ca)
            {
                return nullptr;
            }
            if (waves == 16 && ca + nv <= 8)
            {
                return pick_span_w<16>(shape.threads, shape.ea, nv);
            }
            if (waves == 8)
            {
                return pick_span_w<8>(shape.threads, shape.ea, nv);
            }
            return nullptr;
        }
    } // namespace

    bool span_supported(const RoundShape &shape, int waves, int64_t m, int64_t n)
    {
        if (n < 3 || m < n)
        {
            return false;
        }
        const int64_t m2 = (m + 1) / 2;
        if (shape.ea < 1 || static_cast<int64_t>(shape.ea) * shape.threads < m2)
        {
            return false;
        }
        return pick_span(shape, waves, n) != nullptr;
    }

    int span_max_rounds(int waves, int64_t n)
    {
        const int N = static_cast<int>(n + (n & 1));
        const int H = N / 2;
        return H <= waves ? N - 1 : (waves + 1) / 2;
    }

    void span_partition(int64_t n, int waves, int max_rounds, int &groups, int &own, int &own_end)
    {
        plain::span_partition(n, waves, max_rounds, groups, own, own_end);
    }

    std::vector<int> span_chunks(int64_t n, int max_rounds) { return plain::span_chunks(n, max_rounds); }

    void launch_round_span(const RoundShape &shape, int waves, const double *a_in, double *a_out, int64_t lda,
                           int64_t m, const double *v_in, double *v_out, int64_t ldv, int64_t n, int r0, int rounds,
                           int groups, int own, int own_end, double eps, int *rot_count, unsigned long long *offmax,
                           const int *skip_if_zero, hipStream_t stream, int count_slots)
    {
        if (count_slots < 1 || (count_slots & (count_slots - 1)) != 0)
        {
            throw_invalid("count_slots must be a power of two.");
        }
        SpanFn fn = span_supported(shape, waves, m, n) ? pick_span(shape, waves, n) : nullptr;
        if (fn == nullptr || rounds < 1)
        {
            throw_invalid("round span: unsupported launch shape.");
        }
        const int N = static_cast<int>(n + (n & 1));
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(groups)), dim3(static_cast<unsigned>(64 * waves)), 0,
                           stream, a_in, a_out, lda, m, v_in, v_out, ldv, n, N, r0, rounds, own, own_end, eps,
                           rot_count, offmax, skip_if_zero, count_slots);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_sweep_totals(const int *rot_slots, const unsigned long long *off_slots, int slots, int *rot_total,
                             unsigned long long *off_total, hipStream_t stream)
    {
        hipLaunchKernelGGL(sweep_totals_kernel, dim3(1), dim3(64), 0, stream, rot_slots, off_slots, slots, rot_total,
                           off_total);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_transpose(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                          hipStream_t stream)
    {
        if (rows == 0 || cols == 0)
        {
            return;
        }
        const int64_t tiles_x = (cols + kTile - 1) / kTile;
        const int64_t tiles_y = (rows + kTile - 1) / kTile;
        hipLaunchKernelGGL(transpose_tile_kernel<false>, dim3(static_cast<unsigned>(tiles_x * tiles_y)),
                           dim3(kTile, kTileRows), 0, stream, src, lds, dst, ldd, rows, cols, tiles_x,
                           static_cast<const double *>(nullptr), 0.0);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_u_sigma(const double *a, int64_t lda, int64_t m, int64_t n, double eps, double *u, double *sigma,
                        hipStream_t stream)
    {
        hipLaunchKernelGGL(column_norms_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, stream, a,
                           lda, m, n, sigma, static_cast<double *>(nullptr));
        JSVD_HIP_CHECK(hipGetLastError());
        // A column-major m x n is a row-major n x m matrix (rows = columns of A); transposing it with the
        // per-row (= per-column of A) division yields U row-major m x n.
        const int64_t rows = n;
        const int64_t cols = m;
        const int64_t tiles_x = (cols + kTile - 1) / kTile;
        const int64_t tiles_y = (rows + kTile - 1) / kTile;
        hipLaunchKernelGGL(transpose_tile_kernel<true>, dim3(static_cast<unsigned>(tiles_x * tiles_y)),
                           dim3(kTile, kTileRows), 0, stream, a, lda, u, n, rows, cols, tiles_x,
                           static_cast<const double *>(sigma), eps);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_identity(double *v, int64_t ldv, int64_t n, hipStream_t stream)
    {
        const int64_t total = ldv * n;
        hipLaunchKernelGGL(identity_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, stream,
                           v, ldv, n);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_pair_gram(const double *a, int64_t lda, int64_t m, const int *pairs, int64_t count, double *alpha,
                          double *beta, double *gamma, hipStream_t stream)
    {
        if (count == 0)
        {
            return;
        }
        hipLaunchKernelGGL(pair_gram_kernel, dim3(static_cast<unsigned>(count)), dim3(256), 0, stream, a, lda, m,
                           pairs, alpha, beta, gamma);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_rotation_params(const double *alpha, const double *beta, const double *gamma, int64_t count,
                                double eps, double *c, double *s, int *rotate, hipStream_t stream)
    {
        if (count == 0)
        {
            return;
        }
        hipLaunchKernelGGL(rotation_params_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0,
                           stream, alpha, beta, gamma, count, eps, c, s, rotate);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    int64_t offnorm_workspace_doubles(int64_t n)
    {
        const int64_t tiles = (n + kOffTile - 1) / kOffTile;
        return n + tiles * (tiles + 1) / 2;
    }

    void launch_offnorm(const double *a, int64_t lda, int64_t m, int64_t n, double *workspace, double *out,
                        hipStream_t stream)
    {
        double *norm2 = workspace;
        double *partial = workspace + n;
        hipLaunchKernelGGL(column_norms_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, stream, a,
                           lda, m, n, static_cast<double *>(nullptr), norm2);
        JSVD_HIP_CHECK(hipGetLastError());
        const int64_t tiles = (n + kOffTile - 1) / kOffTile;
        const int64_t blocks = tiles * (tiles + 1) / 2;
        hipLaunchKernelGGL(offnorm_stage1_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, a, lda,
                           m, n, static_cast<const double *>(norm2), tiles, partial);
        JSVD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(offnorm_stage2_kernel, dim3(1), dim3(256), 0, stream, static_cast<const double *>(partial),
                           blocks, out);
        JSVD_HIP_CHECK(hipGetLastError());
    }
} // namespace jsvd
