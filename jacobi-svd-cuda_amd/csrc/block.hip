// block.hip -- block one-sided Jacobi kernels (FP32 and FP64) for the mixed-precision path.
//
// Columns are grouped into blocks of kB = 32; a round pairs blocks (I, J) by the same round-robin circle
// the plain path uses for columns.  For every block pair, one round is:
//
//   block_gram_kernel    G_s = X_s^T X_s for each row slab X_s (kH rows x kW = 64 columns of [A_I A_J]) on
//                        MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64), accumulating in T
//   block_eigen_kernel   G = sum_s G_s, then one cyclic Jacobi sweep over the 64 columns carried out in Gram
//                        space (two-sided 2x2 updates of G), rotation parameters (tau, t, c, s) in FP64 with
//                        the reference skip test; accumulates the 64 x 64 rotation Q
//   block_apply_kernel   [A_I A_J] <- [A_I A_J] Q and [V_I V_J] <- [V_I V_J] Q per row slab on MFMA
//
// Layout: X (A or V) column-major with leading dimension ld (a multiple of kH, padding rows zero); the column
// count is padded to a multiple of kB with zero columns (their Gram entries are zero, so they never rotate).
#include "block.hpp"

namespace jsvd
{
    namespace
    {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        typedef double f64x4 __attribute__((ext_vector_type(4)));

        // One 16x16x4 MFMA step.  Lane l supplies A[l%16][l/16] and B[l/16][l%16]; the four accumulator values
        // of lane l are D[row(l, v)][l%16] with row = 4*(l/16)+v for f32 and row = l/16 + 4*v for f64.
        template <typename T>
        struct Mfma;

        template <>
        struct Mfma<float>
        {
            using Acc = f32x4;
            __device__ static Acc step(float a, float b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
            __device__ static int row(int lane, int v) { return 4 * (lane / 16) + v; }
        };

        template <>
        struct Mfma<double>
        {
            using Acc = f64x4;
            __device__ static Acc step(double a, double b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
            __device__ static int row(int lane, int v) { return lane / 16 + 4 * v; }
        };

        __device__ inline void block_pair(int slot, int round, int Nb, int &I, int &J)
        {
            I = dev::circle_at(slot, round, Nb);
            J = dev::circle_at(Nb - 1 - slot, round, Nb);
        }

        // Physical column of local column c (0..kW-1) of block pair (I, J) under the column map `perm`
        // (logical column k of the sweep order is stored in physical column perm[k]).
        __device__ inline int64_t pair_column(const int *__restrict__ perm, int I, int J, int c)
        {
            return perm[(c < kB) ? (I * kB + c) : (J * kB + c - kB)];
        }

        // Loads rows [r0, r0+kH) of the kW columns [X_I X_J] into LDS as xs[c][r] (column-major, padded).
        template <typename T>
        __device__ inline void load_slab(const T *__restrict__ x, int64_t ld, const int *__restrict__ perm, int I,
                                         int J, int64_t r0, T (*xs)[kH + 1])
        {
            // kW * kH elements, 256 threads: each thread moves kW*kH/256 elements; consecutive threads read
            // consecutive rows of one column (coalesced).  All loads are issued before the LDS stores so
            // they are in flight together.
            constexpr int kPer = kW * kH / 256;
            T tmp[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                const int e = threadIdx.x + 256 * k;
                const int c = e / kH;
                const int r = e % kH;
                tmp[k] = x[pair_column(perm, I, J, c) * ld + r0 + r];
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                const int e = threadIdx.x + 256 * k;
                xs[e / kH][e % kH] = tmp[k];
            }
        }

        // Partial Gram of row slabs [ps * per, min(slabs, (ps + 1) * per)) of block pair `slot`; the partials of a
        // pair are partial[slot][0 .. parts).
        template <typename T>
        __global__ __launch_bounds__(256) void block_gram_kernel(const T *__restrict__ a, int64_t lda,
                                                                 const int *__restrict__ perm, int nb, int Nb,
                                                                 int round, int slabs, int per, int parts,
                                                                 T *__restrict__ partial,
                                                                 const int *__restrict__ skip_if_zero)
        {
            __shared__ T xs[kW][kH + 1];
            if (skip_if_zero != nullptr && *skip_if_zero == 0)
            {
                return;
            }
            const int slot = blockIdx.x;
            const int ps = blockIdx.y;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            if (I >= nb || J >= nb)
            {
                return;
            }
            const int lane = threadIdx.x % 64;
            const int wave = threadIdx.x / 64;
            const int i0 = 16 * wave; // this wave computes G rows i0..i0+15, all kW columns
            typename Mfma<T>::Acc acc[4];
            for (int t = 0; t < 4; ++t)
            {
                acc[t] = typename Mfma<T>::Acc{0, 0, 0, 0};
            }
            const int s_end = min(slabs, (ps + 1) * per);
            for (int s = ps * per; s < s_end; ++s)
            {
                if (s > ps * per)
                {
                    __syncthreads(); // every wave is done with the previous slab
                }
                load_slab(a, lda, perm, I, J, static_cast<int64_t>(s) * kH, xs);
                __syncthreads();
                for (int r0 = 0; r0 < kH; r0 += 4)
                {
                    const int kr = r0 + lane / 16;
                    const T av = xs[i0 + lane % 16][kr];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                    {
                        const T bv = xs[16 * t + lane % 16][kr];
                        acc[t] = Mfma<T>::step(av, bv, acc[t]);
                    }
                }
            }
            T *out = partial + (static_cast<int64_t>(slot) * parts + ps) * (kW * kW);
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int i = i0 + Mfma<T>::row(lane, v);
                    const int j = 16 * t + lane % 16;
                    out[i * kW + j] = acc[t][v];
                }
            }
        }

        // Reduction of the slab partials and one inner cyclic Jacobi sweep in Gram space.
        //
        // G is symmetric and kept only in its upper triangle (g[min][max]).  Each inner round pairs the 64
        // columns by the round-robin circle (pair u of a round = circle positions u and 63 - u).  A round has
        // two phases separated by a barrier:
        //   phase 1: lanes 0..31 of wave 0 form the 32 rotations of the round in FP64 (reference skip test) and
        //            publish them; threads 0..527 load their 2x2 block of G (block (t, u), t <= u, of the 32 x 32
        //            pair grid); the Q waves apply the previous round's rotations to Q (off the critical path);
        //   phase 2: threads 0..527 write J_t^T G J_u back in place.
        // The instruction issue of the 16 waves bounds a round, so the work is packed into as few waves as it
        // needs: G blocks on waves 0..8, Q on waves 12..15.  Q is kept in registers in circle-position order:
        // thread (row k, segment s) of the Q waves holds Q[k][col(p)] for the positions p = 8s..8s+7 and their
        // partners 63-p, so every rotation of a round is local to a thread; after a round the positions move by
        // one along the circle (1..63 cyclically, 0 fixed), which moves one value to each neighbouring segment.
        // Column indices advance incrementally (col(p, r+1) = col(p, r) - 1 on 1..63).  A block pair none of
        // whose 2016 column pairs passes the skip test is left untouched.
        constexpr int kEigenThreads = 1024;
        constexpr int kPairs = kW / 2;
        constexpr int kTriBlocks = kPairs * (kPairs + 1) / 2;   // 2x2 blocks of the upper triangle (528)
        constexpr int kQThreads = 256;                          // 64 rows x 4 segments of 8 position pairs
        constexpr int kQFirst = kEigenThreads - kQThreads;      // Q waves are the last four
        static_assert(kW == 64 && kTriBlocks <= kQFirst, "eigen thread roles assume 64 columns");

        // Column at a circle position in the next inner round.
        __device__ inline int next_column(int c)
        {
            return c == 0 ? 0 : (c == 1 ? kW - 1 : c - 1);
        }

        // One inner round on a thread's Q registers: rotate its 8 position pairs, then move the positions one step
        // along the circle (values crossing a segment boundary come from the neighbouring lanes).
        template <typename T>
        __device__ inline void q_round(T (&lo)[8], T (&hi)[8], const T (*cs)[2], int seg)
        {
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                const T c = cs[8 * seg + i][0], s = cs[8 * seg + i][1];
                const T x = lo[i], z = hi[i];
                lo[i] = c * x - s * z;
                hi[i] = s * x + c * z;
            }
            const T out_lo = lo[7]; // position 8s+7 -> 8s+8 (next segment; segment 3: position 32, own hi[7])
            const T out_hi = hi[0]; // position 63-8s -> 64-8s (previous segment; segment 0: position 1, own lo[1])
            const T from_prev = __shfl_up(out_lo, 1, 64);
            const T from_next = __shfl_down(out_hi, 1, 64);
            const T lo0 = lo[0];
#pragma unroll
            for (int i = 7; i >= 1; --i)
            {
                lo[i] = lo[i - 1];
            }
#pragma unroll
            for (int i = 0; i < 7; ++i)
            {
                hi[i] = hi[i + 1];
            }
            lo[0] = (seg == 0) ? lo0 : from_prev; // position 0 never moves
            hi[7] = (seg == 3) ? out_lo : from_next;
            if (seg == 0)
            {
                lo[1] = out_hi; // position 63 -> 1
            }
        }

        template <typename T>
        __device__ inline bool block_rotation(double app, double aqq, double apq, double eps, double &c, double &s)
        {
            // The FP32 phase applies its rotations in FP32, so the shorter FP64 chain is exact enough there; the
            // FP64 phase keeps the division form.
            if constexpr (sizeof(T) == 4)
            {
                return dev::rotation_params_fast(app, aqq, apq, eps, c, s);
            }
            else
            {
                return dev::rotation_params_block(app, aqq, apq, eps, c, s);
            }
        }

        template <typename T>
        __global__ __launch_bounds__(kEigenThreads) void block_eigen_kernel(
            const T *__restrict__ partial, int slabs, int nb, int Nb, int round, double eps, T *__restrict__ q_out,
            int *__restrict__ apply_flag, int *__restrict__ rot_count, unsigned long long *__restrict__ offmax,
            const int *__restrict__ perm, double *__restrict__ norm2, const int *__restrict__ skip_if_zero)
        {
            __shared__ T g[kW][kW + 1];
            __shared__ T cs[2][kPairs][2];              // rotations of the current and the previous round
            __shared__ unsigned long long rmask[2];     // bit u: pair u rotates
            __shared__ double red[kEigenThreads / 64];
            __shared__ int any;
            __shared__ int needed;
            if (skip_if_zero != nullptr && *skip_if_zero == 0)
            {
                return;
            }
            const int slot = blockIdx.x;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            if (I >= nb || J >= nb)
            {
                return;
            }
            const int tid = threadIdx.x;
            // This thread's 2x2 block (t, u), t <= u, of the upper triangle of the 32 x 32 pair grid.
            int bt = 0, bu = 0;
            if (tid < kTriBlocks)
            {
                int rem = tid;
                while (rem >= kPairs - bt)
                {
                    rem -= kPairs - bt;
                    ++bt;
                }
                bu = bt + rem;
            }
            const T *base = partial + static_cast<int64_t>(slot) * slabs * (kW * kW);
            for (int e = tid; e < kW * kW; e += kEigenThreads)
            {
                const int i = e / kW;
                const int j = e % kW;
                if (i <= j)
                {
                    // eight independent loads in flight per step; the summation order is fixed (slab order)
                    T acc = 0;
                    int s = 0;
                    for (; s + 8 <= slabs; s += 8)
                    {
                        T part[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                        {
                            part[k] = base[static_cast<int64_t>(s + k) * (kW * kW) + e];
                        }
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                        {
                            acc += part[k];
                        }
                    }
                    for (; s < slabs; ++s)
                    {
                        acc += base[static_cast<int64_t>(s) * (kW * kW) + e];
                    }
                    g[i][j] = acc;
                }
            }
            if (tid == 0)
            {
                any = 0;
                needed = 0;
            }
            // Q = I in circle-position order (round 0 positions).
            const bool qthr = tid >= kQFirst;
            const int qk = (tid - kQFirst) / 4;
            const int seg = (tid - kQFirst) % 4;
            T qlo[8], qhi[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                qlo[i] = (dev::circle_at_fast(8 * seg + i, 0, kW) == qk) ? T(1) : T(0);
                qhi[i] = (dev::circle_at_fast(kW - 1 - 8 * seg - i, 0, kW) == qk) ? T(1) : T(0);
            }
            __syncthreads();
            // Off-orthogonality of the pair's 64 columns before the sweep (convergence measure), and whether
            // any pair passes the skip test at all: if none does, no rotation can happen in the sweep (G never
            // changes) and the sweep is skipped outright.
            double best = 0.0;
            bool need = false;
            for (int e = tid; e < kW * kW; e += kEigenThreads)
            {
                const int i = e / kW;
                const int j = e % kW;
                if (i < j)
                {
                    const double gii = static_cast<double>(g[i][i]);
                    const double gjj = static_cast<double>(g[j][j]);
                    const double gij = static_cast<double>(g[i][j]);
                    best = fmax(best, dev::pair_cosine(gii, gjj, gij));
                    const double rhs = eps * sqrt(fmax(gii, 0.0) * fmax(gjj, 0.0));
                    need = need || (fabs(gij) > rhs && fabs(gij) > 1.0e-300);
                }
            }
            best = dev::wave_max(best);
            if ((tid % 64) == 0)
            {
                red[tid / 64] = best;
            }
            if (need)
            {
                needed = 1;
            }
            __syncthreads();
            int my_rotations = 0;
            const bool gthr = tid < kTriBlocks;
            int x1 = dev::circle_at_fast(bt, 0, kW), x2 = dev::circle_at_fast(kW - 1 - bt, 0, kW);
            int z1 = dev::circle_at_fast(bu, 0, kW), z2 = dev::circle_at_fast(kW - 1 - bu, 0, kW);
            const int pu = tid % kPairs; // wave 0: the pair this lane rotates
            int p1 = dev::circle_at_fast(pu, 0, kW), p2 = dev::circle_at_fast(kW - 1 - pu, 0, kW);
            const bool run = needed != 0;
            for (int rr = 0; run && rr < kW - 1; ++rr)
            {
                const int b = rr & 1;
                // phase 1: block loads, the round's rotations, the previous round's Q update
                T g11 = 0, g12 = 0, g21 = 0, g22 = 0;
                int o11 = 0, o12 = 0, o21 = 0, o22 = 0;
                T *gf = &g[0][0];
                if (gthr)
                {
                    o11 = min(x1, z1) * (kW + 1) + max(x1, z1);
                    o12 = min(x1, z2) * (kW + 1) + max(x1, z2);
                    o21 = min(x2, z1) * (kW + 1) + max(x2, z1);
                    o22 = min(x2, z2) * (kW + 1) + max(x2, z2);
                    g11 = gf[o11];
                    g12 = gf[o12];
                    g21 = gf[o21];
                    g22 = gf[o22];
                }
                if (tid < 64)
                {
                    double c, s;
                    const bool rot = block_rotation<T>(static_cast<double>(g[p1][p1]), static_cast<double>(g[p2][p2]),
                                                       static_cast<double>(g[min(p1, p2)][max(p1, p2)]), eps, c, s);
                    const unsigned long long m = __ballot(rot);
                    if (tid < kPairs)
                    {
                        cs[b][tid][0] = static_cast<T>(c);
                        cs[b][tid][1] = static_cast<T>(s);
                        my_rotations += rot ? 1 : 0;
                    }
                    if (tid == 0)
                    {
                        rmask[b] = m;
                    }
                }
                if (qthr && rr > 0)
                {
                    q_round<T>(qlo, qhi, cs[b ^ 1], seg);
                }
                __syncthreads();
                // phase 2: J_t^T G J_u in place (rows by J_t^T, x' = c x - s y on row pairs, then columns by J_u)
                if (gthr)
                {
                    const unsigned long long m = rmask[b];
                    if (((m >> bt) | (m >> bu)) & 1ull)
                    {
                        const T ct = cs[b][bt][0], st = cs[b][bt][1];
                        const T cu = cs[b][bu][0], su = cs[b][bu][1];
                        const T h11 = ct * g11 - st * g21, h12 = ct * g12 - st * g22;
                        const T h21 = st * g11 + ct * g21, h22 = st * g12 + ct * g22;
                        gf[o11] = cu * h11 - su * h12;
                        gf[o12] = su * h11 + cu * h12;
                        gf[o22] = su * h21 + cu * h22;
                        if (bt != bu)
                        {
                            gf[o21] = cu * h21 - su * h22; // on the diagonal block entry 21 aliases entry 12
                        }
                    }
                    x1 = next_column(x1);
                    x2 = next_column(x2);
                    z1 = next_column(z1);
                    z2 = next_column(z2);
                }
                p1 = next_column(p1);
                p2 = next_column(p2);
                __syncthreads();
            }
            if (run && qthr)
            {
                q_round<T>(qlo, qhi, cs[(kW - 2) & 1], seg); // the last round's update
            }
            if (tid < kPairs && my_rotations > 0)
            {
                atomicAdd(&any, my_rotations);
            }
            __syncthreads();
            // After kW - 1 moves the circle is back at its round-0 positions.
            T *qo = q_out + static_cast<int64_t>(slot) * (kW * kW);
            if (qthr)
            {
#pragma unroll
                for (int i = 0; i < 8; ++i)
                {
                    qo[qk * kW + dev::circle_at_fast(8 * seg + i, 0, kW)] = qlo[i]; // row-major Q[c][j]
                    qo[qk * kW + dev::circle_at_fast(kW - 1 - 8 * seg - i, 0, kW)] = qhi[i];
                }
            }
            // squared norms of the pair's columns after the rotation (the diagonal of Q^T G Q), for the
            // norm ordering of the next sweep
            if (norm2 != nullptr && tid < kW)
            {
                norm2[pair_column(perm, I, J, tid)] = static_cast<double>(g[tid][tid]);
            }
            if (tid == 0)
            {
                apply_flag[slot] = any;
                if (any > 0 && rot_count)
                {
                    atomicAdd(rot_count, any);
                }
                if (offmax)
                {
                    double m = 0.0;
                    for (int w = 0; w < kEigenThreads / 64; ++w)
                    {
                        m = fmax(m, red[w]);
                    }
                    dev::atomic_max_nonneg(offmax, m);
                }
            }
        }

        // Y = X Q for one row slab of [X_I X_J]; grid.y spans the slabs of A followed by those of V.
        template <typename T>
        __global__ __launch_bounds__(256) void block_apply_kernel(T *__restrict__ a, int64_t lda, int slabs_a,
                                                                  T *__restrict__ v, int64_t ldv, int nb, int Nb,
                                                                  int round, const T *__restrict__ q_in,
                                                                  const int *__restrict__ apply_flag,
                                                                  const int *__restrict__ perm,
                                                                  const int *__restrict__ skip_if_zero)
        {
            __shared__ T xs[kW][kH + 1];
            __shared__ T qs[kW][kW + 1];
            if (skip_if_zero != nullptr && *skip_if_zero == 0)
            {
                return;
            }
            const int slot = blockIdx.x;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            if (I >= nb || J >= nb || apply_flag[slot] == 0)
            {
                return;
            }
            const int sy = blockIdx.y;
            T *x = (sy < slabs_a) ? a : v;
            const int64_t ld = (sy < slabs_a) ? lda : ldv;
            const int64_t r0 = static_cast<int64_t>((sy < slabs_a) ? sy : sy - slabs_a) * kH;
            load_slab(x, ld, perm, I, J, r0, xs);
            const T *qg = q_in + static_cast<int64_t>(slot) * (kW * kW);
            for (int e = threadIdx.x; e < kW * kW; e += 256)
            {
                qs[e / kW][e % kW] = qg[e];
            }
            __syncthreads();
            const int lane = threadIdx.x % 64;
            const int wave = threadIdx.x / 64;
            // wave computes output rows [32*wave, 32*wave+32) (two 16-row tiles) x all 64 columns (four tiles)
            typename Mfma<T>::Acc acc[2][4];
            for (int a2 = 0; a2 < 2; ++a2)
            {
                for (int t = 0; t < 4; ++t)
                {
                    acc[a2][t] = typename Mfma<T>::Acc{0, 0, 0, 0};
                }
            }
            for (int c0 = 0; c0 < kW; c0 += 4)
            {
                const int kc = c0 + lane / 16;
                const T av0 = xs[kc][32 * wave + lane % 16];
                const T av1 = xs[kc][32 * wave + 16 + lane % 16];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                {
                    const T bv = qs[kc][16 * t + lane % 16];
                    acc[0][t] = Mfma<T>::step(av0, bv, acc[0][t]);
                    acc[1][t] = Mfma<T>::step(av1, bv, acc[1][t]);
                }
            }
            __syncthreads(); // every wave has read xs; overwrite it with the product
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2)
            {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                {
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv)
                    {
                        const int r = 32 * wave + 16 * a2 + Mfma<T>::row(lane, vv);
                        const int c = 16 * t + lane % 16;
                        xs[c][r] = acc[a2][t][vv];
                    }
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < kW * kH; e += 256)
            {
                const int c = e / kH;
                const int r = e % kH;
                x[pair_column(perm, I, J, c) * ld + r0 + r] = xs[c][r];
            }
        }

        template <typename Tin, typename Tout>
        __global__ void convert_kernel(const Tin *__restrict__ src, Tout *__restrict__ dst, int64_t count)
        {
            const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i < count)
            {
                dst[i] = static_cast<Tout>(src[i]);
            }
        }

        // Copies the leading rows x cols block of a column-major src (lds) into dst (ldd), zero-filling
        // dst rows rows..ldd-1 and columns cols..dcols-1 except for a unit diagonal where `identity_pad`.
        template <typename T>
        __global__ void pad_copy_kernel(const T *__restrict__ src, int64_t lds, int64_t rows, int64_t cols,
                                        T *__restrict__ dst, int64_t ldd, int64_t dcols, int identity_pad)
        {
            const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i >= ldd * dcols)
            {
                return;
            }
            const int64_t c = i / ldd;
            const int64_t r = i % ldd;
            T x = 0;
            if (r < rows && c < cols)
            {
                x = src[c * lds + r];
            }
            else if (identity_pad && c >= cols && r == c)
            {
                x = 1;
            }
            dst[i] = x;
        }

        template <typename T>
        __global__ void identity_kernel_t(T *__restrict__ v, int64_t ldv, int64_t cols)
        {
            const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i >= ldv * cols)
            {
                return;
            }
            v[i] = (i / ldv == i % ldv) ? T(1) : T(0);
        }
        // Column order of the next sweep: physical columns sorted by descending squared norm (ties by column
        // index), one workgroup, bitonic network over the next power of two >= n_pad in LDS.
        __global__ __launch_bounds__(1024) void column_order_kernel(const double *__restrict__ norm2, int n_pad,
                                                                    int *__restrict__ perm,
                                                                    const int *__restrict__ skip_if_zero)
        {
            __shared__ double key[kSortMax];
            __shared__ int val[kSortMax];
            if (skip_if_zero != nullptr && *skip_if_zero == 0)
            {
                return;
            }
            int size = 1;
            while (size < n_pad)
            {
                size *= 2;
            }
            for (int i = threadIdx.x; i < size; i += 1024)
            {
                key[i] = (i < n_pad) ? norm2[i] : -1.0;
                val[i] = i;
            }
            __syncthreads();
            for (int k = 2; k <= size; k *= 2)
            {
                for (int j = k / 2; j > 0; j /= 2)
                {
                    for (int i = threadIdx.x; i < size; i += 1024)
                    {
                        const int x = i ^ j;
                        if (x > i)
                        {
                            // does element i belong before element x in the final (descending) order?
                            const bool i_first = key[i] > key[x] || (key[i] == key[x] && val[i] < val[x]);
                            const bool ascending_block = (i & k) == 0;
                            if (ascending_block != i_first)
                            {
                                const double tk = key[i];
                                key[i] = key[x];
                                key[x] = tk;
                                const int tv = val[i];
                                val[i] = val[x];
                                val[x] = tv;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = threadIdx.x; i < n_pad; i += 1024)
            {
                perm[i] = val[i];
            }
        }

        __global__ void iota_kernel(int *__restrict__ perm, int64_t n)
        {
            const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i < n)
            {
                perm[i] = static_cast<int>(i);
            }
        }
    } // namespace

    template <typename T>
    void launch_block_round(const BlockGeometry &g, T *a, T *v, int round, double eps, T *partial, T *q,
                            int *apply_flag, int *rot_count, unsigned long long *offmax, const int *perm,
                            double *norm2, const int *skip_if_zero, hipStream_t stream)
    {
        const int pairs = g.Nb / 2;
        // Slabs per Gram workgroup: fewer partials for the eigen kernel to reduce (its reduction runs on one CU
        // per pair) while the Gram grid still covers the chip.
        int per = 1;
        while ((g.slabs_a + per - 1) / per > 8 && pairs * ((g.slabs_a + 2 * per - 1) / (2 * per)) >= 256)
        {
            per *= 2;
        }
        const int parts = (g.slabs_a + per - 1) / per;
        hipLaunchKernelGGL(block_gram_kernel<T>, dim3(pairs, parts), dim3(256), 0, stream, static_cast<const T *>(a),
                           g.lda, perm, g.nb, g.Nb, round, g.slabs_a, per, parts, partial, skip_if_zero);
        JSVD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(block_eigen_kernel<T>, dim3(pairs), dim3(kEigenThreads), 0, stream,
                           static_cast<const T *>(partial), parts, g.nb, g.Nb, round, eps, q, apply_flag, rot_count,
                           offmax, perm, norm2, skip_if_zero);
        JSVD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(block_apply_kernel<T>, dim3(pairs, g.slabs_a + g.slabs_v), dim3(256), 0, stream, a, g.lda,
                           g.slabs_a, v, g.ldv, g.nb, g.Nb, round, static_cast<const T *>(q),
                           static_cast<const int *>(apply_flag), perm, skip_if_zero);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    template void launch_block_round<float>(const BlockGeometry &, float *, float *, int, double, float *, float *,
                                            int *, int *, unsigned long long *, const int *, double *, const int *,
                                            hipStream_t);
    template void launch_block_round<double>(const BlockGeometry &, double *, double *, int, double, double *,
                                             double *, int *, int *, unsigned long long *, const int *, double *,
                                             const int *, hipStream_t);

    template <typename T>
    void launch_block_eigen(const T *gram, int pairs, double eps, T *q_out, int *rotations, hipStream_t stream)
    {
        // A circle of 2 * pairs blocks at round 0 maps slot p to a valid block pair, so the kernel's pair p reads
        // gram[p] (one partial) and writes q_out[p] and rotations[p]; no column map, norms or sweep counters.
        const int Nb = 2 * pairs;
        hipLaunchKernelGGL(block_eigen_kernel<T>, dim3(pairs), dim3(kEigenThreads), 0, stream, gram, 1, Nb, Nb, 0, eps,
                           q_out, rotations, nullptr, nullptr, nullptr, nullptr, nullptr);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    template void launch_block_eigen<float>(const float *, int, double, float *, int *, hipStream_t);
    template void launch_block_eigen<double>(const double *, int, double, double *, int *, hipStream_t);

    void launch_column_order(const double *norm2, int64_t n_pad, int *perm, const int *skip_if_zero,
                             hipStream_t stream)
    {
        if (n_pad > kSortMax)
        {
            return; // beyond one workgroup's LDS: keep the current order
        }
        hipLaunchKernelGGL(column_order_kernel, dim3(1), dim3(1024), 0, stream, norm2, static_cast<int>(n_pad), perm,
                           skip_if_zero);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_iota(int *perm, int64_t n, hipStream_t stream)
    {
        hipLaunchKernelGGL(iota_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, perm, n);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_convert_f64_to_f32(const double *src, float *dst, int64_t count, hipStream_t stream)
    {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                           0, stream, src, dst, count);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_convert_f32_to_f64(const float *src, double *dst, int64_t count, hipStream_t stream)
    {
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                           0, stream, src, dst, count);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_pad_copy_f64(const double *src, int64_t lds, int64_t rows, int64_t cols, double *dst, int64_t ldd,
                             int64_t dcols, bool identity_pad, hipStream_t stream)
    {
        const int64_t total = ldd * dcols;
        hipLaunchKernelGGL(pad_copy_kernel<double>, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                           stream, src, lds, rows, cols, dst, ldd, dcols, identity_pad ? 1 : 0);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_identity_f32(float *v, int64_t ldv, int64_t cols, hipStream_t stream)
    {
        const int64_t total = ldv * cols;
        hipLaunchKernelGGL(identity_kernel_t<float>, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                           stream, v, ldv, cols);
        JSVD_HIP_CHECK(hipGetLastError());
    }
} // namespace jsvd
