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
        // G is symmetric and kept only in its upper triangle (g[min][max]), in two LDS buffers.  Each inner round
        // pairs the 64 columns by the round-robin circle (pair u of a round = circle positions u and 63 - u).
        // Inner round r is one phase and one barrier:
        //   * the G waves (one 2x2 block (t, u), t <= u, of the 32 x 32 pair grid per thread) write
        //     J_t^T G_r J_u into the other buffer, with the rotations of round r published in the previous phase;
        //   * the Q waves apply the same rotations to Q;
        //   * wave 0 meanwhile forms the rotations of round r+1 (lane l: pair l % 32, FP64, reference skip test)
        //     without waiting for G_{r+1}: the three entries a rotation needs are the two new diagonal entries of
        //     its columns and their new off-diagonal entry, each a 2x2 update of entries of G_r with rotations
        //     of round r that wave 0 already holds (exchanged between its lanes), evaluated by the same
        //     expressions as the G waves use.
        // So the FP64 rotation chain of the next round overlaps the bulk update of this one.  Vector-instruction
        // issue bounds the bulk update, so it is packed into as few waves as it needs.  Q is kept in registers in
        // circle-position order: thread (row k, segment s) of the Q waves holds Q[k][col(p)] for the positions
        // p = 8s..8s+7 and their partners 63-p, so every rotation of a round is local to a thread; after a round
        // the positions move by one along the circle (1..63 cyclically, 0 fixed), which moves one value to each
        // neighbouring segment.  Column indices advance incrementally (col(p, r+1) = col(p, r) - 1 on 1..63).  A
        // block pair none of whose 2016 column pairs passes the skip test is left untouched.
        constexpr int kEigenThreads = 1024;
        constexpr int kPairs = kW / 2;
        constexpr int kTriBlocks = kPairs * (kPairs + 1) / 2;   // 2x2 blocks of the upper triangle (528)
        constexpr int kQThreads = 256;                          // 64 rows x 4 segments of 8 position pairs
        constexpr int kQFirst = kEigenThreads - kQThreads;      // Q waves are the last four (12..15)
        // Waves are spread round-robin over the CU's four SIMDs (wave w on SIMD w % 4).  Wave 0 forms the
        // rotations, so the G blocks go to the waves 1..11 not on its SIMD (1, 2, 3, 5, 6, 7, 9, 10, 11): every
        // SIMD then issues one Q wave and either the rotation wave or a third G wave.
        constexpr int kGWaves = 9;
        static_assert(kW == 64 && kGWaves * 64 >= kTriBlocks, "eigen thread roles assume 64 columns");

        // Index of this thread among the G threads, or -1.
        __device__ inline int g_thread_index(int tid)
        {
            const int w = tid / 64;
            if (w < 1 || w > 11 || (w % 4) == 0)
            {
                return -1;
            }
            const int gi = (w - 1 - w / 4) * 64 + tid % 64;
            return gi < kTriBlocks ? gi : -1;
        }

        // Column at a circle position in the next inner round.
        __device__ inline int next_column(int c)
        {
            return c == 0 ? 0 : (c == 1 ? kW - 1 : c - 1);
        }

        // Circle position, in the previous round, of the column now at position p.
        __device__ inline int previous_position(int p)
        {
            return p == 0 ? 0 : (p == 1 ? kW - 1 : p - 1);
        }

        // J_t^T [g11 g12; g21 g22] J_u for the rotations (ct, st) of the row pair and (cu, su) of the column pair
        // (x' = c x - s y, y' = s x + c y), with explicit fused multiply-adds so that every evaluation of the same
        // inputs gives the same bits.
        template <typename T>
        __device__ inline void rotate_block(T ct, T st, T cu, T su, T g11, T g12, T g21, T g22, T &n11, T &n12,
                                            T &n21, T &n22)
        {
            const T h11 = fma(ct, g11, -(st * g21)), h12 = fma(ct, g12, -(st * g22));
            const T h21 = fma(st, g11, ct * g21), h22 = fma(st, g12, ct * g22);
            n11 = fma(cu, h11, -(su * h12));
            n12 = fma(su, h11, cu * h12);
            n21 = fma(cu, h21, -(su * h22));
            n22 = fma(su, h21, cu * h22);
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
            __shared__ T g[2][kW][kW + 1];              // G of the current round and of the next
            __shared__ T cs[2][kPairs][2];              // rotations of the current round and of the next
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
            const int gi = g_thread_index(tid);
            const bool gthr = gi >= 0;
            int bt = 0, bu = 0;
            if (gthr)
            {
                int rem = gi;
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
                    g[0][i][j] = acc;
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
                    const double gii = static_cast<double>(g[0][i][i]);
                    const double gjj = static_cast<double>(g[0][j][j]);
                    const double gij = static_cast<double>(g[0][i][j]);
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
            int x1 = dev::circle_at_fast(bt, 0, kW), x2 = dev::circle_at_fast(kW - 1 - bt, 0, kW);
            int z1 = dev::circle_at_fast(bu, 0, kW), z2 = dev::circle_at_fast(kW - 1 - bu, 0, kW);
            const bool run = needed != 0;
            // Wave 0: the rotation of pair pu of the current round and the three G entries it was formed from.
            const int pu = tid % kPairs;
            T my_c = 1, my_s = 0, gpp = 0, gqq = 0, gpq = 0;
            if (run && tid < 64)
            {
                const int p1 = dev::circle_at_fast(pu, 0, kW), p2 = dev::circle_at_fast(kW - 1 - pu, 0, kW);
                gpp = g[0][p1][p1];
                gqq = g[0][p2][p2];
                gpq = g[0][min(p1, p2)][max(p1, p2)];
                double c, s;
                const bool rot = block_rotation<T>(static_cast<double>(gpp), static_cast<double>(gqq),
                                                   static_cast<double>(gpq), eps, c, s);
                my_c = static_cast<T>(c);
                my_s = static_cast<T>(s);
                const unsigned long long m = __ballot(rot);
                if (tid < kPairs)
                {
                    cs[0][tid][0] = my_c;
                    cs[0][tid][1] = my_s;
                    my_rotations += rot ? 1 : 0;
                }
                if (tid == 0)
                {
                    rmask[0] = m;
                }
            }
            __syncthreads();
            for (int rr = 0; run && rr < kW - 1; ++rr)
            {
                const int b = rr & 1;
                T(*gc)[kW + 1] = g[b];
                T *gn = &g[b ^ 1][0][0];
                if (gthr)
                {
                    // J_t^T G J_u of this thread's block into the other buffer (copied when neither pair rotates)
                    const int o11 = min(x1, z1) * (kW + 1) + max(x1, z1);
                    const int o12 = min(x1, z2) * (kW + 1) + max(x1, z2);
                    const int o21 = min(x2, z1) * (kW + 1) + max(x2, z1);
                    const int o22 = min(x2, z2) * (kW + 1) + max(x2, z2);
                    const T *gf = &gc[0][0];
                    T n11 = gf[o11], n12 = gf[o12], n21 = gf[o21], n22 = gf[o22];
                    const unsigned long long m = rmask[b];
                    if (((m >> bt) | (m >> bu)) & 1ull)
                    {
                        rotate_block(cs[b][bt][0], cs[b][bt][1], cs[b][bu][0], cs[b][bu][1], n11, n12, n21, n22, n11,
                                     n12, n21, n22);
                    }
                    gn[o11] = n11;
                    gn[o12] = n12;
                    gn[o22] = n22;
                    if (bt != bu)
                    {
                        gn[o21] = n21; // on the diagonal block entry 21 aliases entry 12
                    }
                    x1 = next_column(x1);
                    x2 = next_column(x2);
                    z1 = next_column(z1);
                    z2 = next_column(z2);
                }
                if (qthr)
                {
                    q_round<T>(qlo, qhi, cs[b], seg);
                }
                if (tid < 64 && rr + 1 < kW - 1)
                {
                    // New diagonal entries of this lane's pair (its diagonal block; exact copy when not rotating).
                    T d11, d12, d21, d22;
                    rotate_block(my_c, my_s, my_c, my_s, gpp, gpq, gpq, gqq, d11, d12, d21, d22);
                    // The pair of round rr+1: the columns now at positions pu and 63 - pu came from positions
                    // pa and pb, i.e. from pairs ta (first column when pa < 32) and tb of round rr (ta != tb).
                    const int pa = previous_position(pu), pb = previous_position(kW - 1 - pu);
                    const int ta = pa < kPairs ? pa : kW - 1 - pa, tb = pb < kPairs ? pb : kW - 1 - pb;
                    const bool fa = pa < kPairs, fb = pb < kPairs;
                    const T a1 = __shfl(d11, ta, 64), a2 = __shfl(d22, ta, 64);
                    const T b1 = __shfl(d11, tb, 64), b2 = __shfl(d22, tb, 64);
                    const int t1 = min(ta, tb), t2 = max(ta, tb);
                    const T c1 = __shfl(my_c, t1, 64), s1 = __shfl(my_s, t1, 64);
                    const T c2 = __shfl(my_c, t2, 64), s2 = __shfl(my_s, t2, 64);
                    // Their new off-diagonal entry: block (t1, t2) of round rr, as the G waves compute it.
                    const int r1 = dev::circle_at_fast(t1, rr, kW), r2 = dev::circle_at_fast(kW - 1 - t1, rr, kW);
                    const int k1 = dev::circle_at_fast(t2, rr, kW), k2 = dev::circle_at_fast(kW - 1 - t2, rr, kW);
                    T f11, f12, f21, f22;
                    rotate_block(c1, s1, c2, s2, gc[min(r1, k1)][max(r1, k1)], gc[min(r1, k2)][max(r1, k2)],
                                 gc[min(r2, k1)][max(r2, k1)], gc[min(r2, k2)][max(r2, k2)], f11, f12, f21, f22);
                    const bool row_first = (ta == t1) ? fa : fb; // role of the column from pair t1
                    const bool col_first = (ta == t1) ? fb : fa; // role of the column from pair t2
                    gpp = fa ? a1 : a2;
                    gqq = fb ? b1 : b2;
                    gpq = row_first ? (col_first ? f11 : f12) : (col_first ? f21 : f22);
                    double c, s;
                    const bool rot = block_rotation<T>(static_cast<double>(gpp), static_cast<double>(gqq),
                                                       static_cast<double>(gpq), eps, c, s);
                    my_c = static_cast<T>(c);
                    my_s = static_cast<T>(s);
                    const unsigned long long m = __ballot(rot);
                    if (tid < kPairs)
                    {
                        cs[b ^ 1][tid][0] = my_c;
                        cs[b ^ 1][tid][1] = my_s;
                        my_rotations += rot ? 1 : 0;
                    }
                    if (tid == 0)
                    {
                        rmask[b ^ 1] = m;
                    }
                }
                __syncthreads();
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
                // the last inner round writes buffer (kW - 1) & 1
                norm2[pair_column(perm, I, J, tid)] = static_cast<double>(g[run ? ((kW - 1) & 1) : 0][tid][tid]);
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

        // Row-slab height of the apply kernel: 128 rows in FP32, 64 in FP64 (LDS for two workgroups per CU).
        template <typename T>
        constexpr int apply_rows() { return sizeof(T) == 4 ? 128 : 64; }

        template <typename T>
        struct Vec16; // 16-byte vector of T
        typedef float f32x4v __attribute__((ext_vector_type(4)));
        typedef double f64x2v __attribute__((ext_vector_type(2)));
        template <>
        struct Vec16<float>
        {
            using type = f32x4v;
        };
        template <>
        struct Vec16<double>
        {
            using type = f64x2v;
        };

        // Y = X Q for the row slabs of [X_I X_J] (the slabs of A, then those of V, apply_rows<T>() rows each).
        // Workgroup y of a pair takes slabs y, y + gridDim.y, ...: Q is read once per workgroup (FP32: kept in
        // registers as MFMA B operands; FP64: in LDS), the next slab's global loads are in flight while the
        // current slab's MFMAs run, and the product is stored straight from the accumulators.
        template <typename T>
        __global__ __launch_bounds__(256) void block_apply_kernel(T *__restrict__ a, int64_t lda, int slabs_a,
                                                                  T *__restrict__ v, int64_t ldv, int slabs_v,
                                                                  int nb, int Nb, int round,
                                                                  const T *__restrict__ q_in,
                                                                  const int *__restrict__ apply_flag,
                                                                  const int *__restrict__ perm,
                                                                  const int *__restrict__ skip_if_zero)
        {
            constexpr int H = apply_rows<T>();
            constexpr int kVec = 16 / static_cast<int>(sizeof(T)); // elements per 16-byte vector
            constexpr int kLd = H + kVec;                           // LDS row stride: 16-byte aligned rows
            constexpr int kColVecs = H / kVec;                      // vectors per column segment
            constexpr int kPer = kW * kColVecs / 256;               // vectors per thread per slab
            constexpr int kRT = H / 64;                             // 16-row tiles per wave
            constexpr bool kQRegs = sizeof(T) == 4;
            using V = typename Vec16<T>::type;
            __shared__ __align__(16) T xs[kW][kLd];
            __shared__ T qs[kQRegs ? 1 : kW][kW + 1];
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
            const int total = slabs_a + slabs_v;
            int s = blockIdx.y;
            if (s >= total)
            {
                return;
            }
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            const T *qg = q_in + static_cast<int64_t>(slot) * (kW * kW); // row-major Q[c][j]
            T bq[kQRegs ? kW / 4 : 1][4];
            if constexpr (kQRegs)
            {
#pragma unroll
                for (int kk = 0; kk < kW / 4; ++kk)
                {
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                    {
                        bq[kk][t] = qg[(4 * kk + lane / 16) * kW + 16 * t + lane % 16];
                    }
                }
            }
            else
            {
                for (int e = tid; e < kW * kW; e += 256)
                {
                    qs[e / kW][e % kW] = qg[e];
                }
            }
            // This thread's vectors of a slab: column (tid + 256 k) / kColVecs, rows ((tid + 256 k) % kColVecs) * kVec
            // .. + kVec; the row offset is the same for every k (256 is a multiple of kColVecs).
            static_assert(256 % kColVecs == 0, "apply load layout");
            int col_of[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                col_of[k] = static_cast<int>(pair_column(perm, I, J, (tid + 256 * k) / kColVecs));
            }
            const int vrow = (tid % kColVecs) * kVec;
            V pre[kPer];
            {
                const bool in_a = s < slabs_a;
                const T *x = in_a ? a : v;
                const int64_t ld = in_a ? lda : ldv;
                const int64_t r0 = static_cast<int64_t>(in_a ? s : s - slabs_a) * H + vrow;
#pragma unroll
                for (int k = 0; k < kPer; ++k)
                {
                    pre[k] = *reinterpret_cast<const V *>(x + static_cast<int64_t>(col_of[k]) * ld + r0);
                }
            }
            bool first = true;
            for (; s < total; s += gridDim.y)
            {
                if (!first)
                {
                    __syncthreads(); // every wave has finished reading xs
                }
                first = false;
#pragma unroll
                for (int k = 0; k < kPer; ++k)
                {
                    *reinterpret_cast<V *>(&xs[(tid + 256 * k) / kColVecs][vrow]) = pre[k];
                }
                __syncthreads();
                const bool in_a = s < slabs_a;
                T *x = in_a ? a : v;
                const int64_t ld = in_a ? lda : ldv;
                const int64_t r0 = static_cast<int64_t>(in_a ? s : s - slabs_a) * H;
                {
                    // next slab of this workgroup (the last iteration re-reads the current slab, unused; this keeps
                    // the prefetch unconditional so `pre` stays in registers)
                    const int sn = (s + static_cast<int>(gridDim.y) < total) ? s + static_cast<int>(gridDim.y) : s;
                    const bool n_in_a = sn < slabs_a;
                    const T *xn = n_in_a ? a : v;
                    const int64_t ldn = n_in_a ? lda : ldv;
                    const int64_t r0n = static_cast<int64_t>(n_in_a ? sn : sn - slabs_a) * H + vrow;
#pragma unroll
                    for (int k = 0; k < kPer; ++k)
                    {
                        pre[k] = *reinterpret_cast<const V *>(xn + static_cast<int64_t>(col_of[k]) * ldn + r0n);
                    }
                }
                // wave: output rows [wave * 16 kRT, (wave + 1) * 16 kRT) x all 64 columns
                typename Mfma<T>::Acc acc[kRT][4];
#pragma unroll
                for (int a2 = 0; a2 < kRT; ++a2)
                {
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                    {
                        acc[a2][t] = typename Mfma<T>::Acc{0, 0, 0, 0};
                    }
                }
                const int rb = wave * 16 * kRT;
#pragma unroll
                for (int kk = 0; kk < kW / 4; ++kk)
                {
                    const int kc = 4 * kk + lane / 16;
                    T av[kRT];
#pragma unroll
                    for (int a2 = 0; a2 < kRT; ++a2)
                    {
                        av[a2] = xs[kc][rb + 16 * a2 + lane % 16];
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                    {
                        T bv;
                        if constexpr (kQRegs)
                        {
                            bv = bq[kk][t];
                        }
                        else
                        {
                            bv = qs[kc][16 * t + lane % 16];
                        }
#pragma unroll
                        for (int a2 = 0; a2 < kRT; ++a2)
                        {
                            acc[a2][t] = Mfma<T>::step(av[a2], bv, acc[a2][t]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
                {
                    T *col = x + pair_column(perm, I, J, 16 * t + lane % 16) * ld + r0 + rb;
#pragma unroll
                    for (int a2 = 0; a2 < kRT; ++a2)
                    {
                        if constexpr (sizeof(T) == 4)
                        {
                            // rows 4 (lane / 16) .. + 3 of the tile: one 16-byte store
                            *reinterpret_cast<f32x4v *>(col + 16 * a2 + 4 * (lane / 16)) = acc[a2][t];
                        }
                        else
                        {
#pragma unroll
                            for (int vv = 0; vv < 4; ++vv)
                            {
                                col[16 * a2 + Mfma<T>::row(lane, vv)] = acc[a2][t][vv];
                            }
                        }
                    }
                }
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
        // lda and ldv are multiples of kH, hence of the apply slab height; g.apply_wgs workgroups in total
        const int h = apply_rows<T>();
        const int sa = static_cast<int>(g.lda / h), sv = static_cast<int>(g.ldv / h);
        const int per_pair = std::max(1, std::min(sa + sv, (g.apply_wgs + pairs - 1) / pairs));
        hipLaunchKernelGGL(block_apply_kernel<T>, dim3(pairs, per_pair), dim3(256), 0, stream, a, g.lda, sa, v, g.ldv,
                           sv, g.nb, g.Nb, round, static_cast<const T *>(q), static_cast<const int *>(apply_flag),
                           perm, skip_if_zero);
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
