// block_kernels.hpp -- device code of the block one-sided Jacobi kernels (block.hip): MFMA block Gram, the Gram-space
// inner sweep (full circle or cross ordering), MFMA block apply, column ordering and small helpers.  Kept apart
// from the host launchers so that the same source also builds for the CPU model of a workgroup under
// tests/simt_model.
#pragma once

#include "block.hpp"

namespace jsvd
{
    namespace
    {
        typedef dev::vec_t<float, 4> f32x4;
        typedef dev::vec_t<double, 4> f64x4;

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

        // Block pair of slot `slot` in outer round `round` (the round-robin circle over Nb blocks), or, for the
        // intra-block step (round == kIntraRound), the two halves 2 slot and 2 slot + 1 of one block of twice the
        // kernel's width: the kernels at width kb / 2 then sweep every kb-column block by itself.
        constexpr int kIntraRound = -1;

        __device__ inline void block_pair(int slot, int round, int Nb, int &I, int &J)
        {
            if (round == kIntraRound)
            {
                I = 2 * slot;
                J = 2 * slot + 1;
                return;
            }
            I = dev::circle_at(slot, round, Nb);
            J = dev::circle_at(Nb - 1 - slot, round, Nb);
        }

        // Physical column of local column c (0..2 KB - 1) of block pair (I, J) under the column map `perm`
        // (logical column k of the sweep order is stored in physical column perm[k]).
        template <int KB>
        __device__ inline int64_t pair_column(const int *__restrict__ perm, int I, int J, int c)
        {
            return perm[(c < KB) ? (I * KB + c) : (J * KB + c - KB)];
        }

        // Rows of a kH slab staged in LDS per Gram step: all of them, or half at kb = 64 (LDS for 128 FP64
        // columns, registers for the loads in flight).
        template <typename T, int KB>
        constexpr int gram_rows()
        {
            return (KB == 64) ? kH / 2 : kH;
        }

        // Loads rows [r0, r0+GH) of the 2 KB columns [X_I X_J] (physical columns cols[0 .. 2 KB), in LDS) into LDS
        // as xs[c][r] (column-major, padded).
        template <typename T, int KB, int GH>
        __device__ inline void load_slab(const T *__restrict__ x, int64_t ld, const int *cols, int64_t r0,
                                         T (*xs)[GH + 1])
        {
            // 2 KB * GH elements, 256 threads; consecutive threads read consecutive rows of one column
            // (coalesced).  All loads are issued before the LDS stores so they are in flight together.
            constexpr int kPer = 2 * KB * GH / 256;
            T tmp[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                const int e = threadIdx.x + 256 * k;
                tmp[k] = x[static_cast<int64_t>(cols[e / GH]) * ld + r0 + e % GH];
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                const int e = threadIdx.x + 256 * k;
                xs[e / GH][e % GH] = tmp[k];
            }
        }

        // Partial Gram of kH-row slabs [ps * per, min(slabs, (ps + 1) * per)) of block pair `slot`; the partials
        // of a pair are partial[slot][0 .. parts), each 2 KB x 2 KB row-major.  The (2 KB / 16)^2 output tiles are
        // split evenly over the four waves.
        template <typename T, int KB>
        __global__ __launch_bounds__(256) void block_gram_kernel(const T *__restrict__ a, int64_t lda,
                                                                 const int *__restrict__ perm, int nb, int Nb,
                                                                 int round, int slabs, int per, int parts,
                                                                 T *__restrict__ partial,
                                                                 const int *__restrict__ skip_if_zero)
        {
            constexpr int KW = 2 * KB;
            constexpr int GH = gram_rows<T, KB>();
            constexpr int NT = KW / 16;         // tiles per dimension
            constexpr int TPW = NT * NT / 4;    // tiles per wave
            __shared__ T xs[KW][GH + 1];
            __shared__ int cols[KW]; // physical columns of the pair
            const int slot = blockIdx.x;
            const int ps = blockIdx.y;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            if (I >= nb || J >= nb)
            {
                return;
            }
            // The column map and the skip flag are read together (one memory round trip before the slab loads,
            // not two).
            int col = 0;
            if (threadIdx.x < KW)
            {
                col = static_cast<int>(pair_column<KB>(perm, I, J, threadIdx.x));
            }
            const int go = (skip_if_zero != nullptr) ? *skip_if_zero : 1;
            if (threadIdx.x < KW)
            {
                cols[threadIdx.x] = col;
            }
            if (go == 0)
            {
                return;
            }
            __syncthreads();
            const int lane = threadIdx.x % 64;
            const int wave = threadIdx.x / 64;
            // tile k of this wave: row tile rt(k), column tile ct(k) (k is a compile-time index after unrolling,
            // so tiles that share a row share its operand load)
            auto rt = [&](int k) { return (TPW >= NT) ? wave * (TPW / NT) + k / NT : (wave * TPW + k) / NT; };
            auto ct = [&](int k) { return (TPW >= NT) ? k % NT : (wave * TPW + k) % NT; };
            typename Mfma<T>::Acc acc[TPW];
#pragma unroll
            for (int k = 0; k < TPW; ++k)
            {
                acc[k] = typename Mfma<T>::Acc{0, 0, 0, 0};
            }
            const int s_end = min(slabs, (ps + 1) * per);
            bool first = true;
            for (int s = ps * per; s < s_end; ++s)
            {
                for (int h = 0; h < kH; h += GH)
                {
                    if (!first)
                    {
                        __syncthreads(); // every wave is done with the previous rows
                    }
                    first = false;
                    load_slab<T, KB, GH>(a, lda, cols, static_cast<int64_t>(s) * kH + h, xs);
                    __syncthreads();
                    for (int r0 = 0; r0 < GH; r0 += 4)
                    {
                        const int kr = r0 + lane / 16;
#pragma unroll
                        for (int k = 0; k < TPW; ++k)
                        {
                            const T av = xs[16 * rt(k) + lane % 16][kr];
                            const T bv = xs[16 * ct(k) + lane % 16][kr];
                            acc[k] = Mfma<T>::step(av, bv, acc[k]);
                        }
                    }
                }
            }
            T *out = partial + (static_cast<int64_t>(slot) * parts + ps) * (KW * KW);
#pragma unroll
            for (int k = 0; k < TPW; ++k)
            {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int i = 16 * rt(k) + Mfma<T>::row(lane, v);
                    const int j = 16 * ct(k) + lane % 16;
                    out[i * KW + j] = acc[k][v];
                }
            }
        }

        // Reduction of the slab partials and one inner cyclic Jacobi sweep in Gram space.
        //
        // G is symmetric and kept once per unordered entry, in two LDS buffers.  Each inner round pairs the kw
        // columns by the round-robin circle (pair u of a round = circle positions u and kw - 1 - u).  Inner round
        // r is one phase and one barrier:
        //   * the G threads (2x2 blocks (t, u), t <= u, of the kb x kb pair grid; one or more per thread) write
        //     J_t^T G_r J_u into the other buffer, with the rotations of round r published in the previous phase;
        //   * the Q threads apply the same rotations to Q;
        //   * wave 0 meanwhile forms the rotations of round r+1 (lane l: pair l % kb, FP64, reference skip test)
        //     without waiting for G_{r+1}: the three entries a rotation needs are the two new diagonal entries of
        //     its columns and their new off-diagonal entry, each a 2x2 update of entries of G_r with rotations
        //     of round r that wave 0 already holds (exchanged between its lanes), evaluated by the same
        //     expressions as the G threads use.
        // So the FP64 rotation chain of the next round overlaps the bulk update of this one.  Vector-instruction
        // issue bounds the bulk update, so it is packed into as few waves as it needs.  Q is kept in registers in
        // circle-position order: thread (row k, segment s) of the Q threads holds Q[k][col(p)] for the QP positions
        // p = QP s .. QP s + QP - 1 and their partners kw - 1 - p, so every rotation of a round is local to a
        // thread; after a round the positions move by one along the circle (1..kw-1 cyclically, 0 fixed), which
        // moves one value to each neighbouring segment.  Column indices advance incrementally (col(p, r+1) =
        // col(p, r) - 1 on 1..kw-1).  A block pair none of whose column pairs passes the skip test is left
        // untouched.
        //
        // Thread roles per block width (waves are spread round-robin over the CU's four SIMDs, wave w on SIMD
        // w % 4):
        //   kb = 16: 320 threads; wave 0 rotations, waves 1-3 G (136 blocks), wave 4 Q (32 rows x 2 segments).
        //   kb = 32: 1024 threads; wave 0 rotations, G on the waves 1..11 not on its SIMD (1, 2, 3, 5, 6, 7, 9,
        //            10, 11: 528 blocks) so every SIMD issues one Q wave and either the rotation wave or a third
        //            G wave, waves 12-15 Q (64 rows x 4 segments of 8 position pairs).
        //   kb = 64: 1024 threads; every thread holds Q (128 rows x 8 segments of 8 position pairs), wave 0 also
        //            forms the rotations, waves 1-15 also update G (2080 blocks, up to 3 per thread) -- separate
        //            Q waves would need more registers than a 1024-thread workgroup has.
        template <int KB>
        struct EigenShape;
        template <>
        struct EigenShape<16>
        {
            static constexpr int kThreads = 320, kQP = 8, kGThreads = 192;
        };
        template <>
        struct EigenShape<32>
        {
            static constexpr int kThreads = 1024, kQP = 8, kGThreads = 576;
        };
        template <>
        struct EigenShape<64>
        {
            static constexpr int kThreads = 1024, kQP = 8, kGThreads = 960;
        };

        // LDS image of the symmetric kw x kw G, one slot per unordered entry.  The upper triangle of a square
        // array with padded rows (kw + 2 slots at kb <= 32, kw + 1 at kb = 64; measured: kw + 2 takes 5 % off an
        // inner round at kb = 32 in FP32 and FP64 and costs 4 % at kb = 64), except for FP64 at kb = 64, where two
        // buffers of that would not fit: entry {i, j} at row i, offset d = (j - i) mod kw for d < kw / 2 (row
        // min(i, j) for d = kw / 2), else at row j, offset kw - d, in rows of kw / 2 + 2 slots.  There the even
        // row length keeps both address steps that neighbouring lanes take (+1 along a row, row length - 1 across
        // rows) odd, so a wave's accesses spread over all LDS banks (an odd row length makes the second step a
        // multiple of the bank count).
        template <typename T, int KB, int STRIDE = 0>
        struct GramLds
        {
            static constexpr int KW = 2 * KB;
            static constexpr bool kCirc = KB == 64 && sizeof(T) == 8;
            static constexpr int kStride = STRIDE ? STRIDE : (kCirc ? KW / 2 + 2 : (KB == 64 ? KW + 1 : KW + 2));
            static constexpr int kElems = KW * kStride;
            __device__ static int at(int i, int j)
            {
                if constexpr (kCirc)
                {
                    int d = (j - i) & (KW - 1);
                    if (d > KW / 2 || (d == KW / 2 && i > j))
                    {
                        i = j;
                        d = KW - d;
                    }
                    return i * kStride + d;
                }
                else
                {
                    return min(i, j) * kStride + max(i, j);
                }
            }
        };

        // Index of this thread among the G threads, or -1.
        template <int KB>
        __device__ inline int g_thread_index(int tid)
        {
            const int w = tid / 64;
            if constexpr (KB == 32)
            {
                if (w < 1 || w > 11 || (w % 4) == 0)
                {
                    return -1;
                }
                return (w - 1 - w / 4) * 64 + tid % 64;
            }
            else
            {
                // waves 1 .. up to the first Q wave (all waves when every thread holds Q)
                constexpr int kQFirst = EigenShape<KB>::kThreads - 2 * KB * (KB / EigenShape<KB>::kQP);
                constexpr int kGEnd = (kQFirst > 0) ? kQFirst : EigenShape<KB>::kThreads;
                return (w < 1 || tid >= kGEnd) ? -1 : tid - 64;
            }
        }

        // Column at a circle position in the next inner round.
        template <int KW>
        __device__ inline int next_column(int c)
        {
            return c == 0 ? 0 : (c == 1 ? KW - 1 : c - 1);
        }

        // Circle position, in the previous round, of the column now at position p.
        template <int KW>
        __device__ inline int previous_position(int p)
        {
            return p == 0 ? 0 : (p == 1 ? KW - 1 : p - 1);
        }

        // Columns of a G thread's block after an inner round: along the circle, or in the cross ordering the
        // second-block columns step to kb + (c - kb + 1) mod kb while the first-block columns stay.
        template <int KB, bool CROSS>
        __device__ inline void advance_columns(int &x1, int &x2, int &z1, int &z2)
        {
            constexpr int KW = 2 * KB;
            if constexpr (CROSS)
            {
                x2 = (x2 == KW - 1) ? KB : x2 + 1;
                z2 = (z2 == KW - 1) ? KB : z2 + 1;
            }
            else
            {
                x1 = next_column<KW>(x1);
                x2 = next_column<KW>(x2);
                z1 = next_column<KW>(z1);
                z2 = next_column<KW>(z2);
            }
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

        // v moved by the DPP pattern CTRL (0x138: from lane - 1, 0x130: from lane + 1, 0x134: from lane + 1 mod 64,
        // 0x39: quad_perm [1, 2, 3, 0]) on every lane of the wave: the empty asm keeps the compiler from sinking the
        // move into a branch of a later select (a DPP move reads nothing from an inactive lane).
        template <int CTRL, typename T>
        __device__ inline T full_wave_shift(T v)
        {
            T moved = dev::dpp_move<CTRL, 0xF>(v);
            dev::keep(moved);
            return moved;
        }

        // v moved one lane along the whole wave by DPP (CTRL 0x138: from lane - 1, 0x130: from lane + 1), or the
        // lane's own v where `own`.  A DPP move reads its source lane only while that lane is active, so the
        // move runs on every lane before the select (the empty asm keeps the compiler from sinking it into a
        // branch of the select).
        template <int CTRL, typename T>
        __device__ inline T shifted_or_own(T v, bool own)
        {
            T moved = dev::dpp_move<CTRL, 0xF>(v);
            dev::keep(moved);
            return own ? v : moved;
        }

        // J_t^T G J_u of one 2x2 block with the rotations of its two pairs.  A pair that does not rotate carries
        // exactly (c, s) = (1, 0), with which rotate_block returns its inputs unchanged (each product with 0 is a
        // zero and each fused multiply-add by 1 is exact), so no per-pair rotation flag is read: the LDS reads of
        // the block and of the two rotations are the whole of a G thread's inputs.
        template <typename T>
        __device__ inline void rotate_pair_block(T g11, T g12, T g21, T g22, const T (&rt)[2], const T (&ru)[2], T &n11,
                                                 T &n12, T &n21, T &n22)
        {
            rotate_block(rt[0], rt[1], ru[0], ru[1], g11, g12, g21, g22, n11, n12, n21, n22);
        }

        // One inner round on a thread's Q registers: rotate its QP position pairs, then move the positions one
        // step along the circle (values crossing a segment boundary come from the neighbouring lanes).
        template <typename T, int QP, int NSEG>
        __device__ inline void q_round(T (&lo)[QP], T (&hi)[QP], const T (*cs)[2], int seg)
        {
#pragma unroll
            for (int i = 0; i < QP; ++i)
            {
                const T c = cs[QP * seg + i][0], s = cs[QP * seg + i][1];
                const T x = lo[i], z = hi[i];
                lo[i] = c * x - s * z;
                hi[i] = s * x + c * z;
            }
            // position QP s + QP - 1 -> QP (s + 1) (next segment; last segment: position kb, its own hi[QP-1])
            const T out_lo = lo[QP - 1];
            // position kw - 1 - QP s -> kw - QP s (previous segment; segment 0: position 1, its own lo[1])
            const T out_hi = hi[0];
            // lane - 1 and lane + 1 by whole-wave DPP shifts (the Q waves run this with every lane active; the
            // lanes whose shifted value would come from outside the wave do not use it)
            const T from_prev = full_wave_shift<0x138>(out_lo);
            const T from_next = full_wave_shift<0x130>(out_hi);
            const T lo0 = lo[0];
#pragma unroll
            for (int i = QP - 1; i >= 1; --i)
            {
                lo[i] = lo[i - 1];
            }
#pragma unroll
            for (int i = 0; i < QP - 1; ++i)
            {
                hi[i] = hi[i + 1];
            }
            lo[0] = (seg == 0) ? lo0 : from_prev; // position 0 never moves
            hi[QP - 1] = (seg == NSEG - 1) ? out_lo : from_next;
            if (seg == 0)
            {
                lo[1] = out_hi; // position kw - 1 -> 1
            }
        }

        // One inner round of the cross ordering on a thread's Q registers: lo[i] holds Q[k][u] (the first block's
        // column of slot u = QP seg + i, fixed), hi[i] Q[k][kb + (u + r) mod kb] (the second block's column that
        // slot u pairs with in round r).  Rotate, then every slot takes the second-block column of slot u + 1
        // (mod kb): within the thread, and from the next segment of the same row (the NSEG = 4 segments of a row
        // are one aligned quad of lanes, rotated by a DPP quad permutation).
        template <typename T, int QP, int NSEG>
        __device__ inline void q_round_cross(T (&lo)[QP], T (&hi)[QP], const T (*cs)[2], int seg)
        {
            static_assert(NSEG == 4, "segments of a Q row form one quad of lanes");
#pragma unroll
            for (int i = 0; i < QP; ++i)
            {
                const T c = cs[QP * seg + i][0], s = cs[QP * seg + i][1];
                const T x = lo[i], z = hi[i];
                lo[i] = c * x - s * z;
                hi[i] = s * x + c * z;
            }
            // quad_perm [1, 2, 3, 0]: lane 4 q + j reads lane 4 q + (j + 1) mod 4 (the Q waves run this with every
            // lane active)
            const T from_next = full_wave_shift<0x39>(hi[0]);
#pragma unroll
            for (int i = 0; i < QP - 1; ++i)
            {
                hi[i] = hi[i + 1];
            }
            hi[QP - 1] = from_next;
        }

        template <typename T>
        __device__ inline bool block_rotation(double app, double aqq, double apq, double eps, double &c, double &s)
        {
            // The FP32 phase applies its rotations in FP32 to FP32 Gram entries, so the short FP64 chain (hardware
            // reciprocal square roots with one Newton step, no scaling) is exact enough there; the FP64 phase
            // keeps the division form.
            if constexpr (sizeof(T) == 4)
            {
                return dev::rotation_params_f32(app, aqq, apq, eps, c, s);
            }
            else
            {
                return dev::rotation_params_block(app, aqq, apq, eps, c, s);
            }
        }

        template <typename T, int KB, bool CROSS>
        __global__ __launch_bounds__(EigenShape<KB>::kThreads) void block_eigen_kernel(
            const T *__restrict__ partial, int slabs, int nb, int Nb, int round, double eps, T *__restrict__ q_out,
            int *__restrict__ apply_flag, int *__restrict__ rot_count, unsigned long long *__restrict__ offmax,
            const int *__restrict__ perm, double *__restrict__ norm2, const int *__restrict__ skip_if_zero)
        {
            constexpr int KW = 2 * KB;
            constexpr int kPairs = KB;
            constexpr int kThreads = EigenShape<KB>::kThreads;
            constexpr int kQP = EigenShape<KB>::kQP;
            constexpr int kSeg = kPairs / kQP;
            constexpr int kQThreads = KW * kSeg;
            constexpr int kQFirst = kThreads - kQThreads;
            constexpr int kTri = kPairs * (kPairs + 1) / 2;
            constexpr int kGThreads = EigenShape<KB>::kGThreads;
            constexpr int kBPT = (kTri + kGThreads - 1) / kGThreads; // G blocks per G thread
            static_assert(kQFirst % 64 == 0 && 64 % kSeg == 0 && kPairs <= 64, "eigen thread roles");
            static_assert(!CROSS || KB == 32, "the cross ordering is built for kb = 32");
            // Inner rounds: the full cyclic sweep over the kw columns (circle of kw positions), or the cross
            // ordering (slot u pairs column u of the first block with column kb + (u + r) mod kb of the second).
            constexpr int kRounds = CROSS ? KB : KW - 1;
            // the cross ordering's access pattern: rows of kw + 5 slots (a sweep of kw + 1 .. kw + 9 on the device:
            // kw + 5, kw + 6 and kw + 9 about 2 % faster than kw + 2, kw + 8 10 % slower)
            using L = GramLds<T, KB, CROSS ? KW + 5 : 0>;
            __shared__ T g[2][L::kElems];               // G of the current round and of the next
            __shared__ T cs[2][kPairs][2];              // rotations of the current round and of the next
            __shared__ double red[kThreads / 64];
            __shared__ int any;
            __shared__ int needed;
            // The skip flag is read after the first partials' loads are issued (both in flight together: one
            // memory round trip at the start, not two) and tested once the partials are summed.
            int go = 1;
            const int slot = blockIdx.x;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            if (I >= nb || J >= nb)
            {
                return;
            }
            const int tid = threadIdx.x;
            // This thread's 2x2 blocks (t, u), t <= u, of the upper triangle of the kb x kb pair grid.
            const int gi0 = g_thread_index<KB>(tid);
            const bool gthr = gi0 >= 0;
            int btu[kBPT]; // t | u << 8, or -1 where the thread has no k-th block
#pragma unroll
            for (int k = 0; k < kBPT; ++k)
            {
                const int gi = gi0 + k * kGThreads;
                int t = 0, rem = (gthr && gi < kTri) ? gi : 0;
                while (rem >= kPairs - t)
                {
                    rem -= kPairs - t;
                    ++t;
                }
                btu[k] = (gthr && gi < kTri) ? (t | ((t + rem) << 8)) : -1;
            }
            const T *base = partial + static_cast<int64_t>(slot) * slabs * (KW * KW);
            if constexpr (KW * KW == 4 * kThreads)
            {
                // four consecutive entries of one row per thread, as 4-element vector loads (eight slabs in
                // flight per step); the summation order per entry is the slab order, as below
                using V4 = dev::vec_t<T, 4>;
                const int e0 = 4 * tid;
                V4 acc = V4{0, 0, 0, 0};
                {
                    // the first eight slabs (the last one repeated where there are fewer), then the skip flag
                    V4 part[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        const int sk = min(k, slabs - 1);
                        part[k] = *reinterpret_cast<const V4 *>(base + static_cast<int64_t>(sk) * (KW * KW) + e0);
                    }
                    go = (skip_if_zero != nullptr) ? *skip_if_zero : 1;
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        if (k < slabs)
                        {
                            acc += part[k];
                        }
                    }
                }
                int s = 8;
                for (; s + 8 <= slabs; s += 8)
                {
                    V4 part[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        part[k] = *reinterpret_cast<const V4 *>(base + static_cast<int64_t>(s + k) * (KW * KW) + e0);
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        acc += part[k];
                    }
                }
                for (; s < slabs; ++s)
                {
                    acc += *reinterpret_cast<const V4 *>(base + static_cast<int64_t>(s) * (KW * KW) + e0);
                }
                const int i = e0 / KW;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                {
                    const int j = e0 % KW + k;
                    if (i <= j)
                    {
                        g[0][L::at(i, j)] = acc[k];
                    }
                }
            }
            else
            {
                go = (skip_if_zero != nullptr) ? *skip_if_zero : 1;
                for (int e = tid; e < KW * KW; e += kThreads)
                {
                    const int i = e / KW;
                    const int j = e % KW;
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
                                part[k] = base[static_cast<int64_t>(s + k) * (KW * KW) + e];
                            }
    #pragma unroll
                            for (int k = 0; k < 8; ++k)
                            {
                                acc += part[k];
                            }
                        }
                        for (; s < slabs; ++s)
                        {
                            acc += base[static_cast<int64_t>(s) * (KW * KW) + e];
                        }
                        g[0][L::at(i, j)] = acc;
                    }
                }
            }
            if (go == 0)
            {
                return;
            }
            if (tid == 0)
            {
                any = 0;
                needed = 0;
            }
            // Q = I in circle-position order (round 0 positions).
            const bool qthr = tid >= kQFirst;
            const int qk = (tid - kQFirst) / kSeg;
            const int seg = (tid - kQFirst) % kSeg;
            T qlo[kQP], qhi[kQP];
#pragma unroll
            for (int i = 0; i < kQP; ++i)
            {
                if constexpr (CROSS)
                {
                    qlo[i] = (kQP * seg + i == qk) ? T(1) : T(0);
                    qhi[i] = (KB + kQP * seg + i == qk) ? T(1) : T(0);
                }
                else
                {
                    qlo[i] = (dev::circle_at_fast(kQP * seg + i, 0, KW) == qk) ? T(1) : T(0);
                    qhi[i] = (dev::circle_at_fast(KW - 1 - kQP * seg - i, 0, KW) == qk) ? T(1) : T(0);
                }
            }
            __syncthreads();
            // Off-orthogonality of the pair's columns before the sweep (convergence measure), and whether any
            // pair passes the skip test at all: if none does, no rotation can happen in the sweep (G never
            // changes) and the sweep is skipped outright.
            // Both on squares (no square root per pair): the largest squared cosine gij^2 / (gii gjj), and the skip
            // test gij^2 > eps^2 gii gjj.  The cross ordering checks only the pairs it rotates (one per thread at
            // kb = 32); the pairs inside each block are checked by the intra-block step of the same sweep.
            double best = 0.0;
            bool need = false;
            auto scan = [&](int i, int j) {
                const double gii = fmax(static_cast<double>(g[0][L::at(i, i)]), 0.0);
                const double gjj = fmax(static_cast<double>(g[0][L::at(j, j)]), 0.0);
                const double gij = static_cast<double>(g[0][L::at(i, j)]);
                const double den = gii * gjj;
                const double g2 = gij * gij;
                best = (den > 0.0) ? fmax(best, g2 / den) : best;
                need = need || (g2 > (eps * eps) * den && fabs(gij) > 1.0e-300);
            };
            if constexpr (CROSS)
            {
                static_assert(KB * KB == kThreads, "one cross pair per thread");
                scan(tid / KB, KB + tid % KB);
            }
            else
            {
                for (int e = tid; e < KW * KW; e += kThreads)
                {
                    const int i = e / KW;
                    const int j = e % KW;
                    if (i < j)
                    {
                        scan(i, j);
                    }
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
            // Columns at the blocks' circle positions, advanced incrementally; FP64 at kb = 64 recomputes them per
            // round instead, which keeps the registers for Q.
            constexpr bool kTrack = !(KB == 64 && sizeof(T) == 8);
            constexpr int kTB = kTrack ? kBPT : 1;
            int x1[kTB], x2[kTB], z1[kTB], z2[kTB];
            if constexpr (kTrack)
            {
#pragma unroll
                for (int k = 0; k < kBPT; ++k)
                {
                    const int t = btu[k] & 255, u = (btu[k] >> 8) & 255;
                    if constexpr (CROSS)
                    {
                        x1[k] = t;
                        x2[k] = KB + t;
                        z1[k] = u;
                        z2[k] = KB + u;
                    }
                    else
                    {
                        x1[k] = dev::circle_at_fast(t, 0, KW);
                        x2[k] = dev::circle_at_fast(KW - 1 - t, 0, KW);
                        z1[k] = dev::circle_at_fast(u, 0, KW);
                        z2[k] = dev::circle_at_fast(KW - 1 - u, 0, KW);
                    }
                }
            }
            const bool run = needed != 0;
            // The rotation wave's chain (its reads of G, the FP64 parameters, the write of the next rotations) is
            // the dependent path of every inner round; the other waves' LDS traffic queues behind its requests
            // only if it issues first, so it runs at the highest issue priority.
            if (tid < 64)
            {
                __builtin_amdgcn_s_setprio(3);
            }
            // Wave 0: the rotation of pair pu of the current round and the three G entries it was formed from.
            const int pu = tid % kPairs;
            T my_c = 1, my_s = 0, gpp = 0, gqq = 0, gpq = 0;
            if (run && tid < 64)
            {
                const int p1 = CROSS ? pu : dev::circle_at_fast(pu, 0, KW);
                const int p2 = CROSS ? KB + pu : dev::circle_at_fast(KW - 1 - pu, 0, KW);
                gpp = g[0][L::at(p1, p1)];
                gqq = g[0][L::at(p2, p2)];
                gpq = g[0][L::at(p1, p2)];
                double c, s;
                const bool rot = block_rotation<T>(static_cast<double>(gpp), static_cast<double>(gqq),
                                                   static_cast<double>(gpq), eps, c, s);
                my_c = static_cast<T>(c);
                my_s = static_cast<T>(s);
                if (tid < kPairs)
                {
                    cs[0][tid][0] = my_c;
                    cs[0][tid][1] = my_s;
                    my_rotations += rot ? 1 : 0;
                }
            }
            __syncthreads();
            for (int rr = 0; run && rr < kRounds; ++rr)
            {
                const int b = rr & 1;
                const T *gc = g[b];
                T *gn = g[b ^ 1];
                if (gthr)
                {
                    if constexpr (kBPT == 1)
                    {
                        if (btu[0] >= 0)
                        {
                            const int t = btu[0] & 255, u = btu[0] >> 8;
                            const int c1 = x1[0], c2 = x2[0], c3 = z1[0], c4 = z2[0];
                            // J_t^T G J_u of the block into the other buffer
                            const int o11 = L::at(c1, c3);
                            const int o12 = L::at(c1, c4);
                            const int o21 = L::at(c2, c3);
                            const int o22 = L::at(c2, c4);
                            T n11, n12, n21, n22;
                            rotate_pair_block<T>(gc[o11], gc[o12], gc[o21], gc[o22], cs[b][t], cs[b][u], n11, n12,
                                                 n21, n22);
                            gn[o11] = n11;
                            gn[o12] = n12;
                            gn[o22] = n22;
                            if (t != u)
                            {
                                gn[o21] = n21; // on the diagonal block entry 21 aliases entry 12
                            }
                            advance_columns<KB, CROSS>(x1[0], x2[0], z1[0], z2[0]);
                        }
                    }
                    else
                    {
#pragma unroll
                        for (int k = 0; k < kBPT; ++k)
                        {
                            if (btu[k] >= 0)
                            {
                                const int t = btu[k] & 255, u = btu[k] >> 8;
                                int c1, c2, c3, c4;
                                if constexpr (kTrack)
                                {
                                    c1 = x1[k];
                                    c2 = x2[k];
                                    c3 = z1[k];
                                    c4 = z2[k];
                                }
                                else
                                {
                                    static_assert(!CROSS, "the cross ordering tracks its columns");
                                    c1 = dev::circle_at_fast(t, rr, KW);
                                    c2 = dev::circle_at_fast(KW - 1 - t, rr, KW);
                                    c3 = dev::circle_at_fast(u, rr, KW);
                                    c4 = dev::circle_at_fast(KW - 1 - u, rr, KW);
                                }
                                const int o11 = L::at(c1, c3);
                                const int o12 = L::at(c1, c4);
                                const int o21 = L::at(c2, c3);
                                const int o22 = L::at(c2, c4);
                                T n11, n12, n21, n22;
                                rotate_pair_block<T>(gc[o11], gc[o12], gc[o21], gc[o22], cs[b][t], cs[b][u], n11,
                                                     n12, n21, n22);
                                gn[o11] = n11;
                                gn[o12] = n12;
                                gn[o22] = n22;
                                if (t != u)
                                {
                                    gn[o21] = n21;
                                }
                                if constexpr (kTrack)
                                {
                                    advance_columns<KB, CROSS>(x1[k], x2[k], z1[k], z2[k]);
                                }
                            }
                        }
                    }
                }
                if (qthr)
                {
                    if constexpr (CROSS)
                    {
                        q_round_cross<T, kQP, kSeg>(qlo, qhi, cs[b], seg);
                    }
                    else
                    {
                        q_round<T, kQP, kSeg>(qlo, qhi, cs[b], seg);
                    }
                }
                if (CROSS && tid < 64 && rr + 1 < kRounds)
                {
                    // The pair of slot pu in round rr+1: column pu (its own first-block column, new diagonal entry
                    // d11 of its rotated diagonal block) and the second-block column that slot pu + 1 (mod kb) held
                    // in round rr (that slot's new d22); their new off-diagonal entry is entry 12 of block
                    // (pu, pu + 1) of round rr as the G threads compute it.  The neighbour's values come by a
                    // whole-wave one-lane DPP rotation (wave_rol:1; lanes >= kb repeat lanes tid % kb, so lane kb - 1
                    // reads pair 0).
                    const int ta = pu, tb = (pu + 1) % kPairs;
                    const int r1 = ta, r2 = KB + (ta + rr) % KB, k1 = tb, k2 = KB + (tb + rr) % KB;
                    const T e11 = gc[L::at(r1, k1)], e12 = gc[L::at(r1, k2)];
                    const T e21 = gc[L::at(r2, k1)], e22 = gc[L::at(r2, k2)];
                    T d11, d12, d21, d22;
                    rotate_block(my_c, my_s, my_c, my_s, gpp, gpq, gpq, gqq, d11, d12, d21, d22);
                    const T b2 = full_wave_shift<0x134>(d22);
                    const T c2 = full_wave_shift<0x134>(my_c);
                    const T s2 = full_wave_shift<0x134>(my_s);
                    T f11, f12, f21, f22;
                    rotate_block(my_c, my_s, c2, s2, e11, e12, e21, e22, f11, f12, f21, f22);
                    gpp = d11;
                    gqq = b2;
                    gpq = f12;
                    double c, s;
                    const bool rot = block_rotation<T>(static_cast<double>(gpp), static_cast<double>(gqq),
                                                       static_cast<double>(gpq), eps, c, s);
                    my_c = static_cast<T>(c);
                    my_s = static_cast<T>(s);
                    if (tid < kPairs)
                    {
                        cs[b ^ 1][tid][0] = my_c;
                        cs[b ^ 1][tid][1] = my_s;
                        my_rotations += rot ? 1 : 0;
                    }
                }
                if (!CROSS && tid < 64 && rr + 1 < kRounds)
                {
                    // The pair of round rr+1: the columns now at positions pu and kw - 1 - pu came from positions
                    // pa and pb, i.e. from pairs ta (first column when pa < kb) and tb of round rr (ta != tb).
                    const int pa = previous_position<KW>(pu), pb = previous_position<KW>(KW - 1 - pu);
                    const int ta = pa < kPairs ? pa : KW - 1 - pa, tb = pb < kPairs ? pb : KW - 1 - pb;
                    const bool fa = pa < kPairs, fb = pb < kPairs;
                    // ta = pu - 1 and tb = pu + 1, except ta = pu at pu = 0 and tb = pu at pu = kb - 1, so ta < tb
                    // always.  Their new off-diagonal entry is block (ta, tb) of round rr as the G threads compute
                    // it; its four G entries are read first, so the LDS reads overlap the lane moves below.
                    const int r1 = dev::circle_at_fast(ta, rr, KW), r2 = dev::circle_at_fast(KW - 1 - ta, rr, KW);
                    const int k1 = dev::circle_at_fast(tb, rr, KW), k2 = dev::circle_at_fast(KW - 1 - tb, rr, KW);
                    const T e11 = gc[L::at(r1, k1)], e12 = gc[L::at(r1, k2)];
                    const T e21 = gc[L::at(r2, k1)], e22 = gc[L::at(r2, k2)];
                    // New diagonal entries of this lane's pair (its diagonal block; exact copy when not rotating).
                    T d11, d12, d21, d22;
                    rotate_block(my_c, my_s, my_c, my_s, gpp, gpq, gpq, gqq, d11, d12, d21, d22);
                    // The diagonal entries and rotations of pairs ta and tb come from the neighbouring lanes by
                    // whole-wave one-lane DPP shifts (wave_shr:1, wave_shl:1), not LDS-crossbar shuffles.  Lanes
                    // >= kb repeat lanes pu = tid % kb and take the same values.
                    const bool first = pu == 0, last = pu == kPairs - 1;
                    const T a1 = shifted_or_own<0x138>(d11, first);
                    const T a2 = shifted_or_own<0x138>(d22, first);
                    const T b1 = shifted_or_own<0x130>(d11, last);
                    const T b2 = shifted_or_own<0x130>(d22, last);
                    const T c1 = shifted_or_own<0x138>(my_c, first);
                    const T s1 = shifted_or_own<0x138>(my_s, first);
                    const T c2 = shifted_or_own<0x130>(my_c, last);
                    const T s2 = shifted_or_own<0x130>(my_s, last);
                    T f11, f12, f21, f22;
                    rotate_block(c1, s1, c2, s2, e11, e12, e21, e22, f11, f12, f21, f22);
                    gpp = fa ? a1 : a2;
                    gqq = fb ? b1 : b2;
                    // the column from pair ta indexes the row of the block, the one from pair tb its column
                    gpq = fa ? (fb ? f11 : f12) : (fb ? f21 : f22);
                    double c, s;
                    const bool rot = block_rotation<T>(static_cast<double>(gpp), static_cast<double>(gqq),
                                                       static_cast<double>(gpq), eps, c, s);
                    my_c = static_cast<T>(c);
                    my_s = static_cast<T>(s);
                    if (tid < kPairs)
                    {
                        cs[b ^ 1][tid][0] = my_c;
                        cs[b ^ 1][tid][1] = my_s;
                        my_rotations += rot ? 1 : 0;
                    }
                }
                __syncthreads();
            }
            if (tid < kPairs && my_rotations > 0)
            {
                atomicAdd(&any, my_rotations);
            }
            __syncthreads();
            // After kw - 1 moves the circle is back at its round-0 positions (after kb shifts in the cross
            // ordering every slot holds its round-0 second-block column again).
            T *qo = q_out + static_cast<int64_t>(slot) * (KW * KW);
            if (qthr)
            {
#pragma unroll
                for (int i = 0; i < kQP; ++i)
                {
                    if constexpr (CROSS)
                    {
                        qo[qk * KW + kQP * seg + i] = qlo[i]; // row-major Q[c][j]
                        qo[qk * KW + KB + kQP * seg + i] = qhi[i];
                    }
                    else
                    {
                        qo[qk * KW + dev::circle_at_fast(kQP * seg + i, 0, KW)] = qlo[i]; // row-major Q[c][j]
                        qo[qk * KW + dev::circle_at_fast(KW - 1 - kQP * seg - i, 0, KW)] = qhi[i];
                    }
                }
            }
            // squared norms of the pair's columns after the rotation (the diagonal of Q^T G Q), for the
            // norm ordering of the next sweep
            if (norm2 != nullptr && tid < KW)
            {
                // inner round rr writes buffer (rr + 1) & 1, so the last one (kRounds - 1) leaves kRounds & 1
                norm2[pair_column<KB>(perm, I, J, tid)] = static_cast<double>(g[run ? (kRounds & 1) : 0][L::at(tid, tid)]);
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
                    for (int w = 0; w < kThreads / 64; ++w)
                    {
                        m = fmax(m, red[w]);
                    }
                    dev::atomic_max_nonneg(offmax, sqrt(m)); // the largest cosine
                }
            }
        }

        // Launch shape of the apply kernel per type and block width.  Waves split the slab's rows (kb <= 32:
        // wave w computes rows [w H / 4, (w + 1) H / 4) x all kw columns) or, at kb = 64, its columns (wave w
        // computes all H rows x columns [w kw / 4, (w + 1) kw / 4)), so each wave's B operands (its columns of Q)
        // fit in registers.  H = slab rows: 128 in FP32 (64 at kb = 64), 64 in FP64 (32 at kb = 64), sized for
        // two workgroups per CU (LDS and registers).
        template <typename T, int KB>
        struct ApplyShape
        {
            static constexpr int KW = 2 * KB;
            static constexpr bool kColSplit = KB == 64;
            static constexpr int H = (sizeof(T) == 4) ? (KB == 64 ? 64 : 128) : (KB == 64 ? 32 : 64);
            static constexpr bool kQRegs = sizeof(T) == 4 || kColSplit;
            static constexpr int NT = KW / 16;
            static constexpr int RT = kColSplit ? H / 16 : H / 64; // 16-row tiles per wave
            static constexpr int CT = kColSplit ? NT / 4 : NT;     // 16-column tiles per wave
        };

        template <typename T>
        struct Vec16; // 16-byte vector of T
        typedef dev::vec_t<float, 4> f32x4v;
        typedef dev::vec_t<double, 2> f64x2v;
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

        // Y = X Q for the row slabs of [X_I X_J] (the slabs of A, then those of V, H rows each).  Workgroup y of a
        // pair takes slabs y, y + gridDim.y, ...: Q is read once per workgroup (kept in registers as MFMA B
        // operands, or in LDS for FP64 at kb <= 32), the next slab's global loads are in flight while the current
        // slab's MFMAs run, and the product is stored straight from the accumulators.
        template <typename T, int KB>
        __global__ __launch_bounds__(256) void block_apply_kernel(T *__restrict__ a, int64_t lda, int slabs_a,
                                                                  T *__restrict__ v, int64_t ldv, int slabs_v,
                                                                  int nb, int Nb, int round,
                                                                  const T *__restrict__ q_in,
                                                                  const int *__restrict__ apply_flag,
                                                                  const int *__restrict__ perm,
                                                                  const int *__restrict__ skip_if_zero)
        {
            using S = ApplyShape<T, KB>;
            constexpr int KW = S::KW;
            constexpr int H = S::H;
            constexpr int RT = S::RT;
            constexpr int CT = S::CT;
            constexpr bool kQRegs = S::kQRegs;
            constexpr int kVec = 16 / static_cast<int>(sizeof(T)); // elements per 16-byte vector
            constexpr int kLd = H + kVec;                           // LDS row stride: 16-byte aligned rows
            constexpr int kColVecs = H / kVec;                      // vectors per column segment
            constexpr int kPer = KW * kColVecs / 256;               // vectors per thread per slab
            using V = typename Vec16<T>::type;
            __shared__ __align__(16) T xs[KW][kLd];
            __shared__ T qs[kQRegs ? 1 : KW][KW + 1];
            const int slot = blockIdx.x;
            int I, J;
            block_pair(slot, round, Nb, I, J);
            const int total = slabs_a + slabs_v;
            int s = blockIdx.y;
            if (I >= nb || J >= nb || s >= total)
            {
                return;
            }
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            const int rb = S::kColSplit ? 0 : wave * 16 * RT; // first output row of this wave
            const int cb = S::kColSplit ? wave * CT : 0;      // first output column tile of this wave
            const T *qg = q_in + static_cast<int64_t>(slot) * (KW * KW); // row-major Q[c][j]
            T bq[kQRegs ? KW / 4 : 1][kQRegs ? CT : 1];
            if constexpr (kQRegs)
            {
#pragma unroll
                for (int kk = 0; kk < KW / 4; ++kk)
                {
#pragma unroll
                    for (int t = 0; t < CT; ++t)
                    {
                        bq[kk][t] = qg[(4 * kk + lane / 16) * KW + 16 * (cb + t) + lane % 16];
                    }
                }
            }
            constexpr int kQPer = kQRegs ? 1 : KW * KW / 256; // Q entries per thread on their way to LDS
            T qv[kQPer];
            if constexpr (!kQRegs)
            {
#pragma unroll
                for (int k = 0; k < kQPer; ++k)
                {
                    qv[k] = qg[tid + 256 * k];
                }
            }
            // This thread's vectors of a slab: column (tid + 256 k) / kColVecs, rows ((tid + 256 k) % kColVecs) * kVec
            // .. + kVec; the row offset is the same for every k (256 is a multiple of kColVecs).
            static_assert(256 % kColVecs == 0, "apply load layout");
            static_assert(kQRegs || (KW * KW) % 256 == 0, "apply Q layout");
            int col_of[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                col_of[k] = static_cast<int>(pair_column<KB>(perm, I, J, (tid + 256 * k) / kColVecs));
            }
            // Q, the column map, the skip flag and the pair's rotation flag are read together: one memory round
            // trip before the slab loads, not three.  The flags are loaded unconditionally (with no skip flag: the
            // rotation flag twice), so that neither load has a block of its own to wait in; the empty asm keeps
            // the vector loads ahead of the test.
            const int flag_word = *((skip_if_zero != nullptr) ? skip_if_zero : apply_flag + slot);
            const int rotated = apply_flag[slot];
            if constexpr (kQRegs)
            {
#pragma unroll
                for (int kk = 0; kk < KW / 4; ++kk)
                {
#pragma unroll
                    for (int t = 0; t < CT; ++t)
                    {
                        dev::keep(bq[kk][t]);
                    }
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < kQPer; ++k)
                {
                    dev::keep(qv[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                dev::keep(col_of[k]);
            }
            if (flag_word == 0 || rotated == 0)
            {
                return;
            }
            if constexpr (!kQRegs)
            {
#pragma unroll
                for (int k = 0; k < kQPer; ++k)
                {
                    qs[(tid + 256 * k) / KW][(tid + 256 * k) % KW] = qv[k];
                }
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
                const int sn = s + static_cast<int>(gridDim.y);
                if (sn < total)
                {
                    // next slab of this workgroup, in flight during this slab's products (a workgroup with one
                    // slab issues no second load that its end would have to wait for)
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
                typename Mfma<T>::Acc acc[RT][CT];
#pragma unroll
                for (int a2 = 0; a2 < RT; ++a2)
                {
#pragma unroll
                    for (int t = 0; t < CT; ++t)
                    {
                        acc[a2][t] = typename Mfma<T>::Acc{0, 0, 0, 0};
                    }
                }
#pragma unroll
                for (int kk = 0; kk < KW / 4; ++kk)
                {
                    const int kc = 4 * kk + lane / 16;
                    T av[RT];
#pragma unroll
                    for (int a2 = 0; a2 < RT; ++a2)
                    {
                        av[a2] = xs[kc][rb + 16 * a2 + lane % 16];
                    }
#pragma unroll
                    for (int t = 0; t < CT; ++t)
                    {
                        T bv;
                        if constexpr (kQRegs)
                        {
                            bv = bq[kk][t];
                        }
                        else
                        {
                            bv = qs[kc][16 * (cb + t) + lane % 16];
                        }
#pragma unroll
                        for (int a2 = 0; a2 < RT; ++a2)
                        {
                            acc[a2][t] = Mfma<T>::step(av[a2], bv, acc[a2][t]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < CT; ++t)
                {
                    T *col = x + pair_column<KB>(perm, I, J, 16 * (cb + t) + lane % 16) * ld + r0 + rb;
#pragma unroll
                    for (int a2 = 0; a2 < RT; ++a2)
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

        // cont = 1 while a phase goes on after the sweep whose rotation count is *rot: more than `stop` rotations
        // (one thread; the kernels of the next sweep exit when cont is 0).
        __global__ void sweep_continue_kernel(const int *__restrict__ rot, int stop, int *__restrict__ cont)
        {
            if (threadIdx.x == 0)
            {
                *cont = (*rot > stop) ? 1 : 0;
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
} // namespace jsvd
