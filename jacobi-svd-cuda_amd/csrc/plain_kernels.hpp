// plain_kernels.hpp -- device code of the plain FP64 round kernels (launched from plain_f64.hip): the fused one-round
// kernel and the round-span kernel that runs several rounds of the round-robin circle in one launch.  Kept apart
// from the host launchers so that the same source also builds for the CPU model of a workgroup under
// tests/simt_model.
#pragma once

#include "common.hpp"

#include <vector>

namespace jsvd::plain
{
    // Block-wide sum of three doubles.  Deterministic: a fixed DPP row-move tree inside each wavefront (no LDS
    // crossbar round trips on the reduction's dependent chain), then the per-wave totals are added in wave order.
    // Every thread returns the totals.
    template <int T>
    __device__ inline void block_sum3(double &x, double &y, double &z, double (*scratch)[3])
    {
        x = dev::wave_sum_dpp(x);
        y = dev::wave_sum_dpp(y);
        z = dev::wave_sum_dpp(z);
        if constexpr (T > 64)
        {
            constexpr int W = T / 64;
            const int wave = threadIdx.x / 64;
            if ((threadIdx.x % 64) == 0)
            {
                scratch[wave][0] = x;
                scratch[wave][1] = y;
                scratch[wave][2] = z;
            }
            __syncthreads();
            double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
            for (int w = 0; w < W; ++w)
            {
                sx += scratch[w][0];
                sy += scratch[w][1];
                sz += scratch[w][2];
            }
            x = sx;
            y = sy;
            z = sz;
        }
    }

    // The Gram entries of one double2 row chunk of a column pair added to (spp, sqq, spq), by explicit fused
    // multiply-adds: every kernel that accumulates chunks in the same order gets the same bits.
    __device__ inline void gram_acc(double &spp, double &sqq, double &spq, const double2 &x, const double2 &y)
    {
        spp = fma(x.y, x.y, fma(x.x, x.x, spp));
        sqq = fma(y.y, y.y, fma(y.x, y.x, sqq));
        spq = fma(x.y, y.y, fma(x.x, y.x, spq));
    }

    // (x, y) <- (c x - s y, s x + c y) elementwise, by explicit fused multiply-adds (same bits in every kernel).
    __device__ inline void rotate2(double2 &x, double2 &y, double c, double s)
    {
        const double2 px = x;
        const double2 py = y;
        x.x = fma(c, px.x, -s * py.x);
        x.y = fma(c, px.y, -s * py.y);
        y.x = fma(s, px.x, c * py.x);
        y.y = fma(s, px.y, c * py.y);
    }

    __device__ inline double2 zero2()
    {
        double2 z;
        z.x = 0.0;
        z.y = 0.0;
        return z;
    }

    // A loaded chunk kept ahead of a later branch (dev::keep on both halves).
    __device__ inline void keep2(double2 &v)
    {
        dev::keep(v.x);
        dev::keep(v.y);
    }

    // Chunk idx of a column of len2 double2 chunks (odd_tail: the .y of the last one is padding), as zero where
    // idx >= len2 or !use.  The load itself is unconditional -- chunk 0 of the column where the chunk is not
    // used -- so a run of these issues every load before the first wait: a bounds branch around each load made
    // the compiler wait for every load before issuing the next.  fix_chunk applies the zeroing afterwards.
    __device__ inline double2 raw_chunk(const double2 *col, int64_t idx, int64_t len2, bool use)
    {
        return col[(use && idx < len2) ? idx : 0];
    }

    __device__ inline double2 fix_chunk(double2 v, int64_t idx, int64_t len2, bool odd_tail, bool use)
    {
        const bool in = use && idx < len2;
        double2 r;
        r.x = in ? v.x : 0.0;
        r.y = (in && !(odd_tail && idx == len2 - 1)) ? v.y : 0.0;
        return r;
    }

    // Streams a column pair through registers applying (c, s); the padding half of an odd length is read and
    // written as zero.
    __device__ inline void rotate_stream(double2 *xp, double2 *yp, int64_t len2, bool odd_tail, double c, double s,
                                         int tid, int T)
    {
        for (int64_t idx = tid; idx < len2; idx += T)
        {
            double2 x = xp[idx];
            double2 y = yp[idx];
            if (odd_tail && idx == len2 - 1)
            {
                x.y = 0.0;
                y.y = 0.0;
            }
            rotate2(x, y, c, s);
            xp[idx] = x;
            yp[idx] = y;
        }
    }

    // Register-resident fused round.  Each workgroup owns one slot of the round-robin circle, i.e. one column
    // pair; EA / EV double2 chunks per thread hold the pair's A and V columns.  The grid covers the N/2 slots of
    // the round (N = n rounded up to even).  EA == 0 selects the streaming variant for columns that do not fit in
    // registers: a Gram pass, then an apply pass that re-reads the columns.
    template <int T, int EA, int EV>
    __global__ __launch_bounds__(T) void round_kernel(double *__restrict__ a, int64_t lda, int64_t m,
                                                      double *__restrict__ v, int64_t ldv, int64_t n, int N,
                                                      int round, double eps, int *__restrict__ rot_count,
                                                      unsigned long long *__restrict__ offmax,
                                                      const int *__restrict__ skip_if_zero, int count_slots)
    {
        __shared__ double scratch[T / 64 > 0 ? T / 64 : 1][3];
        // the pair's slot of the rotation counter and the cosine maximum (count_slots is a power of two):
        // spreading the one atomic per workgroup over slots keeps same-address atomics from serialising
        const int cslot = static_cast<int>(blockIdx.x) & (count_slots - 1);
        if (EA == 0 && skip_if_zero != nullptr && *skip_if_zero == 0)
        {
            return;
        }
        const int slot = blockIdx.x;
        const int p = dev::circle_at(slot, round, N);
        const int q = dev::circle_at(N - 1 - slot, round, N);
        if (p >= n || q >= n)
        {
            return; // the dummy column of an odd column count
        }
        const int tid = threadIdx.x;
        const int64_t m2 = (m + 1) / 2;   // double2 chunks holding the m rows
        const bool odd_tail = (m & 1) != 0; // the .y of chunk m2-1 is padding
        const int64_t n2 = (n + 1) / 2;
        const bool odd_tail_v = (n & 1) != 0;
        double2 *colp = reinterpret_cast<double2 *>(a + static_cast<int64_t>(p) * lda);
        double2 *colq = reinterpret_cast<double2 *>(a + static_cast<int64_t>(q) * lda);
        double2 *vp = reinterpret_cast<double2 *>(v + static_cast<int64_t>(p) * ldv);
        double2 *vq = reinterpret_cast<double2 *>(v + static_cast<int64_t>(q) * ldv);

        double spp = 0.0, sqq = 0.0, spq = 0.0;
        if constexpr (EA > 0)
        {
            double2 ap[EA], aq[EA], xv[EV > 0 ? EV : 1], yv[EV > 0 ? EV : 1];
            // every load of the pair's columns in flight at once; V is requested before the reduction so its
            // latency overlaps it (EV == 0: V is streamed after the rotation decision instead, for shapes whose
            // registers are full)
#pragma unroll
            for (int k = 0; k < EA; ++k)
            {
                const int64_t idx = static_cast<int64_t>(k) * T + tid;
                ap[k] = raw_chunk(colp, idx, m2, true);
                aq[k] = raw_chunk(colq, idx, m2, true);
            }
#pragma unroll
            for (int k = 0; k < EV; ++k)
            {
                const int64_t idx = static_cast<int64_t>(k) * T + tid;
                xv[k] = raw_chunk(vp, idx, n2, true);
                yv[k] = raw_chunk(vq, idx, n2, true);
            }
            // The skip flag is read after the column loads are issued, so both are in flight together (one memory
            // round trip at the start of a workgroup, not two), and tested once the loads have landed.
            const int go = (skip_if_zero != nullptr) ? *skip_if_zero : 1;
#pragma unroll
            for (int k = 0; k < EA; ++k)
            {
                const int64_t idx = static_cast<int64_t>(k) * T + tid;
                ap[k] = fix_chunk(ap[k], idx, m2, odd_tail, true);
                aq[k] = fix_chunk(aq[k], idx, m2, odd_tail, true);
            }
#pragma unroll
            for (int k = 0; k < EV; ++k)
            {
                const int64_t idx = static_cast<int64_t>(k) * T + tid;
                xv[k] = fix_chunk(xv[k], idx, n2, odd_tail_v, true);
                yv[k] = fix_chunk(yv[k], idx, n2, odd_tail_v, true);
            }
            // (the empty asm keeps the loads ahead of the test instead of sunk past it)
#pragma unroll
            for (int k = 0; k < EA; ++k)
            {
                keep2(ap[k]);
                keep2(aq[k]);
            }
#pragma unroll
            for (int k = 0; k < EV; ++k)
            {
                keep2(xv[k]);
                keep2(yv[k]);
            }
            if (go == 0)
            {
                return;
            }
#pragma unroll
            for (int k = 0; k < EA; ++k)
            {
                gram_acc(spp, sqq, spq, ap[k], aq[k]);
            }
            block_sum3<T>(spp, sqq, spq, scratch);
            double c, s;
            const bool rotate = dev::rotation_params(spp, sqq, spq, eps, c, s);
            if (tid == 0)
            {
                if (offmax)
                {
                    dev::atomic_max_nonneg(offmax + cslot, dev::pair_cosine(spp, sqq, spq));
                }
                if (rotate && rot_count)
                {
                    atomicAdd(rot_count + cslot, 1);
                }
            }
            if (!rotate)
            {
                return; // c = 1, s = 0 leaves both columns bit-identical: skip the stores
            }
#pragma unroll
            for (int k = 0; k < EA; ++k)
            {
                const int64_t idx = static_cast<int64_t>(k) * T + tid;
                if (idx < m2)
                {
                    rotate2(ap[k], aq[k], c, s);
                    colp[idx] = ap[k];
                    colq[idx] = aq[k];
                }
            }
            if constexpr (EV > 0)
            {
#pragma unroll
                for (int k = 0; k < EV; ++k)
                {
                    const int64_t idx = static_cast<int64_t>(k) * T + tid;
                    if (idx < n2)
                    {
                        rotate2(xv[k], yv[k], c, s);
                        vp[idx] = xv[k];
                        vq[idx] = yv[k];
                    }
                }
            }
            else
            {
                rotate_stream(vp, vq, n2, odd_tail_v, c, s, tid, T);
            }
        }
        else
        {
            for (int64_t idx = tid; idx < m2; idx += T)
            {
                double2 x = colp[idx];
                double2 y = colq[idx];
                if (odd_tail && idx == m2 - 1)
                {
                    x.y = 0.0;
                    y.y = 0.0;
                }
                gram_acc(spp, sqq, spq, x, y);
            }
            block_sum3<T>(spp, sqq, spq, scratch);
            double c, s;
            const bool rotate = dev::rotation_params(spp, sqq, spq, eps, c, s);
            if (tid == 0)
            {
                if (offmax)
                {
                    dev::atomic_max_nonneg(offmax + cslot, dev::pair_cosine(spp, sqq, spq));
                }
                if (rotate && rot_count)
                {
                    atomicAdd(rot_count + cslot, 1);
                }
            }
            if (!rotate)
            {
                return;
            }
            rotate_stream(colp, colq, m2, odd_tail, c, s, tid, T);
            rotate_stream(vp, vq, n2, odd_tail_v, c, s, tid, T);
        }
    }

    // Slots a round-span workgroup owns, [a, b), for `groups` workgroups over the H slots of the circle: workgroup
    // 0 takes own_end slots, each later one `own`, and the last one the rest up to H (at most own_end).  One
    // workgroup owns all slots when groups == 1.
    __host__ __device__ inline void span_owned(int wg, int groups, int H, int own, int own_end, int &a, int &b)
    {
        if (groups == 1)
        {
            a = 0;
            b = H;
            return;
        }
        const int lo_end = own_end + (wg - 1) * own;
        const int hi_end = own_end + wg * own;
        a = (wg == 0) ? 0 : (lo_end < H ? lo_end : H);
        b = (wg == groups - 1) ? H : (hi_end < H ? hi_end : H);
    }

    // Workgroups and owned slots for spans of at most max_rounds rounds over the circle of n columns, with `waves`
    // wavefronts (cone slots) per workgroup; throws when a workgroup's cone would not fit.
    inline void span_partition(int64_t n, int waves, int max_rounds, int &groups, int &own, int &own_end)
    {
        const int N = static_cast<int>(n + (n & 1));
        const int H = N / 2;
        if (H <= waves)
        {
            groups = 1;
            own = H;
            own_end = H;
            return;
        }
        const int halo = max_rounds - 1;
        own = waves - 2 * halo;
        own_end = waves - halo;
        if (own < 1)
        {
            throw_invalid("round span: too many rounds per launch for the workgroup's slots.");
        }
        groups = (H <= 2 * own_end) ? 2 : 2 + (H - 2 * own_end + own - 1) / own;
        for (int w = 0; w < groups; ++w)
        {
            int a = 0, b = 0;
            span_owned(w, groups, H, own, own_end, a, b);
            const int lo = a - halo > 0 ? a - halo : 0;
            const int hi = b + halo < H ? b + halo : H;
            if (b <= a || hi - lo > waves)
            {
                throw_invalid("round span: slot partition does not fit the workgroup.");
            }
        }
    }

    // Rounds per launch over the N - 1 rounds of a sweep: an even number of launches (so every sweep ends in the
    // first of the two buffers) of at most max_rounds rounds, as even as possible.
    inline std::vector<int> span_chunks(int64_t n, int max_rounds)
    {
        const int N = static_cast<int>(n + (n & 1));
        const int R = N - 1;
        if (max_rounds < 2 || R < 3)
        {
            throw_invalid("round span: needs at least 2 rounds per launch and 3 rounds per sweep.");
        }
        int L = (R + max_rounds - 1) / max_rounds;
        if (L % 2 == 1)
        {
            L += 1;
        }
        std::vector<int> chunks(L, R / L);
        for (int i = 0; i < R % L; ++i)
        {
            chunks[i] += 1;
        }
        return chunks;
    }

    // Round span: rounds [r0, r0 + rounds) of one sweep in one launch, by temporal blocking of the round-robin
    // circle, bit-identical to `rounds` launches of round_kernel<TT, EA, *>.
    //
    // Slot k of the circle (k < H = N/2) pairs position k (its top) with position N-1-k (its bottom).  After a
    // round the top of slot k moves to the top of slot k+1, the bottom of slot k to the bottom of slot k-1, the
    // top of slot H-1 turns into the bottom of slot H-1, the bottom of slot 0 into the top of slot 1, and the top
    // of slot 0 (column 0) stays.  Workgroup w owns slots [a, b) (span_owned).  To run `rounds` rounds without
    // hearing from its neighbours it also computes the slots that feed its own: rounds - 1 more slots on each
    // side where the circle does not turn, the cone [lo, hi), one wavefront per slot (hi - lo <= W).  At round t
    // the slots [lo_t, hi_t) hold exact columns -- the cone narrows by one slot per round on each side that is
    // not a turn -- and after the last round [a, b) is exact.  A wavefront holds its slot's two columns of A and
    // V in registers; between rounds the columns move through LDS.  The workgroups read a_in / v_in and write
    // only the columns at their own slots after the last round to a_out / v_out, so they never wait for each
    // other (the host alternates the two buffers).
    //
    // Bits: a wavefront forms its pair's Gram entries in the order a TT-thread round_kernel workgroup uses --
    // VW = TT / 64 virtual wavefronts of 64 lanes, each virtual thread accumulating its EA chunks
    // idx = k TT + w 64 + lane with gram_acc, each virtual wavefront summed by the DPP tree, the VW totals added
    // to 0.0 in order -- and rotates with the same rotation_params and rotate2.  A column pair that does not
    // rotate keeps its bits, as in round_kernel.  Rotations and the largest pair cosine are counted once, by the
    // owner of the slot.
    //
    // Registers per lane: CA = VW EA chunks of each A column (covering m2 = ceil(m / 2) double2 chunks:
    // EA TT >= m2) and NV chunks of each V column (64 NV >= ceil(n / 2)).
    template <int TT, int EA, int NV, int W>
    __global__ __launch_bounds__(64 * W) void round_span_kernel(
        const double *__restrict__ a_in, double *__restrict__ a_out, int64_t lda, int64_t m,
        const double *__restrict__ v_in, double *__restrict__ v_out, int64_t ldv, int64_t n, int N, int r0,
        int rounds, int own, int own_end, double eps, int *__restrict__ rot_count,
        unsigned long long *__restrict__ offmax, const int *__restrict__ skip_if_zero, int count_slots)
    {
        constexpr int VW = TT / 64;
        constexpr int CA = VW * EA; // chunks per lane of an A column
        constexpr int CC = CA + NV; // chunks per lane of a column of A and V
        // two stage buffers (one barrier less per move) where they fit in LDS
        constexpr int NB = (2 * W * CC * 64 * 16 <= 160 * 1024) ? 2 : 1;
        static_assert(W * CC * 64 * 16 <= 160 * 1024, "round span stage exceeds LDS");
        __shared__ double2 stage[NB][W][CC][64];
        const int H = N / 2;
        int a, b;
        span_owned(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), H, own, own_end, a, b);
        const int halo = rounds - 1;
        const int lo = max(0, a - halo);
        const int hi = min(H, b + halo);
        // the wavefront index as a scalar, so that everything derived from the slot stays on the scalar unit
        const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
        const int lane = threadIdx.x % 64;
        const int k = lo + wave; // this wavefront's slot
        const int nc = hi - lo;  // slots in the cone (<= W)
        const bool in_cone = wave < nc;
        const int64_t m2 = (m + 1) / 2;
        const bool odd_m = (m & 1) != 0;
        const int64_t n2 = (n + 1) / 2;
        const bool odd_n = (n & 1) != 0;

        // chunk c < CA of a column: virtual wavefront c / EA, chunk c % EA of its virtual thread; then NV V chunks.
        // Both columns of the slot are requested before the first wait (raw_chunk), then zeroed where unused.
        auto chunk_idx = [&](int c) -> int64_t {
            return c < CA ? static_cast<int64_t>(c % EA) * TT + (c / EA) * 64 + lane
                          : static_cast<int64_t>(c - CA) * 64 + lane;
        };
        auto load_raw2 = [&](int col_t, int col_b, double2 (&x)[CC], double2 (&y)[CC]) {
            const bool ut = col_t < n, ub = col_b < n;
            const double2 *at = reinterpret_cast<const double2 *>(a_in + static_cast<int64_t>(ut ? col_t : 0) * lda);
            const double2 *ab = reinterpret_cast<const double2 *>(a_in + static_cast<int64_t>(ub ? col_b : 0) * lda);
            const double2 *vt = reinterpret_cast<const double2 *>(v_in + static_cast<int64_t>(ut ? col_t : 0) * ldv);
            const double2 *vb = reinterpret_cast<const double2 *>(v_in + static_cast<int64_t>(ub ? col_b : 0) * ldv);
#pragma unroll
            for (int c = 0; c < CC; ++c)
            {
                const int64_t idx = chunk_idx(c);
                x[c] = c < CA ? raw_chunk(at, idx, m2, ut) : raw_chunk(vt, idx, n2, ut);
                y[c] = c < CA ? raw_chunk(ab, idx, m2, ub) : raw_chunk(vb, idx, n2, ub);
            }
        };
        auto fix2 = [&](int col_t, int col_b, double2 (&x)[CC], double2 (&y)[CC]) {
            const bool ut = col_t < n, ub = col_b < n;
#pragma unroll
            for (int c = 0; c < CC; ++c)
            {
                const int64_t idx = chunk_idx(c);
                x[c] = c < CA ? fix_chunk(x[c], idx, m2, odd_m, ut) : fix_chunk(x[c], idx, n2, odd_n, ut);
                y[c] = c < CA ? fix_chunk(y[c], idx, m2, odd_m, ub) : fix_chunk(y[c], idx, n2, odd_n, ub);
            }
        };
        auto store = [&](int col, const double2 (&x)[CC]) {
            if (col >= n)
            {
                return; // the dummy column of an odd column count
            }
#pragma unroll
            for (int c = 0; c < CA; ++c)
            {
                const int64_t idx = static_cast<int64_t>(c % EA) * TT + (c / EA) * 64 + lane;
                if (idx < m2)
                {
                    reinterpret_cast<double2 *>(a_out + static_cast<int64_t>(col) * lda)[idx] = x[c];
                }
            }
#pragma unroll
            for (int c = 0; c < NV; ++c)
            {
                const int64_t idx = static_cast<int64_t>(c) * 64 + lane;
                if (idx < n2)
                {
                    reinterpret_cast<double2 *>(v_out + static_cast<int64_t>(col) * ldv)[idx] = x[CA + c];
                }
            }
        };

        double2 tp[CC], bt[CC]; // the slot's top and bottom columns
        // the columns at the slot's two positions, advanced by one circle step per round (no division)
        int p = dev::circle_at(k, r0, N);
        int q = dev::circle_at(N - 1 - k, r0, N);
        if (in_cone)
        {
            load_raw2(p, q, tp, bt);
        }
        // the skip flag is read after the first column loads are issued (both in flight together) and tested once
        // they have landed
        const int go = (skip_if_zero != nullptr) ? *skip_if_zero : 1;
        if (in_cone)
        {
            fix2(p, q, tp, bt);
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CC; ++c)
            {
                tp[c] = zero2();
                bt[c] = zero2();
            }
        }
#pragma unroll
        for (int c = 0; c < CC; ++c)
        {
            keep2(tp[c]); // the loads stay ahead of the skip test
            keep2(bt[c]);
        }
        if (go == 0)
        {
            return;
        }
        int my_rot = 0;
        double my_cos = 0.0;
        const bool owner = k >= a && k < b;
        for (int t = 0; t < rounds; ++t)
        {
            if (t > 0)
            {
                p = (k == 0) ? 0 : (p == 1 ? N - 1 : p - 1);
                q = (q == 1) ? N - 1 : q - 1;
            }
            const int lo_t = (lo == 0) ? 0 : lo + t;
            const int hi_t = (hi == H) ? H : hi - t;
            if (in_cone && k >= lo_t && k < hi_t && p < n && q < n) // uniform over the wavefront
            {
                double xs[VW], ys[VW], zs[VW];
#pragma unroll
                for (int w = 0; w < VW; ++w)
                {
                    xs[w] = 0.0;
                    ys[w] = 0.0;
                    zs[w] = 0.0;
#pragma unroll
                    for (int e = 0; e < EA; ++e)
                    {
                        gram_acc(xs[w], ys[w], zs[w], tp[w * EA + e], bt[w * EA + e]);
                    }
                }
                double spp, sqq, spq;
                if constexpr (VW == 1)
                {
                    spp = dev::wave_sum_dpp(xs[0]);
                    sqq = dev::wave_sum_dpp(ys[0]);
                    spq = dev::wave_sum_dpp(zs[0]);
                }
                else
                {
                    spp = 0.0;
                    sqq = 0.0;
                    spq = 0.0;
#pragma unroll
                    for (int w = 0; w < VW; ++w)
                    {
                        const double sx = dev::wave_sum_dpp(xs[w]);
                        const double sy = dev::wave_sum_dpp(ys[w]);
                        const double sz = dev::wave_sum_dpp(zs[w]);
                        spp += sx;
                        sqq += sy;
                        spq += sz;
                    }
                }
                double c, s;
                const bool rotate = dev::rotation_params(spp, sqq, spq, eps, c, s);
                if (owner)
                {
                    my_cos = fmax(my_cos, dev::pair_cosine(spp, sqq, spq));
                    my_rot += rotate ? 1 : 0;
                }
                if (rotate)
                {
#pragma unroll
                    for (int e = 0; e < CC; ++e)
                    {
                        rotate2(tp[e], bt[e], c, s);
                    }
                }
            }
            if (t + 1 == rounds)
            {
                break;
            }
            // Bottoms move one slot left; slot H-1 takes its own top (the turn).  Slot 0 keeps its old bottom in
            // the stage through the second phase: it sends no top (column 0 stays), so slot 1 reads slot 0's old
            // bottom as its new top there (the other turn).
            auto &sa = stage[0];
            auto &sb = stage[NB - 1];
            if constexpr (NB == 1)
            {
                __syncthreads(); // the previous round's reads of the single stage are done
            }
            if (in_cone)
            {
#pragma unroll
                for (int e = 0; e < CC; ++e)
                {
                    sa[wave][e][lane] = bt[e];
                }
            }
            __syncthreads();
            if (in_cone)
            {
                if (k == H - 1 && k != 0)
                {
#pragma unroll
                    for (int e = 0; e < CC; ++e)
                    {
                        bt[e] = tp[e];
                    }
                }
                else if (wave + 1 < nc)
                {
#pragma unroll
                    for (int e = 0; e < CC; ++e)
                    {
                        bt[e] = sa[wave + 1][e][lane];
                    }
                }
            }
            if constexpr (NB == 1)
            {
                __syncthreads(); // every bottom is read before the tops overwrite the stage
            }
            // Tops move one slot right.
            if (in_cone && k != 0)
            {
#pragma unroll
                for (int e = 0; e < CC; ++e)
                {
                    sb[wave][e][lane] = tp[e];
                }
            }
            if constexpr (NB == 2)
            {
                // slot 0's old bottom, for slot 1, sits in the first stage; copy it where slot 1 reads
                if (in_cone && k == 0)
                {
#pragma unroll
                    for (int e = 0; e < CC; ++e)
                    {
                        sb[wave][e][lane] = sa[wave][e][lane];
                    }
                }
            }
            __syncthreads();
            if (in_cone && k >= 1 && wave >= 1)
            {
#pragma unroll
                for (int e = 0; e < CC; ++e)
                {
                    tp[e] = sb[wave - 1][e][lane];
                }
            }
        }
        if (in_cone && owner)
        {
            store(p, tp); // the columns at the slot's positions in the last round
            store(q, bt);
        }
        if (lane == 0 && in_cone && owner)
        {
            const int cslot = (static_cast<int>(blockIdx.x) * W + wave) & (count_slots - 1);
            if (offmax && my_cos > 0.0)
            {
                dev::atomic_max_nonneg(offmax + cslot, my_cos);
            }
            if (rot_count && my_rot > 0)
            {
                atomicAdd(rot_count + cslot, my_rot);
            }
        }
    }
} // namespace jsvd::plain
