// autotune.hip -- measured choice of launch parameters per matrix shape.
//
// For an m x n problem the tuner times, on synthetic data of that shape on the device:
//   * every instantiated launch shape of the plain path's fused round kernel that suits the shape (64, 256 or
//     1024 threads per column pair; columns held in registers with the smallest chunk count that fits, V
//     held or streamed, or both streamed), up to 16 back-to-back rounds of a first sweep each;
//   * the block paths' block width (16, 32, 64 columns), up to 8 FP32 block rounds each, compared per sweep
//     (the round time times the rounds of a sweep at that width), then at the chosen width the apply kernel's
//     total workgroup count (256 / 512 / 1024 per launch).
// Every candidate starts from the same seeded data; the best of `repetitions` timings counts.  The fastest
// choices are cached per (m, n) for the life of the process, and plans created or reshaped afterwards for
// that shape use them (unless their layout_transpose_mode forces a launch shape).
#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "autotune.hpp"
#include "block.hpp"

namespace jsvd
{
    namespace
    {
        std::mutex g_mu;
        std::map<std::pair<int64_t, int64_t>, Tuning> g_table;

        __global__ void fill_kernel(double *x, int64_t count, uint64_t seed)
        {
            for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count;
                 i += int64_t{gridDim.x} * blockDim.x)
            {
                // splitmix64 of (seed, i) mapped to [-1, 1)
                uint64_t z = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                x[i] = static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            }
        }

        __global__ void fill_f32_kernel(float *x, const double *src, int64_t count)
        {
            for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count;
                 i += int64_t{gridDim.x} * blockDim.x)
            {
                x[i] = static_cast<float>(src[i]);
            }
        }

        struct DevBuf
        {
            void *p = nullptr;
            explicit DevBuf(size_t bytes) { JSVD_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 256))); }
            ~DevBuf() { (void)hipFree(p); }
            DevBuf(const DevBuf &) = delete;
            DevBuf &operator=(const DevBuf &) = delete;
        };

        struct Events
        {
            hipEvent_t a = nullptr, b = nullptr;
            Events()
            {
                JSVD_HIP_CHECK(hipEventCreate(&a));
                JSVD_HIP_CHECK(hipEventCreate(&b));
            }
            ~Events()
            {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
            double elapsed_us() const
            {
                float ms = 0.0f;
                JSVD_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
                return 1000.0 * ms;
            }
        };

        unsigned grid_for(int64_t count)
        {
            return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 4096)));
        }

        // Launch shapes the round kernel is instantiated for that suit an m x n problem.
        std::vector<RoundShape> round_candidates(int64_t m, int64_t n)
        {
            const int64_t m2 = (m + 1) / 2, n2 = (n + 1) / 2;
            std::vector<RoundShape> out;
            for (int t : {64, 256, 1024})
            {
                RoundShape stream_shape;
                stream_shape.threads = t;
                stream_shape.ea = 0;
                stream_shape.ev = 0;
                out.push_back(stream_shape);
                for (int ea : {1, 2, 4, 8})
                {
                    if (static_cast<int64_t>(ea) * t < m2 || (ea > 1 && static_cast<int64_t>(ea / 2) * t >= m2))
                    {
                        continue; // the smallest register chunk count that holds a column
                    }
                    for (int ev : {0, 1, 2, 4, 8})
                    {
                        if (ev > ea || (ea == 1 && ev != 1) || (ea == 2 && ev == 0) ||
                            (t == 1024 && ea == 8 && ev > 2))
                        {
                            continue;
                        }
                        if (ev > 0 && (static_cast<int64_t>(ev) * t < n2 || (ev > 1 && (ev / 2) * t >= n2)))
                        {
                            continue;
                        }
                        RoundShape s;
                        s.threads = t;
                        s.ea = ea;
                        s.ev = ev;
                        out.push_back(s);
                    }
                }
            }
            return out;
        }

        double time_rounds(const RoundShape &shape, double *a, int64_t lda, int64_t m, double *v, int64_t ldv,
                           int64_t n, int *slots, unsigned long long *off, int rounds, hipStream_t stream)
        {
            Events ev;
            const int per_sweep = static_cast<int>(n + (n & 1) - 1);
            launch_round(shape, a, lda, m, v, ldv, n, 0, 1e-15, slots, off, nullptr, stream, 64); // warm-up
            JSVD_HIP_CHECK(hipEventRecord(ev.a, stream));
            for (int r = 1; r <= rounds; ++r)
            {
                launch_round(shape, a, lda, m, v, ldv, n, r % per_sweep, 1e-15, slots, off, nullptr, stream, 64);
            }
            JSVD_HIP_CHECK(hipEventRecord(ev.b, stream));
            JSVD_HIP_CHECK(hipEventSynchronize(ev.b));
            return ev.elapsed_us() / rounds;
        }

        // Microseconds per sweep of round spans (waves, rounds) of `shape`: one warm-up sweep, then one timed
        // sweep, alternating between (a, v) and (a2, v2) as the plan does.
        double time_span_sweep(const RoundShape &shape, int waves, int rounds, double *a, double *a2, int64_t lda,
                               int64_t m, double *v, double *v2, int64_t ldv, int64_t n, int *slots,
                               unsigned long long *off, hipStream_t stream)
        {
            int groups = 0, own = 0, own_end = 0;
            span_partition(n, waves, rounds, groups, own, own_end);
            const std::vector<int> chunks = span_chunks(n, rounds);
            Events ev;
            for (int rep = 0; rep < 2; ++rep)
            {
                if (rep == 1)
                {
                    JSVD_HIP_CHECK(hipEventRecord(ev.a, stream));
                }
                int r0 = 0;
                for (size_t i = 0; i < chunks.size(); ++i)
                {
                    const bool even = (i % 2) == 0;
                    launch_round_span(shape, waves, even ? a : a2, even ? a2 : a, lda, m, even ? v : v2, even ? v2 : v,
                                      ldv, n, r0, chunks[i], groups, own, own_end, 1e-15, slots, off, nullptr, stream,
                                      64);
                    r0 += chunks[i];
                }
            }
            JSVD_HIP_CHECK(hipEventRecord(ev.b, stream));
            JSVD_HIP_CHECK(hipEventSynchronize(ev.b));
            return ev.elapsed_us();
        }

        // Microseconds per FP32 block round of g on seeded data: up to 8 rounds of a first sweep, best of
        // `repetitions`; with the cross ordering also the intra-block step of a sweep (intra_us).
        double time_block_rounds(const BlockGeometry &g, int repetitions, hipStream_t stream, double &intra_us)
        {
            const size_t a_elems = static_cast<size_t>(g.lda) * g.n_pad;
            const size_t v_elems = static_cast<size_t>(g.ldv) * g.n_pad;
            DevBuf src(sizeof(double) * a_elems), a32(sizeof(float) * a_elems), v32(sizeof(float) * v_elems);
            DevBuf part(sizeof(float) * g.partial_elems()), q(sizeof(float) * g.q_elems());
            DevBuf flags(sizeof(int) * g.flag_elems()), perm(sizeof(int) * g.n_pad);
            hipLaunchKernelGGL(fill_kernel, dim3(grid_for(a_elems)), dim3(256), 0, stream, static_cast<double *>(src.p),
                               static_cast<int64_t>(a_elems), uint64_t{777});
            JSVD_HIP_CHECK(hipGetLastError());
            launch_iota(static_cast<int *>(perm.p), g.n_pad, stream);
            const int rounds = std::min(8, g.rounds());
            double us = -1.0;
            for (int rep = 0; rep < repetitions; ++rep)
            {
                hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_for(a_elems)), dim3(256), 0, stream,
                                   static_cast<float *>(a32.p), static_cast<const double *>(src.p),
                                   static_cast<int64_t>(a_elems));
                JSVD_HIP_CHECK(hipGetLastError());
                launch_identity_f32(static_cast<float *>(v32.p), g.ldv, g.n_pad, stream);
                Events ev;
                JSVD_HIP_CHECK(hipEventRecord(ev.a, stream));
                for (int r = 0; r < rounds; ++r)
                {
                    launch_block_round<float>(g, static_cast<float *>(a32.p), static_cast<float *>(v32.p), r, 1e-7,
                                              static_cast<float *>(part.p), static_cast<float *>(q.p),
                                              static_cast<int *>(flags.p), nullptr, nullptr,
                                              static_cast<const int *>(perm.p), nullptr, nullptr, stream);
                }
                JSVD_HIP_CHECK(hipEventRecord(ev.b, stream));
                JSVD_HIP_CHECK(hipEventSynchronize(ev.b));
                const double t = ev.elapsed_us() / rounds;
                us = (us < 0.0) ? t : std::min(us, t);
            }
            intra_us = 0.0;
            if (g.cross)
            {
                for (int rep = 0; rep < repetitions; ++rep)
                {
                    Events ev;
                    JSVD_HIP_CHECK(hipEventRecord(ev.a, stream));
                    launch_block_intra<float>(g, static_cast<float *>(a32.p), static_cast<float *>(v32.p), 1e-7,
                                              static_cast<float *>(part.p), static_cast<float *>(q.p),
                                              static_cast<int *>(flags.p), nullptr, nullptr,
                                              static_cast<const int *>(perm.p), nullptr, nullptr, stream);
                    JSVD_HIP_CHECK(hipEventRecord(ev.b, stream));
                    JSVD_HIP_CHECK(hipEventSynchronize(ev.b));
                    const double t = ev.elapsed_us();
                    intra_us = (rep == 0) ? t : std::min(intra_us, t);
                }
            }
            return us;
        }
    } // namespace

    bool lookup_tuning(int64_t m, int64_t n, Tuning &out)
    {
        std::lock_guard<std::mutex> lock(g_mu);
        const auto it = g_table.find({m, n});
        if (it == g_table.end())
        {
            return false;
        }
        out = it->second;
        return true;
    }

    namespace
    {
        struct BlockChoice
        {
            int block_width = 0; // 0: the shape leaves fewer than two blocks at every width
            int apply_wgs = 512;
            double sweep_us[3] = {0.0, 0.0, 0.0};
            double round_us_default = 0.0;
            double round_us_best = 0.0;
            int tried = 0;
        };

        // Block width by per-sweep time (round time x rounds per sweep, 512 apply workgroups), then the apply
        // grid size at that width, for block sweeps over an m x n problem.
        BlockChoice tune_block(int64_t m, int64_t n, int repetitions, hipStream_t stream)
        {
            BlockChoice c;
            const int widths[3] = {16, 32, 64};
            double best_sweep = -1.0;
            for (int wi = 0; wi < 3; ++wi)
            {
                BlockGeometry g = BlockGeometry::make(m, n, widths[wi]);
                if (g.nb < 2)
                {
                    continue;
                }
                g.cross = widths[wi] == 32; // the solvers' default inner ordering at this width
                double intra_us = 0.0;
                const double us = time_block_rounds(g, repetitions, stream, intra_us);
                ++c.tried;
                c.sweep_us[wi] = us * g.rounds() + intra_us;
                if (best_sweep < 0.0 || c.sweep_us[wi] < best_sweep)
                {
                    best_sweep = c.sweep_us[wi];
                    c.block_width = widths[wi];
                    c.round_us_default = us;
                }
            }
            if (c.block_width == 0)
            {
                return c;
            }
            BlockGeometry g = BlockGeometry::make(m, n, c.block_width);
            g.cross = c.block_width == 32;
            double best_us = c.round_us_default; // 512 workgroups, measured above
            c.apply_wgs = g.apply_wgs;
            for (int wgs : {256, 1024})
            {
                g.apply_wgs = wgs;
                double intra_us = 0.0;
                const double us = time_block_rounds(g, repetitions, stream, intra_us);
                ++c.tried;
                if (us < best_us)
                {
                    best_us = us;
                    c.apply_wgs = wgs;
                }
            }
            c.round_us_best = best_us;
            return c;
        }

        void store(int64_t m, int64_t n, const Tuning &t)
        {
            std::lock_guard<std::mutex> lock(g_mu);
            Tuning &e = g_table[{m, n}];
            if (t.has_round)
            {
                e.round = t.round;
                e.span_waves = t.span_waves;
                e.span_rounds = t.span_rounds;
                e.has_round = true;
            }
            if (t.has_block)
            {
                e.block_width = t.block_width;
                e.apply_wgs = t.apply_wgs;
                e.has_block = true;
            }
        }
    } // namespace

    void store_batch_shape(int64_t max_rows, int64_t max_columns, int shape)
    {
        std::lock_guard<std::mutex> lock(g_mu);
        Tuning &e = g_table[{max_rows, max_columns}];
        e.batch_shape = shape;
        e.has_batch = true;
    }

    Tuning run_autotune(int64_t m, int64_t n, int repetitions, hipStream_t stream, jsvd_tune_report *report)
    {
        validate_shape(m, n);
        if (repetitions < 1)
        {
            throw_invalid("repetitions must be positive.");
        }
        Tuning best;
        best.round = choose_round_shape(m, n, false);
        jsvd_tune_report rep{};
        const int64_t lda = m + (m & 1), ldv = n + (n & 1);
        // plain path: round-kernel launch shapes
        if (n >= 2)
        {
            DevBuf a0(sizeof(double) * lda * n), a(sizeof(double) * lda * n), v(sizeof(double) * ldv * n);
            DevBuf slots(sizeof(int) * 64), off(sizeof(unsigned long long) * 64);
            hipLaunchKernelGGL(fill_kernel, dim3(grid_for(lda * n)), dim3(256), 0, stream, static_cast<double *>(a0.p),
                               lda * n, uint64_t{12345});
            JSVD_HIP_CHECK(hipGetLastError());
            JSVD_HIP_CHECK(hipMemsetAsync(slots.p, 0, sizeof(int) * 64, stream));
            JSVD_HIP_CHECK(hipMemsetAsync(off.p, 0, sizeof(unsigned long long) * 64, stream));
            const int rounds = 16;
            const int per_sweep = static_cast<int>(n + (n & 1) - 1);
            const bool span_room = lda * n <= int64_t{1024} * 512;
            DevBuf a2(span_room ? sizeof(double) * lda * n : 8), v2(span_room ? sizeof(double) * ldv * n : 8);
            // per candidate shape: one launch per round (16 rounds, scaled to a sweep) and, where the build has
            // span kernels for it, whole sweeps of spans; the fastest (shape, spans) per sweep wins
            double best_sweep = -1.0;
            for (const RoundShape &cand : round_candidates(m, n))
            {
                // same data for every candidate: A from the seeded fill (padding rows zero), V = I
                // (the kernel reads the padding row of an odd m as zero whatever it holds)
                auto reset = [&] {
                    JSVD_HIP_CHECK(hipMemcpyAsync(a.p, a0.p, sizeof(double) * lda * n, hipMemcpyDeviceToDevice, stream));
                    launch_identity(static_cast<double *>(v.p), ldv, n, stream);
                };
                double us = -1.0;
                for (int r = 0; r < repetitions; ++r)
                {
                    reset();
                    const double t = time_rounds(cand, static_cast<double *>(a.p), lda, m, static_cast<double *>(v.p),
                                                 ldv, n, static_cast<int *>(slots.p),
                                                 static_cast<unsigned long long *>(off.p), rounds, stream);
                    us = (us < 0.0) ? t : std::min(us, t);
                }
                ++rep.candidates_tried;
                const RoundShape def = choose_round_shape(m, n, false);
                if (cand.threads == def.threads && cand.ea == def.ea && cand.ev == def.ev)
                {
                    rep.round_us_default = us;
                }
                if (best_sweep < 0.0 || us * per_sweep < best_sweep)
                {
                    best_sweep = us * per_sweep;
                    best.round = cand;
                    best.span_waves = 0;
                    best.span_rounds = 0;
                    rep.round_us_best = us;
                    rep.sweep_us_rounds = us * per_sweep;
                }
                if (!span_room)
                {
                    continue;
                }
                for (int waves : {8, 16})
                {
                    if (!span_supported(cand, waves, m, n))
                    {
                        continue;
                    }
                    const int most = span_max_rounds(waves, n);
                    std::vector<int> cands;
                    for (int r : {2, 3, 4, 6, 8})
                    {
                        if (r < most)
                        {
                            cands.push_back(r);
                        }
                    }
                    cands.push_back(most);
                    for (int r : cands)
                    {
                        double sw = -1.0;
                        for (int rep_i = 0; rep_i < repetitions; ++rep_i)
                        {
                            reset();
                            const double t = time_span_sweep(
                                cand, waves, r, static_cast<double *>(a.p), static_cast<double *>(a2.p), lda, m,
                                static_cast<double *>(v.p), static_cast<double *>(v2.p), ldv, n,
                                static_cast<int *>(slots.p), static_cast<unsigned long long *>(off.p), stream);
                            sw = (sw < 0.0) ? t : std::min(sw, t);
                        }
                        ++rep.candidates_tried;
                        if (rep.sweep_us_span <= 0.0 || sw < rep.sweep_us_span)
                        {
                            rep.sweep_us_span = sw;
                        }
                        if (sw < best_sweep)
                        {
                            best_sweep = sw;
                            best.round = cand;
                            best.span_waves = waves;
                            best.span_rounds = r;
                            rep.round_us_best = us;
                            rep.sweep_us_rounds = us * per_sweep;
                        }
                    }
                }
            }
            best.has_round = true;
            rep.span_waves = best.span_waves;
            rep.span_rounds = best.span_rounds;
        }
        // block paths over m x n (the block FP64 path and the mixed path's direct form)
        const BlockChoice bc = tune_block(m, n, repetitions, stream);
        rep.candidates_tried += bc.tried;
        if (bc.block_width != 0)
        {
            best.block_width = bc.block_width;
            best.apply_wgs = bc.apply_wgs;
            best.has_block = true;
            rep.block_width = bc.block_width;
            for (int k = 0; k < 3; ++k)
            {
                rep.block_sweep_us[k] = bc.sweep_us[k];
            }
            rep.block_round_us_default = bc.round_us_default;
            rep.block_round_us_best = bc.round_us_best;
        }
        // block sweeps over the n x n R^T of the mixed path's QR-first form (the same problem when m == n)
        BlockChoice sq = bc;
        if (m != n)
        {
            sq = tune_block(n, n, repetitions, stream);
            rep.candidates_tried += sq.tried;
        }
        rep.rt_block_width = sq.block_width;
        rep.rt_apply_workgroups = (sq.block_width != 0) ? sq.apply_wgs : 0;
        JSVD_HIP_CHECK(hipStreamSynchronize(stream));
        rep.round_threads = best.round.threads;
        rep.round_ea = best.round.ea;
        rep.round_ev = best.round.ev;
        rep.apply_workgroups = best.apply_wgs;
        store(m, n, best);
        if (m != n && sq.block_width != 0)
        {
            Tuning t;
            t.block_width = sq.block_width;
            t.apply_wgs = sq.apply_wgs;
            t.has_block = true;
            store(n, n, t);
        }
        if (report)
        {
            *report = rep;
        }
        return best;
    }
} // namespace jsvd

extern "C" int jsvd_autotune(int64_t m, int64_t n, int32_t repetitions, void *stream, jsvd_tune_report *report)
{
    return jsvd::guarded([&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        bool own = false;
        if (!s)
        {
            JSVD_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            own = true;
        }
        try
        {
            (void)jsvd::run_autotune(m, n, repetitions, s, report);
        }
        catch (...)
        {
            if (own)
            {
                (void)hipStreamDestroy(s);
            }
            throw;
        }
        if (own)
        {
            JSVD_HIP_CHECK(hipStreamDestroy(s));
        }
    });
}

extern "C" int jsvd_autotune_clear(void)
{
    return jsvd::guarded([&] {
        std::lock_guard<std::mutex> lock(jsvd::g_mu);
        jsvd::g_table.clear();
    });
}
