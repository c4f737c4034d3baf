// abi.hip -- the C ABI of libjsvd (include/jsvd.h): solver plans and the device-resident sweep driver of the
// plain FP64 path (configuration defaults, validation and the error message are in config.cpp).
#include <algorithm>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "autotune.hpp"
#include "cholesky.hpp"
#include "common.hpp"
#include "mixed.hpp"
#include "plain_f64.hpp"

namespace jsvd
{
    // Round-kernel launch shape for a plan: the autotuner's measured choice when it has one for this shape and
    // the layout mode is AUTO, else the built-in rule.
    RoundShape plan_round_shape(int64_t m, int64_t n, const jsvd_config &cfg)
    {
        Tuning t;
        if (cfg.layout_transpose_mode == JSVD_LAYOUT_AUTO && lookup_tuning(m, n, t) && t.has_round)
        {
            return t.round;
        }
        return choose_round_shape(m, n, use_wide_shape(m, n, cfg));
    }

    // Launch shape of the stand-alone round entry points: the autotuner's choice for m x n, else the built-in
    // rule for one wavefront per pair where it fits.
    RoundShape kernel_round_shape(int64_t m, int64_t n)
    {
        Tuning t;
        return (lookup_tuning(m, n, t) && t.has_round) ? t.round : choose_round_shape(m, n, false);
    }

    // Largest m x n whose plans keep the second pair of buffers the round spans alternate with (register-resident
    // columns make spans possible only up to m <= 1024, n <= 512 anyway).
    constexpr int64_t kSpanMaxElems = int64_t{1024} * 512;

    // Round-span setting for a plain plan of m x n: the autotuner's measured choice (round shape and spans) when
    // it has one, else, for m n <= 2^19 where the build has the kernel, spans of up to 4 rounds on 8-wavefront
    // workgroups with the 64-thread round shape (unless a layout mode is forced): one wavefront per pair then
    // reduces in the order of one 64-lane tree.  Measured on one MI355X, spans of 4 rounds on 8 wavefronts beat
    // one launch per round from 64 x 32 (1.7x) to 1000 x 500 (1.14x); 16-wavefront workgroups, whose cones hold
    // 16 slots, were slower on every shape.
    void plan_span(int64_t m, int64_t n, const jsvd_config &cfg, RoundShape &shape, int &waves, int &rounds)
    {
        waves = 0;
        rounds = 0;
        Tuning t;
        if (lookup_tuning(m, n, t) && t.has_round && t.round.threads == shape.threads && t.round.ea == shape.ea &&
            t.round.ev == shape.ev)
        {
            if (t.span_waves > 0 && span_supported(shape, t.span_waves, m, n))
            {
                waves = t.span_waves;
                rounds = t.span_rounds;
            }
            return;
        }
        if (m * n > (int64_t{1} << 19))
        {
            return;
        }
        RoundShape cand = shape;
        if (cfg.layout_transpose_mode == JSVD_LAYOUT_AUTO)
        {
            cand = choose_round_shape(m, n, false);
        }
        if (span_supported(cand, 8, m, n))
        {
            shape = cand;
            waves = 8;
            rounds = std::min(span_max_rounds(8, n), 4);
        }
    }
} // namespace jsvd

// The opaque plan: one shape, its device workspace, a stream, and the pinned mailbox through which the
// host learns each sweep's rotation count without a blocking copy.
struct jsvd_plan
{
    int64_t m = 0;
    int64_t n = 0;
    int64_t lda = 0; // even, >= m
    int64_t ldv = 0; // even, >= n
    int64_t cap_m = 0; // shape the buffers were allocated for (reshape targets fit inside it)
    int64_t cap_n = 0;
    jsvd_config cfg{};
    jsvd::RoundShape shape{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *a_col = nullptr;            // lda x n column-major working copy of A
    double *v_col = nullptr;            // ldv x n column-major V
    // Round spans (several rounds per launch): wavefronts per workgroup (0 = one launch per round), rounds per
    // launch at most, the slot partition and the rounds of each launch of a sweep; the spans alternate between
    // (a_col, v_col) and the second buffers (capacity-sized, allocated when the capacity allows spans).
    int span_waves = 0;
    int span_rounds = 0;
    int span_groups = 0;
    int span_own = 0;
    int span_own_end = 0;
    std::vector<int> span_chunk;
    double *a_alt = nullptr;
    double *v_alt = nullptr;
    int *rot_count = nullptr;           // [max_sweeps] rotated pairs per sweep
    unsigned long long *offmax = nullptr; // [max_sweeps] max pair cosine per sweep (double bits)
    // per-sweep counter slots the round kernels update (folded into rot_count / offmax after each sweep)
    static constexpr int kCountSlots = 64;
    int *rot_slots = nullptr;              // [max_sweeps][kCountSlots]
    unsigned long long *off_slots = nullptr; // [max_sweeps][kCountSlots]
    int *h_rot = nullptr;               // pinned mirror of rot_count
    hipEvent_t sweep_done[2] = {nullptr, nullptr};
    // Host-API staging (allocated on first host solve).
    double *d_a = nullptr;
    double *d_u = nullptr;
    double *d_sigma = nullptr;
    double *d_v = nullptr;
    // Diagnostics of the last solve.
    int last_sweeps = 0;
    std::vector<int> last_rot;
    std::vector<double> last_off;
    // Block / mixed paths (JSVD_PATH_BLOCK_F64, JSVD_PATH_MIXED).
    int path = JSVD_PATH_PLAIN_F64;
    std::unique_ptr<jsvd::BlockSolver> block;
    jsvd_solve_info last_info{};

    ~jsvd_plan()
    {
        block.reset();
        (void)hipFree(a_col);
        (void)hipFree(v_col);
        (void)hipFree(a_alt);
        (void)hipFree(v_alt);
        (void)hipFree(rot_count);
        (void)hipFree(offmax);
        (void)hipFree(rot_slots);
        (void)hipFree(off_slots);
        (void)hipHostFree(h_rot);
        (void)hipFree(d_a);
        (void)hipFree(d_u);
        (void)hipFree(d_sigma);
        (void)hipFree(d_v);
        for (hipEvent_t e : sweep_done)
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
        }
        if (own_stream && stream)
        {
            (void)hipStreamDestroy(stream);
        }
    }

    void create_block(int64_t rows, int64_t cols, const jsvd_config &config, const jsvd_mixed_config &mixed,
                      hipStream_t user_stream)
    {
        m = rows;
        n = cols;
        cap_m = rows;
        cap_n = cols;
        cfg = config;
        path = mixed.path;
        if (user_stream)
        {
            stream = user_stream;
        }
        else
        {
            JSVD_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        block = std::make_unique<jsvd::BlockSolver>(m, n, cfg, mixed, stream);
    }

    void create(int64_t rows, int64_t cols, const jsvd_config &config, hipStream_t user_stream)
    {
        m = rows;
        n = cols;
        cap_m = rows;
        cap_n = cols;
        cfg = config;
        lda = jsvd::round_up(m, 2);
        ldv = jsvd::round_up(n, 2);
        shape = jsvd::plan_round_shape(m, n, cfg);
        if (user_stream)
        {
            stream = user_stream;
        }
        else
        {
            JSVD_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        JSVD_HIP_CHECK(hipMalloc(&a_col, sizeof(double) * lda * n));
        JSVD_HIP_CHECK(hipMalloc(&v_col, sizeof(double) * ldv * n));
        if (lda * n <= jsvd::kSpanMaxElems)
        {
            JSVD_HIP_CHECK(hipMalloc(&a_alt, sizeof(double) * lda * n));
            JSVD_HIP_CHECK(hipMalloc(&v_alt, sizeof(double) * ldv * n));
        }
        JSVD_HIP_CHECK(hipMalloc(&rot_count, sizeof(int) * cfg.max_sweeps));
        JSVD_HIP_CHECK(hipMalloc(&offmax, sizeof(unsigned long long) * cfg.max_sweeps));
        JSVD_HIP_CHECK(hipMalloc(&rot_slots, sizeof(int) * kCountSlots * cfg.max_sweeps));
        JSVD_HIP_CHECK(hipMalloc(&off_slots, sizeof(unsigned long long) * kCountSlots * cfg.max_sweeps));
        JSVD_HIP_CHECK(hipHostMalloc(&h_rot, sizeof(int) * cfg.max_sweeps, hipHostMallocDefault));
        for (hipEvent_t &e : sweep_done)
        {
            JSVD_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        // Padding rows m..lda-1 stay zero for the plan's lifetime: transposes only write rows < m and the
        // round kernel reads/writes the padding row of an odd m as zero.
        JSVD_HIP_CHECK(hipMemsetAsync(a_col, 0, sizeof(double) * lda * n, stream));
        configure_span();
    }

    // Applies the span rule (plan_span) to the current shape; spans need the second buffers.
    void configure_span()
    {
        int waves = 0, rounds = 0;
        if (a_alt != nullptr)
        {
            jsvd::plan_span(m, n, cfg, shape, waves, rounds);
        }
        set_span(waves, rounds);
    }

    // Spans of at most `rounds` rounds on `waves`-wavefront workgroups (waves == 0: one launch per round).
    void set_span(int waves, int rounds)
    {
        span_waves = 0;
        span_rounds = 0;
        span_chunk.clear();
        if (waves == 0)
        {
            return;
        }
        if (a_alt == nullptr || !jsvd::span_supported(shape, waves, m, n))
        {
            jsvd::throw_invalid("round spans are not available for this plan's shape.");
        }
        rounds = std::min(rounds, jsvd::span_max_rounds(waves, n));
        if (rounds < 2)
        {
            jsvd::throw_invalid("round spans need at least 2 rounds per launch.");
        }
        jsvd::span_partition(n, waves, rounds, span_groups, span_own, span_own_end);
        span_chunk = jsvd::span_chunks(n, rounds);
        span_waves = waves;
        span_rounds = rounds;
    }

    // Re-targets the plan at rows x cols with epsilon `eps` without reallocating; false (plan unchanged) when the
    // shape does not fit the buffers allocated at creation.
    bool reshape(int64_t rows, int64_t cols, double eps)
    {
        if (rows < 1 || cols < 1 || rows < cols || rows > cap_m || cols > cap_n)
        {
            return false;
        }
        if (block)
        {
            if (!block->reshape(rows, cols))
            {
                return false;
            }
            block->set_epsilon(eps);
        }
        else if (rows != m || cols != n)
        {
            const int64_t new_lda = jsvd::round_up(rows, 2);
            // padding rows must read as zero: clear the new working extent
            JSVD_HIP_CHECK(hipMemsetAsync(a_col, 0, sizeof(double) * new_lda * cols, stream));
            lda = new_lda;
            ldv = jsvd::round_up(cols, 2);
            shape = jsvd::plan_round_shape(rows, cols, cfg);
        }
        const bool changed = rows != m || cols != n;
        m = rows;
        n = cols;
        cfg.epsilon = eps;
        if (!block && changed)
        {
            configure_span();
        }
        return true;
    }

    // Device-resident solve: transpose in, sweep loop, sigma/U extraction, V transpose out.
    void solve_device(const double *a_rm, double *u_rm, double *sigma, double *v_rm, int *sweeps_out)
    {
        if (block)
        {
            block->solve_device(a_rm, u_rm, sigma, v_rm, last_info);
            last_rot = block->sweep_rotations();
            last_off = block->sweep_offmax();
            last_sweeps = static_cast<int>(last_rot.size());
            if (sweeps_out)
            {
                *sweeps_out = last_info.sweeps_fp64;
            }
            return;
        }
        solve_plain(a_rm, u_rm, sigma, v_rm, sweeps_out);
        last_info = jsvd_solve_info{};
        last_info.path = JSVD_PATH_PLAIN_F64;
        last_info.sweeps_fp64 = last_sweeps;
        last_info.converged = (!last_rot.empty() && last_rot.back() == 0) ? 1 : 0;
        last_info.offmax_fp64 = last_off.empty() ? 0.0 : last_off.back();
    }

    void solve_plain(const double *a_rm, double *u_rm, double *sigma, double *v_rm, int *sweeps_out)
    {
        const int max_sweeps = cfg.max_sweeps;
        const double eps = cfg.epsilon;
        jsvd::launch_transpose(a_rm, n, a_col, lda, m, n, stream);
        jsvd::launch_identity(v_col, ldv, n, stream);
        JSVD_HIP_CHECK(hipMemsetAsync(rot_count, 0, sizeof(int) * max_sweeps, stream));
        JSVD_HIP_CHECK(hipMemsetAsync(offmax, 0, sizeof(unsigned long long) * max_sweeps, stream));
        JSVD_HIP_CHECK(hipMemsetAsync(rot_slots, 0, sizeof(int) * kCountSlots * max_sweeps, stream));
        JSVD_HIP_CHECK(hipMemsetAsync(off_slots, 0, sizeof(unsigned long long) * kCountSlots * max_sweeps, stream));
        std::fill(h_rot, h_rot + max_sweeps, -1);

        const int N = static_cast<int>(n + (n & 1));
        const int rounds = (n >= 2) ? N - 1 : 0;
        int enqueued = 0;
        // Sweep s+1 is enqueued before the host looks at sweep s, so the device never idles on the host;
        // every round of a sweep that follows a rotation-free sweep exits on the device flag.  The host
        // only waits on the event of sweep s-1, which completed when sweep s began.
        for (int s = 0; s < max_sweeps; ++s)
        {
            const int *skip = (s > 0) ? rot_count + (s - 1) : nullptr;
            if (span_waves > 0)
            {
                // an even number of span launches, alternating buffers: every sweep ends in (a_col, v_col), and
                // the launches of a skipped sweep write nothing
                int r0 = 0;
                for (size_t i = 0; i < span_chunk.size(); ++i)
                {
                    const bool even = (i % 2) == 0;
                    jsvd::launch_round_span(shape, span_waves, even ? a_col : a_alt, even ? a_alt : a_col, lda, m,
                                            even ? v_col : v_alt, even ? v_alt : v_col, ldv, n, r0, span_chunk[i],
                                            span_groups, span_own, span_own_end, eps, rot_slots + s * kCountSlots,
                                            off_slots + s * kCountSlots, skip, stream, kCountSlots);
                    r0 += span_chunk[i];
                }
            }
            else
            {
                for (int r = 0; r < rounds; ++r)
                {
                    jsvd::launch_round(shape, a_col, lda, m, v_col, ldv, n, r, eps, rot_slots + s * kCountSlots,
                                       off_slots + s * kCountSlots, skip, stream, kCountSlots);
                }
            }
            jsvd::launch_sweep_totals(rot_slots + s * kCountSlots, off_slots + s * kCountSlots, kCountSlots,
                                      rot_count + s, offmax + s, stream);
            JSVD_HIP_CHECK(hipMemcpyAsync(h_rot + s, rot_count + s, sizeof(int), hipMemcpyDeviceToHost, stream));
            JSVD_HIP_CHECK(hipEventRecord(sweep_done[s & 1], stream));
            enqueued = s + 1;
            if (s >= 1)
            {
                JSVD_HIP_CHECK(hipEventSynchronize(sweep_done[(s - 1) & 1]));
                if (h_rot[s - 1] == 0)
                {
                    break;
                }
            }
            else if (rounds == 0)
            {
                break; // a single column: the first sweep performs no rotation
            }
        }
        jsvd::launch_u_sigma(a_col, lda, m, n, eps, u_rm, sigma, stream);
        jsvd::launch_transpose(v_col, ldv, v_rm, n, n, n, stream);
        last_rot.assign(enqueued, 0);
        std::vector<unsigned long long> off_bits(enqueued, 0);
        JSVD_HIP_CHECK(hipMemcpyAsync(last_rot.data(), rot_count, sizeof(int) * enqueued, hipMemcpyDeviceToHost,
                                      stream));
        JSVD_HIP_CHECK(hipMemcpyAsync(off_bits.data(), offmax, sizeof(unsigned long long) * enqueued,
                                      hipMemcpyDeviceToHost, stream));
        JSVD_HIP_CHECK(hipStreamSynchronize(stream));
        // Sweep count as in src/kernels.cu:976-980: the first rotation-free sweep counts; without one the
        // solver stops at max_sweeps.
        int executed = enqueued;
        for (int s = 0; s < enqueued; ++s)
        {
            if (last_rot[s] == 0)
            {
                executed = s + 1;
                break;
            }
        }
        last_rot.resize(executed);
        last_off.resize(executed);
        for (int s = 0; s < executed; ++s)
        {
            double x;
            std::memcpy(&x, &off_bits[s], sizeof(double));
            last_off[s] = x;
        }
        last_sweeps = executed;
        if (sweeps_out)
        {
            *sweeps_out = executed;
        }
    }

    void ensure_host_staging()
    {
        if (!d_a)
        {
            // sized for the capacity shape: a reshaped plan reuses them
            JSVD_HIP_CHECK(hipMalloc(&d_a, sizeof(double) * cap_m * cap_n));
            JSVD_HIP_CHECK(hipMalloc(&d_u, sizeof(double) * cap_m * cap_n));
            JSVD_HIP_CHECK(hipMalloc(&d_sigma, sizeof(double) * cap_n));
            JSVD_HIP_CHECK(hipMalloc(&d_v, sizeof(double) * cap_n * cap_n));
        }
    }

    void solve_host(const double *a, double *u, double *sigma, double *v, int *sweeps_out)
    {
        ensure_host_staging();
        JSVD_HIP_CHECK(hipMemcpyAsync(d_a, a, sizeof(double) * m * n, hipMemcpyHostToDevice, stream));
        solve_device(d_a, d_u, d_sigma, d_v, sweeps_out);
        JSVD_HIP_CHECK(hipMemcpyAsync(u, d_u, sizeof(double) * m * n, hipMemcpyDeviceToHost, stream));
        JSVD_HIP_CHECK(hipMemcpyAsync(sigma, d_sigma, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
        JSVD_HIP_CHECK(hipMemcpyAsync(v, d_v, sizeof(double) * n * n, hipMemcpyDeviceToHost, stream));
        JSVD_HIP_CHECK(hipStreamSynchronize(stream));
    }
};

namespace
{
    void require(bool ok, const char *message)
    {
        if (!ok)
        {
            jsvd::throw_invalid(message);
        }
    }

    bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
} // namespace

extern "C" {

int jsvd_plan_create(int64_t m, int64_t n, const jsvd_config *cfg, void *stream, jsvd_plan **plan)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan output pointer must not be null.");
        *plan = nullptr;
        jsvd_config c;
        if (cfg)
        {
            c = *cfg;
        }
        else
        {
            jsvd_config_default(&c);
        }
        jsvd::validate_shape(m, n);
        jsvd::validate_config(c);
        auto p = std::make_unique<jsvd_plan>();
        p->create(m, n, c, static_cast<hipStream_t>(stream));
        *plan = p.release();
    });
}

int jsvd_plan_destroy(jsvd_plan *plan)
{
    return jsvd::guarded([&] {
        if (plan)
        {
            if (plan->stream)
            {
                (void)hipStreamSynchronize(plan->stream);
            }
            delete plan;
        }
    });
}

int jsvd_plan_solve_device(jsvd_plan *plan, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                           int32_t *sweeps)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan must not be null.");
        require(d_a && d_u && d_sigma && d_v, "device buffers must not be null.");
        int s = 0;
        plan->solve_device(d_a, d_u, d_sigma, d_v, &s);
        if (sweeps)
        {
            *sweeps = s;
        }
    });
}

int jsvd_plan_solve_host(jsvd_plan *plan, const double *a, double *u, double *sigma, double *v, int32_t *sweeps)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan must not be null.");
        require(a && u && sigma && v, "host buffers must not be null.");
        int s = 0;
        plan->solve_host(a, u, sigma, v, &s);
        if (sweeps)
        {
            *sweeps = s;
        }
    });
}

int jsvd_plan_sweep_stats(const jsvd_plan *plan, int32_t *rotations, double *offmax, int32_t capacity)
{
    int written = 0;
    const int status = jsvd::guarded([&] {
        require(plan != nullptr, "plan must not be null.");
        written = std::min<int>(plan->last_sweeps, std::max(capacity, 0));
        for (int s = 0; s < written; ++s)
        {
            if (rotations)
            {
                rotations[s] = plan->last_rot[s];
            }
            if (offmax)
            {
                offmax[s] = plan->last_off[s];
            }
        }
    });
    return status == JSVD_OK ? written : -status;
}

int jsvd_plain_f64(const double *a, int64_t m, int64_t n, const jsvd_config *cfg, double *u, double *sigma,
                   double *v, int32_t *sweeps)
{
    return jsvd::guarded([&] {
        jsvd_config c;
        if (cfg)
        {
            c = *cfg;
        }
        else
        {
            jsvd_config_default(&c);
        }
        jsvd::validate_shape(m, n);
        jsvd::validate_config(c);
        require(a && u && sigma && v, "host buffers must not be null.");
        jsvd_plan plan;
        plan.create(m, n, c, nullptr);
        int s = 0;
        plan.solve_host(a, u, sigma, v, &s);
        if (sweeps)
        {
            *sweeps = s;
        }
    });
}

namespace
{
    void validate_mixed(const jsvd_mixed_config &mx, int64_t n)
    {
        if (mx.path != JSVD_PATH_PLAIN_F64 && n <= jsvd::kB)
        {
            jsvd::throw_invalid("block and mixed paths need at least two 32-column blocks (n >= 33).");
        }
        if (mx.path < JSVD_PATH_PLAIN_F64 || mx.path > JSVD_PATH_MIXED)
        {
            jsvd::throw_invalid("unknown solver path.");
        }
        if (mx.block_width != 0 && !jsvd::valid_block_width(mx.block_width))
        {
            jsvd::throw_invalid("block_width must be 0, 16, 32 or 64.");
        }
        if (mx.qr_precondition < 0 || mx.qr_precondition > 3)
        {
            jsvd::throw_invalid("qr_precondition must be 0, 1, 2 or 3.");
        }
        if (mx.inner_order != JSVD_INNER_CROSS && mx.inner_order != JSVD_INNER_FULL)
        {
            jsvd::throw_invalid("inner_order must be JSVD_INNER_CROSS or JSVD_INNER_FULL.");
        }
        if (mx.path == JSVD_PATH_MIXED && (mx.fp32_max_sweeps <= 0 || mx.orthogonalize_steps < 1 ||
                                           mx.orthogonalize_steps > 8 || !(mx.fp32_tolerance < 1.0) ||
                                           !(mx.fp32_stop_fraction >= 0.0 && mx.fp32_stop_fraction < 1.0)))
        {
            jsvd::throw_invalid("invalid mixed-precision configuration.");
        }
    }

    std::unique_ptr<jsvd_plan> make_plan(int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed,
                                         hipStream_t stream)
    {
        jsvd_config c;
        if (cfg)
        {
            c = *cfg;
        }
        else
        {
            jsvd_config_default(&c);
        }
        jsvd::validate_shape(m, n);
        jsvd::validate_config(c);
        auto p = std::make_unique<jsvd_plan>();
        if (mixed && mixed->path != JSVD_PATH_PLAIN_F64)
        {
            validate_mixed(*mixed, n);
            p->create_block(m, n, c, *mixed, stream);
        }
        else
        {
            if (mixed)
            {
                validate_mixed(*mixed, n);
            }
            p->create(m, n, c, stream);
        }
        return p;
    }
} // namespace

int jsvd_plan_create_ex(int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed, void *stream,
                        jsvd_plan **plan)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan output pointer must not be null.");
        *plan = nullptr;
        *plan = make_plan(m, n, cfg, mixed, static_cast<hipStream_t>(stream)).release();
    });
}

int jsvd_plan_set_round_span(jsvd_plan *plan, int32_t waves, int32_t max_rounds)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan must not be null.");
        require(!plan->block, "round spans belong to the plain path.");
        require(waves == 0 || waves == 8 || waves == 16, "waves must be 0, 8 or 16.");
        plan->set_span(waves, max_rounds);
    });
}

int jsvd_plan_round_span(const jsvd_plan *plan, int32_t *waves, int32_t *max_rounds, int32_t *launches_per_sweep)
{
    return jsvd::guarded([&] {
        require(plan != nullptr && waves && max_rounds && launches_per_sweep, "plan and outputs must not be null.");
        *waves = plan->span_waves;
        *max_rounds = plan->span_rounds;
        const int64_t N = plan->n + (plan->n & 1);
        *launches_per_sweep = plan->span_waves > 0 ? static_cast<int32_t>(plan->span_chunk.size())
                                                   : static_cast<int32_t>(plan->n >= 2 ? N - 1 : 0);
    });
}

int jsvd_plan_reshape(jsvd_plan *plan, int64_t m, int64_t n, double epsilon)
{
    return jsvd::guarded([&] {
        require(plan != nullptr, "plan must not be null.");
        require(epsilon > 0.0, "epsilon must be positive.");
        jsvd::validate_shape(m, n);
        require(plan->reshape(m, n, epsilon), "shape does not fit the plan's capacity.");
    });
}

int jsvd_cholesky_f64(double *a, int64_t lda, int64_t n, int32_t *info, void *stream)
{
    return jsvd::guarded([&] {
        require(n >= 1 && n <= (int64_t{1} << 20), "n out of range.");
        require(lda >= jsvd::cholesky_padded_size(n), "lda must be at least n rounded up to 64.");
        require(a && info, "a and info must not be null.");
        jsvd::launch_cholesky_upper(a, lda, n, info, static_cast<hipStream_t>(stream));
    });
}

int jsvd_plan_launch_params(const jsvd_plan *plan, int32_t *round_threads, int32_t *round_ea, int32_t *round_ev,
                            int32_t *apply_workgroups, int32_t *block_width)
{
    return jsvd::guarded([&] {
        require(plan != nullptr && round_threads && round_ea && round_ev && apply_workgroups && block_width,
                "plan and outputs must not be null.");
        const bool blk = plan->block != nullptr;
        *round_threads = blk ? 0 : plan->shape.threads;
        *round_ea = blk ? 0 : plan->shape.ea;
        *round_ev = blk ? 0 : plan->shape.ev;
        *apply_workgroups = blk ? plan->block->geometry().apply_wgs : 0;
        *block_width = blk ? plan->block->geometry().kb : 0;
    });
}

int jsvd_plan_last_info(const jsvd_plan *plan, jsvd_solve_info *info)
{
    return jsvd::guarded([&] {
        require(plan != nullptr && info != nullptr, "plan and info must not be null.");
        *info = plan->last_info;
    });
}

int jsvd_solve_f64(const double *a, int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed,
                   double *u, double *sigma, double *v, jsvd_solve_info *info)
{
    return jsvd::guarded([&] {
        require(a && u && sigma && v, "host buffers must not be null.");
        auto p = make_plan(m, n, cfg, mixed, nullptr);
        int s = 0;
        p->solve_host(a, u, sigma, v, &s);
        if (info)
        {
            *info = p->last_info;
        }
    });
}

int jsvd_block_eigen(int32_t precision, int32_t block_width, const void *gram, int64_t pairs, double epsilon,
                     void *q, int32_t *rotations, void *stream)
{
    return jsvd_block_eigen_ordered(precision, block_width, JSVD_INNER_FULL, gram, pairs, epsilon, q, rotations, 1,
                                    stream);
}

int jsvd_block_eigen_ordered(int32_t precision, int32_t block_width, int32_t inner_order, const void *gram,
                             int64_t pairs, double epsilon, void *q, int32_t *rotations, int32_t repeats,
                             void *stream)
{
    return jsvd::guarded([&] {
        require(repeats >= 1 && repeats <= 4096, "repeats must be in 1..4096.");
        require(inner_order == JSVD_INNER_FULL || (inner_order == JSVD_INNER_CROSS && block_width == 32),
                "inner_order must be JSVD_INNER_FULL, or JSVD_INNER_CROSS at block_width 32.");
        require(precision == JSVD_PRECISION_F32 || precision == JSVD_PRECISION_F64, "unknown precision.");
        require(jsvd::valid_block_width(block_width), "block_width must be 16, 32 or 64.");
        require(pairs >= 0 && pairs <= (int64_t{1} << 20), "pairs out of range.");
        require(epsilon > 0.0, "epsilon must be positive.");
        if (pairs == 0)
        {
            return;
        }
        require(gram && q && rotations, "buffers must not be null.");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        for (int32_t k = 0; k < repeats; ++k)
        {
            if (precision == JSVD_PRECISION_F32)
            {
                jsvd::launch_block_eigen(static_cast<const float *>(gram), block_width, static_cast<int>(pairs),
                                         epsilon, static_cast<float *>(q), rotations,
                                         inner_order == JSVD_INNER_CROSS, s);
            }
            else
            {
                jsvd::launch_block_eigen(static_cast<const double *>(gram), block_width, static_cast<int>(pairs),
                                         epsilon, static_cast<double *>(q), rotations,
                                         inner_order == JSVD_INNER_CROSS, s);
            }
        }
    });
}

int jsvd_round_robin_pairs(int64_t n, int32_t round, int32_t *pairs_out)
{
    int count = 0;
    const int status = jsvd::guarded([&] {
        require(n >= 0 && n <= (int64_t{1} << 30), "n out of range.");
        require(pairs_out != nullptr || n < 2, "pairs_out must not be null.");
        if (n < 2)
        {
            return;
        }
        const int N = static_cast<int>(n + (n & 1));
        require(round >= 0 && round < N - 1, "round out of range.");
        for (int slot = 0; slot < N / 2; ++slot)
        {
            const int p = jsvd::dev::circle_at(slot, round, N);
            const int q = jsvd::dev::circle_at(N - 1 - slot, round, N);
            if (p < n && q < n)
            {
                pairs_out[2 * count] = p;
                pairs_out[2 * count + 1] = q;
                ++count;
            }
        }
    });
    return status == JSVD_OK ? count : -status;
}

int jsvd_transpose_f64(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                       void *stream)
{
    return jsvd::guarded([&] {
        require(rows >= 0 && cols >= 0, "rows and cols must be non-negative.");
        require(lds >= cols && ldd >= rows, "leading dimensions too small.");
        require((src && dst) || rows == 0 || cols == 0, "buffers must not be null.");
        jsvd::launch_transpose(src, lds, dst, ldd, rows, cols, static_cast<hipStream_t>(stream));
    });
}

int jsvd_pair_gram_f64(const double *a, int64_t lda, int64_t m, const int32_t *pairs, int64_t count, double *alpha,
                       double *beta, double *gamma, void *stream)
{
    return jsvd::guarded([&] {
        require(m >= 0 && lda >= m && count >= 0, "invalid sizes.");
        require(count == 0 || (a && pairs && alpha && beta && gamma), "buffers must not be null.");
        jsvd::launch_pair_gram(a, lda, m, pairs, count, alpha, beta, gamma, static_cast<hipStream_t>(stream));
    });
}

int jsvd_round_f64(double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n, int32_t round,
                   double epsilon, int32_t *rot_count, uint64_t *offmax, const int32_t *skip_if_zero, void *stream)
{
    return jsvd::guarded([&] {
        require(m > 0 && n > 0 && m >= n, "invalid shape (need m >= n > 0).");
        require(lda >= m && lda % 2 == 0, "lda must be even and >= m.");
        require(ldv >= n && ldv % 2 == 0, "ldv must be even and >= n.");
        require(a && v && aligned16(a) && aligned16(v), "a and v must be non-null 16-byte aligned device pointers.");
        const int64_t N = n + (n & 1);
        require(n < 2 || (round >= 0 && round < N - 1), "round out of range.");
        require(epsilon > 0.0, "epsilon must be positive.");
        const jsvd::RoundShape shape = jsvd::kernel_round_shape(m, n);
        jsvd::launch_round(shape, a, lda, m, v, ldv, n, round, epsilon, rot_count,
                           reinterpret_cast<unsigned long long *>(offmax), skip_if_zero,
                           static_cast<hipStream_t>(stream));
    });
}

int jsvd_rounds_f64(double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n, int32_t first_round,
                    int32_t rounds, double epsilon, int32_t *rot_count, uint64_t *offmax, int32_t count_slots,
                    const int32_t *skip_if_zero, void *stream)
{
    return jsvd::guarded([&] {
        require(count_slots >= 1 && count_slots <= 4096 && (count_slots & (count_slots - 1)) == 0,
                "count_slots must be a power of two in [1, 4096].");
        require(m > 0 && n > 0 && m >= n, "invalid shape (need m >= n > 0).");
        require(lda >= m && lda % 2 == 0, "lda must be even and >= m.");
        require(ldv >= n && ldv % 2 == 0, "ldv must be even and >= n.");
        require(a && v && aligned16(a) && aligned16(v), "a and v must be non-null 16-byte aligned device pointers.");
        const int64_t N = n + (n & 1);
        require(n < 2 || (first_round >= 0 && first_round < N - 1), "first_round out of range.");
        require(rounds >= 0, "rounds must be non-negative.");
        require(epsilon > 0.0, "epsilon must be positive.");
        if (n < 2)
        {
            return;
        }
        const jsvd::RoundShape shape = jsvd::kernel_round_shape(m, n);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        for (int32_t k = 0; k < rounds; ++k)
        {
            const int round = static_cast<int>((first_round + k) % (N - 1));
            jsvd::launch_round(shape, a, lda, m, v, ldv, n, round, epsilon, rot_count,
                               reinterpret_cast<unsigned long long *>(offmax), skip_if_zero, s, count_slots);
        }
    });
}

int jsvd_rotation_params_f64(const double *alpha, const double *beta, const double *gamma, int64_t count,
                             double epsilon, double *c, double *s, int32_t *rotate, void *stream)
{
    return jsvd::guarded([&] {
        require(count >= 0, "count must be non-negative.");
        require(count == 0 || (alpha && beta && gamma && c && s && rotate), "buffers must not be null.");
        jsvd::launch_rotation_params(alpha, beta, gamma, count, epsilon, c, s, rotate,
                                     static_cast<hipStream_t>(stream));
    });
}

int jsvd_offnorm_f64(const double *a, int64_t lda, int64_t m, int64_t n, double *out, void *stream)
{
    return jsvd::guarded([&] {
        require(m > 0 && n > 0 && lda >= m, "invalid sizes.");
        require(a && out, "buffers must not be null.");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        double *workspace = nullptr;
        JSVD_HIP_CHECK(hipMallocAsync(&workspace, sizeof(double) * jsvd::offnorm_workspace_doubles(n), s));
        jsvd::launch_offnorm(a, lda, m, n, workspace, out, s);
        JSVD_HIP_CHECK(hipFreeAsync(workspace, s));
    });
}

int jsvd_u_sigma_f64(const double *a, int64_t lda, int64_t m, int64_t n, double epsilon, double *u, double *sigma,
                     void *stream)
{
    return jsvd::guarded([&] {
        require(m > 0 && n > 0 && lda >= m, "invalid sizes.");
        require(a && u && sigma, "buffers must not be null.");
        jsvd::launch_u_sigma(a, lda, m, n, epsilon, u, sigma, static_cast<hipStream_t>(stream));
    });
}

} // extern "C"
