// batch.hip -- persistent batched one-sided Jacobi: many small matrices of different shapes in one launch.
//
// Each workgroup repeatedly takes the next item index from a device counter and solves that matrix alone: the
// reference algorithm (round-robin pair order, rotation formula and skip test, rotation-free-sweep
// convergence, sigma/U extraction with the absolute epsilon cutoff) in FP32 or FP64 storage and accumulation,
// with the rotation parameters always in FP64 (FP64 items: the reference formula; FP32 items: the same skip test
// and rotation by a shorter chain).  Column pairs of a round are spread over the workgroup's
// wavefronts (pair k -> wave k % waves), each pair reduced within its wavefront by DPP row moves, so an item's
// arithmetic does not depend on the workgroup size, the queue order or the other items.
//
// A (column-major, leading dimension m rounded up to even) and V (column-major n x n, leading dimension n
// rounded up to even) live in LDS when they fit in the workgroup's budget, otherwise in the workgroup's
// slice of a global scratch area (L2/MALL resident).  The workgroup size and LDS budget form the launch
// shape, chosen among a fixed table by measurement (jsvd_batch_autotune).
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "autotune.hpp"
#include "batch_kernel.hpp"
#include "common.hpp"


namespace jsvd
{
    namespace
    {
        using BatchKernel = void (*)(const jsvd_batch_item *, int64_t, jsvd_batch_buffers, unsigned long long *,
                                     double *, int64_t, int, int);

        // The launch shapes the batch autotuner chooses among: threads per workgroup (the wavefronts that share
        // one matrix's column pairs) and the LDS bytes each workgroup stages its matrix in (what does not fit
        // goes to the workgroup's scratch slice).  Every shape runs the same per-item arithmetic, so an item's
        // result is bit-identical whichever shape solves it.
        struct BatchShape
        {
            int threads;
            int lds_bytes;
            BatchKernel kernel;
        };
        constexpr int kKiB = 1024;
        constexpr int64_t kScratchBudgetBytes = int64_t(8) << 30; // global scratch of one batch object
        const BatchShape kShapes[] = {
            {256, 76 * kKiB, batch_kernel<256, 76 * kKiB>},     // default: 2 workgroups x 4 waves per CU
            {128, 38 * kKiB, batch_kernel<128, 38 * kKiB>},     // 4 x 2 waves, more items in flight
            {256, 38 * kKiB, batch_kernel<256, 38 * kKiB>},     // 4 x 4 waves
            {512, 76 * kKiB, batch_kernel<512, 76 * kKiB>},     // 2 x 8 waves
            {512, 152 * kKiB, batch_kernel<512, 152 * kKiB>},   // 1 x 8 waves, most matrices LDS-resident
            {1024, 152 * kKiB, batch_kernel<1024, 152 * kKiB>}, // 1 x 16 waves
        };
        constexpr int kShapeCount = static_cast<int>(sizeof(kShapes) / sizeof(kShapes[0]));
        static_assert(kShapeCount <= JSVD_BATCH_MAX_SHAPES, "report arrays hold every shape");

        int find_shape(int threads, int lds_bytes)
        {
            if (threads == 0 && lds_bytes == 0)
            {
                return 0;
            }
            for (int i = 0; i < kShapeCount; ++i)
            {
                if (kShapes[i].threads == threads && kShapes[i].lds_bytes == lds_bytes)
                {
                    return i;
                }
            }
            return -1;
        }
    } // namespace
} // namespace jsvd

struct jsvd_batch
{
    int64_t max_rows = 0;
    int64_t max_cols = 0;
    int shape = 0;             // index into the shape table
    int grid = 0;              // resident workgroups of the current shape
    int cus = 1;
    int64_t scratch_per_wg = 0; // doubles
    int64_t scratch_wgs = 0;    // workgroup slices the scratch area holds
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *scratch = nullptr;
    unsigned long long *counter = nullptr;
    // device staging of the host-pointer entry point, grown on demand (bytes held in *_cap)
    void *st_items = nullptr, *st_a = nullptr, *st_u = nullptr, *st_s = nullptr, *st_v = nullptr, *st_sw = nullptr;
    size_t cap_items = 0, cap_a = 0, cap_u = 0, cap_s = 0, cap_v = 0, cap_sw = 0;

    ~jsvd_batch()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
        }
        for (void *p : {st_items, st_a, st_u, st_s, st_v, st_sw})
        {
            (void)hipFree(p);
        }
        (void)hipFree(scratch);
        (void)hipFree(counter);
        if (own_stream && stream)
        {
            (void)hipStreamDestroy(stream);
        }
    }

    // Selects launch shape `index`: one resident wave of workgroups (occupancy x CUs) and a scratch slice each.
    void set_shape(int index)
    {
        const jsvd::BatchShape &sh = jsvd::kShapes[index];
        int per_cu = 0;
        JSVD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(sh.kernel), sh.threads, 0));
        // One resident wave of workgroups, fewer where their scratch slices would exceed the budget (large
        // limits: a 4096 x 4096 slice is 268 MB); the persistent workgroups drain the queue either way.
        const int64_t by_budget = std::max<int64_t>(1, jsvd::kScratchBudgetBytes / (sizeof(double) * scratch_per_wg));
        const int g = static_cast<int>(std::min<int64_t>(std::max(per_cu, 1) * cus, by_budget));
        if (g > scratch_wgs)
        {
            JSVD_HIP_CHECK(hipStreamSynchronize(stream)); // launches in flight may still use the old area
            JSVD_HIP_CHECK(hipFree(scratch));
            scratch = nullptr;
            scratch_wgs = 0;
            JSVD_HIP_CHECK(hipMalloc(&scratch, sizeof(double) * scratch_per_wg * g));
            scratch_wgs = g;
        }
        shape = index;
        grid = g;
    }

    // Grows a staging area to hold `bytes` (the stream is idle: callers synchronise before returning).
    void ensure(void *&p, size_t &cap, size_t bytes)
    {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes <= cap)
        {
            return;
        }
        JSVD_HIP_CHECK(hipFree(p));
        p = nullptr;
        cap = 0;
        JSVD_HIP_CHECK(hipMalloc(&p, bytes));
        cap = bytes;
    }

    void enqueue(const jsvd_batch_item *d_items, int64_t count, const jsvd_batch_buffers &buf)
    {
        if (count == 0)
        {
            return;
        }
        JSVD_HIP_CHECK(hipMemsetAsync(counter, 0, sizeof(unsigned long long), stream));
        const int blocks = static_cast<int>(std::min<int64_t>(grid, count));
        const jsvd::BatchShape &sh = jsvd::kShapes[shape];
        hipLaunchKernelGGL(sh.kernel, dim3(blocks), dim3(sh.threads), 0, stream, d_items, count, buf, counter, scratch,
                           scratch_per_wg, static_cast<int>(max_rows), static_cast<int>(max_cols));
        JSVD_HIP_CHECK(hipGetLastError());
    }
};

namespace
{
    void check_buffers(const jsvd_batch_buffers *buf, int64_t count)
    {
        if (count < 0)
        {
            jsvd::throw_invalid("count must be non-negative.");
        }
        if (count > 0 && (!buf || !buf->a || !buf->u || !buf->sigma || !buf->v || !buf->sweeps))
        {
            jsvd::throw_invalid("batch buffers must not be null.");
        }
    }

    // The host side of the item contract (JSVD_BATCH_MIN_COLUMNS in jsvd.h): items the kernel's own check would
    // reject, and items with fewer than JSVD_BATCH_MIN_COLUMNS columns, are refused before anything is launched.
    void check_item(const jsvd_batch_item &it, const jsvd_batch_buffers &lens, int64_t max_rows, int64_t max_cols,
                    int64_t index)
    {
        const int64_t m = it.rows;
        const int64_t n = it.columns;
        const std::string at = "batch item " + std::to_string(index) + ": ";
        if (n < JSVD_BATCH_MIN_COLUMNS || n > max_cols || m < n || m > max_rows)
        {
            jsvd::throw_invalid(at + "needs " + std::to_string(JSVD_BATCH_MIN_COLUMNS) +
                                " <= columns <= rows within the batch limits (" + std::to_string(max_rows) + " x " +
                                std::to_string(max_cols) + "), got " + std::to_string(m) + " x " + std::to_string(n) +
                                ".");
        }
        if (it.max_sweeps < 1 || !(it.tolerance > 0.0) ||
            (it.precision != JSVD_PRECISION_F64 && it.precision != JSVD_PRECISION_F32))
        {
            jsvd::throw_invalid(at + "needs max_sweeps >= 1, tolerance > 0 and a known precision.");
        }
        if (it.a_offset < 0 || it.a_offset + m * n > lens.a_len || it.u_offset < 0 || it.u_offset + m * n > lens.u_len ||
            it.sigma_offset < 0 || it.sigma_offset + n > lens.sigma_len || it.v_offset < 0 ||
            it.v_offset + n * n > lens.v_len)
        {
            jsvd::throw_invalid(at + "offsets outside the packed buffers.");
        }
    }
} // namespace

// Device memory of the allocation-free entry point (jsvd_batch_solve_noalloc): the first 256 bytes hold the
// persistent kernel's queue counter, the rest one scratch slice per resident workgroup.  Every GPU the library
// runs on holds its own copy from the first use until the process ends.  Host code reaches it only through the
// per-device address from hipGetSymbolAddress (static_scratch_address below), never through the name.
constexpr int64_t kStaticScratchBytes = int64_t(128) << 20;
constexpr int64_t kStaticHeaderDoubles = 256 / sizeof(double);
static __device__ __align__(256) double scratch[kStaticScratchBytes / sizeof(double)];

namespace
{
    __global__ void clear_queue_counter_kernel(unsigned long long *counter)
    {
        *counter = 0ull; // a kernel, not hipMemsetAsync, so a caller may capture the call in a graph
    }

    constexpr int kStaticMaxDevices = 64;
    std::mutex g_static_mu;
    void *g_static_address[kStaticMaxDevices] = {};
    int g_static_per_cu[kStaticMaxDevices][JSVD_BATCH_MAX_SHAPES] = {};
    int g_static_cus[kStaticMaxDevices] = {};

    // Device address of `scratch` on the current device (looked up once per device, with that device current),
    // plus the resident workgroups per CU of launch shape `shape` and the CU count.  Allocates nothing.
    double *static_scratch_address(int shape, int &per_cu, int &cus)
    {
        int device = 0;
        JSVD_HIP_CHECK(hipGetDevice(&device));
        if (device < 0 || device >= kStaticMaxDevices)
        {
            jsvd::throw_invalid("device index beyond the allocation-free entry point's table.");
        }
        std::lock_guard<std::mutex> lock(g_static_mu);
        if (!g_static_address[device])
        {
            void *ptr = nullptr;
            JSVD_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            int n_cus = 0;
            JSVD_HIP_CHECK(hipDeviceGetAttribute(&n_cus, hipDeviceAttributeMultiprocessorCount, device));
            g_static_cus[device] = std::max(n_cus, 1);
            g_static_address[device] = ptr;
        }
        if (g_static_per_cu[device][shape] == 0)
        {
            int occ = 0;
            JSVD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, reinterpret_cast<const void *>(jsvd::kShapes[shape].kernel), jsvd::kShapes[shape].threads, 0));
            g_static_per_cu[device][shape] = std::max(occ, 1);
        }
        per_cu = g_static_per_cu[device][shape];
        cus = g_static_cus[device];
        return static_cast<double *>(g_static_address[device]);
    }
} // namespace

extern "C" {

int jsvd_batch_create(int64_t max_rows, int64_t max_columns, void *stream, jsvd_batch **batch)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch output pointer must not be null.");
        }
        *batch = nullptr;
        if (max_columns < 1 || max_rows < max_columns || max_rows > 4096)
        {
            jsvd::throw_invalid("batch limits need 1 <= max_columns <= max_rows <= 4096.");
        }
        auto b = std::make_unique<jsvd_batch>();
        b->max_rows = max_rows;
        b->max_cols = max_columns;
        int device = 0;
        JSVD_HIP_CHECK(hipGetDevice(&device));
        int cus = 0;
        JSVD_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        b->cus = std::max(cus, 1);
        b->scratch_per_wg = (max_rows + 1) * max_columns + (max_columns + 1) * max_columns;
        if (stream)
        {
            b->stream = static_cast<hipStream_t>(stream);
        }
        else
        {
            JSVD_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            b->own_stream = true;
        }
        JSVD_HIP_CHECK(hipMalloc(&b->counter, sizeof(unsigned long long)));
        jsvd::Tuning t;
        const bool tuned = jsvd::lookup_tuning(max_rows, max_columns, t) && t.has_batch;
        b->set_shape(tuned ? t.batch_shape : 0);
        *batch = b.release();
    });
}

int jsvd_batch_destroy(jsvd_batch *batch)
{
    return jsvd::guarded([&] { delete batch; });
}

int jsvd_batch_solve_async(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                           const jsvd_batch_buffers *buf, void *done)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch must not be null.");
        }
        check_buffers(buf, count);
        if (count > 0 && !d_items)
        {
            jsvd::throw_invalid("items must not be null.");
        }
        batch->enqueue(d_items, count, *buf);
        if (done)
        {
            JSVD_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(done), batch->stream));
        }
    });
}

int jsvd_batch_solve_device(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                            const jsvd_batch_buffers *buf)
{
    const int st = jsvd_batch_solve_async(batch, d_items, count, buf, nullptr);
    if (st != JSVD_OK)
    {
        return st;
    }
    return jsvd::guarded([&] { JSVD_HIP_CHECK(hipStreamSynchronize(batch->stream)); });
}

int jsvd_batch_solve_noalloc(const jsvd_batch_item *d_items, int64_t count, const jsvd_batch_buffers *buf,
                             int64_t max_rows, int64_t max_columns, void *stream)
{
    return jsvd::guarded([&] {
        check_buffers(buf, count);
        if (count > 0 && !d_items)
        {
            jsvd::throw_invalid("items must not be null.");
        }
        if (max_columns < 1 || max_rows < max_columns || max_rows > 4096)
        {
            jsvd::throw_invalid("batch limits need 1 <= max_columns <= max_rows <= 4096.");
        }
        // one scratch slice per resident workgroup, as jsvd_batch_create sizes them
        const int64_t slice = (max_rows + 1) * max_columns + (max_columns + 1) * max_columns;
        const int64_t slices = (kStaticScratchBytes / int64_t(sizeof(double)) - kStaticHeaderDoubles) / slice;
        if (slices < 1)
        {
            jsvd::throw_invalid("batch limits " + std::to_string(max_rows) + " x " + std::to_string(max_columns) +
                                " need a scratch slice of " + std::to_string(slice * sizeof(double)) +
                                " bytes per workgroup; the allocation-free entry point holds " +
                                std::to_string(kStaticScratchBytes - 256) + " bytes.");
        }
        if (count == 0)
        {
            return;
        }
        jsvd::Tuning t;
        const int shape = (jsvd::lookup_tuning(max_rows, max_columns, t) && t.has_batch) ? t.batch_shape : 0;
        int per_cu = 1, cus = 1;
        double *base = static_scratch_address(shape, per_cu, cus);
        const int64_t grid = std::min<int64_t>(std::min<int64_t>(int64_t(per_cu) * cus, slices), count);
        hipStream_t s = static_cast<hipStream_t>(stream);
        auto *counter = reinterpret_cast<unsigned long long *>(base);
        hipLaunchKernelGGL(clear_queue_counter_kernel, dim3(1), dim3(1), 0, s, counter);
        JSVD_HIP_CHECK(hipGetLastError());
        const jsvd::BatchShape &sh = jsvd::kShapes[shape];
        hipLaunchKernelGGL(sh.kernel, dim3(static_cast<unsigned>(grid)), dim3(sh.threads), 0, s, d_items, count, *buf,
                           counter, base + kStaticHeaderDoubles, slice, static_cast<int>(max_rows),
                           static_cast<int>(max_columns));
        JSVD_HIP_CHECK(hipGetLastError());
    });
}

int jsvd_batch_set_shape(jsvd_batch *batch, int32_t threads, int32_t lds_bytes)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch must not be null.");
        }
        const int index = jsvd::find_shape(threads, lds_bytes);
        if (index < 0)
        {
            jsvd::throw_invalid("unknown batch launch shape (threads, lds_bytes).");
        }
        batch->set_shape(index);
    });
}

int jsvd_batch_get_shape(const jsvd_batch *batch, int32_t *threads, int32_t *lds_bytes, int32_t *workgroups)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch must not be null.");
        }
        const jsvd::BatchShape &sh = jsvd::kShapes[batch->shape];
        if (threads)
        {
            *threads = sh.threads;
        }
        if (lds_bytes)
        {
            *lds_bytes = sh.lds_bytes;
        }
        if (workgroups)
        {
            *workgroups = batch->grid;
        }
    });
}

int jsvd_batch_autotune(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                        const jsvd_batch_buffers *buf, int32_t repetitions, jsvd_batch_tune_report *report)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch must not be null.");
        }
        check_buffers(buf, count);
        if (count < 1 || !d_items)
        {
            jsvd::throw_invalid("autotuning needs at least one item.");
        }
        if (repetitions < 1)
        {
            jsvd::throw_invalid("repetitions must be positive.");
        }
        jsvd_batch_tune_report rep{};
        rep.candidates = jsvd::kShapeCount;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        JSVD_HIP_CHECK(hipEventCreate(&t0));
        JSVD_HIP_CHECK(hipEventCreate(&t1));
        int best = 0;
        try
        {
            for (int i = 0; i < jsvd::kShapeCount; ++i)
            {
                batch->set_shape(i);
                batch->enqueue(d_items, count, *buf); // warm-up (first launch of this shape)
                JSVD_HIP_CHECK(hipEventRecord(t0, batch->stream));
                for (int r = 0; r < repetitions; ++r)
                {
                    batch->enqueue(d_items, count, *buf);
                }
                JSVD_HIP_CHECK(hipEventRecord(t1, batch->stream));
                JSVD_HIP_CHECK(hipEventSynchronize(t1));
                float ms = 0.0f;
                JSVD_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
                rep.threads[i] = jsvd::kShapes[i].threads;
                rep.lds_bytes[i] = jsvd::kShapes[i].lds_bytes;
                rep.workgroups[i] = batch->grid;
                rep.ms[i] = static_cast<double>(ms) / repetitions;
                if (rep.ms[i] < rep.ms[best])
                {
                    best = i;
                }
            }
        }
        catch (...)
        {
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            throw;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        batch->set_shape(best);
        jsvd::store_batch_shape(batch->max_rows, batch->max_cols, best);
        rep.chosen = best;
        if (report)
        {
            *report = rep;
        }
    });
}

int jsvd_batch_solve_host(jsvd_batch *batch, const jsvd_batch_item *items, int64_t count, const double *a,
                          int64_t a_len, double *u, int64_t u_len, double *sigma, int64_t sigma_len, double *v,
                          int64_t v_len, int32_t *sweeps)
{
    return jsvd::guarded([&] {
        if (!batch)
        {
            jsvd::throw_invalid("batch must not be null.");
        }
        if (count < 0 || a_len < 0 || u_len < 0 || sigma_len < 0 || v_len < 0)
        {
            jsvd::throw_invalid("count and buffer lengths must be non-negative.");
        }
        if (count == 0)
        {
            return;
        }
        if (!items || !a || !u || !sigma || !v || !sweeps)
        {
            jsvd::throw_invalid("batch host buffers must not be null.");
        }
        const jsvd_batch_buffers lens{nullptr, nullptr, nullptr, nullptr, nullptr, a_len, u_len, sigma_len, v_len};
        for (int64_t i = 0; i < count; ++i)
        {
            check_item(items[i], lens, batch->max_rows, batch->max_cols, i);
        }
        const size_t bi = sizeof(jsvd_batch_item) * static_cast<size_t>(count);
        const size_t ba = sizeof(double) * static_cast<size_t>(a_len), bu = sizeof(double) * static_cast<size_t>(u_len);
        const size_t bs = sizeof(double) * static_cast<size_t>(sigma_len), bv = sizeof(double) * static_cast<size_t>(v_len);
        const size_t bw = sizeof(int32_t) * static_cast<size_t>(count);
        JSVD_HIP_CHECK(hipStreamSynchronize(batch->stream)); // staging may still be read by an earlier call
        batch->ensure(batch->st_items, batch->cap_items, bi);
        batch->ensure(batch->st_a, batch->cap_a, ba);
        batch->ensure(batch->st_u, batch->cap_u, bu);
        batch->ensure(batch->st_s, batch->cap_s, bs);
        batch->ensure(batch->st_v, batch->cap_v, bv);
        batch->ensure(batch->st_sw, batch->cap_sw, bw);
        hipStream_t s = batch->stream;
        JSVD_HIP_CHECK(hipMemcpyAsync(batch->st_items, items, bi, hipMemcpyHostToDevice, s));
        JSVD_HIP_CHECK(hipMemcpyAsync(batch->st_a, a, ba, hipMemcpyHostToDevice, s));
        const jsvd_batch_buffers buf{static_cast<const double *>(batch->st_a), static_cast<double *>(batch->st_u),
                                     static_cast<double *>(batch->st_s), static_cast<double *>(batch->st_v),
                                     static_cast<int32_t *>(batch->st_sw), a_len, u_len, sigma_len, v_len};
        batch->enqueue(static_cast<const jsvd_batch_item *>(batch->st_items), count, buf);
        JSVD_HIP_CHECK(hipMemcpyAsync(u, batch->st_u, bu, hipMemcpyDeviceToHost, s));
        JSVD_HIP_CHECK(hipMemcpyAsync(sigma, batch->st_s, bs, hipMemcpyDeviceToHost, s));
        JSVD_HIP_CHECK(hipMemcpyAsync(v, batch->st_v, bv, hipMemcpyDeviceToHost, s));
        JSVD_HIP_CHECK(hipMemcpyAsync(sweeps, batch->st_sw, bw, hipMemcpyDeviceToHost, s));
        JSVD_HIP_CHECK(hipStreamSynchronize(s));
    });
}

} // extern "C"
