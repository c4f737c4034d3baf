// device_stub.cpp -- stand-in for the GPU side of libjsvd in the host sanitizer builds (test infrastructure only).
//
// The HIP runtime calls of hip/hip_runtime.h run on the host (heap "device" memory, stream operations complete
// before they return), and the solver entry points that launch kernels (plans, batches) solve each matrix with
// the oracle's restatement of the reference algorithm (oracle/jacobi_oracle.c), so the host code above them --
// request loop lanes, pipeline reader/solver/writer threads, file I/O, CLI -- runs unchanged and its outputs can
// be checked.  Allocations are exact-sized heap blocks, so AddressSanitizer sees any access past a "device"
// buffer.  JSVD_STUB_FAIL_BATCH_CREATE=1 in the environment makes jsvd_batch_create fail (a lane whose
// initialisation fails).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <string>

#include "common.hpp"

extern "C" int oracle_svd(const double *a_rm, int64_t m, int64_t n, double eps, int max_sweeps, int nthreads,
                          double *u_rm, double *sigma, double *v_rm, int *sweeps_out);

namespace
{
    std::atomic<int> g_streams{0};

    hipError_t host_alloc(void **p, size_t bytes)
    {
        if (!p)
        {
            return hipErrorInvalidValue;
        }
        *p = std::malloc(bytes ? bytes : 1);
        return *p ? hipSuccess : hipErrorOutOfMemory;
    }

    void solve(const double *a, int64_t m, int64_t n, double eps, int max_sweeps, double *u, double *sigma, double *v,
               int *sweeps)
    {
        if (oracle_svd(a, m, n, eps, max_sweeps, 1, u, sigma, v, sweeps) != 0)
        {
            throw jsvd::Error(JSVD_ERR_INTERNAL, "oracle solve failed");
        }
    }

    bool item_fits(const jsvd_batch_item &it, int64_t max_rows, int64_t max_columns, int64_t a_len, int64_t u_len,
                   int64_t s_len, int64_t v_len)
    {
        const int64_t m = it.rows, n = it.columns;
        return n >= JSVD_BATCH_MIN_COLUMNS && n <= max_columns && m >= n && m <= max_rows && it.max_sweeps >= 1 &&
               it.tolerance > 0.0 && (it.precision == JSVD_PRECISION_F64 || it.precision == JSVD_PRECISION_F32) &&
               it.a_offset >= 0 && it.a_offset + m * n <= a_len && it.u_offset >= 0 && it.u_offset + m * n <= u_len &&
               it.sigma_offset >= 0 && it.sigma_offset + n <= s_len && it.v_offset >= 0 &&
               it.v_offset + n * n <= v_len;
    }
} // namespace

hipError_t hipMalloc(void **p, size_t bytes) { return host_alloc(p, bytes); }
hipError_t hipFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMallocAsync(void **p, size_t bytes, hipStream_t) { return host_alloc(p, bytes); }
hipError_t hipFreeAsync(void *p, hipStream_t)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return host_alloc(p, bytes); }
hipError_t hipHostFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind)
{
    if (bytes)
    {
        std::memmove(dst, src, bytes);
    }
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t)
{
    return hipMemcpy(dst, src, bytes, kind);
}
hipError_t hipStreamCreateWithFlags(hipStream_t *stream, unsigned)
{
    *stream = new int(g_streams.fetch_add(1));
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t stream)
{
    delete static_cast<int *>(stream);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *device)
{
    *device = 0;
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t status)
{
    return status == hipSuccess ? "no error" : status == hipErrorOutOfMemory ? "out of memory" : "invalid value";
}

struct jsvd_plan
{
    int64_t cap_m, cap_n, m, n;
    double eps;
    int max_sweeps;
};

struct jsvd_batch
{
    int64_t max_rows, max_columns;
};

extern "C" {

int jsvd_plan_create(int64_t m, int64_t n, const jsvd_config *cfg, void *, jsvd_plan **plan)
{
    return jsvd::guarded([&] {
        if (!plan)
        {
            jsvd::throw_invalid("plan output pointer must not be null.");
        }
        *plan = nullptr;
        jsvd_config c;
        jsvd_config_default(&c);
        if (cfg)
        {
            c = *cfg;
        }
        jsvd::validate_shape(m, n);
        jsvd::validate_config(c);
        *plan = new jsvd_plan{m, n, m, n, c.epsilon, c.max_sweeps};
    });
}

int jsvd_plan_create_ex(int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed, void *stream,
                        jsvd_plan **plan)
{
    if (mixed && mixed->path != JSVD_PATH_PLAIN_F64 && n <= 32)
    {
        return jsvd::guarded([] { jsvd::throw_invalid("block and mixed paths need at least two 32-column blocks."); });
    }
    return jsvd_plan_create(m, n, cfg, stream, plan);
}

int jsvd_plan_destroy(jsvd_plan *plan)
{
    delete plan;
    return JSVD_OK;
}

int jsvd_plan_reshape(jsvd_plan *plan, int64_t m, int64_t n, double epsilon)
{
    return jsvd::guarded([&] {
        if (!plan || m < n || n < 1 || m > plan->cap_m || n > plan->cap_n || !(epsilon > 0.0))
        {
            jsvd::throw_invalid("shape outside the plan's capacity.");
        }
        plan->m = m;
        plan->n = n;
        plan->eps = epsilon;
    });
}

int jsvd_plan_solve_host(jsvd_plan *plan, const double *a, double *u, double *sigma, double *v, int32_t *sweeps)
{
    return jsvd::guarded([&] {
        if (!plan || !a || !u || !sigma || !v)
        {
            jsvd::throw_invalid("null argument.");
        }
        int s = 0;
        solve(a, plan->m, plan->n, plan->eps, plan->max_sweeps, u, sigma, v, &s);
        if (sweeps)
        {
            *sweeps = s;
        }
    });
}

int jsvd_plain_f64(const double *a, int64_t m, int64_t n, const jsvd_config *cfg, double *u, double *sigma, double *v,
                   int32_t *sweeps)
{
    jsvd_plan *plan = nullptr;
    int st = jsvd_plan_create(m, n, cfg, nullptr, &plan);
    if (st == JSVD_OK)
    {
        st = jsvd_plan_solve_host(plan, a, u, sigma, v, sweeps);
        jsvd_plan_destroy(plan);
    }
    return st;
}

int jsvd_batch_create(int64_t max_rows, int64_t max_columns, void *, jsvd_batch **batch)
{
    return jsvd::guarded([&] {
        const char *fail = std::getenv("JSVD_STUB_FAIL_BATCH_CREATE");
        if (fail && std::strcmp(fail, "1") == 0)
        {
            throw jsvd::Error(JSVD_ERR_DEVICE, "batch creation failed (stub)");
        }
        if (!batch || max_columns < 1 || max_rows < max_columns || max_rows > 4096)
        {
            jsvd::throw_invalid("batch limits must satisfy 1 <= max_columns <= max_rows <= 4096.");
        }
        *batch = new jsvd_batch{max_rows, max_columns};
    });
}

int jsvd_batch_destroy(jsvd_batch *batch)
{
    delete batch;
    return JSVD_OK;
}

int jsvd_batch_solve_async(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                           const jsvd_batch_buffers *buf, void *)
{
    return jsvd::guarded([&] {
        if (!batch || !buf || count < 0 || (count > 0 && !d_items))
        {
            jsvd::throw_invalid("null argument.");
        }
        for (int64_t q = 0; q < count; ++q)
        {
            const jsvd_batch_item &it = d_items[q];
            if (!item_fits(it, batch->max_rows, batch->max_columns, buf->a_len, buf->u_len, buf->sigma_len,
                           buf->v_len))
            {
                buf->sweeps[q] = -1;
                continue;
            }
            int s = 0;
            solve(buf->a + it.a_offset, it.rows, it.columns, it.tolerance, it.max_sweeps, buf->u + it.u_offset,
                  buf->sigma + it.sigma_offset, buf->v + it.v_offset, &s);
            buf->sweeps[q] = s;
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
        for (int64_t q = 0; q < count; ++q)
        {
            if (!item_fits(items[q], batch->max_rows, batch->max_columns, a_len, u_len, sigma_len, v_len))
            {
                jsvd::throw_invalid("batch item outside the batch limits or buffers.");
            }
        }
        jsvd_batch_buffers buf{a, u, sigma, v, sweeps, a_len, u_len, sigma_len, v_len};
        const int st = jsvd_batch_solve_async(batch, items, count, &buf, nullptr);
        if (st != JSVD_OK)
        {
            throw jsvd::Error(st, jsvd_last_error());
        }
    });
}

int jsvd_mat_stream_solve(const char *, const char *, const jsvd_config *, const jsvd_mixed_config *,
                          jsvd_mat_stream_report *)
{
    return jsvd::guarded([] { jsvd::throw_invalid("the streamed *.mat path is not part of the host sanitizer build."); });
}

} // extern "C"
