// batch.hip -- persistent batched one-sided Jacobi: many small matrices of different shapes in one launch.
//
// Each 256-thread workgroup repeatedly takes the next item index from a device counter and solves that
// matrix alone: the reference algorithm (round-robin pair order, rotation formula and skip test,
// rotation-free-sweep convergence, sigma/U extraction with the absolute epsilon cutoff) in FP32 or FP64
// storage and accumulation, with the rotation parameters always in FP64.  Column pairs of a round are spread
// over the four wavefronts (pair k -> wave k % 4), each pair reduced within its wavefront by DPP row moves, so
// an item's arithmetic does not depend on the workgroup, the queue order or the other items.
//
// A (column-major, leading dimension m rounded up to even) and V (column-major n x n, leading dimension n
// rounded up to even) live in LDS when they fit in the workgroup's budget, otherwise in the workgroup's
// slice of a global scratch area (L2/MALL resident).
#include <algorithm>
#include <cstring>
#include <memory>

#include "common.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int kBatchThreads = 256;
        constexpr int kBatchWaves = kBatchThreads / 64;
        constexpr int kBatchLdsBytes = 76 * 1024; // two workgroups per CU

        __device__ inline bool item_ok(const jsvd_batch_item &it, const jsvd_batch_buffers &b, int max_rows,
                                       int max_cols)
        {
            const int64_t m = it.rows;
            const int64_t n = it.columns;
            if (n < 1 || m < n || m > max_rows || n > max_cols || it.max_sweeps < 1 || !(it.tolerance > 0.0) ||
                (it.precision != JSVD_PRECISION_F64 && it.precision != JSVD_PRECISION_F32))
            {
                return false;
            }
            return it.a_offset >= 0 && it.a_offset + m * n <= b.a_len && it.u_offset >= 0 &&
                   it.u_offset + m * n <= b.u_len && it.sigma_offset >= 0 && it.sigma_offset + n <= b.sigma_len &&
                   it.v_offset >= 0 && it.v_offset + n * n <= b.v_len;
        }

        template <typename T>
        __device__ void solve_item(const jsvd_batch_item &it, const jsvd_batch_buffers &b, T *a, T *v, int *sweep_rot,
                                   int64_t index)
        {
            const int m = it.rows;
            const int n = it.columns;
            const int lda = m + (m & 1);
            const int ldv = n + (n & 1);
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            const double eps = it.tolerance;
            // A row-major (global, FP64) -> column-major T; V = I.
            const double *src = b.a + it.a_offset;
            for (int e = tid; e < m * n; e += kBatchThreads)
            {
                const int i = e / n;
                const int j = e % n;
                a[j * lda + i] = static_cast<T>(src[e]);
            }
            for (int e = tid; e < n * ldv; e += kBatchThreads)
            {
                v[e] = (e / ldv == e % ldv) ? T(1) : T(0);
            }
            __syncthreads();
            const int N = n + (n & 1);
            const int rounds = (n >= 2) ? N - 1 : 0;
            int executed = 0;
            for (int s = 0; s < it.max_sweeps; ++s)
            {
                if (tid == 0)
                {
                    *sweep_rot = 0;
                }
                __syncthreads();
                for (int r = 0; r < rounds; ++r)
                {
                    for (int k = wave; k < N / 2; k += kBatchWaves)
                    {
                        const int p = dev::circle_at_fast(k, r, N);
                        const int q = dev::circle_at_fast(N - 1 - k, r, N);
                        if (p >= n || q >= n)
                        {
                            continue;
                        }
                        T *cp = a + p * lda;
                        T *cq = a + q * lda;
                        T spp = 0, sqq = 0, spq = 0;
                        for (int i = lane; i < m; i += 64)
                        {
                            const T x = cp[i];
                            const T y = cq[i];
                            spp += x * x;
                            sqq += y * y;
                            spq += x * y;
                        }
                        spp = dev::wave_sum_dpp(spp);
                        sqq = dev::wave_sum_dpp(sqq);
                        spq = dev::wave_sum_dpp(spq);
                        double c, sn;
                        if (dev::rotation_params(static_cast<double>(spp), static_cast<double>(sqq),
                                                 static_cast<double>(spq), eps, c, sn))
                        {
                            const T ct = static_cast<T>(c);
                            const T st = static_cast<T>(sn);
                            for (int i = lane; i < m; i += 64)
                            {
                                const T x = cp[i];
                                const T y = cq[i];
                                cp[i] = ct * x - st * y;
                                cq[i] = st * x + ct * y;
                            }
                            T *vp = v + p * ldv;
                            T *vq = v + q * ldv;
                            for (int i = lane; i < n; i += 64)
                            {
                                const T x = vp[i];
                                const T y = vq[i];
                                vp[i] = ct * x - st * y;
                                vq[i] = st * x + ct * y;
                            }
                            if (lane == 0)
                            {
                                atomicAdd(sweep_rot, 1);
                            }
                        }
                    }
                    __syncthreads();
                }
                executed = s + 1;
                const int rotated = *sweep_rot;
                __syncthreads(); // everyone has read the counter before the next sweep clears it
                if (rotated == 0)
                {
                    break;
                }
            }
            // sigma_j = sqrt(sum a_ij^2) in FP64 over the stored values; U = A / sigma (absolute cutoff eps).
            double *u_out = b.u + it.u_offset;
            double *s_out = b.sigma + it.sigma_offset;
            double *v_out = b.v + it.v_offset;
            for (int j = wave; j < n; j += kBatchWaves)
            {
                const T *cj = a + j * lda;
                double acc = 0.0;
                for (int i = lane; i < m; i += 64)
                {
                    const double x = static_cast<double>(cj[i]);
                    acc += x * x;
                }
                acc = dev::wave_sum(acc);
                const double sj = sqrt(fmax(acc, 0.0));
                if (lane == 0)
                {
                    s_out[j] = sj;
                }
                for (int i = lane; i < m; i += 64)
                {
                    u_out[static_cast<int64_t>(i) * n + j] = (sj > eps) ? (static_cast<double>(cj[i]) / sj) : 0.0;
                }
            }
            for (int e = tid; e < n * n; e += kBatchThreads)
            {
                const int i = e / n;
                const int j = e % n;
                v_out[e] = static_cast<double>(v[j * ldv + i]);
            }
            if (tid == 0)
            {
                b.sweeps[index] = executed;
            }
            __syncthreads(); // LDS / scratch reuse by the next item
        }

        __global__ __launch_bounds__(kBatchThreads) void batch_kernel(const jsvd_batch_item *__restrict__ items,
                                                                     int64_t count, jsvd_batch_buffers b,
                                                                     unsigned long long *__restrict__ counter,
                                                                     double *__restrict__ scratch,
                                                                     int64_t scratch_per_wg, int max_rows, int max_cols)
        {
            __shared__ __align__(16) unsigned char lds[kBatchLdsBytes];
            __shared__ int sweep_rot;
            __shared__ long long next;
            double *my_scratch = scratch + static_cast<int64_t>(blockIdx.x) * scratch_per_wg;
            for (;;)
            {
                if (threadIdx.x == 0)
                {
                    next = static_cast<long long>(atomicAdd(counter, 1ull));
                }
                __syncthreads();
                const long long idx = next;
                __syncthreads();
                if (idx >= count)
                {
                    break; // every workgroup leaves once the queue is drained
                }
                const jsvd_batch_item it = items[idx];
                if (!item_ok(it, b, max_rows, max_cols))
                {
                    if (threadIdx.x == 0)
                    {
                        b.sweeps[idx] = -1;
                    }
                    continue;
                }
                const int m = it.rows;
                const int n = it.columns;
                const int64_t a_elems = static_cast<int64_t>(m + (m & 1)) * n;
                const int64_t v_elems = static_cast<int64_t>(n + (n & 1)) * n;
                if (it.precision == JSVD_PRECISION_F32)
                {
                    const int64_t budget = kBatchLdsBytes / sizeof(float);
                    float *lds_f = reinterpret_cast<float *>(lds);
                    float *scr_f = reinterpret_cast<float *>(my_scratch);
                    float *a = (a_elems <= budget) ? lds_f : scr_f;
                    const int64_t used = (a_elems <= budget) ? a_elems : 0;
                    float *v = (used + v_elems <= budget) ? lds_f + used : scr_f + a_elems;
                    solve_item<float>(it, b, a, v, &sweep_rot, idx);
                }
                else
                {
                    const int64_t budget = kBatchLdsBytes / sizeof(double);
                    double *lds_d = reinterpret_cast<double *>(lds);
                    double *a = (a_elems <= budget) ? lds_d : my_scratch;
                    const int64_t used = (a_elems <= budget) ? a_elems : 0;
                    double *v = (used + v_elems <= budget) ? lds_d + used : my_scratch + a_elems;
                    solve_item<double>(it, b, a, v, &sweep_rot, idx);
                }
            }
        }
    } // namespace
} // namespace jsvd

struct jsvd_batch
{
    int64_t max_rows = 0;
    int64_t max_cols = 0;
    int grid = 0;
    int64_t scratch_per_wg = 0; // doubles
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *scratch = nullptr;
    unsigned long long *counter = nullptr;

    ~jsvd_batch()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
        }
        (void)hipFree(scratch);
        (void)hipFree(counter);
        if (own_stream && stream)
        {
            (void)hipStreamDestroy(stream);
        }
    }

    void enqueue(const jsvd_batch_item *d_items, int64_t count, const jsvd_batch_buffers &buf)
    {
        if (count == 0)
        {
            return;
        }
        JSVD_HIP_CHECK(hipMemsetAsync(counter, 0, sizeof(unsigned long long), stream));
        const int blocks = static_cast<int>(std::min<int64_t>(grid, count));
        hipLaunchKernelGGL(jsvd::batch_kernel, dim3(blocks), dim3(jsvd::kBatchThreads), 0, stream, d_items, count, buf,
                           counter, scratch, scratch_per_wg, static_cast<int>(max_rows), static_cast<int>(max_cols));
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
        b->grid = 2 * std::max(cus, 1); // LDS budget admits two workgroups per CU
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
        JSVD_HIP_CHECK(hipMalloc(&b->scratch, sizeof(double) * b->scratch_per_wg * b->grid));
        JSVD_HIP_CHECK(hipMalloc(&b->counter, sizeof(unsigned long long)));
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

} // extern "C"
