// hbm_copy.hip -- a device-to-device copy kernel of the library's own, the measured practical HBM roof that the
// bandwidth-bound kernels are quoted against (a kernel that reads and writes equal byte counts cannot beat it).
//
// Each workgroup copies contiguous chunks of 256 x U 16-byte vectors (all U loads in flight before the stores;
// consecutive threads take consecutive vectors, so every access of a wavefront is 1 KiB contiguous) and strides
// over the buffer by the grid's chunk count.  A sweep on one MI355X (512 MiB source, grids of 1024 .. 16384
// workgroups): contiguous chunks 5.4-5.7 TB/s, a grid-stride layout 4.1-5.4 TB/s; U = 4 with a grid of 16 workgroups
// per CU the fastest.  Variant 1 issues the stores as non-temporal (streaming: the written lines are not kept in
// the caches) with U = 8.
#include <algorithm>

#include "common.hpp"

namespace
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

    template <int U, bool NT>
    __global__ __launch_bounds__(256) void hbm_copy_kernel(u32x4 *__restrict__ dst, const u32x4 *__restrict__ src,
                                                           int64_t chunks)
    {
        constexpr int64_t kChunk = 256 * U;
        for (int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk; base < chunks;
             base += static_cast<int64_t>(gridDim.x) * kChunk)
        {
            if (base + kChunk <= chunks)
            {
                u32x4 v[U];
#pragma unroll
                for (int k = 0; k < U; ++k)
                {
                    v[k] = src[base + 256 * k + threadIdx.x];
                }
#pragma unroll
                for (int k = 0; k < U; ++k)
                {
                    if constexpr (NT)
                    {
                        __builtin_nontemporal_store(v[k], dst + base + 256 * k + threadIdx.x);
                    }
                    else
                    {
                        dst[base + 256 * k + threadIdx.x] = v[k];
                    }
                }
            }
            else
            {
                for (int64_t i = base + threadIdx.x; i < chunks; i += 256)
                {
                    dst[i] = src[i];
                }
            }
        }
    }
} // namespace

extern "C" int jsvd_hbm_copy(void *dst, const void *src, int64_t bytes, int32_t variant, void *stream)
{
    return jsvd::guarded([&] {
        if (bytes < 0 || (bytes % 16) != 0)
        {
            jsvd::throw_invalid("bytes must be a non-negative multiple of 16.");
        }
        if (bytes > 0 && (!dst || !src || (reinterpret_cast<uintptr_t>(dst) % 16) != 0 ||
                          (reinterpret_cast<uintptr_t>(src) % 16) != 0))
        {
            jsvd::throw_invalid("dst and src must be non-null and 16-byte aligned.");
        }
        if (variant != 0 && variant != 1)
        {
            jsvd::throw_invalid("variant must be 0 (plain stores) or 1 (non-temporal stores).");
        }
        if (bytes == 0)
        {
            return;
        }
        int device = 0, cus = 0;
        JSVD_HIP_CHECK(hipGetDevice(&device));
        JSVD_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        const int64_t chunks = bytes / 16;
        const int64_t blocks =
            std::max<int64_t>(1, std::min<int64_t>(int64_t(std::max(cus, 1)) * 16, (chunks + 1023) / 1024));
        auto *d = static_cast<u32x4 *>(dst);
        const auto *s = static_cast<const u32x4 *>(src);
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (variant == 1)
        {
            hipLaunchKernelGGL((hbm_copy_kernel<8, true>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, d, s,
                               chunks);
        }
        else
        {
            hipLaunchKernelGGL((hbm_copy_kernel<4, false>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, d, s,
                               chunks);
        }
        JSVD_HIP_CHECK(hipGetLastError());
    });
}
