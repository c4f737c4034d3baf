// hbm_copy.hip -- a device-to-device copy kernel of the library's own, the measured practical HBM roof that the
// bandwidth-bound kernels are quoted against (a kernel that reads and writes equal byte counts cannot beat it).
//
// Each thread moves four 16-byte chunks per step with all four loads in flight before the stores; consecutive
// threads take consecutive chunks, so every wavefront reads and writes 1 KiB contiguous per access.  The grid
// is a fixed number of resident workgroups striding over the buffer.  Variant 1 issues the stores as
// non-temporal (streaming: the written lines are not kept in the caches).
#include <algorithm>

#include "common.hpp"

namespace
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

    template <bool NT>
    __global__ __launch_bounds__(256) void hbm_copy_kernel(u32x4 *__restrict__ dst, const u32x4 *__restrict__ src,
                                                           int64_t chunks)
    {
        constexpr int kUnroll = 4;
        const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
        int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        for (; i + (kUnroll - 1) * stride < chunks; i += kUnroll * stride)
        {
            u32x4 v[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                v[k] = src[i + k * stride];
            }
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
            {
                if constexpr (NT)
                {
                    __builtin_nontemporal_store(v[k], dst + i + k * stride);
                }
                else
                {
                    dst[i + k * stride] = v[k];
                }
            }
        }
        for (; i < chunks; i += stride)
        {
            dst[i] = src[i];
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
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(int64_t(std::max(cus, 1)) * 8, (chunks + 255) / 256));
        auto *d = static_cast<u32x4 *>(dst);
        const auto *s = static_cast<const u32x4 *>(src);
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (variant == 1)
        {
            hipLaunchKernelGGL(hbm_copy_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, d, s, chunks);
        }
        else
        {
            hipLaunchKernelGGL(hbm_copy_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, d, s, chunks);
        }
        JSVD_HIP_CHECK(hipGetLastError());
    });
}
