// hip/hip_runtime.h of the CPU model (test infrastructure only): the device-side HIP surface the library's
// kernels use, mapped onto simt.hpp so the kernel sources compile with g++ and run on the host, one fiber per GPU
// thread.  Host-side runtime calls are not modelled; the model drivers launch kernels through simt::launch.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../simt.hpp"

#define JSVD_SIMT_MODEL 1

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define __restrict__ __restrict

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;

struct simt_index3
{
    unsigned x, y, z;
};

inline simt_index3 simt_thread_idx() { return {static_cast<unsigned>(simt::current_thread()->tid), 0, 0}; }
inline simt_index3 simt_block_idx()
{
    const simt::Dim d = simt::current_group()->block_idx();
    return {d.x, d.y, d.z};
}
inline simt_index3 simt_block_dim()
{
    const simt::Dim d = simt::current_group()->block_dim();
    return {d.x, d.y, d.z};
}
inline simt_index3 simt_grid_dim()
{
    const simt::Dim d = simt::current_group()->grid_dim();
    return {d.x, d.y, d.z};
}
#define threadIdx (simt_thread_idx())
#define blockIdx (simt_block_idx())
#define blockDim (simt_block_dim())
#define gridDim (simt_grid_dim())

#define __syncthreads() ::simt::barrier(__FILE__, __LINE__)

// Fibers run one at a time, so plain read-modify-write is atomic here.
template <typename T>
inline T atomicAdd(T *p, T v)
{
    const T old = *p;
    *p = old + v;
    return old;
}
template <typename T>
inline T atomicMax(T *p, T v)
{
    const T old = *p;
    *p = std::max(old, v);
    return old;
}
template <typename T>
inline T atomicMin(T *p, T v)
{
    const T old = *p;
    *p = std::min(old, v);
    return old;
}
template <typename T>
inline T atomicExch(T *p, T v)
{
    const T old = *p;
    *p = v;
    return old;
}
template <typename T>
inline T atomicCAS(T *p, T cmp, T v)
{
    const T old = *p;
    if (old == cmp)
    {
        *p = v;
    }
    return old;
}

inline float __int_as_float(int x)
{
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}
inline int __float_as_int(float x)
{
    int i;
    std::memcpy(&i, &x, 4);
    return i;
}
inline double __longlong_as_double(long long x)
{
    double d;
    std::memcpy(&d, &x, 8);
    return d;
}
inline long long __double_as_longlong(double x)
{
    long long i;
    std::memcpy(&i, &x, 8);
    return i;
}

inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
inline void __builtin_amdgcn_s_setprio(int) {}
// Workgroups run one after another, so every store is visible to the next workgroup without a fence.
inline void __threadfence() {}
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
using std::fabs;
using std::fmax;
using std::fmin;
using std::frexp;
using std::ldexp;
using std::sqrt;

#define SIMT_WAVE(op, in, old, a0, a1, a2, a3) \
    ::simt::wave_op(op, static_cast<uint32_t>(in), static_cast<uint32_t>(old), a0, a1, a2, a3, __FILE__, __LINE__)

#define __builtin_amdgcn_update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) \
    static_cast<int>(SIMT_WAVE(::simt::WaveOp::dpp, (src), (old), (ctrl), (row_mask), (bank_mask), (bound_ctrl) ? 1 : 0))
#define __builtin_amdgcn_mov_dpp(src, ctrl, row_mask, bank_mask, bound_ctrl) \
    __builtin_amdgcn_update_dpp(0, src, ctrl, row_mask, bank_mask, bound_ctrl)
#define __builtin_amdgcn_readlane(v, lane) \
    static_cast<int>(SIMT_WAVE(::simt::WaveOp::readlane, (v), 0, (lane), 0, 0, 0))
#define __builtin_amdgcn_readfirstlane(v) __builtin_amdgcn_readlane(v, 0)

inline unsigned simt_lane() { return static_cast<unsigned>(simt::current_thread()->tid) % simt::kWave; }

// __shfl family (HIP semantics: a source lane outside the caller's width segment reads the caller's own value);
// 64-bit values move as two 32-bit halves.
inline uint32_t simt_shfl32(uint32_t v, int src, const char *file, int line)
{
    return simt::wave_op(simt::WaveOp::shfl, v, static_cast<uint32_t>(src), 0, 0, 0, 0, file, line);
}
template <typename T>
inline T simt_shfl(T v, int src, const char *file, int line)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit shuffle");
    if constexpr (sizeof(T) == 4)
    {
        uint32_t b;
        std::memcpy(&b, &v, 4);
        b = simt_shfl32(b, src, file, line);
        T r;
        std::memcpy(&r, &b, 4);
        return r;
    }
    else
    {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        const uint64_t lo = simt_shfl32(static_cast<uint32_t>(b), src, file, line);
        const uint64_t hi = simt_shfl32(static_cast<uint32_t>(b >> 32), src, file, line);
        const uint64_t r64 = lo | (hi << 32);
        T r;
        std::memcpy(&r, &r64, 8);
        return r;
    }
}
inline int simt_width(int w) { return w == 0 ? simt::kWave : w; }
inline int simt_src_idx(int src, int width)
{
    const int l = static_cast<int>(simt_lane());
    return (l & ~(width - 1)) + (src & (width - 1));
}
inline int simt_src_xor(int mask, int width)
{
    const int l = static_cast<int>(simt_lane());
    const int s = l ^ mask;
    return (s >= ((l + width) & ~(width - 1))) ? l : s;
}
inline int simt_src_down(int delta, int width)
{
    const int l = static_cast<int>(simt_lane());
    return ((l & (width - 1)) + delta >= width) ? l : l + delta;
}
#define __shfl(v, src, ...) simt_shfl((v), simt_src_idx((src), simt_width(__VA_ARGS__ + 0)), __FILE__, __LINE__)
#define __shfl_xor(v, mask, ...) simt_shfl((v), simt_src_xor((mask), simt_width(__VA_ARGS__ + 0)), __FILE__, __LINE__)
#define __shfl_down(v, d, ...) simt_shfl((v), simt_src_down((d), simt_width(__VA_ARGS__ + 0)), __FILE__, __LINE__)

#define __ballot(pred) ::simt::wave_ballot((pred), __FILE__, __LINE__)

typedef float simt_f32x4 __attribute__((vector_size(16)));
typedef double simt_f64x4 __attribute__((vector_size(32)));
inline simt_f32x4 simt_mfma_f32(float a, float b, simt_f32x4 c, const char *file, int line)
{
    const double ci[4] = {c[0], c[1], c[2], c[3]};
    double d[4];
    simt::wave_mfma(simt::WaveOp::mfma_f32, a, b, ci, d, file, line);
    return simt_f32x4{static_cast<float>(d[0]), static_cast<float>(d[1]), static_cast<float>(d[2]),
                      static_cast<float>(d[3])};
}
inline simt_f64x4 simt_mfma_f64(double a, double b, simt_f64x4 c, const char *file, int line)
{
    const double ci[4] = {c[0], c[1], c[2], c[3]};
    double d[4];
    simt::wave_mfma(simt::WaveOp::mfma_f64, a, b, ci, d, file, line);
    return simt_f64x4{d[0], d[1], d[2], d[3]};
}
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, cbsz, abid, blgp) simt_mfma_f32((a), (b), (c), __FILE__, __LINE__)
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, cbsz, abid, blgp) simt_mfma_f64((a), (b), (c), __FILE__, __LINE__)
using std::max;
using std::min;

struct double2
{
    double x, y;
};
struct float2
{
    float x, y;
};
struct float4
{
    float x, y, z, w;
};
struct int2
{
    int x, y;
};
