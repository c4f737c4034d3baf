// hip/hip_runtime.h for the host sanitizer builds (test infrastructure only).  The library's host-side
// sources (request loop, pipeline, I/O, CLI) are compiled with g++ and AddressSanitizer/UBSan or
// ThreadSanitizer against this header: the device-side surface comes from the CPU workgroup model's header
// (so the shared device helpers compile), and the runtime calls below run on the host.  "Device" memory is
// plain heap memory and every stream operation completes before it returns, which is one valid schedule of
// the asynchronous API; host threads, locks and queues run exactly as in the library.
#pragma once

#include <cstddef>

#include "../../simt_model/hip/hip_runtime.h"

enum : int
{
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
};

enum hipMemcpyKind
{
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
};

constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipHostMallocDefault = 0;

hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipFree(void *p);
hipError_t hipMallocAsync(void **p, size_t bytes, hipStream_t stream);
hipError_t hipFreeAsync(void *p, hipStream_t stream);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t stream);
hipError_t hipStreamCreateWithFlags(hipStream_t *stream, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t stream);
hipError_t hipStreamSynchronize(hipStream_t stream);
hipError_t hipSetDevice(int device);
hipError_t hipGetDevice(int *device);
const char *hipGetErrorString(hipError_t status);

template <typename T>
inline hipError_t hipMalloc(T **p, size_t bytes)
{
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}

template <typename T>
inline hipError_t hipHostMalloc(T **p, size_t bytes, unsigned flags)
{
    return hipHostMalloc(reinterpret_cast<void **>(p), bytes, flags);
}
