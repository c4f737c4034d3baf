// selftest.cpp -- checks that the CPU model of a workgroup (simt.hpp) catches what it is used to rule out, and
// that its wavefront operations compute what the library's reductions rely on.  Test infrastructure only.
//
//   selftest sum        wave_sum_dpp / wave_sum over 4 wavefronts: prints the sums (expected 2016 per wave)
//   selftest barrier    __syncthreads reached by one wavefront only       -> model error (exit via abort)
//   selftest divergent  a DPP move inside a branch taken by half the lanes -> model error
//   selftest twosites   two different __syncthreads call sites at once     -> model error
//   selftest lds        a store one element past an LDS array              -> AddressSanitizer error
#include <cstdio>
#include <cstring>

#include "common.hpp"

namespace
{
    double g_out[4];

    __global__ void sum_kernel()
    {
        __shared__ double partial[4];
        const int lane = threadIdx.x % 64;
        const int wave = threadIdx.x / 64;
        const double a = jsvd::dev::wave_sum_dpp(static_cast<double>(lane));
        const double b = jsvd::dev::wave_sum(static_cast<double>(lane));
        if (lane == 0)
        {
            partial[wave] = (a == b) ? a : -1.0;
        }
        __syncthreads();
        if (threadIdx.x < 4)
        {
            g_out[threadIdx.x] = partial[threadIdx.x];
        }
    }

    __global__ void barrier_kernel()
    {
        if (threadIdx.x < 64)
        {
            __syncthreads();
        }
    }

    __global__ void divergent_kernel()
    {
        const int lane = threadIdx.x % 64;
        float x = static_cast<float>(lane);
        if (lane < 32)
        {
            x = jsvd::dev::dpp_move<0x111, 0xF>(x);
        }
        g_out[0] += x;
    }

    __global__ void twosites_kernel()
    {
        if (threadIdx.x < 128)
        {
            __syncthreads();
        }
        else
        {
            __syncthreads();
        }
    }

    __global__ void lds_kernel()
    {
        __shared__ double tile[256];
        double *p = tile; // through a pointer, as the kernels address LDS
        p[threadIdx.x + 1] = 1.0;
        __syncthreads();
    }
} // namespace

int main(int argc, char **argv)
{
    const char *what = argc > 1 ? argv[1] : "sum";
    if (!std::strcmp(what, "sum"))
    {
        simt::launch(1, 256, [] { sum_kernel(); });
        std::printf("%.1f %.1f %.1f %.1f\n", g_out[0], g_out[1], g_out[2], g_out[3]);
        return 0;
    }
    if (!std::strcmp(what, "barrier"))
    {
        simt::launch(1, 256, [] { barrier_kernel(); });
    }
    else if (!std::strcmp(what, "divergent"))
    {
        simt::launch(1, 64, [] { divergent_kernel(); });
    }
    else if (!std::strcmp(what, "twosites"))
    {
        simt::launch(1, 256, [] { twosites_kernel(); });
    }
    else if (!std::strcmp(what, "lds"))
    {
        simt::launch(1, 256, [] { lds_kernel(); });
    }
    std::printf("no error detected\n");
    return 0;
}
