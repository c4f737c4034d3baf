// order_model.cpp -- runs the library's column ordering kernel (csrc/block_kernels.hpp, unchanged source, a
// bitonic network in LDS) on the CPU model of a workgroup (simt.hpp), for tests only.
//
// Usage: order_model input output
// Input (little-endian): int64 n_pad; n_pad doubles (squared column norms).  Output: n_pad int32 (perm).
#include <cstdio>
#include <vector>

#include "block_kernels.hpp"

int main(int argc, char **argv)
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s input output\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    long long n = 0;
    if (!f || std::fread(&n, sizeof(n), 1, f) != 1 || n < 1 || n > jsvd::kSortMax)
    {
        std::fprintf(stderr, "bad input\n");
        return 2;
    }
    std::vector<double> norm2(static_cast<size_t>(n));
    if (std::fread(norm2.data(), sizeof(double), norm2.size(), f) != norm2.size())
    {
        std::fprintf(stderr, "short input\n");
        return 2;
    }
    std::fclose(f);
    std::vector<int> perm(static_cast<size_t>(n), -1);
    simt::launch(1, 1024, [&] {
        jsvd::column_order_kernel(norm2.data(), static_cast<int>(n), perm.data(), nullptr);
    });
    FILE *o = std::fopen(argv[2], "wb");
    std::fwrite(perm.data(), sizeof(int), perm.size(), o);
    std::fclose(o);
    return 0;
}
