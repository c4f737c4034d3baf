// cholesky_model.cpp -- runs the library's blocked Cholesky kernels (csrc/cholesky_kernels.hpp, unchanged source) on
// the CPU model of a workgroup (simt.hpp), for tests only, with the launch sequence of launch_cholesky_upper
// (cholesky.hip): status reset, identity padding to a multiple of 64, then per tile step the diagonal, panel and
// trailing-update kernels.  The matrix is an exact-size heap block, so an AddressSanitizer build reports any access
// outside it.
//
// Usage: cholesky_model input output
// Input (little-endian): int64 n, ld; ld x n64 doubles (column-major, n64 = n rounded up to 64).
// Output: the matrix after the factorisation (same layout), then int64 info.
#include <cstdio>
#include <vector>

#include "cholesky_kernels.hpp"

int main(int argc, char **argv)
{
    using namespace jsvd::chol;
    if (argc != 3)
    {
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    FILE *out = std::fopen(argv[2], "wb");
    int64_t hdr[2];
    if (!in || !out || fread(hdr, sizeof(int64_t), 2, in) != 2)
    {
        return 2;
    }
    const int64_t n = hdr[0], ld = hdr[1];
    const int64_t n64 = (n + kT - 1) / kT * kT;
    std::vector<double> a(ld * n64);
    if (fread(a.data(), sizeof(double), a.size(), in) != a.size())
    {
        return 2;
    }
    int info = -7;
    simt::launch(1, 64, [&] { zero_int_kernel(&info); }); // one thread on the device; the model runs whole waves
    if (n64 != n)
    {
        simt::launch(static_cast<int>((n64 * n64 + 255) / 256), 256,
                     [&] { pad_identity_kernel(a.data(), ld, n, n64); });
    }
    const int nt = static_cast<int>(n64 / kT);
    for (int k = 0; k < nt; ++k)
    {
        simt::launch(1, kT * kDiagWaves, [&] { cholesky_diag_kernel(a.data(), ld, k, &info); });
        const int rest = nt - k - 1;
        if (rest > 0)
        {
            simt::launch(rest, kT * kPanelWaves, [&] { cholesky_panel_kernel(a.data(), ld, k); });
            simt::launch2d(rest, rest, 256, [&] { cholesky_update_kernel(a.data(), ld, k); });
        }
    }
    fwrite(a.data(), sizeof(double), a.size(), out);
    const int64_t info64 = info;
    fwrite(&info64, sizeof(info64), 1, out);
    std::fclose(in);
    std::fclose(out);
    return 0;
}
