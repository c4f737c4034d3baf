// dense_model.cpp -- runs the library's FP64 dense kernels (csrc/dense_kernels.hpp, unchanged source) on the CPU
// model of a workgroup (simt.hpp), for tests only: the product kernel with the launch parameters of the host
// launcher (dense::plan_gemm, including the k-split and its reduction launch) and the triangular solve with the
// launcher's schedule (padded copy of R unless R comes padded, inverted 64 x 64 diagonal blocks, panel ranges and
// their updates).  Every
// buffer is an exact-size heap block, so an AddressSanitizer build reports any access outside it.
//
// Usage: dense_model input output
// Input (little-endian), product: int64 0, trans_a, trans_b, m, n, k, lda, ldb, ldc, ldd, flags, cus, work_elems,
//   a_len, b_len, c_len, d_len; double alpha, beta; then a, b, c, d (initial contents).  Output: d.
// Input, triangular solve: int64 1, rows, n, ldb, ldr, b_len, r_len, cus, work_elems, r_padded; then b, r.
//   Output: b.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dense_kernels.hpp"

namespace
{
    std::vector<int64_t> read_i64(FILE *f, size_t count)
    {
        std::vector<int64_t> v(count);
        if (fread(v.data(), sizeof(int64_t), count, f) != count)
        {
            std::fprintf(stderr, "short input\n");
            std::exit(2);
        }
        return v;
    }

    std::vector<double> read_f64(FILE *f, size_t count)
    {
        std::vector<double> v(count);
        if (count && fread(v.data(), sizeof(double), count, f) != count)
        {
            std::fprintf(stderr, "short input\n");
            std::exit(2);
        }
        return v;
    }

    // launch_gemm_f64 on the model
    void gemm(bool ta, bool tb, const jsvd::dense::GemmParams &p)
    {
        using namespace jsvd::dense;
        auto run = [&](auto kernel) {
            simt::launch2d(p.tiles, p.splits, 256, [&] { kernel(p); });
            if (p.splits > 1)
            {
                simt::launch(p.tiles, 256, [&] { gemm_reduce_kernel(p); });
            }
        };
        if (ta && tb)
        {
            run(gemm_kernel<true, true>);
        }
        else if (ta)
        {
            run(gemm_kernel<true, false>);
        }
        else if (tb)
        {
            run(gemm_kernel<false, true>);
        }
        else
        {
            run(gemm_kernel<false, false>);
        }
    }
} // namespace

int main(int argc, char **argv)
{
    using namespace jsvd::dense;
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: dense_model input output\n");
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    FILE *out = std::fopen(argv[2], "wb");
    if (!in || !out)
    {
        return 2;
    }
    const int64_t op = read_i64(in, 1)[0];
    if (op == 0)
    {
        const auto h = read_i64(in, 16);
        const auto ab = read_f64(in, 2);
        auto a = read_f64(in, h[12]), b = read_f64(in, h[13]), c = read_f64(in, h[14]), d = read_f64(in, h[15]);
        std::vector<double> work(h[11]);
        const GemmParams p = plan_gemm(h[2], h[3], h[4], ab[0], a.data(), h[5], b.data(), h[6], ab[1],
                                       c.empty() ? nullptr : c.data(), h[7], d.data(), h[8], static_cast<int>(h[9]),
                                       static_cast<int>(h[10]), work.empty() ? nullptr : work.data(), h[11]);
        std::fprintf(stderr, "tiles %d splits %d vec %d\n", p.tiles, p.splits, p.vec ? 1 : 0);
        gemm(h[0] != 0, h[1] != 0, p);
        fwrite(d.data(), sizeof(double), d.size(), out);
    }
    else
    {
        const auto h = read_i64(in, 9);
        const int64_t rows = h[0], n = h[1], ldb = h[2], ldr = h[3];
        auto b = read_f64(in, h[4]), r = read_f64(in, h[5]);
        const int cus = static_cast<int>(h[6]);
        std::vector<double> work(h[7]);
        const int64_t np = (n + kSolveBlock - 1) / kSolveBlock * kSolveBlock;
        std::vector<double> rpad, dinv(np * kSolveBlock);
        const double *rr = r.data();
        int64_t ldp = ldr;
        if (h[8] == 0)
        {
            rpad.resize(np * np);
            simt::launch(static_cast<int>((np * np + 255) / 256), 256,
                         [&] { pad_upper_kernel(r.data(), ldr, n, np, rpad.data()); });
            rr = rpad.data();
            ldp = np;
        }
        simt::launch(static_cast<int>(np / kSolveBlock), 64, [&] { diag_inverse_kernel(rr, ldp, dinv.data()); });
        const int panels = static_cast<int>((rows + kPanelRows - 1) / kPanelRows);
        trsm_schedule(
            n,
            [&](int64_t c0, int nblocks) {
                simt::launch(panels, 256, [&] {
                    trsm_panel_kernel(b.data(), ldb, rows, n, c0, nblocks, rr, ldp, dinv.data());
                });
            },
            [&](int64_t c0, int64_t width, int64_t rest) {
                const GemmParams p = plan_gemm(rows, rest, width, -1.0, b.data() + c0 * ldb, ldb,
                                               rr + c0 + (c0 + width) * ldp, ldp, 1.0, b.data() + (c0 + width) * ldb,
                                               ldb, b.data() + (c0 + width) * ldb, ldb, 0, cus,
                                               work.empty() ? nullptr : work.data(), h[7]);
                gemm(false, false, p);
            });
        fwrite(b.data(), sizeof(double), b.size(), out);
    }
    std::fclose(in);
    std::fclose(out);
    return 0;
}
