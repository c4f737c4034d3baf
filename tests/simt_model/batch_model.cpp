// batch_model.cpp -- runs the library's batched kernel (csrc/batch_kernel.hpp, unchanged source) on the CPU model
// of a workgroup (simt.hpp), for tests only.  Every buffer is an exact-size heap block and the scratch area holds
// exactly `grid` workgroup slices, so an AddressSanitizer build reports any access outside them; the model itself
// reports barriers or wavefront operations not reached by the whole workgroup / wavefront.
//
// Input file (little-endian):  int64 threads, lds_bytes, max_rows, max_cols, grid, count, a_len, u_len, s_len,
// v_len; `count` jsvd_batch_item records; a_len doubles of A.
// Output file:  u (u_len doubles), sigma (s_len), v (v_len), sweeps (count int32), then the model's counters
// (int64 barriers, wave operations).
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <vector>

#include "batch_kernel.hpp"

namespace
{
    template <int THREADS, int LDS>
    simt::LaunchStats run_shape(int grid, const jsvd_batch_item *items, int64_t count, const jsvd_batch_buffers &b,
                                unsigned long long *counter, double *scratch, int64_t per_wg, int max_rows, int max_cols)
    {
        std::vector<int> order;
        for (int g = grid - 1; g >= 0; --g)
        {
            order.push_back(g); // the last slice of the scratch area first
        }
        return simt::launch(grid, THREADS,
                            [&] { jsvd::batch_kernel<THREADS, LDS>(items, count, b, counter, scratch, per_wg, max_rows, max_cols); },
                            order);
    }

    void read_exact(FILE *f, void *p, size_t bytes)
    {
        if (bytes && std::fread(p, 1, bytes, f) != bytes)
        {
            throw std::runtime_error("short input file");
        }
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s input output\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f)
    {
        std::perror("input");
        return 2;
    }
    int64_t h[10];
    read_exact(f, h, sizeof(h));
    const int threads = static_cast<int>(h[0]), lds = static_cast<int>(h[1]);
    const int max_rows = static_cast<int>(h[2]), max_cols = static_cast<int>(h[3]), grid = static_cast<int>(h[4]);
    const int64_t count = h[5], a_len = h[6], u_len = h[7], s_len = h[8], v_len = h[9];
    std::unique_ptr<jsvd_batch_item[]> items(new jsvd_batch_item[count]);
    read_exact(f, items.get(), sizeof(jsvd_batch_item) * count);
    std::unique_ptr<double[]> a(new double[a_len]), u(new double[u_len]()), s(new double[s_len]()),
        v(new double[v_len]());
    read_exact(f, a.get(), sizeof(double) * a_len);
    std::fclose(f);
    std::unique_ptr<int32_t[]> sweeps(new int32_t[count]());
    // The host's slice size (batch.hip, jsvd_batch_create) and one slice per resident workgroup.
    const int64_t per_wg = (int64_t(max_rows) + 1) * max_cols + (int64_t(max_cols) + 1) * max_cols;
    std::unique_ptr<double[]> scratch(new double[per_wg * grid]);
    std::unique_ptr<unsigned long long> counter(new unsigned long long(0));
    const jsvd_batch_buffers b{a.get(), u.get(), s.get(), v.get(), sweeps.get(), a_len, u_len, s_len, v_len};
    const int blocks = static_cast<int>(std::min<int64_t>(grid, count));
    simt::LaunchStats st;
#define SHAPE(T, L)                                                                                               \
    else if (threads == T && lds == L * 1024)                                                                    \
    {                                                                                                            \
        st = run_shape<T, L * 1024>(blocks, items.get(), count, b, counter.get(), scratch.get(), per_wg, max_rows, \
                                    max_cols);                                                                   \
    }
    if (count == 0)
    {
    }
    SHAPE(256, 76) SHAPE(128, 38) SHAPE(256, 38) SHAPE(512, 76) SHAPE(512, 152) SHAPE(1024, 152) else
    {
        std::fprintf(stderr, "unknown shape %d x %d\n", threads, lds);
        return 2;
    }
#undef SHAPE
    FILE *o = std::fopen(argv[2], "wb");
    if (!o)
    {
        std::perror("output");
        return 2;
    }
    std::fwrite(u.get(), sizeof(double), u_len, o);
    std::fwrite(s.get(), sizeof(double), s_len, o);
    std::fwrite(v.get(), sizeof(double), v_len, o);
    std::fwrite(sweeps.get(), sizeof(int32_t), count, o);
    const int64_t stats[2] = {static_cast<int64_t>(st.barriers), static_cast<int64_t>(st.wave_ops)};
    std::fwrite(stats, sizeof(stats), 1, o);
    std::fclose(o);
    return 0;
}
