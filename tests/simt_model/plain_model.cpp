// plain_model.cpp -- runs the library's plain round kernels (csrc/plain_kernels.hpp, unchanged source) on the CPU
// model of a workgroup (simt.hpp), for tests only: the same solve once with one round_kernel launch per round and
// once with round_span_kernel launches over spans of rounds (two buffers, alternating), each sweep's rotation count
// and cosine maximum folded from the counter slots as the plan does.  Buffers are exact-size heap blocks, so an
// AddressSanitizer build reports any access outside them.
//
// Usage: plain_model input output
// Input (little-endian): int64 m, n, lda, ldv, threads, ea, ev, waves, max_rounds, sweeps; double eps;
//                        lda x n doubles of A (column-major, padding rows zero).
// Output: for each of the two runs: A (lda x n), V (ldv x n), int64 sweeps run, then per sweep int64 rotations
//         and double cosine maximum; then int64 span launches and the span run's (groups, own, own_end).
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "plain_kernels.hpp"

namespace
{
    using namespace jsvd::plain;

    struct Run
    {
        std::vector<double> a, v;
        std::vector<int64_t> rot;
        std::vector<double> off;
        int64_t launches = 0;
    };

    constexpr int kSlots = 8;

    void fold(const std::vector<int> &rs, const std::vector<unsigned long long> &os, Run &run)
    {
        int64_t c = 0;
        unsigned long long best = 0;
        for (int k = 0; k < kSlots; ++k)
        {
            c += rs[k];
            best = best > os[k] ? best : os[k];
        }
        double x;
        std::memcpy(&x, &best, sizeof(x));
        run.rot.push_back(c);
        run.off.push_back(x);
    }

    template <int T, int EA, int EV>
    void run_rounds(Run &run, int64_t m, int64_t n, int64_t lda, int64_t ldv, double eps, int sweeps)
    {
        const int N = static_cast<int>(n + (n & 1));
        for (int s = 0; s < sweeps; ++s)
        {
            std::vector<int> rs(kSlots, 0);
            std::vector<unsigned long long> os(kSlots, 0);
            for (int r = 0; r < N - 1; ++r)
            {
                simt::launch(N / 2, T, [&] {
                    round_kernel<T, EA, EV>(run.a.data(), lda, m, run.v.data(), ldv, n, N, r, eps, rs.data(),
                                            os.data(), nullptr, kSlots);
                });
                ++run.launches;
            }
            fold(rs, os, run);
            if (run.rot.back() == 0)
            {
                break;
            }
        }
    }

    template <int TT, int EA, int NV, int W>
    void run_spans(Run &run, int64_t m, int64_t n, int64_t lda, int64_t ldv, double eps, int sweeps, int max_rounds,
                   int64_t part[3])
    {
        const int N = static_cast<int>(n + (n & 1));
        int groups = 0, own = 0, own_end = 0;
        span_partition(n, W, max_rounds, groups, own, own_end);
        part[0] = groups;
        part[1] = own;
        part[2] = own_end;
        const std::vector<int> chunks = span_chunks(n, max_rounds);
        std::vector<double> a2(run.a.size(), -7.0), v2(run.v.size(), -7.0); // the second buffers start as garbage
        for (int s = 0; s < sweeps; ++s)
        {
            std::vector<int> rs(kSlots, 0);
            std::vector<unsigned long long> os(kSlots, 0);
            int r0 = 0;
            for (size_t i = 0; i < chunks.size(); ++i)
            {
                const bool even = (i % 2) == 0;
                const double *ain = even ? run.a.data() : a2.data();
                double *aout = even ? a2.data() : run.a.data();
                const double *vin = even ? run.v.data() : v2.data();
                double *vout = even ? v2.data() : run.v.data();
                const int rounds = chunks[i];
                simt::launch(groups, 64 * W, [&] {
                    round_span_kernel<TT, EA, NV, W>(ain, aout, lda, m, vin, vout, ldv, n, N, r0, rounds, own,
                                                     own_end, eps, rs.data(), os.data(), nullptr, kSlots);
                });
                ++run.launches;
                r0 += rounds;
            }
            if (r0 != N - 1)
            {
                throw std::runtime_error("span chunks do not cover the sweep");
            }
            fold(rs, os, run);
            if (run.rot.back() == 0)
            {
                break;
            }
        }
    }

    void read_exact(FILE *f, void *p, size_t bytes)
    {
        if (bytes && std::fread(p, 1, bytes, f) != bytes)
        {
            throw std::runtime_error("short input file");
        }
    }

    void write_run(FILE *f, const Run &r)
    {
        std::fwrite(r.a.data(), sizeof(double), r.a.size(), f);
        std::fwrite(r.v.data(), sizeof(double), r.v.size(), f);
        const int64_t s = static_cast<int64_t>(r.rot.size());
        std::fwrite(&s, sizeof(s), 1, f);
        for (size_t i = 0; i < r.rot.size(); ++i)
        {
            std::fwrite(&r.rot[i], sizeof(int64_t), 1, f);
            std::fwrite(&r.off[i], sizeof(double), 1, f);
        }
    }

    // (threads, ea) of the round kernel and the span kernel's V chunks / wavefronts this model is built for.
    template <int TT, int EA, int NV, int W>
    bool run_shape(int threads, int ea, int nv, int waves, Run &one, Run &span, int64_t m, int64_t n, int64_t lda,
                   int64_t ldv, double eps, int sweeps, int max_rounds, int64_t part[3])
    {
        if (threads != TT || ea != EA || nv != NV || waves != W)
        {
            return false;
        }
        run_rounds<TT, EA, (NV * 64 + TT - 1) / TT>(one, m, n, lda, ldv, eps, sweeps);
        run_spans<TT, EA, NV, W>(span, m, n, lda, ldv, eps, sweeps, max_rounds, part);
        return true;
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s input output\n", argv[0]);
        return 2;
    }
    try
    {
        FILE *f = std::fopen(argv[1], "rb");
        if (!f)
        {
            throw std::runtime_error("cannot open input");
        }
        int64_t h[10];
        read_exact(f, h, sizeof(h));
        double eps = 0;
        read_exact(f, &eps, sizeof(eps));
        const int64_t m = h[0], n = h[1], lda = h[2], ldv = h[3];
        const int threads = static_cast<int>(h[4]), ea = static_cast<int>(h[5]), nv = static_cast<int>(h[6]);
        const int waves = static_cast<int>(h[7]), max_rounds = static_cast<int>(h[8]), sweeps = static_cast<int>(h[9]);
        Run one, span;
        one.a.resize(lda * n);
        read_exact(f, one.a.data(), sizeof(double) * one.a.size());
        std::fclose(f);
        one.v.assign(ldv * n, 0.0);
        for (int64_t j = 0; j < n; ++j)
        {
            one.v[j * ldv + j] = 1.0;
        }
        span.a = one.a;
        span.v = one.v;
        int64_t part[3] = {0, 0, 0};
        const bool ok = run_shape<64, 1, 1, 8>(threads, ea, nv, waves, one, span, m, n, lda, ldv, eps, sweeps,
                                               max_rounds, part) ||
                        run_shape<64, 2, 1, 8>(threads, ea, nv, waves, one, span, m, n, lda, ldv, eps, sweeps,
                                               max_rounds, part) ||
                        run_shape<64, 2, 2, 4>(threads, ea, nv, waves, one, span, m, n, lda, ldv, eps, sweeps,
                                               max_rounds, part) ||
                        run_shape<256, 1, 2, 8>(threads, ea, nv, waves, one, span, m, n, lda, ldv, eps, sweeps,
                                                max_rounds, part) ||
                        run_shape<128, 1, 1, 16>(threads, ea, nv, waves, one, span, m, n, lda, ldv, eps, sweeps,
                                                 max_rounds, part);
        if (!ok)
        {
            throw std::runtime_error("shape not built into the model");
        }
        FILE *o = std::fopen(argv[2], "wb");
        if (!o)
        {
            throw std::runtime_error("cannot open output");
        }
        write_run(o, one);
        write_run(o, span);
        std::fwrite(&span.launches, sizeof(int64_t), 1, o);
        std::fwrite(part, sizeof(int64_t), 3, o);
        std::fclose(o);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "plain_model: %s\n", e.what());
        return 1;
    }
    return 0;
}
