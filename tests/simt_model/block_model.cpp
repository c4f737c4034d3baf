// block_model.cpp -- runs the library's block one-sided Jacobi kernels (csrc/block_kernels.hpp, unchanged source)
// on the CPU model of a workgroup (simt.hpp), for tests only: the sweep loop of the block and mixed paths (column
// ordering by norm before every sweep after the first, the intra-block step of the cross ordering, then the outer
// rounds: Gram partials, inner sweep, apply), with the launch geometry of the host launchers (block.hip).  Every
// buffer is an exact-size heap block, so an AddressSanitizer build reports any access outside it.
//
// Usage: block_model <f32|f64> <cross|full> <kb> <max_sweeps> <eps> input output
// Input (little-endian): int64 m, n; m x n doubles of A, column-major.
// Output: int64 sweeps; per sweep int32 rotations and double offmax; then A (lda x n_pad, column-major, in the
// working precision widened to double) and V (ldv x n_pad), physical column order.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "block_kernels.hpp"

namespace
{
    template <typename T>
    struct Buffers
    {
        std::unique_ptr<T[]> a, v, partial, q;
        std::unique_ptr<int[]> flags, perm, rot;
        std::unique_ptr<unsigned long long[]> off;
        std::unique_ptr<double[]> norm2;
    };

    // One block round over `pairs` pairs (block_round_impl of block.hip, run on the model).
    template <typename T, int KB, bool CROSS>
    void round_impl(const jsvd::BlockGeometry &g, int pairs, int nb, int Nb, Buffers<T> &b, int round, double eps,
                    int *rot_count, unsigned long long *offmax, const int *skip)
    {
        using namespace jsvd;
        int per = 1;
        while ((g.slabs_a + per - 1) / per > 8 && pairs * ((g.slabs_a + 2 * per - 1) / (2 * per)) >= 256)
        {
            per *= 2;
        }
        const int parts = (g.slabs_a + per - 1) / per;
        T *a = b.a.get(), *v = b.v.get(), *partial = b.partial.get(), *q = b.q.get();
        int *flags = b.flags.get();
        const int *perm = b.perm.get();
        double *norm2 = b.norm2.get();
        simt::launch2d(pairs, parts, 256, [&] {
            block_gram_kernel<T, KB>(a, g.lda, perm, nb, Nb, round, g.slabs_a, per, parts, partial, skip);
        });
        simt::launch2d(pairs, 1, EigenShape<KB>::kThreads, [&] {
            block_eigen_kernel<T, KB, CROSS>(partial, parts, nb, Nb, round, eps, q, flags, rot_count, offmax, perm,
                                             norm2, skip);
        });
        const int h = ApplyShape<T, KB>::H;
        const int sa = static_cast<int>(g.lda / h), sv = static_cast<int>(g.ldv / h);
        const int per_pair = std::max(1, std::min(sa + sv, (g.apply_wgs + pairs - 1) / pairs));
        simt::launch2d(pairs, per_pair, 256, [&] {
            block_apply_kernel<T, KB>(a, g.lda, sa, v, g.ldv, sv, nb, Nb, round, q, flags, perm, skip);
        });
    }

    template <typename T, int KB>
    int run(const jsvd::BlockGeometry &g, Buffers<T> &b, int max_sweeps, double eps, std::vector<int> &rots,
            std::vector<double> &offs)
    {
        using namespace jsvd;
        simt::launch(static_cast<int>((g.n_pad + 255) / 256), 256, [&] { iota_kernel(b.perm.get(), g.n_pad); });
        int executed = 0;
        for (int s = 0; s < max_sweeps; ++s)
        {
            const int *skip = (s > 0) ? b.rot.get() + (s - 1) : nullptr;
            if (s > 0)
            {
                simt::launch(1, 1024, [&] {
                    column_order_kernel(b.norm2.get(), static_cast<int>(g.n_pad), b.perm.get(), skip);
                });
            }
            if constexpr (KB == 32)
            {
                if (g.cross)
                {
                    round_impl<T, 16, false>(g, g.nb, 2 * g.nb, 2 * g.nb, b, kIntraRound, eps, b.rot.get() + s,
                                             b.off.get() + s, skip);
                }
            }
            for (int r = 0; r < g.rounds(); ++r)
            {
                if constexpr (KB == 32)
                {
                    if (g.cross)
                    {
                        round_impl<T, KB, true>(g, g.Nb / 2, g.nb, g.Nb, b, r, eps, b.rot.get() + s, b.off.get() + s,
                                                skip);
                        continue;
                    }
                }
                round_impl<T, KB, false>(g, g.Nb / 2, g.nb, g.Nb, b, r, eps, b.rot.get() + s, b.off.get() + s, skip);
            }
            executed = s + 1;
            double x = 0.0;
            std::memcpy(&x, &b.off[s], sizeof(double));
            rots.push_back(b.rot[s]);
            offs.push_back(x);
            if (b.rot[s] == 0)
            {
                break;
            }
        }
        return executed;
    }

    template <typename T>
    int main_t(bool cross, int kb, int max_sweeps, double eps, int64_t m, int64_t n, const std::vector<double> &a_in,
               FILE *out)
    {
        jsvd::BlockGeometry g = jsvd::BlockGeometry::make(m, n, kb);
        g.cross = cross;
        Buffers<T> b;
        const size_t a_elems = static_cast<size_t>(g.lda) * g.n_pad, v_elems = static_cast<size_t>(g.ldv) * g.n_pad;
        b.a.reset(new T[a_elems]());
        b.v.reset(new T[v_elems]());
        b.partial.reset(new T[g.partial_elems()]());
        b.q.reset(new T[g.q_elems()]());
        b.flags.reset(new int[g.flag_elems()]());
        b.perm.reset(new int[g.n_pad]());
        b.rot.reset(new int[max_sweeps]());
        b.off.reset(new unsigned long long[max_sweeps]());
        b.norm2.reset(new double[g.n_pad]());
        for (int64_t j = 0; j < n; ++j)
        {
            for (int64_t i = 0; i < m; ++i)
            {
                b.a[j * g.lda + i] = static_cast<T>(a_in[j * m + i]);
            }
        }
        for (int64_t j = 0; j < g.n_pad; ++j)
        {
            b.v[j * g.ldv + j] = T(1);
        }
        std::vector<int> rots;
        std::vector<double> offs;
        int sweeps = 0;
        switch (kb)
        {
        case 16: sweeps = run<T, 16>(g, b, max_sweeps, eps, rots, offs); break;
        case 32: sweeps = run<T, 32>(g, b, max_sweeps, eps, rots, offs); break;
        default: std::fprintf(stderr, "kb must be 16 or 32\n"); std::exit(2);
        }
        const int64_t hdr[6] = {sweeps, g.lda, g.ldv, g.n_pad, 0, 0};
        std::fwrite(hdr, sizeof(hdr), 1, out);
        std::fwrite(rots.data(), sizeof(int), rots.size(), out);
        std::fwrite(offs.data(), sizeof(double), offs.size(), out);
        std::vector<double> wa(a_elems), wv(v_elems);
        for (size_t i = 0; i < a_elems; ++i)
        {
            wa[i] = static_cast<double>(b.a[i]);
        }
        for (size_t i = 0; i < v_elems; ++i)
        {
            wv[i] = static_cast<double>(b.v[i]);
        }
        std::fwrite(wa.data(), sizeof(double), wa.size(), out);
        std::fwrite(wv.data(), sizeof(double), wv.size(), out);
        return 0;
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc != 8)
    {
        std::fprintf(stderr, "usage: %s f32|f64 cross|full kb max_sweeps eps input output\n", argv[0]);
        return 2;
    }
    const bool f32 = std::string(argv[1]) == "f32";
    const bool cross = std::string(argv[2]) == "cross";
    const int kb = std::atoi(argv[3]);
    const int max_sweeps = std::atoi(argv[4]);
    const double eps = std::atof(argv[5]);
    FILE *f = std::fopen(argv[6], "rb");
    int64_t mn[2];
    if (!f || std::fread(mn, sizeof(mn), 1, f) != 1)
    {
        std::fprintf(stderr, "bad input\n");
        return 2;
    }
    std::vector<double> a(static_cast<size_t>(mn[0] * mn[1]));
    if (std::fread(a.data(), sizeof(double), a.size(), f) != a.size())
    {
        std::fprintf(stderr, "short input\n");
        return 2;
    }
    std::fclose(f);
    FILE *out = std::fopen(argv[7], "wb");
    const int rc = f32 ? main_t<float>(cross, kb, max_sweeps, eps, mn[0], mn[1], a, out)
                       : main_t<double>(cross, kb, max_sweeps, eps, mn[0], mn[1], a, out);
    std::fclose(out);
    return rc;
}
