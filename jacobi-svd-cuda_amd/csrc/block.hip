// block.hip -- block one-sided Jacobi kernels (FP32 and FP64) for the mixed-precision path.
//
// Columns are grouped into blocks of kb = 16, 32 or 64 (a template parameter of every kernel; the autotuner or
// the caller picks it per shape); a round pairs blocks (I, J) by the same round-robin circle the plain path uses
// for columns.  For every block pair (2 kb = kw columns), one round is:
//
//   block_gram_kernel    G_s = X_s^T X_s for each row slab X_s (kH rows x kw columns of [A_I A_J]) on MFMA
//                        (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64), accumulating in T
//   block_eigen_kernel   G = sum_s G_s, then one cyclic Jacobi sweep over the kw columns carried out in Gram
//                        space (two-sided 2x2 updates of G), rotation parameters (tau, t, c, s) in FP64 with
//                        the reference skip test; accumulates the kw x kw rotation Q
//   block_apply_kernel   [A_I A_J] <- [A_I A_J] Q and [V_I V_J] <- [V_I V_J] Q per row slab on MFMA
//
// Layout: X (A or V) column-major with leading dimension ld (a multiple of kH, padding rows zero); the column
// count is padded to a multiple of kb with zero columns (their Gram entries are zero, so they never rotate).
#include "block.hpp"
#include "block_kernels.hpp"

namespace jsvd
{
    namespace
    {
        // One block round over `pairs` block pairs (outer round `round` of the circle over Nb blocks, or the
        // intra-block step for round == kIntraRound) at width KB: Gram partials, reduction + inner sweep (full
        // circle or the cross ordering), apply.
        template <typename T, int KB, bool CROSS>
        void block_round_impl(const BlockGeometry &g, int pairs, int nb, int Nb, T *a, T *v, int round, double eps,
                              T *partial, T *q, int *apply_flag, int *rot_count, unsigned long long *offmax,
                              const int *perm, double *norm2, const int *skip_if_zero, hipStream_t stream)
        {
            // Slabs per Gram workgroup: fewer partials for the eigen kernel to reduce (its reduction runs on one
            // CU per pair) while the Gram grid still covers the chip.
            int per = 1;
            while ((g.slabs_a + per - 1) / per > 8 && pairs * ((g.slabs_a + 2 * per - 1) / (2 * per)) >= 256)
            {
                per *= 2;
            }
            const int parts = (g.slabs_a + per - 1) / per;
            hipLaunchKernelGGL((block_gram_kernel<T, KB>), dim3(pairs, parts), dim3(256), 0, stream,
                               static_cast<const T *>(a), g.lda, perm, nb, Nb, round, g.slabs_a, per, parts,
                               partial, skip_if_zero);
            JSVD_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL((block_eigen_kernel<T, KB, CROSS>), dim3(pairs), dim3(EigenShape<KB>::kThreads), 0,
                               stream, static_cast<const T *>(partial), parts, nb, Nb, round, eps, q, apply_flag,
                               rot_count, offmax, perm, norm2, skip_if_zero);
            JSVD_HIP_CHECK(hipGetLastError());
            // lda and ldv are multiples of kH, hence of the apply slab height; g.apply_wgs workgroups in total
            const int h = ApplyShape<T, KB>::H;
            const int sa = static_cast<int>(g.lda / h), sv = static_cast<int>(g.ldv / h);
            const int per_pair = std::max(1, std::min(sa + sv, (g.apply_wgs + pairs - 1) / pairs));
            hipLaunchKernelGGL((block_apply_kernel<T, KB>), dim3(pairs, per_pair), dim3(256), 0, stream, a, g.lda, sa,
                               v, g.ldv, sv, nb, Nb, round, static_cast<const T *>(q),
                               static_cast<const int *>(apply_flag), perm, skip_if_zero);
            JSVD_HIP_CHECK(hipGetLastError());
        }

        template <typename T, int KB>
        void block_round_kb(const BlockGeometry &g, T *a, T *v, int round, double eps, T *partial, T *q,
                            int *apply_flag, int *rot_count, unsigned long long *offmax, const int *perm,
                            double *norm2, const int *skip_if_zero, hipStream_t stream)
        {
            if constexpr (KB == 32)
            {
                if (g.cross)
                {
                    block_round_impl<T, KB, true>(g, g.Nb / 2, g.nb, g.Nb, a, v, round, eps, partial, q, apply_flag,
                                                  rot_count, offmax, perm, norm2, skip_if_zero, stream);
                    return;
                }
            }
            block_round_impl<T, KB, false>(g, g.Nb / 2, g.nb, g.Nb, a, v, round, eps, partial, q, apply_flag,
                                           rot_count, offmax, perm, norm2, skip_if_zero, stream);
        }

        template <typename T, int KB, bool CROSS>
        void block_eigen_kb(const T *gram, int pairs, double eps, T *q_out, int *rotations, hipStream_t stream)
        {
            // A circle of 2 * pairs blocks at round 0 maps slot p to a valid block pair, so the kernel's pair p
            // reads gram[p] (one partial) and writes q_out[p] and rotations[p]; no column map, norms or sweep
            // counters.
            const int Nb = 2 * pairs;
            hipLaunchKernelGGL((block_eigen_kernel<T, KB, CROSS>), dim3(pairs), dim3(EigenShape<KB>::kThreads), 0,
                               stream, gram, 1, Nb, Nb, 0, eps, q_out, rotations, nullptr, nullptr, nullptr, nullptr,
                               nullptr);
            JSVD_HIP_CHECK(hipGetLastError());
        }
    } // namespace

    template <typename T>
    void launch_block_round(const BlockGeometry &g, T *a, T *v, int round, double eps, T *partial, T *q,
                            int *apply_flag, int *rot_count, unsigned long long *offmax, const int *perm,
                            double *norm2, const int *skip_if_zero, hipStream_t stream)
    {
        switch (g.kb)
        {
        case 16:
            block_round_kb<T, 16>(g, a, v, round, eps, partial, q, apply_flag, rot_count, offmax, perm, norm2,
                                  skip_if_zero, stream);
            break;
        case 32:
            block_round_kb<T, 32>(g, a, v, round, eps, partial, q, apply_flag, rot_count, offmax, perm, norm2,
                                  skip_if_zero, stream);
            break;
        case 64:
            block_round_kb<T, 64>(g, a, v, round, eps, partial, q, apply_flag, rot_count, offmax, perm, norm2,
                                  skip_if_zero, stream);
            break;
        default: throw_invalid("block width must be 16, 32 or 64.");
        }
    }

    template void launch_block_round<float>(const BlockGeometry &, float *, float *, int, double, float *, float *,
                                            int *, int *, unsigned long long *, const int *, double *, const int *,
                                            hipStream_t);

    template <typename T>
    void launch_block_intra(const BlockGeometry &g, T *a, T *v, double eps, T *partial, T *q, int *apply_flag,
                            int *rot_count, unsigned long long *offmax, const int *perm, double *norm2,
                            const int *skip_if_zero, hipStream_t stream)
    {
        if (!g.cross)
        {
            return;
        }
        if (g.kb != 32)
        {
            throw_invalid("the cross ordering runs at block width 32.");
        }
        // every 32-column block is one "pair" of two 16-column halves; one full cyclic sweep over its columns
        block_round_impl<T, 16, false>(g, g.nb, 2 * g.nb, 2 * g.nb, a, v, kIntraRound, eps, partial, q, apply_flag,
                                       rot_count, offmax, perm, norm2, skip_if_zero, stream);
    }

    template void launch_block_intra<float>(const BlockGeometry &, float *, float *, double, float *, float *, int *,
                                            int *, unsigned long long *, const int *, double *, const int *,
                                            hipStream_t);
    template void launch_block_intra<double>(const BlockGeometry &, double *, double *, double, double *, double *,
                                             int *, int *, unsigned long long *, const int *, double *, const int *,
                                             hipStream_t);
    template void launch_block_round<double>(const BlockGeometry &, double *, double *, int, double, double *,
                                             double *, int *, int *, unsigned long long *, const int *, double *,
                                             const int *, hipStream_t);

    template <typename T>
    void launch_block_eigen(const T *gram, int kb, int pairs, double eps, T *q_out, int *rotations, bool cross,
                            hipStream_t stream)
    {
        if (cross)
        {
            if (kb != 32)
            {
                throw_invalid("the cross ordering runs at block width 32.");
            }
            block_eigen_kb<T, 32, true>(gram, pairs, eps, q_out, rotations, stream);
            return;
        }
        switch (kb)
        {
        case 16: block_eigen_kb<T, 16, false>(gram, pairs, eps, q_out, rotations, stream); break;
        case 32: block_eigen_kb<T, 32, false>(gram, pairs, eps, q_out, rotations, stream); break;
        case 64: block_eigen_kb<T, 64, false>(gram, pairs, eps, q_out, rotations, stream); break;
        default: throw_invalid("block width must be 16, 32 or 64.");
        }
    }

    template void launch_block_eigen<float>(const float *, int, int, double, float *, int *, bool, hipStream_t);
    template void launch_block_eigen<double>(const double *, int, int, double, double *, int *, bool, hipStream_t);

    void launch_column_order(const double *norm2, int64_t n_pad, int *perm, const int *skip_if_zero,
                             hipStream_t stream)
    {
        if (n_pad > kSortMax)
        {
            return; // beyond one workgroup's LDS: keep the current order
        }
        hipLaunchKernelGGL(column_order_kernel, dim3(1), dim3(1024), 0, stream, norm2, static_cast<int>(n_pad), perm,
                           skip_if_zero);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_sweep_continue(const int *rot, int stop, int *cont, hipStream_t stream)
    {
        hipLaunchKernelGGL(sweep_continue_kernel, dim3(1), dim3(64), 0, stream, rot, stop, cont);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_iota(int *perm, int64_t n, hipStream_t stream)
    {
        hipLaunchKernelGGL(iota_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, perm, n);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_convert_f64_to_f32(const double *src, float *dst, int64_t count, hipStream_t stream)
    {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                           0, stream, src, dst, count);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_convert_f32_to_f64(const float *src, double *dst, int64_t count, hipStream_t stream)
    {
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                           0, stream, src, dst, count);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_pad_copy_f64(const double *src, int64_t lds, int64_t rows, int64_t cols, double *dst, int64_t ldd,
                             int64_t dcols, bool identity_pad, hipStream_t stream)
    {
        const int64_t total = ldd * dcols;
        hipLaunchKernelGGL(pad_copy_kernel<double>, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                           stream, src, lds, rows, cols, dst, ldd, dcols, identity_pad ? 1 : 0);
        JSVD_HIP_CHECK(hipGetLastError());
    }

    void launch_identity_f32(float *v, int64_t ldv, int64_t cols, hipStream_t stream)
    {
        const int64_t total = ldv * cols;
        hipLaunchKernelGGL(identity_kernel_t<float>, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                           stream, v, ldv, cols);
        JSVD_HIP_CHECK(hipGetLastError());
    }
} // namespace jsvd
