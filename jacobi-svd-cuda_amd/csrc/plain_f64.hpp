// plain_f64.hpp -- launchers of the plain FP64 kernels (plain_f64.hip).
#pragma once

#include "common.hpp"

namespace jsvd
{
    // Launch shape of the fused round kernel: `threads` per workgroup (one workgroup per column pair) and the
    // number of double2 chunks per thread holding an A column (ea) and a V column (ev); ea == 0 selects the
    // streaming variant for columns longer than 16 * threads rows.
    struct RoundShape
    {
        int threads = 256;
        int ea = 1;
        int ev = 1;
    };

    // `wide` selects a 256-thread workgroup per pair where one wavefront per pair would also fit
    // (m <= 1024); longer columns always take 256- or 1024-thread workgroups.
    RoundShape choose_round_shape(int64_t m, int64_t n, bool wide);

    // One fused round (Gram + rotation + apply to A and V) over the N/2 slots of round `round`.  rot_count and
    // offmax (each may be null) hold count_slots entries (a power of two): the workgroup of slot k adds its
    // rotation to rot_count[k % count_slots] and its pair cosine to offmax[k % count_slots];
    // launch_sweep_totals folds the slots.
    void launch_round(const RoundShape &shape, double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n,
                      int round, double eps, int *rot_count, unsigned long long *offmax, const int *skip_if_zero,
                      hipStream_t stream, int count_slots = 1);

    // rot_total = sum of rot_slots[0..slots), off_total = max of off_slots[0..slots) (one workgroup).
    void launch_sweep_totals(const int *rot_slots, const unsigned long long *off_slots, int slots, int *rot_total,
                             unsigned long long *off_total, hipStream_t stream);

    // Row-major rows x cols (lds) -> row-major cols x rows (ldd).
    void launch_transpose(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                          hipStream_t stream);

    // sigma and row-major U from the column-major A (lda).
    void launch_u_sigma(const double *a, int64_t lda, int64_t m, int64_t n, double eps, double *u, double *sigma,
                        hipStream_t stream);

    void launch_identity(double *v, int64_t ldv, int64_t n, hipStream_t stream);

    void launch_pair_gram(const double *a, int64_t lda, int64_t m, const int *pairs, int64_t count, double *alpha,
                          double *beta, double *gamma, hipStream_t stream);

    void launch_rotation_params(const double *alpha, const double *beta, const double *gamma, int64_t count,
                                double eps, double *c, double *s, int *rotate, hipStream_t stream);

    int64_t offnorm_workspace_doubles(int64_t n);
    void launch_offnorm(const double *a, int64_t lda, int64_t m, int64_t n, double *workspace, double *out,
                        hipStream_t stream);
} // namespace jsvd
