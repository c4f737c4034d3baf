// plain_f64.hpp -- launchers of the plain FP64 kernels (plain_f64.hip).
#pragma once

#include "common.hpp"

#include <vector>

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

    // Round span (several rounds of a sweep per launch, bit-identical to one launch per round with `shape`):
    // whether the build has a span kernel for `shape` at `waves` wavefronts per workgroup (8 or 16) and this
    // m x n (register-resident columns: m <= 1024, n <= 512, n >= 3).
    bool span_supported(const RoundShape &shape, int waves, int64_t m, int64_t n);
    // Most rounds one span launch can run: the whole sweep when one workgroup holds every slot, else the rounds
    // whose halo still leaves each workgroup a slot of its own.
    int span_max_rounds(int waves, int64_t n);
    // Workgroups and owned slots per workgroup for spans of at most max_rounds rounds (throws when the cone of a
    // workgroup would not fit).
    void span_partition(int64_t n, int waves, int max_rounds, int &groups, int &own, int &own_end);
    // Rounds per launch over one sweep: an even number of launches of at most max_rounds rounds.
    std::vector<int> span_chunks(int64_t n, int max_rounds);
    // Rounds [r0, r0 + rounds) from (a_in, v_in) into (a_out, v_out); the columns must be in both buffers'
    // layout (lda, ldv).  Counter slots as launch_round.
    void launch_round_span(const RoundShape &shape, int waves, const double *a_in, double *a_out, int64_t lda,
                           int64_t m, const double *v_in, double *v_out, int64_t ldv, int64_t n, int r0, int rounds,
                           int groups, int own, int own_end, double eps, int *rot_count, unsigned long long *offmax,
                           const int *skip_if_zero, hipStream_t stream, int count_slots);

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
