// autotune.hpp -- measured launch parameters per matrix shape (autotune.hip).
#pragma once

#include "block.hpp"
#include "plain_f64.hpp"

namespace jsvd
{
    struct Tuning
    {
        RoundShape round{};   // plain path: round-kernel launch shape (valid when has_round)
        int span_waves = 0;   // plain path: round spans on this many wavefronts per workgroup (0: one launch per round)
        int span_rounds = 0;  // plain path: rounds per span launch at most
        int apply_wgs = 512;  // block sweeps over m x n: total workgroups of one apply launch (has_block)
        int block_width = 32; // block sweeps over m x n: columns per block (has_block)
        int batch_shape = 0;  // batched solver with limits m x n: index of its launch shape (has_batch)
        bool has_round = false;
        bool has_block = false;
        bool has_batch = false;
    };

    // The cached choices for m x n, if jsvd_autotune measured any in this process (jsvd_autotune(m, n) also
    // records the block choice for n x n, the sweep shape of the mixed path's QR-first form).
    bool lookup_tuning(int64_t m, int64_t n, Tuning &out);

    // Records the batched solver's measured launch shape for batches with limits max_rows x max_columns.
    void store_batch_shape(int64_t max_rows, int64_t max_columns, int shape);

    // Measures the candidates for m x n on `stream`, caches and returns the fastest.
    Tuning run_autotune(int64_t m, int64_t n, int repetitions, hipStream_t stream, jsvd_tune_report *report);
} // namespace jsvd
