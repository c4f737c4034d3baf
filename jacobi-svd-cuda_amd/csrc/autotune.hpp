// autotune.hpp -- measured launch parameters per matrix shape (autotune.hip).
#pragma once

#include "block.hpp"
#include "plain_f64.hpp"

namespace jsvd
{
    struct Tuning
    {
        RoundShape round{};  // plain path: round-kernel launch shape
        int apply_wgs = 512; // block paths: total workgroups of one apply launch
        int block_width = 32; // block paths: columns per block
    };

    // The cached choice for m x n, if jsvd_autotune measured it in this process.
    bool lookup_tuning(int64_t m, int64_t n, Tuning &out);

    // Measures the candidates for m x n on `stream`, caches and returns the fastest.
    Tuning run_autotune(int64_t m, int64_t n, int repetitions, hipStream_t stream, jsvd_tune_report *report);
} // namespace jsvd
