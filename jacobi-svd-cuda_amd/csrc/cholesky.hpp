// cholesky.hpp -- blocked FP64 Cholesky factorisation (cholesky.hip).
#pragma once

#include "common.hpp"

namespace jsvd
{
    // n rounded up to the 64 x 64 tile size of launch_cholesky_upper.
    int64_t cholesky_padded_size(int64_t n);

    // A = R^T R for the symmetric positive definite n x n matrix a (column-major, leading dimension ld >=
    // cholesky_padded_size(n)); the upper triangle of a is overwritten by R.  The padding rows and columns
    // n..cholesky_padded_size(n)-1 are set to the identity first, so a must have room for them.  *info (device)
    // becomes 0, or 1 + the first row whose pivot was not positive.
    void launch_cholesky_upper(double *a, int64_t ld, int64_t n, int *info, hipStream_t stream);
} // namespace jsvd
