// kernels.hpp -- C++ entry points of the one-sided Jacobi SVD, source-compatible with the reference's
// include/kernels.hpp (kleedaisuki/jacobi-svd-cuda, namespace jacobi::svd).  The solver behind them is
// libjsvd.so (HIP for gfx950) reached through the C ABI of jsvd.h.
#pragma once

#include <cstddef>
#include <span>
#include <stdexcept>
#include <vector>

namespace jacobi::svd
{
    // Layout policy knob (reference include/kernels.hpp:13-29).  The MI355X solver always works on a
    // column-major device copy of A (coalesced column-pair access); the knob is validated and reported.
    enum class LayoutTransposeMode
    {
        auto_select,
        force_enable,
        force_disable
    };

    // Result of auto_tune_layout_transpose_threshold (reference include/kernels.hpp:34-60).
    struct LayoutTransposeAutoTuneReport final
    {
        bool executed = false;
        int recommended_min_columns = 16;
        std::size_t recommended_min_elements = 4096;
        double estimated_best_speedup = 1.0;
        std::size_t sample_count = 0;
    };

    // Device/runtime failure (reference include/kernels.hpp:65-73).  The name is kept for source
    // compatibility; the message reports the failing HIP call.
    class CudaError final : public std::runtime_error
    {
    public:
        explicit CudaError(const char *message);
    };

    // Solver configuration (reference include/kernels.hpp:78-124): same fields, order and defaults, so
    // designated initializers written against the reference compile unchanged.
    struct JacobiSvdConfig final
    {
        double epsilon = 1.0e-9;
        int max_sweeps = 128;
        int threads_per_block = 256;
        LayoutTransposeMode layout_transpose_mode = LayoutTransposeMode::auto_select;
        int layout_transpose_min_columns = 16;
        std::size_t layout_transpose_min_elements = 4096;
        bool layout_transpose_auto_tune = false;
        int layout_transpose_benchmark_repetitions = 2;
        int layout_transpose_benchmark_sweeps = 8;
    };

    // Move-only owner of a row-major FP64 matrix in device memory (reference include/kernels.hpp:130-244).
    class DeviceMatrix final
    {
    public:
        DeviceMatrix() = default;
        DeviceMatrix(std::size_t rows, std::size_t columns);
        ~DeviceMatrix();

        DeviceMatrix(const DeviceMatrix &) = delete;
        DeviceMatrix &operator=(const DeviceMatrix &) = delete;
        DeviceMatrix(DeviceMatrix &&other) noexcept;
        DeviceMatrix &operator=(DeviceMatrix &&other) noexcept;

        void reset(std::size_t rows, std::size_t columns);

        [[nodiscard]] std::size_t rows() const noexcept;
        [[nodiscard]] std::size_t columns() const noexcept;
        [[nodiscard]] std::size_t size() const noexcept;
        [[nodiscard]] std::size_t bytes() const noexcept;
        [[nodiscard]] double *data() noexcept;
        [[nodiscard]] const double *data() const noexcept;

        void copy_from_host(std::span<const double> host_values);
        [[nodiscard]] std::vector<double> copy_to_host() const;

    private:
        std::size_t rows_ = 0;
        std::size_t columns_ = 0;
        double *data_ = nullptr;
    };

    // Host result (reference include/kernels.hpp:249-280): U m x n and V n x n row-major, sigma in
    // Jacobi column order (not sorted), executed sweeps (== max_sweeps when not converged).
    struct JacobiSvdResult final
    {
        std::size_t rows = 0;
        std::size_t columns = 0;
        std::vector<double> u;
        std::vector<double> sigma;
        std::vector<double> v;
        int sweeps = 0;
    };

    // Thin SVD of the row-major m x n matrix host_input (m >= n) by cyclic one-sided Jacobi with the
    // round-robin pair ordering (reference include/kernels.hpp:295-298).  Throws std::invalid_argument on
    // bad shapes or configuration and CudaError on device failures.
    [[nodiscard]] JacobiSvdResult one_sided_jacobi_svd(std::span<const double> host_input,
                                                       std::size_t rows,
                                                       std::size_t columns,
                                                       const JacobiSvdConfig &config = {});

    // Micro-benchmark of the solver over the reference's sample sizes (n in {8,...,96}, m = 2n), reported in
    // the reference's report struct (include/kernels.hpp:306-307).  On MI355X the "transpose path" and the
    // "direct path" are the two launch shapes the solver can take for small matrices (one wavefront per
    // column pair versus a 256-thread workgroup per pair); the recommended thresholds say from which size
    // the wider shape wins.
    [[nodiscard]] LayoutTransposeAutoTuneReport auto_tune_layout_transpose_threshold(
        const JacobiSvdConfig &config = {});
} // namespace jacobi::svd
