// kernels.cpp -- the jacobi::svd C++ entry points (include/kernels.hpp) over the C ABI of libjsvd.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <limits>
#include <sstream>
#include <string>
#include <utility>

#include "jsvd.h"

namespace jacobi::svd
{
    namespace
    {
        void throw_if_hip_failed(hipError_t status, const char *expression)
        {
            if (status == hipSuccess)
            {
                return;
            }
            std::ostringstream stream;
            stream << "HIP call failed: " << expression << ", code=" << static_cast<int>(status)
                   << ", message=" << hipGetErrorString(status);
            throw CudaError(stream.str().c_str());
        }

#define JACOBI_HIP_CHECK(EXPR) throw_if_hip_failed((EXPR), #EXPR)

        // Status code of the C ABI -> the reference's exception types (std::invalid_argument, CudaError,
        // std::runtime_error).
        void raise_status(int status)
        {
            if (status == JSVD_OK)
            {
                return;
            }
            const std::string message = jsvd_last_error();
            if (status == JSVD_ERR_INVALID_ARGUMENT)
            {
                throw std::invalid_argument(message);
            }
            if (status == JSVD_ERR_DEVICE)
            {
                throw CudaError(message.c_str());
            }
            throw std::runtime_error(message);
        }

        jsvd_config to_c_config(const JacobiSvdConfig &config)
        {
            jsvd_config c;
            jsvd_config_default(&c);
            c.epsilon = config.epsilon;
            c.max_sweeps = config.max_sweeps;
            c.threads_per_block = config.threads_per_block;
            c.layout_transpose_mode = static_cast<int32_t>(config.layout_transpose_mode);
            c.layout_transpose_min_columns = config.layout_transpose_min_columns;
            c.layout_transpose_min_elements = config.layout_transpose_min_elements;
            c.layout_transpose_auto_tune = config.layout_transpose_auto_tune ? 1 : 0;
            c.layout_transpose_benchmark_repetitions = config.layout_transpose_benchmark_repetitions;
            c.layout_transpose_benchmark_sweeps = config.layout_transpose_benchmark_sweeps;
            return c;
        }

        // The reference's autotuner sample (src/kernels.cu:1012-1025): a smooth deterministic 2n x n matrix.
        std::vector<double> benchmark_matrix(std::size_t rows, std::size_t columns)
        {
            std::vector<double> values(rows * columns);
            for (std::size_t r = 0; r < rows; ++r)
            {
                for (std::size_t c = 0; c < columns; ++c)
                {
                    values[r * columns + c] = std::sin(0.013 * static_cast<double>((r + 1) * (c + 1))) +
                                              std::cos(0.017 * static_cast<double>((r + 3) * (c + 5)));
                }
            }
            return values;
        }

        // One warm-up, then the mean wall time of `reps` solves through a reusable plan.
        double time_solves_ms(const std::vector<double> &a, std::size_t rows, std::size_t columns,
                              const jsvd_config &cfg, int reps)
        {
            jsvd_plan *plan = nullptr;
            raise_status(jsvd_plan_create(static_cast<int64_t>(rows), static_cast<int64_t>(columns), &cfg, nullptr,
                                          &plan));
            std::vector<double> u(rows * columns), s(columns), v(columns * columns);
            int32_t sweeps = 0;
            double total = 0.0;
            int status = jsvd_plan_solve_host(plan, a.data(), u.data(), s.data(), v.data(), &sweeps);
            for (int i = 0; status == JSVD_OK && i < reps; ++i)
            {
                const auto t0 = std::chrono::steady_clock::now();
                status = jsvd_plan_solve_host(plan, a.data(), u.data(), s.data(), v.data(), &sweeps);
                const auto t1 = std::chrono::steady_clock::now();
                total += std::chrono::duration<double, std::milli>(t1 - t0).count();
            }
            if (status != JSVD_OK)
            {
                const std::string message = jsvd_last_error();
                (void)jsvd_plan_destroy(plan);
                if (status == JSVD_ERR_INVALID_ARGUMENT)
                {
                    throw std::invalid_argument(message);
                }
                throw CudaError(message.c_str());
            }
            raise_status(jsvd_plan_destroy(plan));
            return total / static_cast<double>(reps);
        }
    } // namespace

    CudaError::CudaError(const char *message) : std::runtime_error(message) {}

    DeviceMatrix::DeviceMatrix(std::size_t rows, std::size_t columns) { reset(rows, columns); }

    DeviceMatrix::~DeviceMatrix()
    {
        if (data_ != nullptr)
        {
            (void)hipFree(data_);
        }
    }

    DeviceMatrix::DeviceMatrix(DeviceMatrix &&other) noexcept
        : rows_(std::exchange(other.rows_, 0)), columns_(std::exchange(other.columns_, 0)),
          data_(std::exchange(other.data_, nullptr))
    {
    }

    DeviceMatrix &DeviceMatrix::operator=(DeviceMatrix &&other) noexcept
    {
        if (this != &other)
        {
            if (data_ != nullptr)
            {
                (void)hipFree(data_);
            }
            rows_ = std::exchange(other.rows_, 0);
            columns_ = std::exchange(other.columns_, 0);
            data_ = std::exchange(other.data_, nullptr);
        }
        return *this;
    }

    void DeviceMatrix::reset(std::size_t rows, std::size_t columns)
    {
        if (rows == 0 || columns == 0)
        {
            throw std::invalid_argument("DeviceMatrix dimensions must be positive.");
        }
        if (columns > std::numeric_limits<std::size_t>::max() / sizeof(double) / rows)
        {
            throw std::invalid_argument("DeviceMatrix dimensions overflow.");
        }
        if (data_ != nullptr)
        {
            JACOBI_HIP_CHECK(hipFree(data_));
            data_ = nullptr;
        }
        rows_ = rows;
        columns_ = columns;
        JACOBI_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&data_), bytes()));
    }

    std::size_t DeviceMatrix::rows() const noexcept { return rows_; }
    std::size_t DeviceMatrix::columns() const noexcept { return columns_; }
    std::size_t DeviceMatrix::size() const noexcept { return rows_ * columns_; }
    std::size_t DeviceMatrix::bytes() const noexcept { return size() * sizeof(double); }
    double *DeviceMatrix::data() noexcept { return data_; }
    const double *DeviceMatrix::data() const noexcept { return data_; }

    void DeviceMatrix::copy_from_host(std::span<const double> host_values)
    {
        if (host_values.size() != size())
        {
            throw std::invalid_argument("Host data size does not match device matrix shape.");
        }
        JACOBI_HIP_CHECK(hipMemcpy(data_, host_values.data(), bytes(), hipMemcpyHostToDevice));
    }

    std::vector<double> DeviceMatrix::copy_to_host() const
    {
        std::vector<double> host(size());
        if (!host.empty())
        {
            JACOBI_HIP_CHECK(hipMemcpy(host.data(), data_, bytes(), hipMemcpyDeviceToHost));
        }
        return host;
    }

    JacobiSvdResult one_sided_jacobi_svd(std::span<const double> host_input, std::size_t rows, std::size_t columns,
                                         const JacobiSvdConfig &config)
    {
        // Shape checks first, in the reference's order (src/kernels.cu:807-833).
        if (rows == 0 || columns == 0)
        {
            throw std::invalid_argument("Input matrix shape must be non-zero.");
        }
        if (rows < columns)
        {
            throw std::invalid_argument("One-sided Jacobi SVD requires rows >= columns in this implementation.");
        }
        if (host_input.size() != rows * columns)
        {
            throw std::invalid_argument("Input buffer size mismatch.");
        }
        const jsvd_config c = to_c_config(config);
        JacobiSvdResult result;
        result.rows = rows;
        result.columns = columns;
        result.u.resize(rows * columns);
        result.sigma.resize(columns);
        result.v.resize(columns * columns);
        int32_t sweeps = 0;
        raise_status(jsvd_plain_f64(host_input.data(), static_cast<int64_t>(rows), static_cast<int64_t>(columns), &c,
                                    result.u.data(), result.sigma.data(), result.v.data(), &sweeps));
        result.sweeps = sweeps;
        return result;
    }

    LayoutTransposeAutoTuneReport auto_tune_layout_transpose_threshold(const JacobiSvdConfig &config)
    {
        // Same validation and order as the reference (src/kernels.cu:1078-1085): layout knobs, then sweeps.
        if (config.layout_transpose_min_columns <= 0)
        {
            throw std::invalid_argument("layout_transpose_min_columns must be positive.");
        }
        if (config.layout_transpose_min_elements == 0)
        {
            throw std::invalid_argument("layout_transpose_min_elements must be positive.");
        }
        if (config.layout_transpose_benchmark_repetitions <= 0)
        {
            throw std::invalid_argument("layout_transpose_benchmark_repetitions must be positive.");
        }
        if (config.layout_transpose_benchmark_sweeps <= 0)
        {
            throw std::invalid_argument("layout_transpose_benchmark_sweeps must be positive.");
        }
        if (config.max_sweeps <= 0)
        {
            throw std::invalid_argument("max_sweeps must be positive.");
        }
        jsvd_config base = to_c_config(config);
        base.layout_transpose_auto_tune = 0;
        base.max_sweeps = std::min(config.max_sweeps, std::max(1, config.layout_transpose_benchmark_sweeps));

        constexpr std::array<int, 8> sample_columns = {8, 12, 16, 24, 32, 48, 64, 96};
        LayoutTransposeAutoTuneReport report{};
        report.executed = true;
        report.recommended_min_columns = config.layout_transpose_min_columns;
        report.recommended_min_elements = config.layout_transpose_min_elements;
        report.estimated_best_speedup = 1.0;
        report.sample_count = sample_columns.size();
        bool selected = false;
        for (const int n : sample_columns)
        {
            const std::size_t cols = static_cast<std::size_t>(n);
            const std::size_t rows = 2 * cols;
            const std::vector<double> a = benchmark_matrix(rows, cols);
            jsvd_config narrow = base;
            narrow.layout_transpose_mode = JSVD_LAYOUT_FORCE_DISABLE;
            jsvd_config wide = base;
            wide.layout_transpose_mode = JSVD_LAYOUT_FORCE_ENABLE;
            const int reps = config.layout_transpose_benchmark_repetitions;
            const double narrow_ms = time_solves_ms(a, rows, cols, narrow, reps);
            const double wide_ms = time_solves_ms(a, rows, cols, wide, reps);
            const double speedup = narrow_ms / std::max(wide_ms, std::numeric_limits<double>::min());
            report.estimated_best_speedup = std::max(report.estimated_best_speedup, speedup);
            if (!selected && wide_ms < narrow_ms)
            {
                report.recommended_min_columns = n;
                report.recommended_min_elements = rows * cols;
                selected = true;
            }
        }
        return report;
    }
} // namespace jacobi::svd
