// config.cpp -- host-only part of the C ABI (include/jsvd.h): the thread-local error message, the HIP error
// message, validation and normalisation of the solver configuration and matrix shape, and the configuration
// defaults.  No device code: the host sanitizer builds compile this file unchanged.
#include <algorithm>
#include <cstring>
#include <sstream>
#include <string>

#include "common.hpp"

namespace jsvd
{
    namespace
    {
        thread_local std::string t_last_error;
    }

    void set_last_error(const std::string &message) { t_last_error = message; }
    void clear_last_error() { t_last_error.clear(); }

    void throw_hip(hipError_t status, const char *expression, const char *file, int line)
    {
        std::ostringstream stream;
        stream << "HIP call failed: " << expression << " @ " << file << ':' << line
               << ", code=" << static_cast<int>(status) << ", message=" << hipGetErrorString(status);
        throw Error(JSVD_ERR_DEVICE, stream.str());
    }

    // src/kernels.cu:338-356 (layout knobs) and 823-832 (epsilon, max_sweeps).
    void validate_config(const jsvd_config &cfg)
    {
        if (!(cfg.epsilon > 0.0))
        {
            throw_invalid("epsilon must be positive.");
        }
        if (cfg.max_sweeps <= 0)
        {
            throw_invalid("max_sweeps must be positive.");
        }
        if (cfg.layout_transpose_min_columns <= 0)
        {
            throw_invalid("layout_transpose_min_columns must be positive.");
        }
        if (cfg.layout_transpose_min_elements == 0)
        {
            throw_invalid("layout_transpose_min_elements must be positive.");
        }
        if (cfg.layout_transpose_benchmark_repetitions <= 0)
        {
            throw_invalid("layout_transpose_benchmark_repetitions must be positive.");
        }
        if (cfg.layout_transpose_benchmark_sweeps <= 0)
        {
            throw_invalid("layout_transpose_benchmark_sweeps must be positive.");
        }
        if (cfg.layout_transpose_mode < JSVD_LAYOUT_AUTO || cfg.layout_transpose_mode > JSVD_LAYOUT_FORCE_DISABLE)
        {
            throw_invalid("layout_transpose_mode is not a valid LayoutTransposeMode.");
        }
    }

    // src/kernels.cu:813-820.
    void validate_shape(int64_t m, int64_t n)
    {
        if (m <= 0 || n <= 0)
        {
            throw_invalid("Input matrix shape must be non-zero.");
        }
        if (m < n)
        {
            throw_invalid("One-sided Jacobi SVD requires rows >= columns in this implementation.");
        }
        if (m > (int64_t{1} << 31) || n > (int64_t{1} << 30))
        {
            throw_invalid("Input matrix shape exceeds the supported range.");
        }
    }

    // The reference's layout policy (src/kernels.cu:365-380) decides between its strided and its transposed
    // path; here it decides between the two launch shapes of short columns.
    bool use_wide_shape(int64_t m, int64_t n, const jsvd_config &cfg)
    {
        if (cfg.layout_transpose_mode == JSVD_LAYOUT_FORCE_ENABLE)
        {
            return true;
        }
        if (cfg.layout_transpose_mode == JSVD_LAYOUT_FORCE_DISABLE)
        {
            return false;
        }
        return n >= cfg.layout_transpose_min_columns &&
               static_cast<uint64_t>(m) * static_cast<uint64_t>(n) >= cfg.layout_transpose_min_elements;
    }

    int normalize_threads_per_block(int raw)
    {
        const int clamped = std::clamp(raw, 32, 1024);
        return std::min(((clamped + 31) / 32) * 32, 1024);
    }
} // namespace jsvd

extern "C" {

int jsvd_config_validate(const jsvd_config *cfg)
{
    return jsvd::guarded([&] {
        if (!cfg)
        {
            jsvd::throw_invalid("config must not be null.");
        }
        jsvd::validate_config(*cfg);
    });
}

void jsvd_config_default(jsvd_config *cfg)
{
    if (!cfg)
    {
        return;
    }
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->epsilon = 1.0e-9;
    cfg->max_sweeps = 128;
    cfg->threads_per_block = 256;
    cfg->layout_transpose_mode = JSVD_LAYOUT_AUTO;
    cfg->layout_transpose_min_columns = 16;
    cfg->layout_transpose_min_elements = 4096;
    cfg->layout_transpose_auto_tune = 0;
    cfg->layout_transpose_benchmark_repetitions = 2;
    cfg->layout_transpose_benchmark_sweeps = 8;
}

const char *jsvd_last_error(void) { return jsvd::t_last_error.c_str(); }

const char *jsvd_version(void) { return "jsvd-gfx950 0.1.0"; }

void jsvd_mixed_config_default(jsvd_mixed_config *mixed)
{
    if (!mixed)
    {
        return;
    }
    std::memset(mixed, 0, sizeof(*mixed));
    mixed->path = JSVD_PATH_MIXED;
    mixed->fp32_max_sweeps = 40;
    mixed->fp32_tolerance = 0.0;
    mixed->orthogonalize_steps = 2;
    mixed->qr_precondition = 3;
    mixed->fp32_stop_fraction = 0.05;
}

} // extern "C"
