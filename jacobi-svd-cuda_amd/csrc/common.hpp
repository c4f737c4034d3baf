// common.hpp -- shared helpers of libjsvd: status/error plumbing for the C ABI, HIP error checks,
// device reductions and the closed-form round-robin schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "jsvd.h"

namespace jsvd
{
    // Exception carrying a C-ABI status code.  Thrown inside the library, caught at the ABI edge and
    // turned into (status, thread-local message).
    class Error : public std::runtime_error
    {
    public:
        Error(int status, const std::string &message) : std::runtime_error(message), status_(status) {}
        [[nodiscard]] int status() const noexcept { return status_; }

    private:
        int status_;
    };

    [[noreturn]] inline void throw_invalid(const std::string &message)
    {
        throw Error(JSVD_ERR_INVALID_ARGUMENT, message);
    }

    // Message shape follows the reference's CUDA check (src/kernels.cu:29-40), with HIP names.
    [[noreturn]] void throw_hip(hipError_t status, const char *expression, const char *file, int line);

    void set_last_error(const std::string &message);
    void clear_last_error();

    // Runs `body` and converts any exception into a status code plus the thread-local message.
    template <typename F>
    int guarded(F &&body) noexcept
    {
        try
        {
            clear_last_error();
            body();
            return JSVD_OK;
        }
        catch (const Error &e)
        {
            set_last_error(e.what());
            return e.status();
        }
        catch (const std::invalid_argument &e)
        {
            set_last_error(e.what());
            return JSVD_ERR_INVALID_ARGUMENT;
        }
        catch (const std::exception &e)
        {
            set_last_error(e.what());
            return JSVD_ERR_INTERNAL;
        }
        catch (...)
        {
            set_last_error("unknown exception");
            return JSVD_ERR_INTERNAL;
        }
    }

    // Validation of the solver configuration, restating src/kernels.cu:338-356 and 807-833.
    void validate_config(const jsvd_config &cfg);
    void validate_shape(int64_t m, int64_t n);
    // Reference normalisation of threads_per_block (src/kernels.cu:685-690).
    int normalize_threads_per_block(int raw);
    bool use_wide_shape(int64_t m, int64_t n, const jsvd_config &cfg);

    inline int64_t round_up(int64_t x, int64_t k) { return (x + k - 1) / k * k; }
} // namespace jsvd

#define JSVD_HIP_CHECK(EXPR)                                                                  \
    do                                                                                        \
    {                                                                                         \
        const hipError_t jsvd_status_ = (EXPR);                                               \
        if (jsvd_status_ != hipSuccess)                                                       \
        {                                                                                     \
            ::jsvd::throw_hip(jsvd_status_, #EXPR, __FILE__, __LINE__);                        \
        }                                                                                     \
    } while (0)

namespace jsvd::dev
{
#ifndef JSVD_SIMT_MODEL
    // N-element register vector of T (MFMA operands and accumulators, 16-byte loads).
    template <typename T, int N>
    using vec_t = T __attribute__((ext_vector_type(N)));

    // Keeps v as computed at this point: an empty asm the compiler cannot see through, so a cross-lane move is
    // not sunk into a later branch (where it would run with lanes inactive).
    template <typename T>
    __device__ inline void keep(T &v)
    {
        asm volatile("" : "+v"(v));
    }
#else
    // The CPU model of a workgroup (tests/simt_model) builds the kernels with g++: GCC vector types, and no
    // register constraint to keep.
    template <typename T, int N>
    struct vec_of;
    template <>
    struct vec_of<float, 4>
    {
        typedef float type __attribute__((vector_size(16)));
    };
    template <>
    struct vec_of<double, 4>
    {
        typedef double type __attribute__((vector_size(32)));
    };
    template <>
    struct vec_of<double, 2>
    {
        typedef double type __attribute__((vector_size(16)));
    };
    template <typename T, int N>
    using vec_t = typename vec_of<T, N>::type;

    template <typename T>
    inline void keep(T &)
    {
    }
#endif

    // Position `slot` of the round-robin circle at round `r` for an even circle of size N.
    // The circle starts as 0..N-1 (value N-1 is the dummy when the column count is odd) and after every
    // round positions 1..N-1 rotate right by one, so position i >= 1 holds 1 + ((i-1-r) mod (N-1)).
    // This is the closed form of the host schedule of src/kernels.cu:637-678.
    __host__ __device__ inline int circle_at(int slot, int r, int N)
    {
        if (slot == 0)
        {
            return 0;
        }
        const int period = N - 1;
        int k = (slot - 1 - r) % period;
        if (k < 0)
        {
            k += period;
        }
        return 1 + k;
    }

    // Rotation parameters and skip test in FP64 (src/kernels.cu:500-511).  Contraction is disabled so
    // the result is bit-identical to an IEEE evaluation of the same formula on the host.
    __host__ __device__ inline bool rotation_params(double app, double aqq, double apq, double eps, double &c,
                                                    double &s)
    {
#pragma clang fp contract(off)
        const double rhs = eps * sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
        if (fabs(apq) > rhs && fabs(apq) > 1.0e-300)
        {
            const double tau = (aqq - app) / (2.0 * apq);
            const double root = sqrt(1.0 + tau * tau);
            const double t = (tau >= 0.0) ? (1.0 / (tau + root)) : (-1.0 / (-tau + root));
            const double cosine = 1.0 / sqrt(1.0 + t * t);
            c = cosine;
            s = t * cosine;
            return true;
        }
        c = 1.0;
        s = 0.0;
        return false;
    }

    // circle_at for 0 <= r < N-1 without an integer modulo (inner loops).
    __device__ inline int circle_at_fast(int slot, int r, int N)
    {
        int k = slot - 1 - r;
        k += (k < 0) ? (N - 1) : 0;
        return (slot == 0) ? 0 : 1 + k;
    }

    // Same skip test as rotation_params, shorter dependent chain for the block paths' inner sweeps:
    // with d = beta - alpha and g = 2 gamma scaled by 1 / max(|d|, |g|) (no overflow for any finite input),
    // t = sign(d) g / (|d| + hypot(d, g)) and c = (|d| + hypot) / hypot(|d| + hypot, g), the same rotation as
    // tau = d / g, t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2).
    __device__ inline bool rotation_params_block(double app, double aqq, double apq, double eps, double &c,
                                                 double &s)
    {
        const double rhs = eps * sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
        if (fabs(apq) > rhs && fabs(apq) > 1.0e-300)
        {
            const double d0 = aqq - app;
            const double g0 = 2.0 * apq;
            const double inv = 1.0 / fmax(fabs(d0), fabs(g0));
            const double d = d0 * inv;
            const double g = g0 * inv;
            const double hyp = sqrt(d * d + g * g);
            const double den = fabs(d) + hyp;
            const double t = ((d >= 0.0) ? g : -g) / den;
            const double cosine = den / sqrt(den * den + g * g);
            c = cosine;
            s = t * cosine;
            return true;
        }
        c = 1.0;
        s = 0.0;
        return false;
    }

    // rotation_params_block with a shorter chain: d and g are scaled by an exact power of two (no division),
    // and with w = 1 / sqrt(den^2 + g^2) the rotation is c = den w, s = sign(d) g w (s = t c), so the chain is
    // one sqrt and one rsqrt.  Branch-free: the parameters of a skipped pair are replaced by (1, 0).
    __device__ inline bool rotation_params_fast(double app, double aqq, double apq, double eps, double &c,
                                                double &s)
    {
        const double rhs = eps * sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
        const bool rot = fabs(apq) > rhs && fabs(apq) > 1.0e-300;
        const double d0 = aqq - app;
        const double g0 = 2.0 * apq;
        int e = 0;
        (void)frexp(fmax(fabs(d0), fabs(g0)), &e);
        const double d = ldexp(d0, -e);
        const double g = ldexp(g0, -e);
        const double hyp = sqrt(d * d + g * g);
        const double den = fabs(d) + hyp;
        const double w = rsqrt(den * den + g * g);
        c = rot ? den * w : 1.0;
        s = rot ? ((d >= 0.0) ? g : -g) * w : 0.0;
        return rot;
    }

    // 1 / sqrt(x) for x > 0: the hardware estimate (v_rsq_f64, about 2^-23 relative) refined by one Newton step
    // y (3 - x y^2) / 2, which leaves about 2^-45 relative.
    __device__ inline double rsqrt_newton(double x)
    {
        const double y = __builtin_amdgcn_rsq(x);
        const double e = fma(-x * y, y, 1.0); // 1 - x y^2
        return fma(0.5 * y, e, y);
    }

    // Rotation parameters for Gram entries that are FP32 values (the FP32 phase of the block paths, whose
    // rotations are applied in FP32): the skip test of rotation_params on squares (apq^2 > eps^2 app aqq, no square
    // root), then the same rotation as rotation_params_block (t = sign(d) g / (|d| + hypot(d, g)),
    // c = 1 / sqrt(1 + t^2)) as c = den w, s = sign(d) g w with den = |d| + hypot(d, g), w = 1 / hypot(den, g), the
    // two square roots by the hardware estimate and one Newton step (about 2^-45 relative, far below FP32
    // rounding).  FP32 inputs have no overflow or underflow in these FP64 squares, so no scaling is needed.
    // Branch-free: a skipped pair gets (1, 0).
    __device__ inline bool rotation_params_f32(double app, double aqq, double apq, double eps, double &c, double &s)
    {
        const double pq = fmax(app, 0.0) * fmax(aqq, 0.0);
        const bool rot = (apq * apq > (eps * eps) * pq) && fabs(apq) > 1.0e-300;
        const double d = aqq - app;
        const double g = 2.0 * apq;
        const double g2 = g * g;
        const double h2 = fma(d, d, g2);
        const double hyp = h2 * rsqrt_newton(h2);
        const double den = fabs(d) + hyp;
        const double w = rsqrt_newton(fma(den, den, g2));
        c = rot ? den * w : 1.0;
        s = rot ? ((d >= 0.0) ? g : -g) * w : 0.0;
        return rot;
    }

    // Off-orthogonality of a pair, |gamma| / sqrt(alpha beta), 0 when a column is zero.
    __host__ __device__ inline double pair_cosine(double app, double aqq, double apq)
    {
        const double denom = sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
        return denom > 0.0 ? fabs(apq) / denom : 0.0;
    }

    // Butterfly sum over the 64 lanes of a wavefront; every lane ends with the total.
    __device__ inline double wave_sum(double x)
    {
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            x += __shfl_xor(x, offset, 64);
        }
        return x;
    }

    // Lane values moved by a DPP pattern (CTRL, ROW_MASK: compile-time), 0 where the pattern reads outside a
    // row or the row is masked off.  64-bit values move as two 32-bit halves with the same pattern.
    template <int CTRL, int ROW_MASK>
    __device__ inline float dpp_move(float v)
    {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true));
    }

    template <int CTRL, int ROW_MASK>
    __device__ inline double dpp_move(double v)
    {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROW_MASK, 0xF, true);
        const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROW_MASK, 0xF, true);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    }

    __device__ inline float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }

    __device__ inline double lane63(double v)
    {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), 63);
        const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    }

    // Sum over the 64 lanes of a wavefront with DPP row moves instead of LDS-crossbar shuffles: an inclusive
    // scan inside each 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into row 3
    // (row_bcast:15), then rows 0-1 into rows 2-3 (row_bcast:31); lane 63 holds the total, which every lane
    // reads back (v_readlane).  A fixed summation order, so the result does not depend on anything but the
    // lane values.
    template <typename T>
    __device__ inline T wave_sum_dpp(T x)
    {
        x += dpp_move<0x111, 0xF>(x); // row_shr:1
        x += dpp_move<0x112, 0xF>(x); // row_shr:2
        x += dpp_move<0x114, 0xF>(x); // row_shr:4
        x += dpp_move<0x118, 0xF>(x); // row_shr:8
        x += dpp_move<0x142, 0xA>(x); // row_bcast:15 into rows 1 and 3
        x += dpp_move<0x143, 0xC>(x); // row_bcast:31 into rows 2 and 3
        return lane63(x);
    }

    __device__ inline double wave_max(double x)
    {
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            x = fmax(x, __shfl_xor(x, offset, 64));
        }
        return x;
    }

    // atomicMax of a non-negative double through its bit pattern (IEEE order == unsigned order for x >= 0).
    __device__ inline void atomic_max_nonneg(unsigned long long *slot, double x)
    {
        atomicMax(slot, static_cast<unsigned long long>(__double_as_longlong(x)));
    }
} // namespace jsvd::dev
