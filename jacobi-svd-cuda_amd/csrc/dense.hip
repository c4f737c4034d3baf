// dense.hip -- launchers of the FP64 dense kernels (dense_kernels.hpp): tile grids, k-splitting, and the column
// ranges of the triangular solve.
#include <algorithm>
#include <mutex>
#include <vector>

#include "dense.hpp"
#include "dense_kernels.hpp"

namespace jsvd
{
    namespace
    {
        constexpr int64_t kSplitTilesMax = 1024; // partial tiles in flight: about four per CU of a 256-CU device

        int device_cus()
        {
            static std::mutex mu;
            static int cached[64] = {};
            int dev = 0;
            JSVD_HIP_CHECK(hipGetDevice(&dev));
            std::lock_guard<std::mutex> lock(mu);
            if (dev < 0 || dev >= 64)
            {
                return 256;
            }
            if (cached[dev] == 0)
            {
                int cus = 0;
                JSVD_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
                cached[dev] = std::max(cus, 1);
            }
            return cached[dev];
        }

    } // namespace

    int64_t dense_work_elems() { return kSplitTilesMax * 16 * 256; }

    void launch_gemm_f64(bool trans_a, bool trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                         int64_t lda, const double *b, int64_t ldb, double beta, const double *c, int64_t ldc,
                         double *d, int64_t ldd, int flags, const DenseWorkspace &ws, hipStream_t stream)
    {
        using namespace dense;
        if (m <= 0 || n <= 0)
        {
            return;
        }
        if ((flags & kUpperOut) && m != n)
        {
            throw_invalid("an upper-triangle product needs a square result.");
        }
        const dense::GemmParams p = plan_gemm(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, d, ldd, flags, device_cus(),
                                              ws.work, ws.work_elems);
        const dim3 grid(static_cast<unsigned>(p.tiles), static_cast<unsigned>(p.splits));
        if (trans_a && trans_b)
        {
            hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, stream, p);
        }
        else if (trans_a)
        {
            hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, stream, p);
        }
        else if (trans_b)
        {
            hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, stream, p);
        }
        else
        {
            hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, stream, p);
        }
        JSVD_HIP_CHECK(hipGetLastError());
        if (p.splits > 1)
        {
            hipLaunchKernelGGL(gemm_reduce_kernel, dim3(static_cast<unsigned>(p.tiles)), dim3(256), 0, stream, p);
            JSVD_HIP_CHECK(hipGetLastError());
        }
    }

    int64_t trsm_padded_size(int64_t n) { return (n + dense::kSolveBlock - 1) / dense::kSolveBlock * dense::kSolveBlock; }
    int64_t trsm_rp_elems(int64_t n) { return trsm_padded_size(n) * trsm_padded_size(n); }
    int64_t trsm_dinv_elems(int64_t n) { return trsm_padded_size(n) * dense::kSolveBlock; }

    void launch_trsm_right_upper_f64(double *b, int64_t ldb, int64_t rows, int64_t n, const double *r, int64_t ldr,
                                     bool r_padded, double *rp, double *dinv, const DenseWorkspace &ws,
                                     hipStream_t stream)
    {
        using namespace dense;
        if (rows <= 0 || n <= 0)
        {
            return;
        }
        const int64_t np = trsm_padded_size(n);
        const double *rr = r;
        int64_t ldp = ldr;
        if (!r_padded)
        {
            hipLaunchKernelGGL(pad_upper_kernel, dim3(static_cast<unsigned>((np * np + 255) / 256)), dim3(256), 0,
                               stream, r, ldr, n, np, rp);
            JSVD_HIP_CHECK(hipGetLastError());
            rr = rp;
            ldp = np;
        }
        hipLaunchKernelGGL(diag_inverse_kernel, dim3(static_cast<unsigned>(np / kSolveBlock)), dim3(64), 0, stream, rr,
                           ldp, dinv);
        JSVD_HIP_CHECK(hipGetLastError());
        const unsigned panels = static_cast<unsigned>((rows + kPanelRows - 1) / kPanelRows);
        trsm_schedule(
            n,
            [&](int64_t c0, int nblocks) {
                hipLaunchKernelGGL(trsm_panel_kernel, dim3(panels), dim3(256), 0, stream, b, ldb, rows, n, c0, nblocks,
                                   rr, ldp, static_cast<const double *>(dinv));
                JSVD_HIP_CHECK(hipGetLastError());
            },
            [&](int64_t c0, int64_t width, int64_t rest) {
                // B[:, c0 + width :] -= X[:, c0 : c0 + width] R[c0 : c0 + width, c0 + width :]
                launch_gemm_f64(false, false, rows, rest, width, -1.0, b + c0 * ldb, ldb, rr + c0 + (c0 + width) * ldp,
                                ldp, 1.0, b + (c0 + width) * ldb, ldb, b + (c0 + width) * ldb, ldb, 0, ws, stream);
            });
    }
} // namespace jsvd

namespace
{
    // Per-call scratch of the stand-alone entry points (stream-ordered allocation, freed after the work).
    struct CallScratch
    {
        hipStream_t stream;
        std::vector<void *> blocks;
        explicit CallScratch(hipStream_t s) : stream(s) {}
        template <typename T>
        T *get(int64_t count)
        {
            void *p = nullptr;
            JSVD_HIP_CHECK(hipMallocAsync(&p, sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 1)), stream));
            blocks.push_back(p);
            return static_cast<T *>(p);
        }
        jsvd::DenseWorkspace workspace()
        {
            jsvd::DenseWorkspace ws;
            ws.work_elems = jsvd::dense_work_elems();
            ws.work = get<double>(ws.work_elems);
            return ws;
        }
        ~CallScratch()
        {
            for (void *p : blocks)
            {
                (void)hipFreeAsync(p, stream);
            }
            (void)hipStreamSynchronize(stream);
        }
    };
} // namespace

extern "C" {

int jsvd_gemm_f64(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                  int64_t lda, const double *b, int64_t ldb, double beta, const double *c, int64_t ldc, double *d,
                  int64_t ldd, int32_t flags, void *stream)
{
    return jsvd::guarded([&] {
        using namespace jsvd::dense;
        const int known = kUpperOut | kMirror | kTriA | kTriB | kZeroLower | kRowMajorCD;
        if (m < 0 || n < 0 || k < 0 || (flags & ~known) != 0)
        {
            jsvd::throw_invalid("sizes must be non-negative and flags a combination of JSVD_GEMM_*.");
        }
        if (((flags & kMirror) && !(flags & kUpperOut)) || ((flags & kMirror) && (flags & kRowMajorCD)) ||
            ((flags & kUpperOut) && m != n))
        {
            jsvd::throw_invalid("JSVD_GEMM_MIRROR needs JSVD_GEMM_UPPER_OUT and column-major output; "
                                "JSVD_GEMM_UPPER_OUT needs m == n.");
        }
        const int64_t rows_a = trans_a ? k : m, rows_b = trans_b ? n : k;
        const int64_t ld_out = (flags & kRowMajorCD) ? n : m;
        if (m == 0 || n == 0)
        {
            return;
        }
        if (!a || !b || !d || (beta != 0.0 && !c) || lda < std::max<int64_t>(1, rows_a) ||
            ldb < std::max<int64_t>(1, rows_b) || ldd < ld_out || (beta != 0.0 && ldc < ld_out))
        {
            jsvd::throw_invalid("null operand or leading dimension too small.");
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        CallScratch scratch(st);
        const jsvd::DenseWorkspace ws = scratch.workspace();
        jsvd::launch_gemm_f64(trans_a != 0, trans_b != 0, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, d, ldd, flags,
                              ws, st);
    });
}

int jsvd_trsm_right_upper_f64(double *b, int64_t ldb, int64_t rows, int64_t n, const double *r, int64_t ldr,
                              void *stream)
{
    return jsvd::guarded([&] {
        if (rows < 0 || n < 0)
        {
            jsvd::throw_invalid("sizes must be non-negative.");
        }
        if (rows == 0 || n == 0)
        {
            return;
        }
        if (!b || !r || ldb < rows || ldr < n)
        {
            jsvd::throw_invalid("null operand or leading dimension too small.");
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        CallScratch scratch(st);
        const jsvd::DenseWorkspace ws = scratch.workspace();
        double *rp = scratch.get<double>(jsvd::trsm_rp_elems(n));
        double *dinv = scratch.get<double>(jsvd::trsm_dinv_elems(n));
        jsvd::launch_trsm_right_upper_f64(b, ldb, rows, n, r, ldr, false, rp, dinv, ws, st);
    });
}

} // extern "C"
