// mat_stream.hip -- streamed solve of a *.mat file on one GPU.
//
// The reference reads each record with an ifstream into a pinned buffer, byte-swaps it on a CPU worker and
// copies it to the device synchronously inside the solver (src/io.cu:1101-1168, src/pipeline.cpp:711-739).
// Here the file is memory-mapped and a loader thread moves record k+1 while record k is being solved:
//
//   mmap (page cache) --memcpy--> pinned staging[k % 2] --hipMemcpyAsync, copy stream--> raw HBM[k % 2]
//   --decode kernel (big-endian u64 -> double), copy stream--> A[k % 2]   ... event ready[k % 2]
//
// The solver thread waits for the slot, solves it through a plan on the compute stream (one plan per path,
// re-targeted at each record's shape) into one of two result slots, and hands the slot to an output stream that encodes U, sigma (1 x n)
// and V big-endian on the device and copies them to pinned memory; a writer thread appends the three records
// to the output mapping in input order (the reference pipeline's output packets) while the next record is
// solved.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "common.hpp"
#include "mapped_file.hpp"

namespace jsvd
{
    namespace
    {
        constexpr std::size_t kHeaderBytes = 2 * sizeof(std::uint64_t);
        constexpr std::size_t kValueBytes = sizeof(std::uint64_t);

        struct MatRecord
        {
            std::size_t payload_offset = 0;
            std::size_t rows = 0;
            std::size_t columns = 0;
            [[nodiscard]] std::size_t count() const { return rows * columns; }
        };

        std::uint64_t load_be_u64(const std::byte *p)
        {
            std::uint64_t raw;
            std::memcpy(&raw, p, sizeof(raw));
            return mapped::be64(raw);
        }

        // Record index of a mapped *.mat file; the same truncation errors as the record reader
        // (src/io.cu:1180-1244).
        std::vector<MatRecord> index_records(std::span<const std::byte> bytes)
        {
            std::vector<MatRecord> out;
            std::size_t cursor = 0;
            while (cursor < bytes.size())
            {
                if (bytes.size() - cursor < kHeaderBytes)
                {
                    throw std::runtime_error("Truncated *.mat header.");
                }
                const std::uint64_t rows = load_be_u64(bytes.data() + cursor);
                const std::uint64_t cols = load_be_u64(bytes.data() + cursor + kValueBytes);
                cursor += kHeaderBytes;
                const std::size_t payload = mapped::mul_checked(mapped::mul_checked(rows, cols), kValueBytes);
                if (bytes.size() - cursor < payload)
                {
                    throw std::runtime_error("Truncated *.mat payload.");
                }
                out.push_back(MatRecord{cursor, static_cast<std::size_t>(rows), static_cast<std::size_t>(cols)});
                cursor += payload;
            }
            return out;
        }

        __global__ void decode_be_kernel(const unsigned long long *__restrict__ raw, double *__restrict__ out,
                                         int64_t count)
        {
            for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count;
                 i += int64_t{gridDim.x} * blockDim.x)
            {
                out[i] = __longlong_as_double(static_cast<long long>(__builtin_bswap64(raw[i])));
            }
        }

        __global__ void encode_be_kernel(const double *__restrict__ in, unsigned long long *__restrict__ out,
                                         int64_t count)
        {
            for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count;
                 i += int64_t{gridDim.x} * blockDim.x)
            {
                out[i] = __builtin_bswap64(static_cast<unsigned long long>(__double_as_longlong(in[i])));
            }
        }

        unsigned grid_for(int64_t count)
        {
            return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
        }

        struct PinnedBytes
        {
            void *p = nullptr;
            explicit PinnedBytes(std::size_t bytes)
            {
                JSVD_HIP_CHECK(hipHostMalloc(&p, std::max<std::size_t>(bytes, 64), hipHostMallocDefault));
            }
            ~PinnedBytes() { (void)hipHostFree(p); }
            PinnedBytes(const PinnedBytes &) = delete;
            PinnedBytes &operator=(const PinnedBytes &) = delete;
        };

        struct DeviceBytes
        {
            void *p = nullptr;
            explicit DeviceBytes(std::size_t bytes)
            {
                JSVD_HIP_CHECK(hipMalloc(&p, std::max<std::size_t>(bytes, 256)));
            }
            ~DeviceBytes() { (void)hipFree(p); }
            DeviceBytes(const DeviceBytes &) = delete;
            DeviceBytes &operator=(const DeviceBytes &) = delete;
        };

        struct Stream
        {
            hipStream_t s = nullptr;
            Stream() { JSVD_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
            ~Stream() { (void)hipStreamDestroy(s); }
            Stream(const Stream &) = delete;
            Stream &operator=(const Stream &) = delete;
        };

        struct Event
        {
            hipEvent_t e = nullptr;
            Event() { JSVD_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
            ~Event() { (void)hipEventDestroy(e); }
            Event(const Event &) = delete;
            Event &operator=(const Event &) = delete;
        };

        struct PlanDeleter
        {
            void operator()(jsvd_plan *p) const { (void)jsvd_plan_destroy(p); }
        };
        using PlanPtr = std::unique_ptr<jsvd_plan, PlanDeleter>;

        // Rethrows a failed C-ABI call of this library with its status kept.
        void check_status(int status)
        {
            if (status != JSVD_OK)
            {
                throw Error(status, jsvd_last_error());
            }
        }

        double seconds_since(std::chrono::steady_clock::time_point t0)
        {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }

        class MatStreamSolver
        {
        public:
            MatStreamSolver(const char *input, const jsvd_config &cfg, const jsvd_mixed_config &mixed)
                : file_(input), cfg_(cfg), mixed_(mixed)
            {
                records_ = index_records(file_.bytes());
                for (const MatRecord &r : records_)
                {
                    validate_shape(static_cast<int64_t>(r.rows), static_cast<int64_t>(r.columns));
                    max_count_ = std::max(max_count_, r.count());
                    max_out_ = std::max(max_out_, r.count() + r.columns + r.columns * r.columns);
                }
            }

            void run(const char *output, jsvd_mat_stream_report &rep)
            {
                const auto t0 = std::chrono::steady_clock::now();
                std::unique_ptr<mapped::MappedAppendOutput> out;
                if (output != nullptr)
                {
                    out = std::make_unique<mapped::MappedAppendOutput>(output);
                }
                rep = jsvd_mat_stream_report{};
                if (records_.empty())
                {
                    return;
                }
                if (out)
                {
                    // every output record's size is known from the input's record index: map the whole output
                    // once (a growing mapping flushes its prefix to the disk at every doubling)
                    std::size_t total = 0;
                    for (const MatRecord &r : records_)
                    {
                        const std::size_t values = r.rows * r.columns + r.columns + r.columns * r.columns;
                        total = mapped::add_checked(total, 3 * 2 * sizeof(std::uint64_t) + values * kValueBytes);
                    }
                    out->reserve(total);
                }
                const std::size_t in_bytes = max_count_ * kValueBytes;
                PinnedBytes staging0(in_bytes), staging1(in_bytes);
                DeviceBytes raw0(in_bytes), raw1(in_bytes), a0(in_bytes), a1(in_bytes);
                void *staging[2] = {staging0.p, staging1.p};
                void *raw[2] = {raw0.p, raw1.p};
                double *a[2] = {static_cast<double *>(a0.p), static_cast<double *>(a1.p)};
                Stream copy, compute, outs;
                Event ready[2];
                // Result slots, two alternating: U / sigma / V of record k in slot k % 2 on the device, encoded
                // big-endian and copied to pinned memory on the output stream, appended by the writer thread
                // while record k + 1 is solved.
                const std::size_t ub = max_count_ * kValueBytes, ob = max_out_ * kValueBytes;
                DeviceBytes d_u0(ub), d_u1(ub), d_s0(ob), d_s1(ob), d_v0(ob), d_v1(ob), d_enc0(ob), d_enc1(ob);
                std::unique_ptr<PinnedBytes> result0, result1;
                if (out)
                {
                    result0 = std::make_unique<PinnedBytes>(ob);
                    result1 = std::make_unique<PinnedBytes>(ob);
                }
                double *d_u[2] = {static_cast<double *>(d_u0.p), static_cast<double *>(d_u1.p)};
                double *d_s[2] = {static_cast<double *>(d_s0.p), static_cast<double *>(d_s1.p)};
                double *d_v[2] = {static_cast<double *>(d_v0.p), static_cast<double *>(d_v1.p)};
                unsigned long long *d_enc[2] = {static_cast<unsigned long long *>(d_enc0.p),
                                                static_cast<unsigned long long *>(d_enc1.p)};
                void *result[2] = {result0 ? result0->p : nullptr, result1 ? result1->p : nullptr};
                Event solved[2], copied[2];

                // Slot hand-over between the loader and this thread.
                std::mutex mu;
                std::condition_variable cv;
                long long loaded[2] = {-1, -1}; // record held by a slot, -1 = free
                bool stop = false;
                std::exception_ptr loader_error;
                double load_seconds = 0.0;

                std::thread loader([&] {
                    try
                    {
                        for (std::size_t k = 0; k < records_.size(); ++k)
                        {
                            const int slot = static_cast<int>(k & 1);
                            {
                                std::unique_lock<std::mutex> lock(mu);
                                cv.wait(lock, [&] { return stop || loaded[slot] < 0; });
                                if (stop)
                                {
                                    return;
                                }
                            }
                            const auto tl = std::chrono::steady_clock::now();
                            const MatRecord &r = records_[k];
                            const std::size_t bytes = r.count() * kValueBytes;
                            std::memcpy(staging[slot], file_.bytes().data() + r.payload_offset, bytes);
                            JSVD_HIP_CHECK(hipMemcpyAsync(raw[slot], staging[slot], bytes, hipMemcpyHostToDevice,
                                                          copy.s));
                            hipLaunchKernelGGL(decode_be_kernel, dim3(grid_for(r.count())), dim3(256), 0, copy.s,
                                               static_cast<const unsigned long long *>(raw[slot]), a[slot],
                                               static_cast<int64_t>(r.count()));
                            JSVD_HIP_CHECK(hipGetLastError());
                            JSVD_HIP_CHECK(hipEventRecord(ready[slot].e, copy.s));
                            // the staging buffer is reused two records later: wait for its copy here, off the
                            // solver's path
                            JSVD_HIP_CHECK(hipEventSynchronize(ready[slot].e));
                            load_seconds += seconds_since(tl);
                            {
                                std::lock_guard<std::mutex> lock(mu);
                                loaded[slot] = static_cast<long long>(k);
                            }
                            cv.notify_all();
                        }
                    }
                    catch (...)
                    {
                        std::lock_guard<std::mutex> lock(mu);
                        loader_error = std::current_exception();
                        stop = true;
                        cv.notify_all();
                    }
                });

                struct Joiner
                {
                    std::thread &t;
                    std::mutex &mu;
                    std::condition_variable &cv;
                    bool &stop;
                    ~Joiner()
                    {
                        {
                            std::lock_guard<std::mutex> lock(mu);
                            stop = true;
                        }
                        cv.notify_all();
                        t.join();
                    }
                } joiner{loader, mu, cv, stop};

                // Writer thread: appends the results of each record in input order once its copy to pinned
                // memory is done, then frees the slot.
                std::deque<std::size_t> to_write;
                bool out_busy[2] = {false, false};
                std::size_t written = 0;
                bool writer_stop = false;
                std::exception_ptr writer_error;
                double output_seconds = 0.0;
                std::thread writer([&] {
                    try
                    {
                        for (;;)
                        {
                            std::size_t k = 0;
                            {
                                std::unique_lock<std::mutex> lock(mu);
                                cv.wait(lock, [&] { return writer_stop || !to_write.empty(); });
                                if (to_write.empty())
                                {
                                    return; // stopped and drained
                                }
                                k = to_write.front();
                                to_write.pop_front();
                            }
                            const int j = static_cast<int>(k & 1);
                            const auto tw = std::chrono::steady_clock::now();
                            JSVD_HIP_CHECK(hipEventSynchronize(copied[j].e));
                            const MatRecord &r = records_[k];
                            const std::size_t nu = r.count(), ns = r.columns;
                            const auto *bytes = static_cast<const std::byte *>(result[j]);
                            append_record(*out, r.rows, r.columns, bytes);
                            append_record(*out, 1, r.columns, bytes + nu * kValueBytes);
                            append_record(*out, r.columns, r.columns, bytes + (nu + ns) * kValueBytes);
                            {
                                std::lock_guard<std::mutex> lock(mu);
                                output_seconds += seconds_since(tw);
                                out_busy[j] = false;
                                ++written;
                            }
                            cv.notify_all();
                        }
                    }
                    catch (...)
                    {
                        std::lock_guard<std::mutex> lock(mu);
                        writer_error = std::current_exception();
                        cv.notify_all();
                    }
                });
                struct WriterJoiner
                {
                    std::thread &t;
                    std::mutex &mu;
                    std::condition_variable &cv;
                    bool &stop;
                    ~WriterJoiner()
                    {
                        {
                            std::lock_guard<std::mutex> lock(mu);
                            stop = true;
                        }
                        cv.notify_all();
                        t.join();
                    }
                } writer_joiner{writer, mu, cv, writer_stop};

                struct PlanSlot
                {
                    PlanPtr plan;
                    int64_t m = 0;
                    int64_t n = 0;
                };
                PlanSlot plain_plan, path_plan;
                for (std::size_t k = 0; k < records_.size(); ++k)
                {
                    const int slot = static_cast<int>(k & 1);
                    {
                        // the input slot holds record k, and the result slot's previous record (k - 2) is written
                        std::unique_lock<std::mutex> lock(mu);
                        cv.wait(lock, [&] {
                            return loader_error || writer_error ||
                                   (loaded[slot] == static_cast<long long>(k) && !out_busy[slot]);
                        });
                        if (loader_error)
                        {
                            std::rethrow_exception(loader_error);
                        }
                        if (writer_error)
                        {
                            std::rethrow_exception(writer_error);
                        }
                    }
                    const MatRecord &r = records_[k];
                    const int64_t m = static_cast<int64_t>(r.rows), n = static_cast<int64_t>(r.columns);
                    const auto ts = std::chrono::steady_clock::now();
                    // The block paths need two 32-column blocks; narrower records take the plain path.  One plan
                    // per path, sized for the largest record it has met and re-targeted at each record
                    // (jsvd_plan_reshape: no allocation per shape, results equal to a plan of the exact shape).
                    const bool narrow = n < 33 || mixed_.path == JSVD_PATH_PLAIN_F64;
                    PlanSlot &ps = narrow ? plain_plan : path_plan;
                    if (!ps.plan || m > ps.m || n > ps.n)
                    {
                        jsvd_mixed_config mx = mixed_;
                        if (narrow)
                        {
                            mx.path = JSVD_PATH_PLAIN_F64;
                        }
                        const int64_t cm = std::max(m, ps.m), cn = std::max(n, ps.n); // cm >= cn
                        ps.plan.reset();
                        jsvd_plan *p = nullptr;
                        check_status(jsvd_plan_create_ex(cm, cn, &cfg_, &mx, compute.s, &p));
                        ps.plan.reset(p);
                        ps.m = cm;
                        ps.n = cn;
                    }
                    check_status(jsvd_plan_reshape(ps.plan.get(), m, n, cfg_.epsilon));
                    PlanPtr &plan = ps.plan;
                    JSVD_HIP_CHECK(hipStreamWaitEvent(compute.s, ready[slot].e, 0));
                    double *u = d_u[slot];
                    double *s = d_s[slot];
                    double *v = d_v[slot];
                    int32_t sweeps = 0;
                    check_status(jsvd_plan_solve_device(plan.get(), a[slot], u, s, v, &sweeps));
                    jsvd_solve_info info{};
                    check_status(jsvd_plan_last_info(plan.get(), &info));
                    rep.seconds_solve += seconds_since(ts);
                    {
                        std::lock_guard<std::mutex> lock(mu);
                        loaded[slot] = -1; // A[slot] is no longer read: the loader may refill it
                    }
                    cv.notify_all();

                    if (out)
                    {
                        // encode and copy back on the output stream after the solve; the writer appends
                        const std::size_t nu = r.count(), ns = r.columns, nv = r.columns * r.columns;
                        unsigned long long *enc = d_enc[slot];
                        JSVD_HIP_CHECK(hipEventRecord(solved[slot].e, compute.s));
                        JSVD_HIP_CHECK(hipStreamWaitEvent(outs.s, solved[slot].e, 0));
                        hipLaunchKernelGGL(encode_be_kernel, dim3(grid_for(nu)), dim3(256), 0, outs.s, u, enc,
                                           static_cast<int64_t>(nu));
                        hipLaunchKernelGGL(encode_be_kernel, dim3(grid_for(ns)), dim3(256), 0, outs.s, s,
                                           enc + nu, static_cast<int64_t>(ns));
                        hipLaunchKernelGGL(encode_be_kernel, dim3(grid_for(nv)), dim3(256), 0, outs.s, v,
                                           enc + nu + ns, static_cast<int64_t>(nv));
                        JSVD_HIP_CHECK(hipGetLastError());
                        JSVD_HIP_CHECK(hipMemcpyAsync(result[slot], enc, (nu + ns + nv) * kValueBytes,
                                                      hipMemcpyDeviceToHost, outs.s));
                        JSVD_HIP_CHECK(hipEventRecord(copied[slot].e, outs.s));
                        {
                            std::lock_guard<std::mutex> lock(mu);
                            out_busy[slot] = true;
                            to_write.push_back(k);
                        }
                        cv.notify_all();
                    }
                    rep.records += 1;
                    rep.payload_bytes += static_cast<int64_t>(r.count() * kValueBytes);
                    rep.sweeps_fp32_last = info.sweeps_fp32;
                    rep.sweeps_fp64_last = info.sweeps_fp64;
                    rep.total_sweeps += info.sweeps_fp32 + info.sweeps_fp64;
                }
                if (out)
                {
                    std::unique_lock<std::mutex> lock(mu);
                    cv.wait(lock, [&] { return writer_error || written == records_.size(); });
                    if (writer_error)
                    {
                        std::rethrow_exception(writer_error);
                    }
                }
                if (out)
                {
                    out->flush();
                }
                {
                    std::lock_guard<std::mutex> lock(mu);
                    rep.seconds_load = load_seconds;
                    rep.seconds_output = output_seconds;
                }
                rep.seconds_total = seconds_since(t0);
            }

        private:
            static void append_record(mapped::MappedAppendOutput &out, std::size_t rows, std::size_t cols,
                                      const std::byte *payload_be)
            {
                const std::uint64_t header[2] = {mapped::be64(rows), mapped::be64(cols)};
                out.append(header, sizeof(header));
                out.append(payload_be, rows * cols * kValueBytes);
            }

            mapped::MappedInput file_;
            jsvd_config cfg_;
            jsvd_mixed_config mixed_;
            std::vector<MatRecord> records_;
            std::size_t max_count_ = 0;
            std::size_t max_out_ = 0;
        };
    } // namespace
} // namespace jsvd

extern "C" int jsvd_mat_stream_solve(const char *input_path, const char *output_path, const jsvd_config *config,
                                     const jsvd_mixed_config *mixed, jsvd_mat_stream_report *report)
{
    return jsvd::guarded([&] {
        if (input_path == nullptr)
        {
            jsvd::throw_invalid("input_path must not be null.");
        }
        jsvd_config cfg;
        if (config)
        {
            cfg = *config;
        }
        else
        {
            jsvd_config_default(&cfg);
        }
        jsvd::validate_config(cfg);
        jsvd_mixed_config mx;
        jsvd_mixed_config_default(&mx);
        if (mixed)
        {
            mx = *mixed;
        }
        jsvd_mat_stream_report rep{};
        jsvd::MatStreamSolver solver(input_path, cfg, mx);
        solver.run(output_path, rep);
        if (report)
        {
            *report = rep;
        }
    });
}
