// pipeline.cpp -- ordered file-to-file SVD pipeline (include/pipeline.hpp).
//
// Shape: the calling thread reads cases (pinned *.mat dispatch or *.txt), a pool of solver lanes solves
// them, and one writer thread emits U, Sigma(1 x n), V per case strictly in input order.  Each lane owns
// a HIP stream and a solver plan reused while the shape repeats, so several cases run on the GPU at once
// (the reference's workers all share the legacy default stream and serialise).  The FIFO of pending
// results holds at most max_queued_results entries; the reader blocks when it is full.  With
// device_batch_cases > 0, runs of consecutive small cases go to a lane as one group, solved by one persistent
// batched launch, each case keeping its own place in the output order.
#include "pipeline.hpp"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <future>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <variant>
#include <vector>

#include "io.hpp"
#include "jsvd.h"

namespace jacobi::svd::pipeline
{
    namespace
    {
        struct Packet
        {
            std::size_t index = 0;
            int sweeps = 0;
            io::Matrix u;
            io::Matrix sigma;
            io::Matrix v;
        };

        MatrixFileFormat resolve_format(MatrixFileFormat requested, const std::filesystem::path &path)
        {
            if (requested != MatrixFileFormat::auto_detect)
            {
                return requested;
            }
            std::string ext = path.extension().string();
            std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return std::tolower(c); });
            if (ext == ".mat")
            {
                return MatrixFileFormat::mat;
            }
            if (ext == ".txt")
            {
                return MatrixFileFormat::txt;
            }
            throw std::invalid_argument("Cannot auto-detect file format for path: " + path.string());
        }

        void check_case(const io::Matrix &m, std::size_t index)
        {
            if (m.rows == 0 || m.columns == 0)
            {
                throw std::invalid_argument("Testcase[" + std::to_string(index) + "] has zero dimension.");
            }
            if (m.columns > std::numeric_limits<std::size_t>::max() / m.rows ||
                m.values.size() != m.rows * m.columns)
            {
                throw std::invalid_argument("Testcase[" + std::to_string(index) +
                                            "] payload size does not match rows*columns.");
            }
        }

        jsvd_config c_config(const JacobiSvdConfig &k)
        {
            jsvd_config c;
            jsvd_config_default(&c);
            c.epsilon = k.epsilon;
            c.max_sweeps = k.max_sweeps;
            c.threads_per_block = k.threads_per_block;
            c.layout_transpose_mode = static_cast<int32_t>(k.layout_transpose_mode);
            c.layout_transpose_min_columns = k.layout_transpose_min_columns;
            c.layout_transpose_min_elements = k.layout_transpose_min_elements;
            c.layout_transpose_auto_tune = 0;
            c.layout_transpose_benchmark_repetitions = k.layout_transpose_benchmark_repetitions;
            c.layout_transpose_benchmark_sweeps = k.layout_transpose_benchmark_sweeps;
            return c;
        }

        [[noreturn]] void raise(int status)
        {
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

        // A solver lane: one HIP stream (owned by its plan), a plan sized for the largest case it has seen, and
        // (for batched groups) a batch object.
        class SolverLane
        {
        public:
            explicit SolverLane(const JacobiSvdConfig &config) : cfg_(c_config(config)) {}
            ~SolverLane()
            {
                if (plan_)
                {
                    (void)jsvd_plan_destroy(plan_);
                }
                if (batch_)
                {
                    (void)jsvd_batch_destroy(batch_);
                }
            }
            SolverLane(const SolverLane &) = delete;
            SolverLane &operator=(const SolverLane &) = delete;

            Packet solve(const io::Matrix &a, std::size_t index)
            {
                check_case(a, index);
                if (a.rows < a.columns)
                {
                    throw std::invalid_argument("One-sided Jacobi SVD requires rows >= columns in this implementation.");
                }
                const auto m = static_cast<int64_t>(a.rows);
                const auto n = static_cast<int64_t>(a.columns);
                // One plan per lane, sized for the largest case seen so far and re-targeted at each case
                // (jsvd_plan_reshape: no allocation, results bit-equal to a plan of the exact shape).
                if (!plan_ || m > m_ || n > n_)
                {
                    if (plan_)
                    {
                        (void)jsvd_plan_destroy(plan_);
                        plan_ = nullptr;
                    }
                    const int64_t cm = std::max(m, m_), cn = std::max(n, n_); // cm >= m >= n and cm >= m_ >= n_
                    const int st = jsvd_plan_create(cm, cn, &cfg_, nullptr, &plan_);
                    if (st != JSVD_OK)
                    {
                        plan_ = nullptr;
                        raise(st);
                    }
                    m_ = cm;
                    n_ = cn;
                }
                {
                    const int st = jsvd_plan_reshape(plan_, m, n, cfg_.epsilon);
                    if (st != JSVD_OK)
                    {
                        raise(st);
                    }
                }
                Packet p;
                p.index = index;
                p.u = io::Matrix{a.rows, a.columns, std::vector<double>(a.rows * a.columns)};
                p.sigma = io::Matrix{1, a.columns, std::vector<double>(a.columns)};
                p.v = io::Matrix{a.columns, a.columns, std::vector<double>(a.columns * a.columns)};
                int32_t sweeps = 0;
                const int st = jsvd_plan_solve_host(plan_, a.values.data(), p.u.values.data(), p.sigma.values.data(),
                                                    p.v.values.data(), &sweeps);
                if (st != JSVD_OK)
                {
                    raise(st);
                }
                p.sweeps = sweeps;
                return p;
            }

            // Small cases (batchable(), validated by the caller) in one persistent batched launch; one packet per
            // case, in the given order.
            std::vector<Packet> solve_batch(const std::vector<std::shared_ptr<const io::Matrix>> &cases,
                                            const std::vector<std::size_t> &indices)
            {
                if (!batch_)
                {
                    const int st = jsvd_batch_create(kBatchRows, kBatchColumns, nullptr, &batch_);
                    if (st != JSVD_OK)
                    {
                        batch_ = nullptr;
                        raise(st);
                    }
                }
                const std::size_t k = cases.size();
                std::vector<jsvd_batch_item> items(k);
                int64_t na = 0, nv = 0, ns = 0;
                for (std::size_t i = 0; i < k; ++i)
                {
                    const io::Matrix &a = *cases[i];
                    jsvd_batch_item &it = items[i];
                    it.a_offset = na;
                    it.u_offset = na;
                    it.sigma_offset = ns;
                    it.v_offset = nv;
                    it.rows = static_cast<int32_t>(a.rows);
                    it.columns = static_cast<int32_t>(a.columns);
                    it.precision = JSVD_PRECISION_F64;
                    it.max_sweeps = cfg_.max_sweeps;
                    it.tolerance = cfg_.epsilon;
                    na += static_cast<int64_t>(a.rows * a.columns);
                    ns += static_cast<int64_t>(a.columns);
                    nv += static_cast<int64_t>(a.columns * a.columns);
                }
                std::vector<double> ha(static_cast<std::size_t>(na)), hu(ha.size()), hs(static_cast<std::size_t>(ns)),
                    hv(static_cast<std::size_t>(nv));
                for (std::size_t i = 0; i < k; ++i)
                {
                    std::copy(cases[i]->values.begin(), cases[i]->values.end(), ha.begin() + items[i].a_offset);
                }
                std::vector<int32_t> sweeps(k);
                const int st = jsvd_batch_solve_host(batch_, items.data(), static_cast<int64_t>(k), ha.data(), na,
                                                     hu.data(), na, hs.data(), ns, hv.data(), nv, sweeps.data());
                if (st != JSVD_OK)
                {
                    raise(st);
                }
                std::vector<Packet> out(k);
                for (std::size_t i = 0; i < k; ++i)
                {
                    const std::size_t m = cases[i]->rows, n = cases[i]->columns;
                    if (sweeps[i] < 0)
                    {
                        throw std::runtime_error("Testcase[" + std::to_string(indices[i]) +
                                                 "] was rejected by the batched solver.");
                    }
                    Packet &p = out[i];
                    p.index = indices[i];
                    p.sweeps = sweeps[i];
                    const auto u0 = hu.begin() + items[i].u_offset;
                    const auto s0 = hs.begin() + items[i].sigma_offset;
                    const auto v0 = hv.begin() + items[i].v_offset;
                    p.u = io::Matrix{m, n, std::vector<double>(u0, u0 + static_cast<std::ptrdiff_t>(m * n))};
                    p.sigma = io::Matrix{1, n, std::vector<double>(s0, s0 + static_cast<std::ptrdiff_t>(n))};
                    p.v = io::Matrix{n, n, std::vector<double>(v0, v0 + static_cast<std::ptrdiff_t>(n * n))};
                }
                return out;
            }

            static constexpr int64_t kBatchRows = 256;
            static constexpr int64_t kBatchColumns = 128;

        private:
            jsvd_config cfg_;
            jsvd_plan *plan_ = nullptr;
            jsvd_batch *batch_ = nullptr;
            int64_t m_ = 0;
            int64_t n_ = 0;
        };

        // Cases the batched launch takes: well-formed, rows >= columns >= JSVD_BATCH_MIN_COLUMNS, within the batch
        // limits.  Narrower cases go through the lane's plan.
        bool batchable(std::size_t rows, std::size_t columns)
        {
            return columns >= static_cast<std::size_t>(JSVD_BATCH_MIN_COLUMNS) && rows >= columns &&
                   rows <= static_cast<std::size_t>(SolverLane::kBatchRows) &&
                   columns <= static_cast<std::size_t>(SolverLane::kBatchColumns);
        }

        // Fixed pool of solver lanes; each worker thread owns one lane for its whole life.
        class LanePool
        {
        public:
            using Job = std::function<void(SolverLane &)>;

            LanePool(std::size_t lanes, const JacobiSvdConfig &config)
            {
                for (std::size_t i = 0; i < lanes; ++i)
                {
                    threads_.emplace_back([this, config] { worker(config); });
                }
            }

            ~LanePool()
            {
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    stop_ = true;
                }
                cv_.notify_all();
                for (auto &t : threads_)
                {
                    t.join();
                }
            }

            std::future<Packet> submit(std::function<Packet(SolverLane &)> fn)
            {
                auto task = std::make_shared<std::packaged_task<Packet(SolverLane &)>>(std::move(fn));
                std::future<Packet> fut = task->get_future();
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    jobs_.emplace_back([task](SolverLane &lane) { (*task)(lane); });
                }
                cv_.notify_one();
                return fut;
            }

            // A job that reports through its own promises (a batched group of cases).
            void submit_job(Job job)
            {
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    jobs_.emplace_back(std::move(job));
                }
                cv_.notify_one();
            }

        private:
            void worker(JacobiSvdConfig config)
            {
                SolverLane lane(config);
                for (;;)
                {
                    Job job;
                    {
                        std::unique_lock<std::mutex> lock(mu_);
                        cv_.wait(lock, [this] { return stop_ || !jobs_.empty(); });
                        if (jobs_.empty())
                        {
                            return;
                        }
                        job = std::move(jobs_.front());
                        jobs_.pop_front();
                    }
                    job(lane);
                }
            }

            std::mutex mu_;
            std::condition_variable cv_;
            std::deque<Job> jobs_;
            bool stop_ = false;
            std::vector<std::thread> threads_;
        };

        // Single writer thread consuming a bounded FIFO of result futures in submission order.
        class OrderedWriter
        {
        public:
            OrderedWriter(const std::filesystem::path &path, MatrixFileFormat format, std::size_t capacity)
                : capacity_(std::max<std::size_t>(capacity, 1))
            {
                const std::filesystem::path parent = path.parent_path();
                if (!parent.empty())
                {
                    std::filesystem::create_directories(parent);
                }
                if (format == MatrixFileFormat::mat)
                {
                    stream_.emplace<io::MatOutputStream>(path);
                }
                else
                {
                    stream_.emplace<io::TxtOutputStream>(path);
                }
                thread_ = std::thread([this] { consume(); });
            }

            ~OrderedWriter()
            {
                if (thread_.joinable())
                {
                    abort();
                }
            }

            void submit(std::future<Packet> f)
            {
                std::unique_lock<std::mutex> lock(mu_);
                not_full_.wait(lock, [this] { return error_ || queue_.size() < capacity_; });
                if (error_)
                {
                    std::rethrow_exception(error_);
                }
                queue_.push_back(std::move(f));
                not_empty_.notify_one();
            }

            void close_success(std::size_t expected)
            {
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    closing_ = true;
                }
                not_empty_.notify_all();
                thread_.join();
                if (error_)
                {
                    std::rethrow_exception(error_);
                }
                if (written_ != expected)
                {
                    throw std::runtime_error("Output stage wrote " + std::to_string(written_) + " of " +
                                             std::to_string(expected) + " test cases.");
                }
            }

            void abort() noexcept
            {
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    closing_ = true;
                    aborted_ = true;
                }
                not_empty_.notify_all();
                not_full_.notify_all();
                if (thread_.joinable())
                {
                    thread_.join();
                }
            }

            std::size_t total_sweeps() const { return sweeps_; }

        private:
            void consume()
            {
                for (;;)
                {
                    std::future<Packet> f;
                    {
                        std::unique_lock<std::mutex> lock(mu_);
                        not_empty_.wait(lock, [this] { return closing_ || !queue_.empty(); });
                        if (aborted_ || (queue_.empty() && closing_))
                        {
                            break;
                        }
                        f = std::move(queue_.front());
                        queue_.pop_front();
                        not_full_.notify_one();
                    }
                    try
                    {
                        Packet p = f.get();
                        std::visit(
                            [&](auto &s) {
                                if constexpr (!std::is_same_v<std::decay_t<decltype(s)>, std::monostate>)
                                {
                                    s.write_one(p.u);
                                    s.write_one(p.sigma);
                                    s.write_one(p.v);
                                }
                            },
                            stream_);
                        sweeps_ += static_cast<std::size_t>(p.sweeps);
                        ++written_;
                    }
                    catch (...)
                    {
                        std::lock_guard<std::mutex> lock(mu_);
                        error_ = std::current_exception();
                        not_full_.notify_all();
                        break;
                    }
                }
                try
                {
                    std::visit(
                        [&](auto &s) {
                            if constexpr (!std::is_same_v<std::decay_t<decltype(s)>, std::monostate>)
                            {
                                s.flush();
                            }
                        },
                        stream_);
                }
                catch (...)
                {
                    std::lock_guard<std::mutex> lock(mu_);
                    if (!error_)
                    {
                        error_ = std::current_exception();
                    }
                }
            }

            std::size_t capacity_;
            std::variant<std::monostate, io::MatOutputStream, io::TxtOutputStream> stream_;
            std::mutex mu_;
            std::condition_variable not_full_;
            std::condition_variable not_empty_;
            std::deque<std::future<Packet>> queue_;
            bool closing_ = false;
            bool aborted_ = false;
            std::exception_ptr error_;
            std::size_t written_ = 0;
            std::size_t sweeps_ = 0;
            std::thread thread_;
        };

        std::size_t lane_count()
        {
            const unsigned hw = std::thread::hardware_concurrency();
            return std::clamp<std::size_t>(hw == 0 ? 4 : hw, 1, 8);
        }
    } // namespace

    JacobiSvdPipeline::JacobiSvdPipeline(PipelineConfig config) : config_(std::move(config)) {}

    PipelineReport JacobiSvdPipeline::run() const
    {
        if (config_.input_path.empty())
        {
            throw std::invalid_argument("Pipeline input path is empty.");
        }
        if (config_.output_path.empty())
        {
            throw std::invalid_argument("Pipeline output path is empty.");
        }
        const MatrixFileFormat in_format = resolve_format(config_.input_format, config_.input_path);
        const MatrixFileFormat out_format = resolve_format(config_.output_format, config_.output_path);

        JacobiSvdConfig kernel = config_.kernel_config;
        LayoutTransposeAutoTuneReport tuning{};
        if (kernel.layout_transpose_auto_tune && kernel.layout_transpose_mode == LayoutTransposeMode::auto_select)
        {
            tuning = auto_tune_layout_transpose_threshold(kernel);
            kernel.layout_transpose_min_columns = tuning.recommended_min_columns;
            kernel.layout_transpose_min_elements = tuning.recommended_min_elements;
            kernel.layout_transpose_auto_tune = false;
        }

        OrderedWriter writer(config_.output_path, out_format, config_.max_queued_results);
        std::size_t submitted = 0;
        try
        {
            LanePool pool(lane_count(), kernel);
            // Groups of consecutive small cases for one batched launch each (device_batch_cases > 0 and a valid
            // kernel config; with an invalid one every case goes to a lane, which reports the error in order).
            const jsvd_config ckernel = c_config(kernel);
            const std::size_t group_max =
                (config_.device_batch_cases > 0 && jsvd_config_validate(&ckernel) == JSVD_OK)
                    ? config_.device_batch_cases
                    : 0;
            std::vector<std::shared_ptr<const io::Matrix>> group;
            std::vector<std::size_t> group_idx;
            // The group's job goes to the pool before its results are queued on the writer, so a full writer
            // queue always waits on work that is already running.
            auto flush = [&] {
                if (group.empty())
                {
                    return;
                }
                auto promises = std::make_shared<std::vector<std::promise<Packet>>>(group.size());
                std::vector<std::future<Packet>> futures;
                for (auto &pr : *promises)
                {
                    futures.push_back(pr.get_future());
                }
                auto cases = std::make_shared<std::vector<std::shared_ptr<const io::Matrix>>>(std::move(group));
                auto idx = std::make_shared<std::vector<std::size_t>>(std::move(group_idx));
                group = {};
                group_idx = {};
                pool.submit_job([cases, idx, promises](SolverLane &lane) {
                    try
                    {
                        std::vector<Packet> out = lane.solve_batch(*cases, *idx);
                        for (std::size_t i = 0; i < out.size(); ++i)
                        {
                            (*promises)[i].set_value(std::move(out[i]));
                        }
                    }
                    catch (...)
                    {
                        const std::exception_ptr e = std::current_exception();
                        for (auto &pr : *promises)
                        {
                            try
                            {
                                pr.set_exception(e);
                            }
                            catch (...)
                            {
                                // already satisfied
                            }
                        }
                    }
                });
                for (auto &f : futures)
                {
                    writer.submit(std::move(f));
                }
            };
            auto add_to_group = [&](std::shared_ptr<const io::Matrix> a, std::size_t index) {
                group.push_back(std::move(a));
                group_idx.push_back(index);
                if (group.size() >= group_max)
                {
                    flush();
                }
            };
            if (in_format == MatrixFileFormat::mat)
            {
                io::MatDispatchReader reader(config_.input_path);
                io::MatDispatchTask task;
                while (reader.read_next(task))
                {
                    auto shared = std::make_shared<io::MatDispatchTask>(std::move(task));
                    task = io::MatDispatchTask{};
                    const std::size_t index = shared->sequence_index;
                    if (group_max > 0 && batchable(shared->rows, shared->columns))
                    {
                        std::shared_ptr<const io::Matrix> a;
                        try
                        {
                            a = std::make_shared<const io::Matrix>(io::decode_dispatch_task_matrix(*shared));
                        }
                        catch (...)
                        {
                            a = nullptr; // the lane path decodes it again and reports the error in order
                        }
                        if (a && a->values.size() == a->rows * a->columns)
                        {
                            add_to_group(std::move(a), index);
                            ++submitted;
                            continue;
                        }
                    }
                    flush();
                    writer.submit(pool.submit([shared, index](SolverLane &lane) {
                        const io::Matrix a = io::decode_dispatch_task_matrix(*shared);
                        return lane.solve(a, index);
                    }));
                    ++submitted;
                }
            }
            else
            {
                io::TxtInputStream reader(config_.input_path);
                io::Matrix a;
                while (reader.read_one(a))
                {
                    auto shared = std::make_shared<io::Matrix>(std::move(a));
                    a = io::Matrix{};
                    const std::size_t index = submitted;
                    if (group_max > 0 && batchable(shared->rows, shared->columns) &&
                        shared->values.size() == shared->rows * shared->columns)
                    {
                        add_to_group(std::move(shared), index);
                        ++submitted;
                        continue;
                    }
                    flush();
                    writer.submit(pool.submit([shared, index](SolverLane &lane) { return lane.solve(*shared, index); }));
                    ++submitted;
                }
            }
            flush();
            writer.close_success(submitted);
        }
        catch (...)
        {
            writer.abort();
            throw;
        }

        PipelineReport report;
        report.testcase_count = submitted;
        report.emitted_matrix_count = submitted * 3;
        report.total_sweeps = writer.total_sweeps();
        report.layout_transpose_mode = kernel.layout_transpose_mode;
        report.layout_transpose_min_columns = kernel.layout_transpose_min_columns;
        report.layout_transpose_min_elements = kernel.layout_transpose_min_elements;
        report.layout_transpose_auto_tuned = tuning.executed;
        report.layout_transpose_estimated_best_speedup = tuning.estimated_best_speedup;
        return report;
    }

    PipelineReport run_pipeline(const PipelineConfig &config) { return JacobiSvdPipeline(config).run(); }
} // namespace jacobi::svd::pipeline
