// host_driver.cpp -- scenarios for the host sanitizer builds (test infrastructure only): the request loop
// (csrc/server.cpp), the ordered pipeline (csrc/pipeline.cpp) and the file I/O (csrc/io.cpp, mapped_file.hpp)
// run unchanged over the host stand-in of the device (device_stub.cpp), with several threads submitting,
// waiting and polling at once.  Every solve is checked bit for bit against a direct oracle solve of the same
// matrix with the same tolerance.  Usage: host_driver <scenario> <scratch directory>; exit 0 = pass.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "io.hpp"
#include "jsvd.h"
#include "pipeline.hpp"

extern "C" int oracle_svd(const double *a_rm, int64_t m, int64_t n, double eps, int max_sweeps, int nthreads,
                          double *u_rm, double *sigma, double *v_rm, int *sweeps_out);

namespace
{
    namespace fs = std::filesystem;
    using jacobi::svd::io::Matrix;

    [[noreturn]] void fail(const std::string &what)
    {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        std::fflush(stderr);
        std::exit(1);
    }

    void expect(bool ok, const std::string &what)
    {
        if (!ok)
        {
            fail(what);
        }
    }

    Matrix random_matrix(std::mt19937_64 &rng, size_t m, size_t n)
    {
        std::normal_distribution<double> nd;
        Matrix a;
        a.rows = m;
        a.columns = n;
        a.values.resize(m * n);
        for (double &x : a.values)
        {
            x = nd(rng);
        }
        return a;
    }

    struct Solved
    {
        std::vector<double> u, s, v;
    };

    Solved oracle(const Matrix &a, double eps, int max_sweeps)
    {
        Solved r;
        r.u.resize(a.rows * a.columns);
        r.s.resize(a.columns);
        r.v.resize(a.columns * a.columns);
        int sw = 0;
        expect(oracle_svd(a.values.data(), static_cast<int64_t>(a.rows), static_cast<int64_t>(a.columns), eps,
                          max_sweeps, 1, r.u.data(), r.s.data(), r.v.data(), &sw) == 0,
               "oracle solve");
        return r;
    }

    // ---------------------------------------------------------------------------------------------------
    // Request loop: producers submit while waiters collect and a poller reads the statistics.
    // ---------------------------------------------------------------------------------------------------
    template <typename T>
    class Channel
    {
    public:
        void push(T v)
        {
            {
                std::lock_guard<std::mutex> lock(mu_);
                q_.push_back(std::move(v));
            }
            cv_.notify_one();
        }
        T pop()
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return !q_.empty(); });
            T v = std::move(q_.front());
            q_.pop_front();
            return v;
        }

    private:
        std::mutex mu_;
        std::condition_variable cv_;
        std::deque<T> q_;
    };

    double server_tolerance(size_t m, double relerr)
    {
        return relerr >= 1e-7 ? 1e-6 : std::max(1e-14, 2.0 * std::sqrt(static_cast<double>(m)) * std::ldexp(1.0, -53));
    }

    void check_result(const Matrix &a, double relerr, int max_sweeps, const std::vector<double> &u,
                      const std::vector<double> &s, const std::vector<double> &v, const std::string &tag)
    {
        const Solved want = oracle(a, server_tolerance(a.rows, relerr), max_sweeps);
        expect(u == want.u && s == want.s && v == want.v, tag + ": result differs from the direct solve");
    }

    void scenario_server(bool small_budget)
    {
        jsvd_server_config cfg;
        jsvd_server_config_default(&cfg);
        cfg.lanes = 4;
        cfg.small_max_rows = 64;
        cfg.small_max_columns = 32;
        cfg.max_sweeps = 30;
        cfg.large_path = JSVD_PATH_PLAIN_F64;
        if (small_budget)
        {
            cfg.hbm_budget_bytes = 8 * (3 * 96 * 48 + 2 * 48 * 48 + 48) * 2; // about two of the largest requests
        }
        jsvd_server *srv = nullptr;
        expect(jsvd_server_create(&cfg, &srv) == JSVD_OK, "server create");

        constexpr int kProducers = 3, kPerProducer = 30;
        struct Sent
        {
            uint64_t id;
            Matrix a;
            double relerr;
        };
        std::vector<Channel<Sent>> channels(kProducers);
        std::atomic<bool> done{false};
        std::atomic<int> verified{0};
        std::vector<std::thread> threads;
        for (int p = 0; p < kProducers; ++p)
        {
            threads.emplace_back([&, p] {
                std::mt19937_64 rng(100 + p);
                for (int k = 0; k < kPerProducer; ++k)
                {
                    const size_t m = 4 + rng() % 93;
                    const size_t n = 1 + rng() % std::min<size_t>(m, 48);
                    Matrix a = random_matrix(rng, m, n);
                    const double relerr = (k % 2) ? 1e-12 : 1e-6;
                    uint64_t id = 0;
                    expect(jsvd_server_submit(srv, a.values.data(), static_cast<int64_t>(m), static_cast<int64_t>(n),
                                              relerr, &id) == JSVD_OK,
                           "submit");
                    channels[p].push(Sent{id, std::move(a), relerr});
                    if (k % 7 == 0)
                    {
                        std::this_thread::yield();
                    }
                }
            });
            threads.emplace_back([&, p] {
                for (int k = 0; k < kPerProducer; ++k)
                {
                    Sent s = channels[p].pop();
                    std::vector<double> u(s.a.rows * s.a.columns), sig(s.a.columns), v(s.a.columns * s.a.columns);
                    jsvd_request_info info{};
                    expect(jsvd_server_wait(srv, s.id, u.data(), sig.data(), v.data(), &info) == JSVD_OK, "wait");
                    expect(info.status == JSVD_OK, "request status");
                    check_result(s.a, s.relerr, cfg.max_sweeps, u, sig, v, "server request");
                    ++verified;
                }
            });
        }
        std::thread poller([&] {
            while (!done.load())
            {
                jsvd_server_stats st{};
                expect(jsvd_server_get_stats(srv, &st) == JSVD_OK, "stats");
                expect(st.completed <= st.submitted, "stats consistency");
                if (small_budget)
                {
                    expect(st.budget_in_use <= cfg.hbm_budget_bytes || st.max_in_flight <= 1, "budget respected");
                }
                std::this_thread::yield();
            }
        });
        for (auto &t : threads)
        {
            t.join();
        }
        done = true;
        poller.join();
        jsvd_server_stats st{};
        expect(jsvd_server_get_stats(srv, &st) == JSVD_OK, "final stats");
        expect(st.completed == kProducers * kPerProducer && verified.load() == kProducers * kPerProducer,
               "every request completed and verified");
        expect(st.batches > 0 && st.plans > 0, "both the batched and the plan class ran");
        expect(st.budget_in_use == 0, "budget released");
        expect(jsvd_server_destroy(srv) == JSVD_OK, "server destroy");
    }

    // Lanes whose initialisation fails: every request must come back with an error, none may hang.
    void scenario_server_lane_failure()
    {
        jsvd_server_config cfg;
        jsvd_server_config_default(&cfg);
        cfg.lanes = 4;
        jsvd_server *srv = nullptr;
        expect(jsvd_server_create(&cfg, &srv) == JSVD_OK, "server create");
        std::vector<std::thread> threads;
        std::atomic<int> failed{0};
        for (int p = 0; p < 4; ++p)
        {
            threads.emplace_back([&, p] {
                std::mt19937_64 rng(7 + p);
                for (int k = 0; k < 6; ++k)
                {
                    Matrix a = random_matrix(rng, 40, 10 + k);
                    uint64_t id = 0;
                    expect(jsvd_server_submit(srv, a.values.data(), 40, static_cast<int64_t>(a.columns), 1e-12, &id) ==
                               JSVD_OK,
                           "submit");
                    std::vector<double> u(400), s(a.columns), v(a.columns * a.columns);
                    u.resize(40 * a.columns);
                    jsvd_request_info info{};
                    const int st = jsvd_server_wait(srv, id, u.data(), s.data(), v.data(), &info);
                    expect(st != JSVD_OK && info.status != JSVD_OK, "a failed lane reports an error");
                    expect(std::strstr(jsvd_last_error(), "lane initialisation failed") != nullptr,
                           std::string("error message: ") + jsvd_last_error());
                    ++failed;
                }
            });
        }
        for (auto &t : threads)
        {
            t.join();
        }
        expect(failed.load() == 24, "all requests failed cleanly");
        expect(jsvd_server_destroy(srv) == JSVD_OK, "server destroy");
    }

    // ---------------------------------------------------------------------------------------------------
    // Pipeline: reader, solver lanes and the ordered writer over *.mat and *.txt files.
    // ---------------------------------------------------------------------------------------------------
    std::vector<Matrix> pipeline_cases(uint64_t seed, size_t count)
    {
        std::mt19937_64 rng(seed);
        std::vector<Matrix> cases;
        for (size_t i = 0; i < count; ++i)
        {
            size_t m = 2 + rng() % 69;
            size_t n = 1 + rng() % std::min<size_t>(m, 40);
            if (i % 5 == 0)
            {
                m = 40; // runs of one shape: the lanes reuse their plans
                n = 12;
            }
            cases.push_back(random_matrix(rng, m, n));
        }
        return cases;
    }

    void check_pipeline_output(const std::vector<Matrix> &cases, const std::vector<Matrix> &out, double eps,
                               int max_sweeps, bool exact)
    {
        expect(out.size() == 3 * cases.size(), "three output matrices per case");
        for (size_t i = 0; i < cases.size(); ++i)
        {
            const Solved want = oracle(cases[i], eps, max_sweeps);
            const Matrix &u = out[3 * i], &s = out[3 * i + 1], &v = out[3 * i + 2];
            expect(u.rows == cases[i].rows && u.columns == cases[i].columns && s.rows == 1 &&
                       s.columns == cases[i].columns && v.rows == cases[i].columns && v.columns == cases[i].columns,
                   "output shapes of case " + std::to_string(i));
            if (exact)
            {
                expect(u.values == want.u && s.values == want.s && v.values == want.v,
                       "case " + std::to_string(i) + " differs from the direct solve");
            }
            else
            {
                double worst = 0.0;
                for (size_t k = 0; k < want.s.size(); ++k)
                {
                    worst = std::max(worst, std::abs(s.values[k] - want.s[k]) / std::max(want.s[0], 1e-300));
                }
                expect(worst < 1e-10, "case " + std::to_string(i) + " sigma differs from the direct solve");
            }
        }
    }

    void scenario_pipeline(const fs::path &dir)
    {
        const std::vector<Matrix> cases = pipeline_cases(21, 48);
        const fs::path in_mat = dir / "cases.mat", in_txt = dir / "cases.txt";
        jacobi::svd::io::write_mat_file(in_mat, cases);
        jacobi::svd::io::write_txt_file(in_txt, cases);
        for (size_t batch : {size_t{0}, size_t{6}})
        {
            for (size_t queued : {size_t{1}, size_t{4}})
            {
                jacobi::svd::pipeline::PipelineConfig pc{};
                pc.input_path = in_mat;
                pc.output_path = dir / "out.mat";
                pc.max_queued_results = queued;
                pc.device_batch_cases = batch;
                const auto report = jacobi::svd::pipeline::run_pipeline(pc);
                expect(report.testcase_count == cases.size() && report.emitted_matrix_count == 3 * cases.size(),
                       "pipeline report counts");
                // grouped small cases keep the reference algorithm but not necessarily every rounding
                check_pipeline_output(cases, jacobi::svd::io::read_mat_file(pc.output_path),
                                      pc.kernel_config.epsilon, pc.kernel_config.max_sweeps, batch == 0);
            }
        }
        jacobi::svd::pipeline::PipelineConfig pc{};
        pc.input_path = in_txt;
        pc.output_path = dir / "out.txt";
        (void)jacobi::svd::pipeline::run_pipeline(pc);
        check_pipeline_output(cases, jacobi::svd::io::read_txt_file(pc.output_path), pc.kernel_config.epsilon,
                              pc.kernel_config.max_sweeps, false);
    }

    // A case the solver rejects (rows < columns) in the middle of the file: run() throws, nothing hangs.
    void scenario_pipeline_error(const fs::path &dir)
    {
        std::vector<Matrix> cases = pipeline_cases(5, 20);
        std::mt19937_64 rng(9);
        cases[11] = random_matrix(rng, 3, 5);
        const fs::path in = dir / "bad.mat";
        jacobi::svd::io::write_mat_file(in, cases);
        for (size_t queued : {size_t{1}, size_t{3}})
        {
            jacobi::svd::pipeline::PipelineConfig pc{};
            pc.input_path = in;
            pc.output_path = dir / "bad_out.mat";
            pc.max_queued_results = queued;
            bool threw = false;
            try
            {
                (void)jacobi::svd::pipeline::run_pipeline(pc);
            }
            catch (const std::exception &e)
            {
                threw = true;
                expect(std::string(e.what()).find("rows >= columns") != std::string::npos,
                       std::string("pipeline error message: ") + e.what());
            }
            expect(threw, "pipeline rethrows the failing case's error");
        }
    }

    // ---------------------------------------------------------------------------------------------------
    // File I/O edge cases.
    // ---------------------------------------------------------------------------------------------------
    void write_bytes(const fs::path &p, const std::string &bytes)
    {
        std::ofstream f(p, std::ios::binary | std::ios::trunc);
        f.write(bytes.data(), static_cast<std::streamsize>(bytes.size()));
    }

    template <typename F>
    void expect_throws(F &&f, const std::string &what)
    {
        try
        {
            f();
        }
        catch (const std::exception &)
        {
            return;
        }
        fail(what + " did not throw");
    }

    void scenario_io(const fs::path &dir)
    {
        namespace io = jacobi::svd::io;
        // round trips, including 1 x 1 and a wide block of values
        std::mt19937_64 rng(3);
        std::vector<Matrix> ms{random_matrix(rng, 1, 1), random_matrix(rng, 7, 3), random_matrix(rng, 300, 200)};
        io::write_mat_file(dir / "rt.mat", ms);
        const auto back = io::read_mat_file(dir / "rt.mat");
        expect(back.size() == ms.size(), "mat round trip count");
        for (size_t i = 0; i < ms.size(); ++i)
        {
            expect(back[i].rows == ms[i].rows && back[i].values == ms[i].values, "mat round trip values");
        }
        io::write_txt_file(dir / "rt.txt", ms);
        const auto tback = io::read_txt_file(dir / "rt.txt");
        expect(tback.size() == ms.size(), "txt round trip count");
        for (size_t i = 0; i < ms.size(); ++i)
        {
            expect(tback[i].rows == ms[i].rows && tback[i].columns == ms[i].columns, "txt round trip shape");
        }
        // empty files hold no cases
        write_bytes(dir / "empty.mat", "");
        write_bytes(dir / "empty.txt", "");
        expect(io::read_mat_file(dir / "empty.mat").empty(), "empty mat");
        expect(io::read_txt_file(dir / "empty.txt").empty(), "empty txt");
        // a *.mat record cut inside its header and inside its payload
        std::string whole;
        {
            std::ifstream f(dir / "rt.mat", std::ios::binary);
            whole.assign(std::istreambuf_iterator<char>(f), {});
        }
        for (size_t cut : {size_t{5}, size_t{16 + 3}, whole.size() - 1})
        {
            write_bytes(dir / "cut.mat", whole.substr(0, cut));
            expect_throws([&] { (void)io::read_mat_file(dir / "cut.mat"); }, "truncated mat at " + std::to_string(cut));
            io::MatDispatchReader reader(dir / "cut.mat");
            io::MatDispatchTask task;
            expect_throws(
                [&] {
                    while (reader.read_next(task, 0))
                    {
                        (void)io::decode_dispatch_task_matrix(task);
                    }
                },
                "truncated mat through the dispatch reader at " + std::to_string(cut));
        }
        // *.txt (one row per line, matrices separated by blank lines): ragged rows and a bad token throw;
        // blocks of different widths, extra blank lines and whitespace-only lines are fine
        write_bytes(dir / "ragged.txt", "1 2 3\n4 5\n");
        expect_throws([&] { (void)io::read_txt_file(dir / "ragged.txt"); }, "ragged txt");
        write_bytes(dir / "token.txt", "1 x\n3 4\n");
        expect_throws([&] { (void)io::read_txt_file(dir / "token.txt"); }, "bad token txt");
        write_bytes(dir / "blocks.txt", "\n\n1 2 3\n4 5 6\n \t\n\n7\n8\n9\n\n");
        const auto blocks = io::read_txt_file(dir / "blocks.txt");
        expect(blocks.size() == 2 && blocks[0].rows == 2 && blocks[0].columns == 3 && blocks[1].rows == 3 &&
                   blocks[1].columns == 1 && blocks[1].values == std::vector<double>{7, 8, 9},
               "txt blocks");
        // the dispatch reader over a file of many records, with a pinned buffer that has to grow
        std::vector<Matrix> many;
        for (int i = 0; i < 40; ++i)
        {
            many.push_back(random_matrix(rng, 1 + i * 5, 1 + i % 9));
        }
        io::write_mat_file(dir / "many.mat", many);
        io::MatDispatchReader reader(dir / "many.mat");
        io::MatDispatchTask task;
        size_t k = 0;
        while (reader.read_next(task, 64))
        {
            const Matrix got = io::decode_dispatch_task_matrix(task);
            expect(k < many.size() && got.rows == many[k].rows && got.values == many[k].values,
                   "dispatch reader record " + std::to_string(k));
            ++k;
        }
        expect(k == many.size(), "dispatch reader count");
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc < 3)
    {
        std::fprintf(stderr, "usage: host_driver <scenario> <scratch directory>\n");
        return 2;
    }
    const std::string scenario = argv[1];
    const fs::path dir = argv[2];
    try
    {
        if (scenario == "server")
        {
            scenario_server(false);
        }
        else if (scenario == "server_budget")
        {
            scenario_server(true);
        }
        else if (scenario == "server_lane_failure")
        {
            scenario_server_lane_failure();
        }
        else if (scenario == "pipeline")
        {
            scenario_pipeline(dir);
        }
        else if (scenario == "pipeline_error")
        {
            scenario_pipeline_error(dir);
        }
        else if (scenario == "io")
        {
            scenario_io(dir);
        }
        else
        {
            std::fprintf(stderr, "unknown scenario %s\n", scenario.c_str());
            return 2;
        }
    }
    catch (const std::exception &e)
    {
        fail(std::string("uncaught exception: ") + e.what());
    }
    std::printf("ok %s\n", scenario.c_str());
    return 0;
}
