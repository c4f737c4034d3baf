// test_dropin.cpp -- C++ tests of the drop-in entry points (include/kernels.hpp, io.hpp, pipeline.hpp),
// written the way a user of the reference's headers would call them.
//
//   test_dropin cpu <tmpdir> <golden_dir>   host-only checks (I/O codecs, validation): no GPU needed
//   test_dropin gpu <tmpdir>                solver and pipeline on the GPU
//
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "io.hpp"
#include "kernels.hpp"
#include "pipeline.hpp"

namespace fs = std::filesystem;
using namespace jacobi::svd;

namespace
{
    int g_checks = 0;

    void check(bool ok, const std::string &what)
    {
        ++g_checks;
        if (!ok)
        {
            std::cout << "FAIL " << what << std::endl;
            std::exit(1);
        }
        std::cout << "ok   " << what << std::endl;
    }

    template <typename E>
    void expect_throw(const std::function<void()> &fn, const std::string &what, const std::string &needle = "")
    {
        try
        {
            fn();
        }
        catch (const E &e)
        {
            check(needle.empty() || std::string(e.what()).find(needle) != std::string::npos,
                  what + " (message: " + e.what() + ")");
            return;
        }
        catch (const std::exception &e)
        {
            check(false, what + " threw the wrong type: " + e.what());
        }
        check(false, what + " did not throw");
    }

    std::vector<char> read_bytes(const fs::path &p)
    {
        std::ifstream in(p, std::ios::binary);
        return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    }

    io::Matrix gaussian(std::size_t m, std::size_t n, unsigned seed)
    {
        std::mt19937_64 gen(seed);
        std::normal_distribution<double> d;
        io::Matrix a{m, n, std::vector<double>(m * n)};
        for (double &x : a.values)
        {
            x = d(gen);
        }
        return a;
    }

    // ||A - U S V^T||_F / max(1, ||A||_F), ||U^T U - I||_F, ||V^T V - I||_F
    void invariants(const io::Matrix &a, const JacobiSvdResult &r, double &ra, double &ru, double &rv)
    {
        const std::size_t m = a.rows, n = a.columns;
        double num = 0.0, den = 0.0;
        for (std::size_t i = 0; i < m; ++i)
        {
            for (std::size_t j = 0; j < n; ++j)
            {
                double s = 0.0;
                for (std::size_t k = 0; k < n; ++k)
                {
                    s += r.u[i * n + k] * r.sigma[k] * r.v[j * n + k];
                }
                const double d = a.values[i * n + j] - s;
                num += d * d;
                den += a.values[i * n + j] * a.values[i * n + j];
            }
        }
        ra = std::sqrt(num) / std::max(1.0, std::sqrt(den));
        auto orth = [](const std::vector<double> &x, std::size_t rows, std::size_t cols) {
            double acc = 0.0;
            for (std::size_t p = 0; p < cols; ++p)
            {
                for (std::size_t q = 0; q < cols; ++q)
                {
                    double s = 0.0;
                    for (std::size_t i = 0; i < rows; ++i)
                    {
                        s += x[i * cols + p] * x[i * cols + q];
                    }
                    const double d = s - (p == q ? 1.0 : 0.0);
                    acc += d * d;
                }
            }
            return std::sqrt(acc);
        };
        ru = orth(r.u, m, n);
        rv = orth(r.v, n, n);
    }

    void cpu_checks(const fs::path &tmp, const fs::path &golden)
    {
        // *.mat known answer: [1.0, -2.0] as a 1 x 2 record.
        const fs::path mat = tmp / "one.mat";
        io::write_mat_file(mat, std::vector<io::Matrix>{io::Matrix{1, 2, {1.0, -2.0}}});
        check(read_bytes(mat) == read_bytes(golden / "mat_1x2.bin"), "mat writer matches the hand-computed record bytes");
        const auto back = io::read_mat_file(golden / "mat_1x2.bin");
        check(back.size() == 1 && back[0].rows == 1 && back[0].columns == 2 && back[0].values[0] == 1.0 &&
                  back[0].values[1] == -2.0,
              "mat reader decodes the known-answer record");

        // Round trip of several records, including a 1 x 1 and special values.
        std::vector<io::Matrix> many = {gaussian(5, 3, 1), io::Matrix{1, 1, {-0.0}}, gaussian(2, 7, 2)};
        many[2].values[3] = std::numeric_limits<double>::denorm_min();
        many[2].values[4] = std::numeric_limits<double>::infinity();
        const fs::path rt = tmp / "many.mat";
        io::write_mat_file(rt, many);
        const auto got = io::read_mat_file(rt);
        bool same = got.size() == many.size();
        for (std::size_t i = 0; same && i < many.size(); ++i)
        {
            same = got[i].rows == many[i].rows && got[i].columns == many[i].columns &&
                   std::memcmp(got[i].values.data(), many[i].values.data(), 8 * many[i].values.size()) == 0;
        }
        check(same, "mat write/read round trip is bit-exact");
        check(fs::file_size(rt) == (16 + 15 * 8) + (16 + 8) + (16 + 14 * 8), "mat file size is header + payload");

        // Truncated header / payload.
        {
            const auto bytes = read_bytes(rt);
            const fs::path t1 = tmp / "trunc_header.mat";
            std::ofstream(t1, std::ios::binary).write(bytes.data(), 10);
            expect_throw<std::runtime_error>([&] { (void)io::read_mat_file(t1); }, "truncated header throws",
                                             "Truncated *.mat header");
            const fs::path t2 = tmp / "trunc_payload.mat";
            std::ofstream(t2, std::ios::binary).write(bytes.data(), 16 + 8 * 7);
            expect_throw<std::runtime_error>([&] { (void)io::read_mat_file(t2); }, "truncated payload throws",
                                             "Truncated *.mat payload");
        }
        // Overflowing dimensions.
        {
            const fs::path t3 = tmp / "overflow.mat";
            const unsigned char hdr[16] = {0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff,
                                           0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff};
            std::ofstream(t3, std::ios::binary).write(reinterpret_cast<const char *>(hdr), 16);
            expect_throw<std::overflow_error>([&] { (void)io::read_mat_file(t3); }, "size overflow throws");
        }

        // Text: max_digits10 round trip, blank-line separated blocks, ragged rows.
        const fs::path txt = tmp / "many.txt";
        std::vector<io::Matrix> finite = {gaussian(4, 3, 3), io::Matrix{1, 1, {1e-300}}, gaussian(3, 3, 4)};
        io::write_txt_file(txt, finite);
        const auto tback = io::read_txt_file(txt);
        bool tsame = tback.size() == finite.size();
        for (std::size_t i = 0; tsame && i < finite.size(); ++i)
        {
            tsame = tback[i].rows == finite[i].rows && tback[i].values == finite[i].values;
        }
        check(tsame, "txt write/read round trip is exact at max_digits10");
        const fs::path ragged = tmp / "ragged.txt";
        std::ofstream(ragged) << "1 2 3\n4 5\n";
        expect_throw<std::invalid_argument>([&] { (void)io::read_txt_file(ragged); }, "ragged txt rows throw",
                                            "Inconsistent column count");
        const fs::path badtok = tmp / "badtok.txt";
        std::ofstream(badtok) << "1 x 3\n";
        expect_throw<std::invalid_argument>([&] { (void)io::read_txt_file(badtok); }, "bad txt token throws",
                                            "Invalid numeric token");
        const fs::path spaced = tmp / "spaced.txt";
        std::ofstream(spaced) << "\n\n1 2\n3 4\n\n\n\n5\n";
        const auto sp = io::read_txt_file(spaced);
        check(sp.size() == 2 && sp[0].rows == 2 && sp[0].columns == 2 && sp[1].values == std::vector<double>{5.0},
              "txt blocks separated by several blank lines");

        // Solver validation happens before any device work.
        const std::vector<double> a34(12, 1.0);
        expect_throw<std::invalid_argument>([&] { (void)one_sided_jacobi_svd(a34, 3, 4); }, "m < n rejected",
                                            "requires rows >= columns");
        expect_throw<std::invalid_argument>([&] { (void)one_sided_jacobi_svd(a34, 4, 4); }, "size mismatch rejected",
                                            "Input buffer size mismatch");
        expect_throw<std::invalid_argument>([&] { (void)one_sided_jacobi_svd(a34, 0, 3); }, "zero shape rejected",
                                            "shape must be non-zero");
        JacobiSvdConfig bad{.epsilon = 0.0};
        expect_throw<std::invalid_argument>([&] { (void)one_sided_jacobi_svd(a34, 4, 3, bad); }, "eps <= 0 rejected",
                                            "epsilon must be positive");
        JacobiSvdConfig bad2{.epsilon = 1e-9, .max_sweeps = 0};
        expect_throw<std::invalid_argument>([&] { (void)one_sided_jacobi_svd(a34, 4, 3, bad2); },
                                            "max_sweeps <= 0 rejected", "max_sweeps must be positive");
        JacobiSvdConfig bad3{};
        bad3.layout_transpose_min_columns = 0;
        expect_throw<std::invalid_argument>([&] { (void)auto_tune_layout_transpose_threshold(bad3); },
                                            "autotune validates layout knobs", "layout_transpose_min_columns");

        // Pipeline argument errors (no device work).
        pipeline::PipelineConfig pc;
        expect_throw<std::invalid_argument>([&] { (void)pipeline::run_pipeline(pc); }, "empty input path",
                                            "input path is empty");
        pc.input_path = tmp / "x.bin";
        pc.output_path = tmp / "y.mat";
        expect_throw<std::invalid_argument>([&] { (void)pipeline::run_pipeline(pc); }, "unknown extension",
                                            "Cannot auto-detect");
    }

    void gpu_checks(const fs::path &tmp)
    {
        // Dispatch reader: pinned (hipHostMalloc) buffers + sequence numbers + decode; needs the HIP runtime.
        std::vector<io::Matrix> many = {gaussian(5, 3, 1), io::Matrix{1, 1, {-0.0}}, gaussian(2, 7, 2)};
        const fs::path rt = tmp / "dispatch.mat";
        io::write_mat_file(rt, many);
        {
            io::MatDispatchReader reader(rt);
            io::MatDispatchTask task;
            std::size_t k = 0;
            bool ok = true;
            while (reader.read_next(task, 128))
            {
                const io::Matrix m = io::decode_dispatch_task_matrix(task);
                ok = ok && task.sequence_index == k && m.rows == many[k].rows &&
                     std::memcmp(m.values.data(), many[k].values.data(), 8 * m.values.size()) == 0 &&
                     task.buffer.workspace_size_bytes() == 128 && task.buffer.capacity_bytes() >= 64 * 1024;
                ++k;
            }
            check(ok && k == many.size(), "dispatch reader yields every record in order into pinned buffers");
        }

        {
            const auto bytes = read_bytes(rt);
            const fs::path t2 = tmp / "dispatch_trunc.mat";
            std::ofstream(t2, std::ios::binary).write(bytes.data(), 16 + 8 * 7);
            io::MatDispatchReader d(t2);
            io::MatDispatchTask task;
            expect_throw<std::runtime_error>([&] { (void)d.read_next(task); }, "dispatch reader truncated payload",
                                             "Truncated *.mat payload while dispatch reading");
        }

        // The reference report's suggested CI cases (4 x 3 and 8 x 8) plus an odd shape.
        for (auto [m, n] : {std::pair<std::size_t, std::size_t>{4, 3}, {8, 8}, {37, 20}})
        {
            const io::Matrix a = gaussian(m, n, static_cast<unsigned>(m * 100 + n));
            JacobiSvdConfig cfg{.epsilon = 1e-15, .max_sweeps = 60};
            const JacobiSvdResult r = one_sided_jacobi_svd(a.values, m, n, cfg);
            double ra, ru, rv;
            invariants(a, r, ra, ru, rv);
            std::ostringstream what;
            what << "solve " << m << "x" << n << " sweeps=" << r.sweeps << " rhoA=" << ra << " rhoU=" << ru
                 << " rhoV=" << rv;
            check(r.rows == m && r.columns == n && ra < 1e-14 && ru < 1e-13 && rv < 1e-13, what.str());
        }

        // DeviceMatrix round trip.
        DeviceMatrix dm(3, 5);
        std::vector<double> host(15);
        for (std::size_t i = 0; i < host.size(); ++i)
        {
            host[i] = static_cast<double>(i) * 0.5 - 3.0;
        }
        dm.copy_from_host(host);
        DeviceMatrix moved = std::move(dm);
        check(moved.copy_to_host() == host && moved.rows() == 3 && moved.bytes() == 120 && dm.data() == nullptr,
              "DeviceMatrix copy/move");

        // Pipeline, *.mat -> *.mat, several shapes, order preserved.
        std::vector<io::Matrix> cases;
        for (unsigned i = 0; i < 13; ++i)
        {
            const std::size_t m = 6 + 7 * (i % 5);
            const std::size_t n = 2 + (i % 4) * 3;
            cases.push_back(gaussian(m, std::min(n, m), 1000 + i));
        }
        const fs::path in = tmp / "cases.mat";
        const fs::path out = tmp / "out" / "result.mat";
        io::write_mat_file(in, cases);
        pipeline::PipelineConfig pc;
        pc.input_path = in;
        pc.output_path = out;
        pc.max_queued_results = 2;
        pc.kernel_config = JacobiSvdConfig{.epsilon = 1e-14, .max_sweeps = 60};
        const pipeline::PipelineReport rep = pipeline::run_pipeline(pc);
        const auto results = io::read_mat_file(out);
        bool ordered = results.size() == 3 * cases.size() && rep.testcase_count == cases.size() &&
                       rep.emitted_matrix_count == 3 * cases.size() && rep.total_sweeps > 0;
        for (std::size_t i = 0; ordered && i < cases.size(); ++i)
        {
            const auto &u = results[3 * i];
            const auto &s = results[3 * i + 1];
            const auto &v = results[3 * i + 2];
            ordered = u.rows == cases[i].rows && u.columns == cases[i].columns && s.rows == 1 &&
                      s.columns == cases[i].columns && v.rows == cases[i].columns;
            if (ordered)
            {
                const JacobiSvdResult r = one_sided_jacobi_svd(cases[i].values, cases[i].rows, cases[i].columns,
                                                               pc.kernel_config);
                ordered = r.sigma == s.values && r.v == v.values && r.u == u.values;
            }
        }
        check(ordered, "pipeline mat->mat emits U, Sigma(1xn), V per case in input order, equal to direct solves");

        // Batched small cases (device_batch_cases): same order, agreement with direct solves to rounding (the
        // same reference algorithm through the persistent batched kernel), larger cases on the lanes between.
        {
            std::vector<io::Matrix> mix;
            for (unsigned i = 0; i < 45; ++i)
            {
                const std::size_t m = 8 + (i * 37) % 249;
                const std::size_t n = std::min<std::size_t>(m, 4 + (i * 53) % 125);
                mix.push_back(gaussian(m, n, 5000 + i));
                if (i == 11)
                {
                    mix.push_back(gaussian(300, 40, 77)); // beyond the batch rows: a lane
                }
                if (i == 30)
                {
                    mix.push_back(gaussian(260, 140, 78)); // beyond the batch columns: a lane
                }
            }
            const fs::path bin = tmp / "batched.mat";
            const fs::path bout = tmp / "batched_out.mat";
            io::write_mat_file(bin, mix);
            pipeline::PipelineConfig bc = pc;
            bc.input_path = bin;
            bc.output_path = bout;
            bc.device_batch_cases = 16;
            const pipeline::PipelineReport brep = pipeline::run_pipeline(bc);
            const auto bres = io::read_mat_file(bout);
            bool agree = bres.size() == 3 * mix.size() && brep.testcase_count == mix.size() && brep.total_sweeps > 0;
            double worst_s = 0.0, worst_uv = 0.0;
            for (std::size_t i = 0; agree && i < mix.size(); ++i)
            {
                const auto &u = bres[3 * i];
                const auto &sg = bres[3 * i + 1];
                const auto &v = bres[3 * i + 2];
                agree = u.rows == mix[i].rows && u.columns == mix[i].columns && sg.columns == mix[i].columns &&
                        v.rows == mix[i].columns;
                if (!agree)
                {
                    break;
                }
                const JacobiSvdResult r = one_sided_jacobi_svd(mix[i].values, mix[i].rows, mix[i].columns,
                                                               bc.kernel_config);
                for (std::size_t j = 0; j < r.sigma.size(); ++j)
                {
                    worst_s = std::max(worst_s, std::abs(r.sigma[j] - sg.values[j]) / r.sigma[j]);
                }
                for (std::size_t j = 0; j < r.u.size(); ++j)
                {
                    worst_uv = std::max(worst_uv, std::abs(r.u[j] - u.values[j]));
                }
                for (std::size_t j = 0; j < r.v.size(); ++j)
                {
                    worst_uv = std::max(worst_uv, std::abs(r.v[j] - v.values[j]));
                }
            }
            check(agree && worst_s < 1e-12 && worst_uv < 1e-9,
                  "pipeline with device_batch_cases: input order kept, sigma within 1e-12 (worst " +
                      std::to_string(worst_s) + "), U and V within 1e-9 of direct solves");
        }

        // Pipeline txt -> txt.
        const fs::path tin = tmp / "cases.txt";
        const fs::path tout = tmp / "result.txt";
        io::write_txt_file(tin, std::vector<io::Matrix>(cases.begin(), cases.begin() + 4));
        pc.input_path = tin;
        pc.output_path = tout;
        const auto rep2 = pipeline::run_pipeline(pc);
        check(rep2.testcase_count == 4 && io::read_txt_file(tout).size() == 12, "pipeline txt->txt");

        // A failing case (m < n) surfaces from run() as std::invalid_argument.
        std::vector<io::Matrix> bad = {cases[0], gaussian(2, 5, 9), cases[1]};
        io::write_mat_file(tmp / "bad.mat", bad);
        pc.input_path = tmp / "bad.mat";
        pc.output_path = tmp / "bad_out.mat";
        expect_throw<std::invalid_argument>([&] { (void)pipeline::run_pipeline(pc); }, "pipeline rethrows a worker error",
                                            "rows >= columns");

        // Autotune report (reference sample sizes).
        JacobiSvdConfig tune{};
        tune.layout_transpose_benchmark_repetitions = 1;
        const LayoutTransposeAutoTuneReport t = auto_tune_layout_transpose_threshold(tune);
        check(t.executed && t.sample_count == 8 && t.estimated_best_speedup >= 1.0 && t.recommended_min_columns > 0,
              "auto_tune_layout_transpose_threshold report (min_columns=" + std::to_string(t.recommended_min_columns) +
                  ", speedup=" + std::to_string(t.estimated_best_speedup) + ")");
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc < 3)
    {
        std::cerr << "usage: test_dropin cpu|gpu <tmpdir> [golden_dir]\n";
        return 2;
    }
    const std::string mode = argv[1];
    const fs::path tmp = argv[2];
    fs::create_directories(tmp);
    try
    {
        if (mode == "cpu")
        {
            cpu_checks(tmp, argc > 3 ? fs::path(argv[3]) : fs::path("tests/golden"));
        }
        else
        {
            gpu_checks(tmp);
        }
    }
    catch (const std::exception &e)
    {
        std::cout << "FAIL uncaught exception: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "passed " << g_checks << " checks" << std::endl;
    return 0;
}
