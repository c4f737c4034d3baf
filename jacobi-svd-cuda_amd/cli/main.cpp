// main.cpp -- command-line front end of the MI355X Jacobi SVD: the reference CLI's options, defaults, output
// naming, reports and exit codes (reference src/main.cpp:877-1300), running the pipeline of include/pipeline.hpp
// (plain FP64 path, reference semantics) or, with --solver block|mixed, the streamed *.mat path of
// jsvd_mat_stream_solve.
//
// Exit codes: 0 success (also --help, --version, --dry-run), 2 argument error, 1 execution failure.
#include "jsvd.h"
#include "pipeline.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <exception>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <system_error>
#include <vector>

namespace
{
    namespace fs = std::filesystem;
    using jacobi::svd::JacobiSvdConfig;
    using jacobi::svd::LayoutTransposeMode;
    using jacobi::svd::pipeline::MatrixFileFormat;

    // Thrown for anything wrong with the command line (exit code 2).
    struct UsageError : std::invalid_argument
    {
        using std::invalid_argument::invalid_argument;
    };

    enum class Solver
    {
        plain,
        block,
        mixed
    };

    struct Options
    {
        fs::path input;
        fs::path output;
        MatrixFileFormat input_format = MatrixFileFormat::auto_detect;
        MatrixFileFormat output_format = MatrixFileFormat::auto_detect;
        JacobiSvdConfig kernel{};
        std::size_t queue_capacity = 4;
        Solver solver = Solver::plain;
        std::size_t device_batch = 0; // 0: every case on a solver lane
        bool force = false;
        bool dry_run = false;
        bool print_config = false;
        bool json_report = false;
        bool quiet = false;
        bool help = false;
        bool version = false;
    };

    // One row of the option table: long name, short name ('\0' for none), value hint (empty: a flag) and the
    // help text.  Names, hints and texts follow the reference's table so scripts and muscle memory carry over.
    struct Opt
    {
        std::string_view name;
        char letter;
        std::string_view hint;
        std::string_view text;
    };

    const std::vector<Opt> &table()
    {
        static const std::vector<Opt> t = {
            {"input", 'i', "PATH", "Input matrix stream file path."},
            {"output", 'o', "PATH", "Output matrix stream file path."},
            {"input-format", '\0', "FMT", "Input format: auto|mat|txt."},
            {"output-format", '\0', "FMT", "Output format: auto|mat|txt."},
            {"format", 'f', "FMT", "Set both input/output format: auto|mat|txt."},
            {"epsilon", 'e', "NUM", "Convergence epsilon (>0)."},
            {"max-sweeps", 's', "N", "Maximum sweep count (>0)."},
            {"threads-per-block", 't', "N", "Threads per block (>0; validated, launch shapes are per matrix)."},
            {"layout-transpose-mode", '\0', "MODE", "Layout-transpose mode: auto|on|off."},
            {"layout-transpose-min-cols", '\0', "N", "Auto-mode threshold: minimum columns (>0)."},
            {"layout-transpose-min-elems", '\0', "N", "Auto-mode threshold: minimum elements (>0)."},
            {"layout-transpose-auto-tune", '\0', "", "Run micro-benchmark to auto-tune layout thresholds before pipeline."},
            {"layout-transpose-bench-reps", '\0', "N", "Auto-tune repetitions per scanned matrix size (>0)."},
            {"layout-transpose-bench-sweeps", '\0', "N", "Auto-tune sweep cap per benchmark run (>0)."},
            {"queue-capacity", 'c', "N", "In-flight/reorder window size (>0)."},
            {"solver", '\0', "PATH", "Solver: plain (reference algorithm, default) | block | mixed (*.mat only)."},
            {"device-batch", '\0', "N", "Plain solver: group up to N consecutive small cases (m<=256, n<=128) per batched launch (>0)."},
            {"force", 'y', "", "Overwrite existing output file."},
            {"dry-run", '\0', "", "Validate arguments and print config only."},
            {"print-config", '\0', "", "Print effective config before execution."},
            {"json-report", '\0', "", "Print execution report in JSON."},
            {"quiet", 'q', "", "Suppress text report."},
            {"help", 'h', "", "Show this help message."},
            {"version", 'v', "", "Show version information."},
        };
        return t;
    }

    std::string lower(std::string_view s)
    {
        std::string out(s);
        for (char &c : out)
        {
            c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        }
        return out;
    }

    MatrixFileFormat parse_format(std::string_view v, std::string_view name)
    {
        const std::string s = lower(v);
        if (s == "auto")
        {
            return MatrixFileFormat::auto_detect;
        }
        if (s == "mat")
        {
            return MatrixFileFormat::mat;
        }
        if (s == "txt")
        {
            return MatrixFileFormat::txt;
        }
        throw UsageError("Invalid value for --" + std::string(name) + ": expected one of {auto, mat, txt}.");
    }

    LayoutTransposeMode parse_mode(std::string_view v, std::string_view name)
    {
        const std::string s = lower(v);
        if (s == "auto")
        {
            return LayoutTransposeMode::auto_select;
        }
        if (s == "on")
        {
            return LayoutTransposeMode::force_enable;
        }
        if (s == "off")
        {
            return LayoutTransposeMode::force_disable;
        }
        throw UsageError("Invalid value for --" + std::string(name) + ": expected one of {auto, on, off}.");
    }

    Solver parse_solver(std::string_view v, std::string_view name)
    {
        const std::string s = lower(v);
        if (s == "plain")
        {
            return Solver::plain;
        }
        if (s == "block")
        {
            return Solver::block;
        }
        if (s == "mixed")
        {
            return Solver::mixed;
        }
        throw UsageError("Invalid value for --" + std::string(name) + ": expected one of {plain, block, mixed}.");
    }

    // Whole-token numeric parses; the messages are the reference's.
    template <typename T, typename F>
    T parse_number(std::string_view v, std::string_view name, const char *kind, F convert)
    {
        const std::string text(v);
        std::size_t used = 0;
        T value{};
        try
        {
            value = convert(text, &used);
        }
        catch (const std::exception &)
        {
            throw UsageError(std::string("Invalid ") + kind + " value for --" + std::string(name) + ".");
        }
        if (used != text.size())
        {
            throw UsageError(std::string("Invalid ") + kind + " value for --" + std::string(name) + ".");
        }
        return value;
    }

    int positive_int(std::string_view v, std::string_view name)
    {
        const int x = parse_number<int>(v, name, "integer",
                                        [](const std::string &t, std::size_t *u) { return std::stoi(t, u, 10); });
        if (x <= 0)
        {
            throw UsageError("Option --" + std::string(name) + " must be positive.");
        }
        return x;
    }

    std::size_t positive_size(std::string_view v, std::string_view name)
    {
        const unsigned long long x = parse_number<unsigned long long>(
            v, name, "integer", [](const std::string &t, std::size_t *u) { return std::stoull(t, u, 10); });
        if (x == 0ULL)
        {
            throw UsageError("Option --" + std::string(name) + " must be positive.");
        }
        return static_cast<std::size_t>(x);
    }

    double positive_double(std::string_view v, std::string_view name)
    {
        const double x = parse_number<double>(v, name, "floating-point",
                                              [](const std::string &t, std::size_t *u) { return std::stod(t, u); });
        if (!std::isfinite(x) || x <= 0.0)
        {
            throw UsageError("Option --" + std::string(name) + " must be a positive finite number.");
        }
        return x;
    }

    void apply(Options &o, const Opt &opt, std::string_view v)
    {
        const std::string_view n = opt.name;
        if (n == "input") o.input = fs::path(std::string(v));
        else if (n == "output") o.output = fs::path(std::string(v));
        else if (n == "input-format") o.input_format = parse_format(v, n);
        else if (n == "output-format") o.output_format = parse_format(v, n);
        else if (n == "format") o.input_format = o.output_format = parse_format(v, n);
        else if (n == "epsilon") o.kernel.epsilon = positive_double(v, n);
        else if (n == "max-sweeps") o.kernel.max_sweeps = positive_int(v, n);
        else if (n == "threads-per-block") o.kernel.threads_per_block = positive_int(v, n);
        else if (n == "layout-transpose-mode") o.kernel.layout_transpose_mode = parse_mode(v, n);
        else if (n == "layout-transpose-min-cols") o.kernel.layout_transpose_min_columns = positive_int(v, n);
        else if (n == "layout-transpose-min-elems") o.kernel.layout_transpose_min_elements = positive_size(v, n);
        else if (n == "layout-transpose-auto-tune") o.kernel.layout_transpose_auto_tune = true;
        else if (n == "layout-transpose-bench-reps") o.kernel.layout_transpose_benchmark_repetitions = positive_int(v, n);
        else if (n == "layout-transpose-bench-sweeps") o.kernel.layout_transpose_benchmark_sweeps = positive_int(v, n);
        else if (n == "queue-capacity") o.queue_capacity = positive_size(v, n);
        else if (n == "solver") o.solver = parse_solver(v, n);
        else if (n == "device-batch") o.device_batch = positive_size(v, n);
        else if (n == "force") o.force = true;
        else if (n == "dry-run") o.dry_run = true;
        else if (n == "print-config") o.print_config = true;
        else if (n == "json-report") o.json_report = true;
        else if (n == "quiet") o.quiet = true;
        else if (n == "help") o.help = true;
        else if (n == "version") o.version = true;
    }

    const Opt *find_long(std::string_view name)
    {
        for (const Opt &o : table())
        {
            if (o.name == name)
            {
                return &o;
            }
        }
        return nullptr;
    }

    const Opt *find_short(char letter)
    {
        for (const Opt &o : table())
        {
            if (o.letter != '\0' && o.letter == letter)
            {
                return &o;
            }
        }
        return nullptr;
    }

    std::optional<MatrixFileFormat> format_of(const fs::path &p)
    {
        const std::string ext = lower(p.extension().string());
        if (ext == ".mat")
        {
            return MatrixFileFormat::mat;
        }
        if (ext == ".txt")
        {
            return MatrixFileFormat::txt;
        }
        return std::nullopt;
    }

    const char *extension_of(MatrixFileFormat f) { return f == MatrixFileFormat::txt ? ".txt" : ".mat"; }

    // Fills in the output path and formats the way the reference does: an omitted output becomes
    // <input-stem>.svd.<ext> beside the input, an output without a recognised extension takes the input's format.
    void complete(Options &o)
    {
        if (o.input.empty())
        {
            throw UsageError("Missing input file. Use --input <path>.");
        }
        const std::optional<MatrixFileFormat> in_ext = format_of(o.input);
        if (o.input_format == MatrixFileFormat::auto_detect && !in_ext)
        {
            throw UsageError("Cannot infer input format from extension. Use --input-format {mat|txt}.");
        }
        const MatrixFileFormat in_fmt = (o.input_format != MatrixFileFormat::auto_detect) ? o.input_format : *in_ext;
        if (o.output.empty())
        {
            if (o.output_format == MatrixFileFormat::auto_detect)
            {
                o.output_format = in_fmt;
            }
            std::string stem = o.input.stem().string();
            if (stem.empty())
            {
                stem = "result";
            }
            o.output = o.input.parent_path() / (stem + ".svd" + extension_of(o.output_format));
        }
        if (o.output_format == MatrixFileFormat::auto_detect && !format_of(o.output))
        {
            if (o.output.extension().empty())
            {
                o.output += extension_of(in_fmt);
            }
            o.output_format = in_fmt;
        }
    }

    fs::path comparable(const fs::path &p)
    {
        std::error_code ec;
        const fs::path abs = fs::absolute(p, ec);
        return ec ? p.lexically_normal() : abs.lexically_normal();
    }

    void check_files(const Options &o)
    {
        if (!fs::exists(o.input))
        {
            throw UsageError("Input file does not exist: " + o.input.string());
        }
        if (fs::is_directory(o.input))
        {
            throw UsageError("Input path is a directory, expected a file: " + o.input.string());
        }
        if (fs::is_directory(o.output))
        {
            throw UsageError("Output path points to a directory: " + o.output.string());
        }
        if (comparable(o.input) == comparable(o.output))
        {
            throw UsageError("Input and output paths must be different files.");
        }
        if (fs::exists(o.output) && !o.force)
        {
            throw UsageError("Output file already exists. Use --force to overwrite: " + o.output.string());
        }
        if (o.solver != Solver::plain)
        {
            const MatrixFileFormat in = (o.input_format != MatrixFileFormat::auto_detect) ? o.input_format
                                                                                          : *format_of(o.input);
            const MatrixFileFormat out = (o.output_format != MatrixFileFormat::auto_detect) ? o.output_format
                                                                                            : *format_of(o.output);
            if (in != MatrixFileFormat::mat || out != MatrixFileFormat::mat)
            {
                throw UsageError("--solver block|mixed reads and writes *.mat files only.");
            }
            if (o.device_batch > 0)
            {
                throw UsageError("--device-batch applies to the plain solver only.");
            }
        }
    }

    Options parse(int argc, char **argv)
    {
        Options o;
        std::vector<std::string_view> positional;
        for (int k = 1; k < argc; ++k)
        {
            const std::string_view tok = argv[k] ? argv[k] : "";
            if (tok.empty())
            {
                continue;
            }
            auto take_value = [&](const Opt &opt) -> std::string_view {
                if (k + 1 >= argc || argv[k + 1] == nullptr)
                {
                    throw UsageError("Option --" + std::string(opt.name) + " requires a value.");
                }
                return argv[++k];
            };
            if (tok == "--")
            {
                for (++k; k < argc; ++k)
                {
                    positional.emplace_back(argv[k] ? argv[k] : "");
                }
                break;
            }
            if (tok.substr(0, 2) == "--")
            {
                const std::string_view body = tok.substr(2);
                const std::size_t eq = body.find('=');
                const std::string_view name = body.substr(0, eq);
                const Opt *opt = find_long(name);
                if (opt == nullptr)
                {
                    throw UsageError("Unknown option: --" + std::string(name));
                }
                if (opt->hint.empty())
                {
                    if (eq != std::string_view::npos)
                    {
                        throw UsageError("Option --" + std::string(opt->name) + " does not take a value.");
                    }
                    apply(o, *opt, {});
                }
                else
                {
                    apply(o, *opt, (eq != std::string_view::npos) ? body.substr(eq + 1) : take_value(*opt));
                }
                continue;
            }
            if (tok.size() > 1 && tok[0] == '-')
            {
                // bundled flags (-yq), a value attached (-e1e-10) or in the next argument (-e 1e-10)
                for (std::size_t p = 1; p < tok.size(); ++p)
                {
                    const Opt *opt = find_short(tok[p]);
                    if (opt == nullptr)
                    {
                        throw UsageError(std::string("Unknown short option: -") + tok[p]);
                    }
                    if (opt->hint.empty())
                    {
                        apply(o, *opt, {});
                        continue;
                    }
                    apply(o, *opt, (p + 1 < tok.size()) ? tok.substr(p + 1) : take_value(*opt));
                    break;
                }
                continue;
            }
            positional.push_back(tok);
        }
        std::size_t used = 0;
        if (o.input.empty() && used < positional.size())
        {
            o.input = fs::path(std::string(positional[used++]));
        }
        if (o.output.empty() && used < positional.size())
        {
            o.output = fs::path(std::string(positional[used++]));
        }
        if (used < positional.size())
        {
            throw UsageError("Unexpected positional argument: " + std::string(positional[used]));
        }
        if (!o.help && !o.version)
        {
            complete(o);
            check_files(o);
        }
        return o;
    }

    std::string help(const std::string &exe)
    {
        std::ostringstream s;
        s << "Jacobi SVD CLI (MI355X, HIP)\n\nUsage:\n  " << exe << " [OPTIONS] <input> [output]\n\nOptions:\n";
        for (const Opt &o : table())
        {
            std::string names = (o.letter != '\0') ? std::string("-") + o.letter + ", " : std::string("    ");
            names += "--" + std::string(o.name);
            if (!o.hint.empty())
            {
                names += " " + std::string(o.hint);
            }
            s << "  " << names << std::string(names.size() < 32 ? 32 - names.size() : 1, ' ') << o.text << '\n';
        }
        s << "\nExamples:\n"
          << "  " << exe << " -i experiments/inputs/a.mat -o experiments/outputs/r.mat\n"
          << "  " << exe << " experiments/inputs/a.mat --print-config\n"
          << "  " << exe << " input.txt output.txt --format txt --epsilon 1e-10\n"
          << "  " << exe << " --input a.mat --output b.txt --output-format txt --json-report --force\n"
          << "  " << exe << " --input a.mat --solver mixed --epsilon 1e-14 --max-sweeps 60\n"
          << "\nNotes:\n"
          << "  - When [output] is omitted, default output is <input-stem>.svd.{mat|txt}.\n"
          << "  - Existing output file requires --force to overwrite.\n";
        return s.str();
    }

    const char *format_name(MatrixFileFormat f)
    {
        return f == MatrixFileFormat::mat ? "mat" : f == MatrixFileFormat::txt ? "txt" : "auto";
    }

    const char *mode_name(LayoutTransposeMode m)
    {
        return m == LayoutTransposeMode::force_enable ? "on" : m == LayoutTransposeMode::force_disable ? "off" : "auto";
    }

    const char *solver_name(Solver s) { return s == Solver::block ? "block" : s == Solver::mixed ? "mixed" : "plain"; }

    void print_config(const Options &o)
    {
        const JacobiSvdConfig &k = o.kernel;
        std::cout << "input           : " << o.input.string() << '\n'
                  << "output          : " << o.output.string() << '\n'
                  << "input-format    : " << format_name(o.input_format) << '\n'
                  << "output-format   : " << format_name(o.output_format) << '\n'
                  << "epsilon         : " << k.epsilon << '\n'
                  << "max-sweeps      : " << k.max_sweeps << '\n'
                  << "threads-per-blk : " << k.threads_per_block << '\n'
                  << "layout-mode     : " << mode_name(k.layout_transpose_mode) << '\n'
                  << "layout-min-cols : " << k.layout_transpose_min_columns << '\n'
                  << "layout-min-elem : " << k.layout_transpose_min_elements << '\n'
                  << "layout-auto-tune: " << (k.layout_transpose_auto_tune ? "true" : "false") << '\n'
                  << "layout-bench-rep: " << k.layout_transpose_benchmark_repetitions << '\n'
                  << "layout-bench-swp: " << k.layout_transpose_benchmark_sweeps << '\n'
                  << "queue-capacity  : " << o.queue_capacity << '\n'
                  << "solver          : " << solver_name(o.solver) << '\n'
                  << "device-batch    : " << o.device_batch << '\n'
                  << "force-overwrite : " << (o.force ? "true" : "false") << '\n';
    }

    void print_text(const jacobi::svd::pipeline::PipelineReport &r, double ms)
    {
        std::cout << "Pipeline completed.\n"
                  << "testcases       : " << r.testcase_count << '\n'
                  << "emitted-matrices: " << r.emitted_matrix_count << '\n'
                  << "total-sweeps    : " << r.total_sweeps << '\n'
                  << "layout-mode     : " << mode_name(r.layout_transpose_mode) << '\n'
                  << "layout-min-cols : " << r.layout_transpose_min_columns << '\n'
                  << "layout-min-elem : " << r.layout_transpose_min_elements << '\n'
                  << "layout-auto-tune: " << (r.layout_transpose_auto_tuned ? "true" : "false") << '\n'
                  << "layout-best-spd : " << std::fixed << std::setprecision(3)
                  << r.layout_transpose_estimated_best_speedup << '\n'
                  << "elapsed-ms      : " << std::fixed << std::setprecision(3) << ms << '\n';
    }

    void print_json(const jacobi::svd::pipeline::PipelineReport &r, double ms)
    {
        std::cout << "{\n"
                  << "  \"testcase_count\": " << r.testcase_count << ",\n"
                  << "  \"emitted_matrix_count\": " << r.emitted_matrix_count << ",\n"
                  << "  \"total_sweeps\": " << r.total_sweeps << ",\n"
                  << "  \"layout_transpose_mode\": \"" << mode_name(r.layout_transpose_mode) << "\",\n"
                  << "  \"layout_transpose_min_columns\": " << r.layout_transpose_min_columns << ",\n"
                  << "  \"layout_transpose_min_elements\": " << r.layout_transpose_min_elements << ",\n"
                  << "  \"layout_transpose_auto_tuned\": " << (r.layout_transpose_auto_tuned ? "true" : "false")
                  << ",\n"
                  << "  \"layout_transpose_estimated_best_speedup\": " << std::fixed << std::setprecision(3)
                  << r.layout_transpose_estimated_best_speedup << ",\n"
                  << "  \"elapsed_ms\": " << std::fixed << std::setprecision(3) << ms << '\n'
                  << "}\n";
    }

    // The block / mixed solvers through the streamed *.mat path; the report carries the same fields.
    jacobi::svd::pipeline::PipelineReport run_streamed(const Options &o)
    {
        jsvd_config cfg;
        jsvd_config_default(&cfg);
        cfg.epsilon = o.kernel.epsilon;
        cfg.max_sweeps = o.kernel.max_sweeps;
        cfg.threads_per_block = o.kernel.threads_per_block;
        cfg.layout_transpose_mode = static_cast<int32_t>(o.kernel.layout_transpose_mode);
        cfg.layout_transpose_min_columns = o.kernel.layout_transpose_min_columns;
        cfg.layout_transpose_min_elements = o.kernel.layout_transpose_min_elements;
        jsvd_mixed_config mx;
        jsvd_mixed_config_default(&mx);
        mx.path = (o.solver == Solver::block) ? JSVD_PATH_BLOCK_F64 : JSVD_PATH_MIXED;
        jsvd_mat_stream_report rep{};
        const std::string in = o.input.string(), out = o.output.string();
        if (jsvd_mat_stream_solve(in.c_str(), out.c_str(), &cfg, &mx, &rep) != JSVD_OK)
        {
            throw std::runtime_error(jsvd_last_error());
        }
        jacobi::svd::pipeline::PipelineReport r{};
        r.testcase_count = static_cast<std::size_t>(rep.records);
        r.emitted_matrix_count = static_cast<std::size_t>(3 * rep.records);
        r.total_sweeps = static_cast<std::size_t>(rep.total_sweeps);
        r.layout_transpose_mode = o.kernel.layout_transpose_mode;
        r.layout_transpose_min_columns = o.kernel.layout_transpose_min_columns;
        r.layout_transpose_min_elements = o.kernel.layout_transpose_min_elements;
        return r;
    }
} // namespace

int main(int argc, char **argv)
{
    try
    {
        const std::string exe = (argc > 0 && argv[0]) ? fs::path(argv[0]).filename().string() : "jacobi-svd";
        const Options o = parse(argc, argv);
        if (o.help)
        {
            std::cout << help(exe);
            return 0;
        }
        if (o.version)
        {
            std::cout << "jacobi-svd-cuda CLI v0.1.0 (MI355X gfx950 build)\n";
            return 0;
        }
        if (o.dry_run)
        {
            std::cout << "Dry run: pipeline was not executed.\n";
            print_config(o);
            return 0;
        }
        if (o.print_config)
        {
            std::cout << "Effective configuration:\n";
            print_config(o);
            std::cout << '\n';
        }
        const auto t0 = std::chrono::steady_clock::now();
        jacobi::svd::pipeline::PipelineReport report;
        if (o.solver == Solver::plain)
        {
            jacobi::svd::pipeline::PipelineConfig pc{};
            pc.input_path = o.input;
            pc.output_path = o.output;
            pc.input_format = o.input_format;
            pc.output_format = o.output_format;
            pc.max_queued_results = o.queue_capacity;
            pc.kernel_config = o.kernel;
            pc.device_batch_cases = o.device_batch;
            report = jacobi::svd::pipeline::run_pipeline(pc);
        }
        else
        {
            report = run_streamed(o);
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!o.quiet)
        {
            print_text(report, ms);
        }
        if (o.json_report)
        {
            print_json(report, ms);
        }
        return 0;
    }
    catch (const UsageError &e)
    {
        std::cerr << "Argument error: " << e.what() << '\n' << "Use --help for usage.\n";
        return 2;
    }
    catch (const std::exception &e)
    {
        std::cerr << "Execution failed: " << e.what() << '\n';
        return 1;
    }
}
