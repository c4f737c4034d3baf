// pipeline.hpp -- file-to-file SVD pipeline, source-compatible with the reference's include/pipeline.hpp
// (kleedaisuki/jacobi-svd-cuda, namespace jacobi::svd::pipeline).
//
// Reads test cases from a *.mat or *.txt file, solves each one on the GPU and writes, in input order,
// three matrices per case: U (m x n), Sigma (1 x n), V (n x n).
#pragma once

#include "kernels.hpp"

#include <cstddef>
#include <filesystem>

namespace jacobi::svd::pipeline
{
    // Matrix file format (reference pipeline.hpp:13-29); auto_detect resolves by extension (.mat / .txt).
    enum class MatrixFileFormat
    {
        auto_detect,
        mat,
        txt
    };

    // Pipeline configuration (reference pipeline.hpp:34-65).  max_queued_results bounds the number of
    // solved-but-unwritten cases (backpressure on the reader).
    struct PipelineConfig final
    {
        std::filesystem::path input_path;
        std::filesystem::path output_path;
        MatrixFileFormat input_format = MatrixFileFormat::auto_detect;
        MatrixFileFormat output_format = MatrixFileFormat::auto_detect;
        std::size_t max_queued_results = 4;
        JacobiSvdConfig kernel_config{};
        // Appended (no reference counterpart): when > 0, runs of consecutive small cases (rows <= 256,
        // columns <= 128, rows >= columns) are grouped, up to this many, into one persistent batched launch
        // (jsvd_batch_solve_host) instead of one solver-lane solve each.  Both follow the reference algorithm
        // (pair order, rotation and skip test, convergence rule), so results agree to rounding with the same
        // column order and signs; 0 (default) keeps every case on a solver lane, bit-equal to
        // one_sided_jacobi_svd.
        std::size_t device_batch_cases = 0;
    };

    // Run summary (reference pipeline.hpp:70-111).
    struct PipelineReport final
    {
        std::size_t testcase_count = 0;
        std::size_t emitted_matrix_count = 0;
        std::size_t total_sweeps = 0;
        LayoutTransposeMode layout_transpose_mode = LayoutTransposeMode::auto_select;
        int layout_transpose_min_columns = 16;
        std::size_t layout_transpose_min_elements = 4096;
        bool layout_transpose_auto_tuned = false;
        double layout_transpose_estimated_best_speedup = 1.0;
    };

    // The pipeline (reference pipeline.hpp:117-138).  run() throws the first reader, solver or writer
    // error (std::invalid_argument, CudaError, std::runtime_error, std::system_error).
    class JacobiSvdPipeline final
    {
    public:
        explicit JacobiSvdPipeline(PipelineConfig config);
        [[nodiscard]] PipelineReport run() const;

    private:
        PipelineConfig config_;
    };

    [[nodiscard]] PipelineReport run_pipeline(const PipelineConfig &config);
} // namespace jacobi::svd::pipeline
