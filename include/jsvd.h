/*
 * jsvd.h -- C ABI of the MI355X (gfx950) one-sided Jacobi SVD library, libjsvd.so.
 *
 * This is the drop-in boundary for the reference's plain FP64 solver (kleedaisuki/jacobi-svd-cuda).
 * The C++ entry points of include/kernels.hpp (namespace jacobi::svd) are thin wrappers over these
 * functions; a ctypes / FFI binding binds these symbols directly.  All pointers are plain host or
 * device pointers, all sizes are int64_t, and every function returns a status code:
 *
 *   JSVD_OK (0)                    success
 *   JSVD_ERR_INVALID_ARGUMENT (1)  precondition violated (maps to std::invalid_argument)
 *   JSVD_ERR_DEVICE (2)            HIP runtime failure      (maps to jacobi::svd::CudaError)
 *   JSVD_ERR_INTERNAL (3)          anything else            (maps to std::runtime_error)
 *
 * jsvd_last_error() returns the thread-local message of the last failing call on this thread.
 *
 * Matrices crossing the boundary are row-major, exactly as in the reference: A is m x n, U is
 * m x n, V is n x n, sigma has n entries in Jacobi column order (not sorted), m >= n.
 */
#ifndef JSVD_H
#define JSVD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JSVD_OK 0
#define JSVD_ERR_INVALID_ARGUMENT 1
#define JSVD_ERR_DEVICE 2
#define JSVD_ERR_INTERNAL 3

/* layout_transpose_mode values, mirroring jacobi::svd::LayoutTransposeMode (include/kernels.hpp:13-29). */
#define JSVD_LAYOUT_AUTO 0
#define JSVD_LAYOUT_FORCE_ENABLE 1
#define JSVD_LAYOUT_FORCE_DISABLE 2

/*
 * Solver configuration.  Field-for-field the reference's JacobiSvdConfig (include/kernels.hpp:78-124),
 * same defaults (jsvd_config_default).  threads_per_block is validated and normalised as the reference
 * does (clamped to [32,1024], rounded up to a multiple of 32; src/kernels.cu:685-690) and reported, but the
 * launch shape itself is chosen per matrix shape in wave64 units.
 */
typedef struct jsvd_config
{
    double epsilon;                                 /* kernels.hpp:83, default 1e-9 */
    int32_t max_sweeps;                             /* kernels.hpp:88, default 128 */
    int32_t threads_per_block;                      /* kernels.hpp:93, default 256 */
    int32_t layout_transpose_mode;                  /* kernels.hpp:98, JSVD_LAYOUT_* */
    int32_t layout_transpose_min_columns;           /* kernels.hpp:103, default 16 */
    uint64_t layout_transpose_min_elements;         /* kernels.hpp:108, default 4096 */
    int32_t layout_transpose_auto_tune;             /* kernels.hpp:113, default 0 */
    int32_t layout_transpose_benchmark_repetitions; /* kernels.hpp:118, default 2 */
    int32_t layout_transpose_benchmark_sweeps;      /* kernels.hpp:123, default 8 */
    int32_t reserved;
} jsvd_config;

/* Fills *cfg with the reference defaults (include/kernels.hpp:83-123). */
void jsvd_config_default(jsvd_config *cfg);

/* Checks *cfg as every solve does (the reference's validation, src/kernels.cu:807-833: epsilon > 0,
 * max_sweeps > 0, layout and benchmark knobs), without a matrix: JSVD_OK or JSVD_ERR_INVALID_ARGUMENT with the
 * reference's message.  For callers that route work before solving (the pipeline's batched cases). */
int jsvd_config_validate(const jsvd_config *cfg);

/* Thread-local message of the last failing call on the calling thread ("" if none). */
const char *jsvd_last_error(void);

/* Library version string. */
const char *jsvd_version(void);

/*
 * Plain FP64 one-sided Jacobi SVD of a host row-major m x n matrix.
 * Replaces jacobi::svd::one_sided_jacobi_svd (include/kernels.hpp:295-298, src/kernels.cu:1066-1076):
 * same validation (src/kernels.cu:807-833), pair ordering (src/kernels.cu:637-678), rotation and skip
 * test (src/kernels.cu:485-521), convergence rule and sweep count (src/kernels.cu:889-981) and sigma/U
 * extraction (src/kernels.cu:586-629).  u (m*n), sigma (n) and v (n*n) are host buffers owned by the
 * caller; sweeps receives the executed sweep count (== max_sweeps when not converged, not an error).
 */
int jsvd_plain_f64(const double *a, int64_t m, int64_t n, const jsvd_config *cfg,
                   double *u, double *sigma, double *v, int32_t *sweeps);

/*
 * Reusable solver plan for one shape: owns the device workspace and a HIP stream, so repeated solves of
 * the same shape allocate nothing.  No reference counterpart (the reference allocates per call,
 * src/kernels.cu:854-885); it backs jsvd_plain_f64 and the device-resident entry point below.
 * stream may be NULL (the plan then creates its own non-blocking stream) or a hipStream_t.
 */
typedef struct jsvd_plan jsvd_plan;
int jsvd_plan_create(int64_t m, int64_t n, const jsvd_config *cfg, void *stream, jsvd_plan **plan);
int jsvd_plan_destroy(jsvd_plan *plan);

/*
 * Device-resident solve: d_a (row-major m x n), d_u (m x n), d_sigma (n), d_v (n x n) are device
 * pointers.  d_a is not modified.  Work is enqueued on the plan's stream; the call returns after the
 * solve completes (the host must learn the sweep count).  Same semantics as jsvd_plain_f64.
 */
int jsvd_plan_solve_device(jsvd_plan *plan, const double *d_a, double *d_u, double *d_sigma, double *d_v,
                           int32_t *sweeps);

/* Host-pointer solve through a plan (H2D copy, device solve, D2H copies). */
int jsvd_plan_solve_host(jsvd_plan *plan, const double *a, double *u, double *sigma, double *v, int32_t *sweeps);

/*
 * Per-sweep diagnostics of the last solve on this plan: rotated-pair count and the largest
 * |gamma|/sqrt(alpha*beta) over all pairs of each sweep (the off-orthogonality measure of the
 * convergence test).  Block and mixed plans list their FP32 sweeps first, then their FP64 sweeps.
 * Arrays of length at least `capacity`; returns the number of sweeps written.
 */
int jsvd_plan_sweep_stats(const jsvd_plan *plan, int32_t *rotations, double *offmax, int32_t capacity);

/* ---------------------------------------------------------------------------------------------------
 * Block and mixed-precision paths.  No reference counterpart: they add precision options beside the
 * reference's plain FP64 solver (one_sided_jacobi_svd, include/kernels.hpp:295-298) and keep its output
 * contract (row-major U, sigma in Jacobi column order, row-major V, sweep counts, the same convergence
 * rule "stop after a sweep without rotations", applied to block sweeps).
 * ------------------------------------------------------------------------------------------------- */
#define JSVD_PATH_PLAIN_F64 0 /* column-pair cyclic Jacobi, the reference algorithm */
#define JSVD_PATH_BLOCK_F64 1 /* block Jacobi (16/32/64-column blocks), FP64 MFMA Gram and rotate */
#define JSVD_PATH_MIXED 2     /* FP64 QR, FP32 MFMA block sweeps, FP64 re-orthogonalisation, FP64 finishing */

typedef struct jsvd_mixed_config
{
    int32_t path;                /* JSVD_PATH_* (default JSVD_PATH_MIXED) */
    int32_t fp32_max_sweeps;     /* cap of the FP32 block sweeps (default 40) */
    double fp32_tolerance;       /* FP32 phase threshold; <= 0 selects max(1e-6, 4 sqrt(m) 2^-24) */
    int32_t orthogonalize_steps; /* Newton-Schulz steps applied to the FP32 V in FP64 (default 2) */
    int32_t block_width;         /* columns per block: 16, 32 or 64; 0 (default) = jsvd_autotune's choice for the
                                    shape when it measured one, else 32; halved while n <= width */
    int32_t qr_precondition;     /* mixed path: 1 = factor A = Q R in FP64 first and sweep the n x n
                                    R^T (sigma_j = |x_j|, V from the normalised columns X, U = Q V_M); 2 = also
                                    factor R^T = Q2 R2 and sweep R2^T (an LQ step of R: the same singular values,
                                    columns closer to orthogonal; U = Q Y, V = Q2 V_N): 3-4 fewer FP32 sweeps on
                                    kappa = 1e8 inputs with random singular vectors (4096 x 1024: 19.6-20.1 ms
                                    against 21.1-21.3 ms), none saved on well-conditioned ones, and the second
                                    QR costs about 2.4 ms at n = 1024; 3 (default) = 1, plus the second QR when
                                    n >= 512 and max|r_ii| / min|r_ii| of the first R is >= 100 (results equal
                                    bit for bit those of the form chosen); 0 = sweep A.  A factor that breaks
                                    down falls back to the next simpler form. */
    int32_t inner_order;         /* block rounds at width 32: JSVD_INNER_CROSS (default) = only the pairs across
                                    the two blocks (32 inner rounds), plus one step per sweep that sweeps every
                                    block by itself; JSVD_INNER_FULL = one cyclic sweep over all 64 columns of the
                                    block pair (63 inner rounds).  Other widths always use the full sweep. */
    double fp32_stop_fraction;   /* mixed path: the FP32 phase (a preconditioner of the FP64 phase) ends after the
                                    first sweep that rotates at most this fraction of the column pairs; 0 = only
                                    after a rotation-free sweep.  Default 0.05: the rotation count falls from
                                    about a third of the pairs to 1-2 % within one sweep near the FP32 limit, so
                                    every fraction from 0.03 to 0.2 stops at that sweep, without a longer FP64
                                    phase on the measured workloads.  The FP64 phase always ends on a
                                    rotation-free sweep (the reference's rule). */
} jsvd_mixed_config;

#define JSVD_INNER_CROSS 0
#define JSVD_INNER_FULL 1

/* Per-solve report of a plan. */
typedef struct jsvd_solve_info
{
    int32_t path;        /* path that ran */
    int32_t sweeps_fp32; /* FP32 block sweeps (0 outside the mixed path) */
    int32_t sweeps_fp64; /* FP64 sweeps (column-pair or block); also the value returned as `sweeps` */
    int32_t converged;   /* 1 when the FP64 phase ended on a rotation-free sweep */
    double offmax_fp32;  /* max pair cosine seen in the last FP32 sweep */
    double offmax_fp64;  /* max pair cosine seen in the last FP64 sweep */
} jsvd_solve_info;

void jsvd_mixed_config_default(jsvd_mixed_config *mixed);

/*
 * Plan for the path selected by mixed->path (NULL: the plain FP64 path, identical to jsvd_plan_create).
 * The block paths need at least two 32-column blocks (n >= 33); smaller n is rejected with
 * JSVD_ERR_INVALID_ARGUMENT.
 */
int jsvd_plan_create_ex(int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed, void *stream,
                        jsvd_plan **plan);

/* Report of the last solve on this plan. */
int jsvd_plan_last_info(const jsvd_plan *plan, jsvd_solve_info *info);

/*
 * Re-targets a plan at an m x n problem (m >= n) and epsilon without reallocating: the shape must fit the
 * buffers allocated for the shape the plan was created with (m <= created m, n <= created n; block paths also
 * need n >= 33), else JSVD_ERR_INVALID_ARGUMENT and the plan is unchanged.  Lets one plan per stream serve
 * requests of many shapes (the request loop keeps one such plan per lane).  No reference counterpart: the
 * reference allocates per call (src/kernels.cu:844-885).
 */
int jsvd_plan_reshape(jsvd_plan *plan, int64_t m, int64_t n, double epsilon);

/*
 * Round spans of a plain FP64 plan: the sweep's rounds run as an even number of launches of up to max_rounds
 * rounds each instead of one launch per round (src/kernels.cu:889-981 launches the two kernels of every round
 * from the host).  Each workgroup owns a range of the round-robin circle's slots and also computes the slots that
 * feed them (temporal blocking with redundant halos), one wavefront per slot, the columns in registers, so
 * workgroups never wait for each other; the launches alternate between two buffers.  Results are bit-identical to
 * one launch per round.  waves = 8 or 16 wavefronts per workgroup, or 0 for one launch per round.  Plans choose
 * spans by themselves (the autotuner's measured choice, else spans for m n <= 2^19 where the build has the
 * kernel); this call overrides that until the next reshape.  Fails (JSVD_ERR_INVALID_ARGUMENT) when the shape
 * has no span kernel (m <= 1024, n <= 512, n >= 3 with the plan's round shape), when the plan's capacity is above
 * 1024 x 512 (no second buffers), or for block / mixed plans.
 */
int jsvd_plan_set_round_span(jsvd_plan *plan, int32_t waves, int32_t max_rounds);

/* The plan's current round-span setting: waves (0 = one launch per round), rounds per launch at most, and the
 * launches one sweep takes. */
int jsvd_plan_round_span(const jsvd_plan *plan, int32_t *waves, int32_t *max_rounds, int32_t *launches_per_sweep);

/*
 * The launch parameters a plan currently uses: for the plain path the fused round kernel's threads per column
 * pair and register chunk counts (0 = streamed), block fields 0; for the block paths the apply kernel's total
 * workgroups and the block width, round fields 0.  They come from the mixed config's block_width, from
 * jsvd_autotune's cache when it measured the plan's shape, else from the built-in rules.  No reference
 * counterpart.
 */
int jsvd_plan_launch_params(const jsvd_plan *plan, int32_t *round_threads, int32_t *round_ea, int32_t *round_ev,
                            int32_t *apply_workgroups, int32_t *block_width);

/* One-shot host-pointer solve through the path selected by `mixed` (NULL: plain FP64). */
int jsvd_solve_f64(const double *a, int64_t m, int64_t n, const jsvd_config *cfg, const jsvd_mixed_config *mixed,
                   double *u, double *sigma, double *v, jsvd_solve_info *info);

/*
 * Cholesky factorisation A = R^T R of the symmetric positive definite n x n FP64 matrix a (device, column-major,
 * leading dimension lda >= n rounded up to 64): the upper triangle of a is overwritten by R, the padding rows
 * and columns up to the next multiple of 64 are set to the identity (so a needs that room), the strictly lower
 * triangle is left as it was.  *info (device int32) becomes 0, or 1 + the first row whose pivot was not
 * positive.  Tiled right-looking factorisation (64 x 64 tiles, FP64 MFMA trailing updates); the building block
 * of the mixed path's Cholesky QR.  No reference counterpart.
 */
int jsvd_cholesky_f64(double *a, int64_t lda, int64_t n, int32_t *info, void *stream);

/* ---------------------------------------------------------------------------------------------------
 * Batched path: many independent matrices of different shapes solved by one persistent launch.  No
 * reference counterpart as a batch (the reference solves one matrix per call, pipeline.cpp:711-739); per
 * matrix it runs the reference algorithm (pair ordering src/kernels.cu:637-678, rotation and skip test
 * src/kernels.cu:485-521, convergence rule src/kernels.cu:889-981, sigma/U extraction
 * src/kernels.cu:586-629) in the requested precision, rotation parameters always in FP64.
 * ------------------------------------------------------------------------------------------------- */
#define JSVD_PRECISION_F64 0
#define JSVD_PRECISION_F32 1

/* Smallest column count a batch item may have.  Every item needs JSVD_BATCH_MIN_COLUMNS <= columns <=
 * max_columns and columns <= rows <= max_rows (the batch object's limits), max_sweeps >= 1, tolerance > 0, a
 * known precision and offsets inside the packed buffers.  The host-pointer entry point checks every item and
 * fails with JSVD_ERR_INVALID_ARGUMENT before anything is launched; for the device-pointer entry points this is
 * the caller's precondition (the kernel still writes sweeps = -1 for an item outside the limits and skips it).
 * Matrices with fewer columns go through the plan entry points (jsvd_plan_*). */
#define JSVD_BATCH_MIN_COLUMNS 4

/* One matrix of a batch.  Offsets are element offsets into the batch's packed FP64 buffers: A row-major
 * rows x columns at a_offset, U row-major rows x columns at u_offset, sigma (columns) at sigma_offset,
 * V row-major columns x columns at v_offset.  tolerance is the matrix's epsilon. */
typedef struct jsvd_batch_item
{
    int64_t a_offset;
    int64_t u_offset;
    int64_t sigma_offset;
    int64_t v_offset;
    int32_t rows;
    int32_t columns;
    int32_t precision;  /* JSVD_PRECISION_* */
    int32_t max_sweeps;
    double tolerance;
} jsvd_batch_item;

/* Sizes (in doubles) of the packed buffers, checked by the kernel for every item. */
typedef struct jsvd_batch_buffers
{
    const double *a;
    double *u;
    double *sigma;
    double *v;
    int32_t *sweeps; /* one per item; -1 for an item rejected by the kernel's bounds checks */
    int64_t a_len;
    int64_t u_len;
    int64_t sigma_len;
    int64_t v_len;
} jsvd_batch_buffers;

typedef struct jsvd_batch jsvd_batch;

/* Batch solver for matrices up to max_rows x max_columns (max_columns <= max_rows <= 4096).  Items whose A or
 * V do not fit the workgroup's LDS use a global scratch slice per resident workgroup; that area is capped at
 * 8 GiB per batch object (very large limits run fewer resident workgroups). */
int jsvd_batch_create(int64_t max_rows, int64_t max_columns, void *stream, jsvd_batch **batch);
int jsvd_batch_destroy(jsvd_batch *batch);

/*
 * Solves `count` items whose descriptors are in device memory (d_items), with device buffers `buf`.
 * One persistent launch: each workgroup takes the next item from a device queue until none is left, so
 * items are processed in descriptor order (put expensive ones first for load balance).  Each item's result
 * is bit-identical whatever else is in the batch.  Returns after completion.
 */
int jsvd_batch_solve_device(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                            const jsvd_batch_buffers *buf);

/*
 * Enqueues the same work on the batch's stream and returns at once; `done` (optional, a hipEvent_t) is
 * recorded after it.  For the request loop, which overlaps several batches.
 */
int jsvd_batch_solve_async(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                           const jsvd_batch_buffers *buf, void *done);

/*
 * Host-pointer form of jsvd_batch_solve_device: item descriptors and the packed buffers (lengths in elements,
 * as in jsvd_batch_buffers) live in host memory; the batch object copies them through device staging it keeps
 * and grows, solves in one persistent launch and copies U, sigma, V and the sweep counts back.  Synchronous.
 * No reference counterpart; the pipeline's batched cases use it.
 */
int jsvd_batch_solve_host(jsvd_batch *batch, const jsvd_batch_item *items, int64_t count, const double *a,
                          int64_t a_len, double *u, int64_t u_len, double *sigma, int64_t sigma_len, double *v,
                          int64_t v_len, int32_t *sweeps);

/*
 * Allocation-free form of jsvd_batch_solve_async.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No batch object: the same persistent kernel runs with its queue counter and workgroup scratch
 * slices in a fixed device array of 128 MiB inside the library (one copy per GPU, held from the first call on
 * that GPU until the process ends); the caller passes the item descriptors and packed buffers in device memory
 * (items under the contract of JSVD_BATCH_MIN_COLUMNS), the limits max_rows x max_columns its items stay within,
 * and a stream.  Nothing is allocated and nothing is waited for: the work is enqueued on `stream` (a
 * hipStream_t; null = the default stream) and the call returns.  The limits fix the scratch slice per
 * workgroup, ((max_rows + 1) + (max_columns + 1)) * max_columns doubles; limits whose slice does not fit the
 * array are refused with JSVD_ERR_INVALID_ARGUMENT, and the number of resident workgroups is capped by the
 * slices it holds (e.g. 339 at 256 x 128).  The launch shape is the one jsvd_batch_autotune recorded for these
 * limits, else the default.  Results are bit-identical to jsvd_batch_solve_device's.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (on different streams, or from different threads without
 * ordering) race on it and give undefined results; calls ordered on one stream are safe.
 */
int jsvd_batch_solve_noalloc(const jsvd_batch_item *d_items, int64_t count, const jsvd_batch_buffers *buf,
                             int64_t max_rows, int64_t max_columns, void *stream);

/*
 * Launch shape of the persistent batched kernel (no reference counterpart): threads per workgroup (the
 * wavefronts sharing one matrix's column pairs) and the LDS bytes each workgroup stages its matrix in.  The
 * shapes are (256, 77824) (the default), (128, 38912), (256, 38912), (512, 77824), (512, 155648) and
 * (1024, 155648); (0, 0) selects the default.  Every shape gives bit-identical results; only speed differs.
 * A batch created for limits that jsvd_batch_autotune has measured in this process starts with that choice.
 */
#define JSVD_BATCH_MAX_SHAPES 8
int jsvd_batch_set_shape(jsvd_batch *batch, int32_t threads, int32_t lds_bytes);
int jsvd_batch_get_shape(const jsvd_batch *batch, int32_t *threads, int32_t *lds_bytes, int32_t *workgroups);

/* Per-shape timings of one jsvd_batch_autotune call. */
typedef struct jsvd_batch_tune_report
{
    int32_t candidates;                       /* shapes measured */
    int32_t chosen;                           /* index of the fastest, now set on the batch */
    int32_t threads[JSVD_BATCH_MAX_SHAPES];
    int32_t lds_bytes[JSVD_BATCH_MAX_SHAPES];
    int32_t workgroups[JSVD_BATCH_MAX_SHAPES]; /* resident workgroups of the launch */
    int32_t reserved;
    double ms[JSVD_BATCH_MAX_SHAPES];         /* mean time of one solve of the whole batch */
} jsvd_batch_tune_report;

/*
 * Batch autotuner (no reference counterpart): solves the given batch `repetitions` times (after one warm-up)
 * with every launch shape, keeps the fastest on `batch` and records it for its limits (max_rows, max_columns),
 * so batches created later with the same limits (the request loop's) start with it.  The buffers receive the
 * batch's results, as after jsvd_batch_solve_device.  Synchronous.
 */
int jsvd_batch_autotune(jsvd_batch *batch, const jsvd_batch_item *d_items, int64_t count,
                        const jsvd_batch_buffers *buf, int32_t repetitions, jsvd_batch_tune_report *report);

/* ---------------------------------------------------------------------------------------------------
 * Request loop (continuous batching).  No reference counterpart (the reference pipeline solves a file of
 * cases with one thread per case, pipeline.cpp:754-863).  Requests are admitted against an HBM budget;
 * `lanes` worker lanes (a host thread and a HIP stream each) take jobs: one large request through a solver
 * plan, or every admitted small request at once through one persistent batched launch.  Requests that arrive
 * while jobs run are admitted as budget frees up and form the next job, so work of different requests
 * overlaps on the device.  A request's result does not depend on what it was batched with.
 * ------------------------------------------------------------------------------------------------- */
typedef struct jsvd_server_config
{
    int32_t lanes;             /* concurrent jobs (default 8; measured at 2,500 requests/s of the serving mix: 4
                                  lanes 1,520/s, 8 lanes 1,631/s, 12 lanes 1,644/s) */
    int32_t small_max_rows;    /* requests with rows <= this and columns <= small_max_columns are batched (256) */
    int32_t small_max_columns; /* (128) */
    int32_t max_batch;         /* requests per batched launch (4096) */
    int64_t hbm_budget_bytes;  /* admission budget for request state on the device (64 GiB) */
    int32_t max_sweeps;        /* per request (60) */
    int32_t large_path;        /* JSVD_PATH_* for requests above the batch limits (JSVD_PATH_PLAIN_F64) */
} jsvd_server_config;

typedef struct jsvd_request_info
{
    int32_t status;     /* JSVD_OK, or the status of this request's failure */
    int32_t sweeps;     /* executed sweeps (FP64 sweeps on the plan paths) */
    int32_t precision;  /* JSVD_PRECISION_* chosen from the requested relative error */
    int32_t batch_size; /* requests in the launch that served this one (1 for a plan) */
    double queued_s;    /* submission -> start of its job */
    double latency_s;   /* submission -> results ready on the host */
} jsvd_request_info;

typedef struct jsvd_server_stats
{
    int64_t submitted;
    int64_t completed;
    int64_t batches;        /* batched launches */
    int64_t plans;          /* plan solves */
    int64_t budget_in_use;  /* bytes currently reserved */
    int64_t budget_peak;    /* highest reservation seen */
    int64_t max_in_flight;  /* most requests admitted at once */
} jsvd_server_stats;

typedef struct jsvd_server jsvd_server;

void jsvd_server_config_default(jsvd_server_config *cfg);
int jsvd_server_create(const jsvd_server_config *cfg, jsvd_server **server);
/* Waits for every submitted request, then releases the lanes. */
int jsvd_server_destroy(jsvd_server *server);

/*
 * Submits a row-major m x n host matrix (copied before the call returns) with a requested relative error
 * of sigma: relerr >= 1e-7 is served in FP32 (tolerance 1e-6), smaller in FP64 (tolerance
 * max(1e-14, 2 sqrt(m) 2^-53)).  Returns at once with the request id.
 */
int jsvd_server_submit(jsvd_server *server, const double *a, int64_t m, int64_t n, double relerr, uint64_t *id);

/* Blocks until request `id` is done, copies U (m x n), sigma (n), V (n x n) out and forgets the request. */
int jsvd_server_wait(jsvd_server *server, uint64_t id, double *u, double *sigma, double *v, jsvd_request_info *info);

int jsvd_server_get_stats(jsvd_server *server, jsvd_server_stats *stats);

/* ---------------------------------------------------------------------------------------------------
 * Streamed *.mat solve (file in, file out) on one GPU.
 * ------------------------------------------------------------------------------------------------- */

typedef struct jsvd_mat_stream_report
{
    int64_t records;          /* matrices solved */
    int64_t payload_bytes;    /* input payload bytes moved to HBM */
    double seconds_total;     /* open to last result written */
    double seconds_load;      /* loader thread: mmap -> pinned copy, H2D copy and device decode */
    double seconds_solve;     /* solver thread: solves (the load of the next record overlaps these) */
    double seconds_output;    /* writer thread: wait for a record's encoded results, append them to the output
                                 file (overlaps the next record's solve) */
    int32_t sweeps_fp32_last; /* sweeps of the last record */
    int32_t sweeps_fp64_last;
    int64_t total_sweeps;     /* FP32 + FP64 sweeps over all records */
} jsvd_mat_stream_report;

/*
 * Solves every record of the *.mat file input_path (u64_be rows | u64_be cols | rows*cols f64_be, row-major,
 * to end of file) and, when output_path is not NULL, writes U (m x n), sigma (1 x n) and V (n x n) per record,
 * in input order, as *.mat records: the reference pipeline's *.mat-to-*.mat run (src/pipeline.cpp:754-863 with
 * the record reader of src/io.cu:1101-1168 and the packets of src/pipeline.cpp:711-739).  The file is
 * memory-mapped; a loader thread copies record k+1 into a pinned staging buffer (two, alternating), sends it to
 * HBM on a copy stream and decodes it there (big-endian to double) while record k is solved on a compute
 * stream; record k's U, sigma and V are encoded big-endian and copied back on an output stream and appended by
 * a writer thread while record k + 1 is solved (two result slots).  config / mixed select the solver (NULL:
 * defaults, plain FP64); records narrower than 33 columns take the plain path.  Every record's shape is
 * validated before the first solve (m >= n > 0, else JSVD_ERR_INVALID_ARGUMENT); truncated files fail with the
 * reader's messages.
 */
int jsvd_mat_stream_solve(const char *input_path, const char *output_path, const jsvd_config *config,
                          const jsvd_mixed_config *mixed, jsvd_mat_stream_report *report);

/* Result of jsvd_autotune: the launch parameters chosen for one shape and the timings behind them. */
typedef struct jsvd_tune_report
{
    int32_t candidates_tried;      /* timed candidates (round-kernel shapes + apply grid sizes) */
    int32_t round_threads;         /* plain path: threads per column pair of the fused round kernel */
    int32_t round_ea;              /* plain path: register chunks per thread for an A column (0 = streamed) */
    int32_t round_ev;              /* plain path: register chunks per thread for a V column (0 = streamed) */
    int32_t apply_workgroups;      /* block paths: total workgroups of one apply launch */
    int32_t block_width;           /* block sweeps over m x n: chosen block width (16, 32 or 64; 0 if n <= 16) */
    int32_t rt_block_width;        /* block sweeps over n x n (the mixed path's R^T): chosen block width */
    int32_t rt_apply_workgroups;   /* block sweeps over n x n: chosen apply workgroups (0 if n <= 16) */
    double round_us_default;       /* us per round of the shape chosen by the built-in rule */
    double round_us_best;          /* us per round of the chosen shape */
    double block_sweep_us[3];      /* estimated us per FP32 block sweep at widths 16, 32, 64 (512 apply
                                      workgroups; 0 where the width leaves fewer than two blocks) */
    double block_round_us_default; /* us per FP32 block round at the chosen width, 512 apply workgroups */
    double block_round_us_best;    /* us per FP32 block round at the chosen width and workgroup count */
    int32_t span_waves;            /* plain path: wavefronts per round-span workgroup chosen (0 = one launch
                                      per round) */
    int32_t span_rounds;           /* plain path: rounds per span launch at most (0 = one launch per round) */
    double sweep_us_rounds;        /* plain path: us per sweep with one launch per round (chosen shape) */
    double sweep_us_span;          /* plain path: us per sweep with the chosen spans (0 when none was timed) */
} jsvd_tune_report;

/*
 * Autotuner of the launch parameters for m x n (m >= n >= 1): times on the device, on seeded data of that
 * shape, every round-kernel launch shape that suits it (16 rounds of a first sweep each), the block width
 * (16 / 32 / 64 columns: up to 8 FP32 block rounds each, compared per sweep, i.e. per round times the rounds a
 * sweep takes at that width) and, at the chosen width, the block apply kernel's grid size (256 / 512 / 1024
 * workgroups), keeping the best of `repetitions` timings per candidate.  The block choice is made for block
 * sweeps over m x n and, when m > n, also over n x n (the mixed path sweeps the n x n R^T of A = Q R).  The winners are cached for the process: plans created or reshaped
 * to m x n afterwards use them, except that layout_transpose_mode FORCE_ENABLE / FORCE_DISABLE still fix the
 * round kernel's shape.  stream may be NULL (a private stream).  No counterpart in the reference, whose only
 * tuning is the layout-transpose threshold search (src/kernels.cu:1078-1127; auto_tune_layout_transpose_threshold
 * of include/kernels.hpp here).
 */
int jsvd_autotune(int64_t m, int64_t n, int32_t repetitions, void *stream, jsvd_tune_report *report);

/* Forgets every cached jsvd_autotune result (plans created afterwards use the built-in rules). */
int jsvd_autotune_clear(void);

/*
 * Host evaluation of the closed-form round-robin schedule the kernels use: writes the (p, q) pairs of
 * round `round` of n columns into pairs_out (2 * (n+1)/2 int32 capacity), in slot order, dropping the
 * pair of the dummy column when n is odd, and returns the pair count (negative status on error).
 * Equals build_round_robin_schedule (src/kernels.cu:637-678) round by round, orientation included.
 */
int jsvd_round_robin_pairs(int64_t n, int32_t round, int32_t *pairs_out);

/*
 * The inner step of a block round alone (block and mixed paths): for each of `pairs` independent kw x kw
 * Gram matrices gram[p] (kw = 2 block_width, block_width 16, 32 or 64; row-major, device, upper triangle read;
 * float when precision is JSVD_PRECISION_F32, double when JSVD_PRECISION_F64), one cyclic Jacobi sweep in Gram
 * space in the round-robin order of kw columns, rotation parameters in FP64 with the reference's skip test
 * (src/kernels.cu:503-511 applied to the Gram entries).  q[p] (row-major kw x kw, same type) receives the
 * accumulated rotation Q (G <- Q^T G Q), rotations[p] (device int32) the number of rotated column pairs.
 * No reference counterpart: the reference has no block path.
 */
int jsvd_block_eigen(int32_t precision, int32_t block_width, const void *gram, int64_t pairs, double epsilon,
                     void *q, int32_t *rotations, void *stream);

/*
 * jsvd_block_eigen in a chosen inner ordering: JSVD_INNER_FULL = the full cyclic sweep above; JSVD_INNER_CROSS
 * (block_width 32 only) = the ordering of the block rounds' default: only the 32 x 32 pairs (i, 32 + j) across
 * the two halves, in 32 inner rounds (round r pairs column u with column 32 + (u + r) mod 32).  The same launch
 * is enqueued `repeats` times back to back (1..4096; each writes the same q and rotations), so it can be timed
 * without host gaps.  No reference counterpart.
 */
int jsvd_block_eigen_ordered(int32_t precision, int32_t block_width, int32_t inner_order, const void *gram,
                             int64_t pairs, double epsilon, void *q, int32_t *rotations, int32_t repeats,
                             void *stream);

/* ---------------------------------------------------------------------------------------------------
 * Kernel-level entry points (device pointers, enqueued on `stream`, asynchronous).  They expose the
 * four kernels of the plain path for parity tests and roofline measurement.
 * ------------------------------------------------------------------------------------------------- */

/*
 * Out-of-place transpose of a row-major rows x cols matrix src (leading dimension lds >= cols) into
 * dst as row-major cols x rows (leading dimension ldd >= rows).  LDS-tiled, coalesced both ways.
 * Replaces transpose_row_major_kernel and its launchers (src/kernels.cu:246-321).
 */
int jsvd_transpose_f64(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                       void *stream);

/*
 * Column-pair Gram statistics: for each k < count, with (p,q) = pairs[2k], pairs[2k+1],
 * alpha[k] = |a_p|^2, beta[k] = |a_q|^2, gamma[k] = a_p . a_q over the m rows of the column-major
 * matrix a (leading dimension lda).  Replaces pair_stats_kernel (src/kernels.cu:412-472).
 * pairs is a device int32 array.
 */
int jsvd_pair_gram_f64(const double *a, int64_t lda, int64_t m, const int32_t *pairs, int64_t count,
                       double *alpha, double *beta, double *gamma, void *stream);

/*
 * One fused round of the plain path: for every pair (p,q) of round `round` of the round-robin schedule
 * of n columns, compute (alpha, beta, gamma), the rotation (c, s) in FP64 with the reference's skip
 * test, and apply it to columns p, q of the column-major a (m rows, lda) and v (n rows, ldv).
 * Replaces pair_stats_kernel + compute_rotation_params_kernel + apply_rotation_kernel of one round
 * (src/kernels.cu:908-965).  rot_count (device int32, may be NULL) is incremented once per rotated
 * pair; offmax (device uint64 holding a non-negative double's bits, may be NULL) receives the max of
 * |gamma|/sqrt(alpha*beta).  If skip_if_zero is non-NULL and *skip_if_zero == 0 on entry the kernel
 * does nothing (device-side early exit used after convergence).  The launch shape is jsvd_autotune's choice
 * for m x n when it measured one, else the built-in rule.
 */
int jsvd_round_f64(double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n, int32_t round,
                   double epsilon, int32_t *rot_count, uint64_t *offmax, const int32_t *skip_if_zero, void *stream);

/*
 * `rounds` consecutive fused rounds (rounds first_round, first_round + 1, ... modulo the n - 1 (even n) or n
 * (odd n) rounds of a sweep), launched back to back on `stream` with the arguments of jsvd_round_f64; rounds =
 * n - 1 from round 0 is one sweep of the round loop of src/kernels.cu:893-966.  rot_count and offmax (each may
 * be NULL) hold count_slots entries (a power of two, 1..4096): the pair of slot k of a round counts into entry
 * k % count_slots, as the solver does to keep same-address atomics apart; the caller adds the entries (takes
 * their maximum for offmax).
 */
int jsvd_rounds_f64(double *a, int64_t lda, int64_t m, double *v, int64_t ldv, int64_t n, int32_t first_round,
                    int32_t rounds, double epsilon, int32_t *rot_count, uint64_t *offmax, int32_t count_slots,
                    const int32_t *skip_if_zero, void *stream);

/*
 * Rotation parameters and skip test for `count` pairs: rotate[k] = 1 and (c[k], s[k]) from alpha, beta,
 * gamma in FP64 when |gamma| > epsilon*sqrt(max(alpha,0)*max(beta,0)) and |gamma| > 1e-300, else
 * rotate[k] = 0, c = 1, s = 0.  Replaces compute_rotation_params_kernel (src/kernels.cu:485-521); the
 * per-pair flag array replaces its single atomicExch flag.  Evaluated without FMA contraction, so the
 * result equals an IEEE double evaluation of the same formula bit for bit.
 */
int jsvd_rotation_params_f64(const double *alpha, const double *beta, const double *gamma, int64_t count,
                             double epsilon, double *c, double *s, int32_t *rotate, void *stream);

/*
 * Two-stage off-orthogonality reduction: out[0] = max over pairs i<j of |a_i . a_j| / (|a_i| |a_j|) for
 * the column-major a (m x n, lda) (pairs with a zero column are skipped).  Stage one reduces row
 * tiles into per-block partial maxima, stage two reduces the partials.  No reference counterpart as a
 * kernel (the reference only sets a flag, src/kernels.cu:514); it is the convergence measure of the
 * skip test.
 */
int jsvd_offnorm_f64(const double *a, int64_t lda, int64_t m, int64_t n, double *out, void *stream);

/*
 * sigma/U extraction: sigma_j = sqrt(sum_i a_ij^2), U[:,j] = a_j / sigma_j if sigma_j > epsilon else 0,
 * from the column-major a (m x n, lda) into row-major u (m x n).  Replaces build_u_sigma_kernel
 * (src/kernels.cu:586-629) fused with the final column-to-row transpose (src/kernels.cu:983-988).
 */
int jsvd_u_sigma_f64(const double *a, int64_t lda, int64_t m, int64_t n, double epsilon, double *u,
                     double *sigma, void *stream);

/*
 * Device-to-device copy of `bytes` (a multiple of 16; both pointers 16-byte aligned) by the library's own
 * streaming kernel, enqueued on `stream`: the measured practical HBM roof for the bandwidth-bound kernels
 * (read + write of equal byte counts).  variant 0 = plain stores, 1 = non-temporal stores.  No reference
 * counterpart (the reference quotes a cost model, tex/sections/03-gpu-parallel-reduction.tex:72-100).
 */
int jsvd_hbm_copy(void *dst, const void *src, int64_t bytes, int32_t variant, void *stream);

/*
 * The FP64 dense kernels of the mixed path's QR-first preconditioning and final products, as stand-alone
 * device-pointer entry points (enqueued on `stream`, a hipStream_t; they return once the work is done).  No
 * reference counterpart: the reference has no dense linear algebra (it sweeps A itself).
 *
 * jsvd_gemm_f64: d = alpha op(a) op(b) + beta c, op(x) = x or x^T (trans_a / trans_b nonzero), op(a) m x k, op(b)
 * k x n, column-major a (lda), b (ldb); c and d column-major m x n (ldc, ldd) unless JSVD_GEMM_ROW_MAJOR_CD.  c may
 * equal d; it is not read when beta == 0.  flags restrict the work to the structure:
 *   JSVD_GEMM_UPPER_OUT   only entries (i, j) in 64 x 64 tiles on or above the diagonal are written (m == n)
 *   JSVD_GEMM_MIRROR      with UPPER_OUT: the tiles above the diagonal are also written transposed below it
 *                         (a symmetric product in full)
 *   JSVD_GEMM_TRI_A       op(a) is upper triangular: its entries below the diagonal are not read
 *   JSVD_GEMM_TRI_B       op(b) is upper triangular: likewise
 *   JSVD_GEMM_ZERO_LOWER  entries of d below the diagonal are written as zero
 *   JSVD_GEMM_ROW_MAJOR_CD c and d row-major (entry (i, j) at i * ld + j)
 * Results do not depend on the launch's timing (split-k partials are summed in a fixed order).
 *
 * jsvd_trsm_right_upper_f64: b (rows x n, column-major, ldb) <- b r^-1 for the upper triangle of r (n x n,
 * column-major, ldr; its strictly lower triangle is not read).
 */
#define JSVD_GEMM_UPPER_OUT 1
#define JSVD_GEMM_MIRROR 2
#define JSVD_GEMM_TRI_A 4
#define JSVD_GEMM_TRI_B 8
#define JSVD_GEMM_ZERO_LOWER 16
#define JSVD_GEMM_ROW_MAJOR_CD 32
int jsvd_gemm_f64(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                  int64_t lda, const double *b, int64_t ldb, double beta, const double *c, int64_t ldc, double *d,
                  int64_t ldd, int32_t flags, void *stream);
int jsvd_trsm_right_upper_f64(double *b, int64_t ldb, int64_t rows, int64_t n, const double *r, int64_t ldr,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif /* JSVD_H */
