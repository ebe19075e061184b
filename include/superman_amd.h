/*
 * superman_amd.h -- C ABI of the MI355X exact-permanent library
 * (libsuperman_amd.so).  Plain pointers and sizes only; every call is
 * reentrant (no process globals beyond a mutex-guarded per-device cache).
 *
 * Each entry point replaces a reference GPU driver (file:line cited per
 * function).  The reference's drivers return the permanent as a double inside
 * `Result{time, permanent}` (CommonFiles/Result.h:29-37) and call exit() on a
 * CUDA error (Algorithms/GPU/GPUHelpers.cuh:77-85); here every call returns a
 * status code and fills an extended `spm_result`, whose `re` field is exactly
 * what the reference's Result.permanent carries.  The C++ drop-in layer
 * (the headers under include/superman/) maps these back onto Result /
 * runtime_error.
 *
 * Matrices are row-major n x n (a[i*n + j] = a_ij), 1 <= n <= 63
 * (reference limit: CommonFiles/DecomposePerman.h:185-188).  Complex
 * matrices are interleaved (re, im) pairs, i.e. std::complex<double>.
 */
#ifndef SUPERMAN_AMD_H
#define SUPERMAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SPM_OK            0
#define SPM_EINVAL       -1   /* bad argument (n out of range, null pointer, ...) */
#define SPM_EHIP         -2   /* HIP runtime error */
#define SPM_ENOMEM       -3   /* device allocation failed / over budget */
#define SPM_ERANGE       -4   /* exact integer result cannot be guaranteed */
#define SPM_ECOMPILE     -5   /* hiprtc compilation of a generated kernel failed */
#define SPM_ENOTINT      -6   /* integer path given non-integer entries */
#define SPM_ECANCELLED   -7   /* request cancelled before it ran */
#define SPM_EAGAIN       -8   /* admission refused (budget), try later */
#define SPM_ETIMEDOUT    -9   /* request exceeded its deadline */
#define SPM_EFAULTINJ   -10   /* fault injected through spm_batch_inject_fault */

/* ---- precision modes (Settings.h:59-65 `Precision`) ---------------------- */
#define SPM_PREC_PLAIN        0   /* reference "dd": plain FP64 sums          */
#define SPM_PREC_COMPENSATED  1   /* reference "kahan" (its default, IO.h:178):
                                     TwoSum-compensated double-double sums   */
#define SPM_PREC_F32_PRODUCT  2   /* real, nonnegative entries only (else
                                     SPM_EINVAL): rows scaled by powers of two,
                                     FP64 row sums, FP32 products, compensated
                                     FP64 accumulation; err_est adds the FP32
                                     product error to the FP64 estimate (no
                                     reference counterpart).  Non-production on
                                     MI355X: slower and less accurate than
                                     SPM_PREC_COMPENSATED there. */

/* Extended result.  `re` (and `im` for complex) is the value the reference
 * returns; hi/lo carry the double-double sum; `abs_sum` = 2 * sum |term|
 * (compensated modes) so that cond = abs_sum / |value|; `err_est` is an
 * estimate of the absolute error, about 2n roundings of u = 2^-53 per term
 * (2n u abs_sum real, 4n u abs_sum complex): it counts each factor's
 * rounding once, not the drift of the up to 2^L incremental updates a
 * factor carries since its chunk's exact start, so it is no guaranteed
 * bound (-1 in PLAIN mode, 0 for exact results); `exact_*` hold the exact integer
 * value (two's complement 128-bit) when `is_exact` is set. */
typedef struct spm_result {
    double re, im;
    double hi_re, lo_re, hi_im, lo_im;
    double abs_sum;
    double err_est;
    int64_t exact_lo, exact_hi;
    int32_t is_exact;
    int32_t status;
    double seconds;       /* host wall time of the call                     */
    double kernel_ms;     /* device time of the permanent kernels (events)  */
    uint64_t subsets;     /* 2^(n-1) Gray-code subsets walked               */
    int32_t n;
    int32_t launches;
} spm_result;

const char* spm_version(void);
const char* spm_strerror(int status);
/* Number of visible HIP devices (0 if none).  Returns status. */
int spm_device_count(int* count);
/* Block until all work queued on `device` has finished. */
int spm_device_synchronize(int device);
/* Device description: compute units and engine clock (kHz). */
int spm_device_info(int device, int* compute_units, int* clock_khz, char* name, int name_len);

/* ---- dense real FP64 ------------------------------------------------------
 * Replaces gpuDPSingleGPU<double,S> (Algorithms/GPU/GPUKernelWrappers.cu:
 * 208-268) -> dpSingleGPU::permanentFunction (Algorithms/GPU/Dense/
 * dpSingleGPU.cuh:44-219) with kernels xRegisterMShared / xRegisterMSharedKahan
 * (Algorithms/GPU/Dense/DenseKernelDefinitions.cuh:190-405).
 * `stream` is a hipStream_t (NULL = the device's null stream); the call is
 * synchronous with respect to it. */
int spm_dense_f64(const double* a, int n, int precision, int device, void* stream, spm_result* out);

/* spm_dense_f64 without any allocation per call: no hipMalloc / hipFree /
 * hipHostMalloc.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Same arguments, kernels, geometry and
 * bit-identical result as spm_dense_f64.  Device memory is one fixed array
 * per device (191 KiB, held until the process ends: the +A/-A planes,
 * x(empty set) and up to 4096 per-block records), set up at the first call
 * on each device.  Limits: 1 <= n <= 63 (else SPM_EINVAL); a launch needing
 * more than 4096 blocks returns SPM_ENOMEM.  Calls that overlap on one
 * device share that array and race: serialize them. */
int spm_dense_f64_noalloc(const double* a, int n, int precision, int device, void* stream, spm_result* out);

/* ---- dense complex128 ----------------------------------------------------
 * Replaces gpuComputeComplex (Algorithms/Complex/gpuComputeComplex.cu:199-418,
 * kernels :39-197), reached through complexPerman (Algorithms/AllAlgorithms.h:
 * 93-110).  `a` is row-major interleaved (re, im), i.e. std::complex<double>
 * storage; the result is in re/im (the reference prints it and returns
 * void).  1 <= n <= 63. */
int spm_dense_c128(const double* a, int n, int precision, int device, void* stream, spm_result* out);

/* ---- exact integer ---------------------------------------------------------
 * Integer matrices (|a_ij| <= 2^30, row sums of |a_ij| < 2^30) computed
 * exactly: y = 2x in int32 -- x itself on rows with an even sum, whose y
 * is always even -- T = sum (-1)^k prod y (mod 2^128), perm =
 * (-1)^(n-1) T / 2^(n-1-e) with e such rows, returned in exact_lo/exact_hi
 * (is_exact = 1) and rounded into `re`.  Returns SPM_ERANGE unless
 * 2^(n-1-e) times a bound on |perm| (Bregman-Minc for 0/1 matrices, the
 * product of absolute row sums otherwise) stays below 2^126.
 * The reference has no integer path: its sparse drivers gpuSPSingleGPU
 * (Algorithms/GPU/GPUKernelWrappers.cu:65-113) -> spSingleGPU
 * (Algorithms/GPU/Sparse/spSingleGPU.cuh:44-234) and the generated-kernel
 * driver KernelGenSingleGPU (Algorithms/GPU/CodeGeneration/
 * KernelGenSingleGPU.cuh:47-197) compute 0/1 and integer inputs in FP64;
 * this entry point replaces them for integer-valued matrices. */
int spm_int_exact(const int64_t* a, int n, int device, void* stream, spm_result* out);

/* ---- SpaRyser with per-pattern generated kernels ----------------------------
 * Replaces the reference's code-generation path: generateKernels
 * (Algorithms/GPU/CodeGeneration/KernelGenerator.cuh:114-176) with
 * generateUTOrderedKernelCode (UTOrderedKernelCodeGen.cuh:34-583) or
 * generateNaiveKernelCode (NaiveKernelCodeGen.cuh:34-365), the wrapper's
 * rebuild-and-rerun protocol (wrapper.cpp:271-344) and the driver
 * KernelGenSingleGPU (KernelGenSingleGPU.cuh:47-197).  The kernel source is
 * emitted for the matrix's pattern and exact integer values and compiled
 * in-process with hiprtc; compiled kernels are cached per (device, hash)
 * in the process and as code objects on disk (directory $SPM_KERNEL_CACHE,
 * "" or "0" = no disk cache; default $XDG_CACHE_HOME/superman_amd or
 * ~/.cache/superman_amd), so a pattern compiles once per machine.
 * Integer matrices only, 5 <= n <= 63, same exactness bound as
 * spm_int_exact. */
#define SPM_CG_REGISTER_EFFICIENT 1  /* rows/columns in IO::UTOrder order
                                        (REGEFFICIENTCODEGENERATION); without
                                        it the input order (NAIVECODEGENERATION) */
int spm_sparse_codegen(const int64_t* a, int n, int options, int device, void* stream, spm_result* out);
/* The generated source, NUL-terminated, into buf (needed = full size + 1). */
int spm_codegen_source(const int64_t* a, int n, int options, char* buf, size_t buflen, size_t* needed);
/* hiprtc-compiles the generated kernel without loading it (no device
 * needed): compile time and code-object size. */
int spm_codegen_compile_only(const int64_t* a, int n, int options, double* seconds, size_t* code_bytes);
/* Generated SpaRyser kernels for real (FP64) matrices.  Same geometry and
 * generator options as spm_sparse_codegen; `precision` is SPM_PREC_PLAIN or
 * SPM_PREC_COMPENSATED.  Literals are the entries' exact binary64 values
 * (hex floats; the reference prints 6-decimal std::to_string literals,
 * UTOrderedKernelCodeGen.cuh:292-295), each flip updates only the rows of
 * the flipped column, and the product reuses cached products of the rows a
 * flip cannot have changed (the reference's register-area split of rows
 * kept in registers and a cached globalProduct, UTOrderedKernelCodeGen.cuh:
 * 68-132, 365-454).  Replaces KernelGenSingleGPU (KernelGenSingleGPU.cuh:
 * 47-197) for non-integer matrices; err_est / abs_sum as for spm_dense_f64. */
int spm_sparse_codegen_real(const double* a, int n, int options, int precision, int device, void* stream,
                            spm_result* out);
int spm_codegen_source_real(const double* a, int n, int options, int precision, char* buf, size_t buflen,
                            size_t* needed);
/* IO::UTOrder (IO/IO.h:926-1002) on the pattern of `a`: the permuted matrix
 * is b[row_perm[i]][col_perm[j]] = a[i][j] (IO::applyPermutations,
 * IO/IO.h:1095-1112).  Host only. */
int spm_ut_order(const double* a, int n, int* row_perm, int* col_perm);

/* Device-resident plans: the matrix is uploaded once, then spm_plan_run
 * computes with inputs already in HBM (what the benchmark times). */
typedef struct spm_plan spm_plan;
int spm_plan_dense_f64(const double* a, int n, int precision, int device, spm_plan** plan);
int spm_plan_dense_c128(const double* a, int n, int precision, int device, spm_plan** plan);
int spm_plan_int(const int64_t* a, int n, int device, spm_plan** plan);
/* The same with the integer walk forced: SPM_INT_WALK_INT32 (any entries in
 * range), SPM_INT_WALK_FP64 (exact FP64 walk with grouped products; only
 * when every absolute row sum is below 2^b, b = bit width of n -- all 0/1
 * matrices -- else SPM_ERANGE) or SPM_INT_WALK_SPARSE (compressed columns:
 * per step only the flipped column's rows, a zero-row count, the product
 * only when no row is zero; the reference's runtime-CSC sparse kernels,
 * SparseKernelDefinitions.cuh:34-509).  SPM_INT_WALK_AUTO takes the sparse
 * walk for sparse matrices and otherwise the FP64 walk whenever it
 * applies. */
#define SPM_INT_WALK_AUTO   0
#define SPM_INT_WALK_INT32  1
#define SPM_INT_WALK_FP64   2
#define SPM_INT_WALK_SPARSE 3
int spm_plan_int_walk(const int64_t* a, int n, int walk, int device, spm_plan** plan);
/* Compiles (or takes from the cache) the generated kernel for `a`. */
int spm_plan_codegen(const int64_t* a, int n, int options, int device, spm_plan** plan);
/* The same for a real matrix (spm_sparse_codegen_real). */
int spm_plan_codegen_real(const double* a, int n, int options, int precision, int device, spm_plan** plan);
int spm_plan_codegen_info(const spm_plan* plan, double* compile_seconds, int* cached, uint64_t* hash);
int spm_plan_run(spm_plan* plan, void* stream, spm_result* out);
int spm_plan_destroy(spm_plan* plan);
/* Launch geometry of a plan: grid blocks, block threads, L (Gray steps per
 * chunk = 2^L), number of wave tasks. */
int spm_plan_info(const spm_plan* plan, int* grid, int* block, int* log2_chunk, uint64_t* tasks);
/* The Gray-walk geometry the dense kernels use for order n when
 * `resident_waves` wavefronts fit on the device: each lane walks 2^log2_chunk
 * consecutive Gray indices; chunks = 2^(n-1-log2_chunk); tasks = chunks / 64
 * rounded up; the kernel unrolls blocks of 2^unroll_bits steps (may be NULL).
 * Pure host computation. */
int spm_gray_geometry(int n, int64_t resident_waves, int* log2_chunk, uint64_t* chunks, uint64_t* tasks,
                      int* unroll_bits);
/* Name of the permanent kernel a plan launches (for profiler lookups). */
const char* spm_plan_kernel_name(const spm_plan* plan);
/* Dense plans only: walk chunks of 2^log2_chunk Gray steps instead of the
 * automatic length (U + 1 <= log2_chunk <= n - 1, U = the kernel's unroll
 * bits).  A different chunk length starts x from an exact rebuild at other
 * places, so the rounding differs while the permanent does not (used by
 * accuracy tests and tuning sweeps).  No reference counterpart: the
 * reference's chunk follows from its thread count (dpSingleGPU.cuh:148-199). */
int spm_plan_set_log2_chunk(spm_plan* plan, int log2_chunk);

/* ---- automatic path choice ------------------------------------------------
 * One real matrix by the path whose estimated time (rates measured on
 * MI355X per Gray subset, plus the hiprtc compile of a generated kernel;
 * for an integer matrix, a load instead where that kernel is already
 * compiled, in the process or on disk) is smallest -- the MI355X counterpart of the reference's automatic
 * algorithm choice (CommonFiles/AlgorithmSelector.h:52-67), which weighs
 * nnz per row and component size on its own GPUs:
 *   integer entries        exact walk or exact generated kernel;
 *   rows s_i * integers    exact integer permanent times prod s_i, rounded
 *                          once (err_est = 2^-52 |value|);
 *   other real entries     generated FP64 kernel or the dense FP64 walk.
 * `path_taken` (may be NULL) receives the SPM_PATH_* value that ran;
 * spm_choose_path gives the choice for kernels not yet compiled, without
 * running it (host only; deterministic, so the drop-in `auto` does not
 * depend on what a machine has compiled before), and spm_path_estimate one
 * path's estimated seconds (compile included). */
#define SPM_PATH_INT        1
#define SPM_PATH_GEN_INT    2
#define SPM_PATH_SCALED_INT 3
#define SPM_PATH_GEN_REAL   4
#define SPM_PATH_DENSE      5
int spm_permanent_auto(const double* a, int n, int precision, int device, void* stream, spm_result* out,
                       int* path_taken);
int spm_choose_path(const double* a, int n, int* path);
int spm_path_estimate(const double* a, int n, int path, double* seconds);

/* ---- structural preprocessing (sparse inputs) ---------------------------------
 * The reference's pipeline in front of its kernels, with the same
 * semantics: fine Dulmage-Mendelsohn pruning of entries in no perfect
 * matching (IO/IO.h:477-508 with IO/dm.cpp:372-501), connected components
 * whose permanents multiply (IO/dm.h:43-107, main.cpp:73-111), and
 * DecomposePerman's compressions of blocks larger than 30 -- degree-1
 * expansion, degree-2 column merging, degree-3/4 splitting into two
 * matrices whose permanents add (CommonFiles/DecomposePerman.h:119-557).
 * The permanent is unchanged: perm(a) = prod over components of the sum of
 * the component's leaves' permanents.  Host only. */
#define SPM_DECOMP_DM          1  /* prune entries outside all perfect matchings */
#define SPM_DECOMP_COMPONENTS  2  /* split into connected components              */
#define SPM_DECOMP_COMPRESS    4  /* DecomposePerman compressions above order 30  */
#define SPM_DECOMP_CODEGEN     8  /* code-generation mode: no component split and
                                     degree-1/2 compressions only (one leaf), as
                                     the reference does for its generated kernels */
#define SPM_DECOMP_SMALL      16 /* compress above order 8 instead of 30 (lets tests
                                     exercise every compression on small inputs) */
#define SPM_DECOMP_ALL_FLAGS  31
#define SPM_DECOMP_DEFAULT     7  /* DM | COMPONENTS | COMPRESS */
typedef struct spm_decomp spm_decomp;
int spm_decompose(const double* a, int n, int options, spm_decomp** out);
/* Components, leaves, average leaf order and nonzeros (the reference's
 * "Submatrix count / Average N / Average NNZ" line, main.cpp:115-118), and
 * whether the permanent is structurally zero. */
int spm_decomp_stats(const spm_decomp* d, int* components, int* leaves, double* avg_n, double* avg_nnz,
                     int* structurally_zero);
/* Leaf `leaf`: its component, order and (if out != NULL) its n x n entries. */
int spm_decomp_leaf(const spm_decomp* d, int leaf, int* component, int* n, double* out);
int spm_decomp_destroy(spm_decomp* d);
/* The permanent through the decomposition: leaves of integer matrices run
 * exactly (spm_int_exact) and combine in 128-bit integers (is_exact = 1
 * while that range holds), others through spm_dense_f64 with `precision`.
 * Replaces the DecomposePerman<...>::computePermanentRecursively drivers
 * (CommonFiles/DecomposePerman.h:76-117) around the GPU algorithms. */
int spm_permanent_decomposed(const double* a, int n, int options, int precision, int device, void* stream,
                             spm_result* out);
/* The same with the reference's `scaling=true` (Settings scaling,
 * scalingIterationNo, scaleInto): every leaf is balanced first -- IO::scale
 * (IO/IO.h:109-161) as DecomposePerman::addQueue applies it
 * (CommonFiles/DecomposePerman.h:189-193) -- and its permanent divided by
 * the row and column scales afterwards (DecomposePerman.h:97-107).  Scaled
 * leaves are real, so the result is FP64 (is_exact = 0). */
int spm_permanent_decomposed_scaled(const double* a, int n, int options, int precision, int scaling_iterations,
                                    double scale_into, int device, void* stream, spm_result* out);
/* IO::scale on its own (host only): `iterations` rounds of row then column
 * balancing towards line sums `target`, then a is scaled in place
 * (a_ij r_i c_j); row_scale / col_scale receive r and c. */
int spm_sinkhorn_scale(double* a, int n, int iterations, double target, double* row_scale, double* col_scale);

/* ---- batched and served permanents -----------------------------------------
 * No single counterpart in the reference, which computes one matrix per
 * driver call (main.cpp:69-115 calls the selected Algorithm once per
 * matrix/component, each call allocating, launching a chunk loop and
 * synchronising: dpSingleGPU.cuh:134-213).  Here many matrices of orders
 * 1..28 share one device arena and one persistent launch fed by a device
 * work queue; each matrix is split into 64-chunk work items whose chunk size
 * comes from the engine's tuning table (set by spm_batch_autotune).  A
 * matrix's result depends only on its data and its chunk size, never on the
 * other matrices of its batch.
 *
 * Kinds: SPM_KIND_REAL (compensated FP64), SPM_KIND_COMPLEX (compensated
 * complex128, `a` interleaved re/im), SPM_KIND_INT (exact; `a` holds
 * integer values as doubles, same range rules as spm_int_exact). */
#define SPM_KIND_REAL     1
#define SPM_KIND_COMPLEX  2
#define SPM_KIND_INT      3
#define SPM_BATCH_MAX_N   28

#define SPM_EXEC_DEVICE   0  /* run on the GPU                                    */
#define SPM_EXEC_DRYRUN   1  /* no device at all: batches are formed, "run" for a
                                time proportional to their work and complete with
                                NaN values; for testing the admission, deadline,
                                cancellation and fault paths of the serving loop */

typedef struct spm_batch_config {
    int device;
    int slots;              /* batches in flight, one HIP stream each (>= 1)       */
    int max_requests;       /* matrices per batch                                   */
    int max_items;          /* work items per batch                                 */
    size_t arena_bytes;     /* device bytes per batch for matrix data               */
    int max_pending;        /* admission queue length; submit beyond it: SPM_EAGAIN */
    double max_batch_work;  /* work budget per batch (FP64-op units, 0 = none)      */
    int use_graph;          /* launch each batch through a captured hipGraph        */
    int executor;           /* SPM_EXEC_DEVICE or SPM_EXEC_DRYRUN                   */
} spm_batch_config;

typedef struct spm_batch_tuning {
    int fused;              /* 1: one kernel walks every kind; 0: one per kind      */
    int waves_per_cu;       /* persistent grid size (0 = all resident waves)        */
    int log2_chunk[4];      /* [kind]: target Gray steps per lane chunk (2^L)       */
} spm_batch_tuning;

typedef struct spm_request {
    int kind;
    int n;
    const double* a;        /* row-major n x n (complex: 2 n^2 doubles)             */
    uint64_t tag;           /* returned with the completion                         */
    double deadline_s;      /* relative to submission; 0 = none                     */
} spm_request;

typedef struct spm_completion {
    uint64_t ticket, tag;
    spm_result result;      /* result.status: SPM_OK or the request's error        */
    double t_submit, t_launch, t_done;  /* engine clock, seconds                    */
} spm_completion;

typedef struct spm_serve_stats {
    uint64_t completed;     /* with status SPM_OK                                   */
    uint64_t failed;
    uint64_t batches;
    double wall_s;          /* first arrival to last completion                     */
    double perms_per_s;     /* completed / wall_s                                   */
    double p50_s, p99_s, mean_s, max_s;  /* latency: arrival -> result on host      */
    double mean_batch;      /* matrices per batch                                   */
    double device_busy;     /* fraction of wall_s with a batch in flight            */
    uint64_t late_publish;  /* results whose flag reached host memory after their
                               batch's completion was seen (waited for, not lost) */
    uint64_t retried;       /* requests run again after their result failed a check
                               (stale item records, integer sum not a multiple of
                               2^(n-1)); spm_batch_last_error names the check */
} spm_serve_stats;

typedef struct spm_batch spm_batch;
void spm_batch_config_default(spm_batch_config* cfg);
int spm_batch_create(const spm_batch_config* cfg, spm_batch** out);
int spm_batch_destroy(spm_batch* b);
int spm_batch_get_tuning(const spm_batch* b, spm_batch_tuning* t);
int spm_batch_set_tuning(spm_batch* b, const spm_batch_tuning* t);
/* Times `sample` (all submitted at once, best of `reps`) under candidate
 * tunings -- fused or per-kind kernels, persistent grid size, per-kind chunk
 * size -- and keeps the fastest, except where latency decides: the fused
 * kernel and the smallest chunk are kept while within 5 % of the fastest
 * time.  The engine must be idle. */
int spm_batch_autotune(spm_batch* b, int count, const spm_request* sample, int reps, spm_batch_tuning* best,
                       double* best_seconds);
/* Computes `count` permanents (as many batches as admission needs) and
 * returns when all are done; out[i] belongs to reqs[i].  Per-request errors
 * are in out[i].status; the call returns the first one (or SPM_OK). */
int spm_batch_run(spm_batch* b, int count, const spm_request* reqs, spm_result* out);
/* Serving: submit copies and prepares the matrix and queues it (SPM_EAGAIN
 * when the queue is full, SPM_ENOMEM when it can never fit a batch);
 * progress retires finished batches and launches new ones while a slot is
 * free; poll returns finished requests (it calls progress first). */
int spm_serve_submit(spm_batch* b, const spm_request* req, uint64_t* ticket);
int spm_serve_progress(spm_batch* b, int* in_flight);
int spm_serve_poll(spm_batch* b, spm_completion* out, int max, int* got);
/* Cancels a queued or running request: it completes with SPM_ECANCELLED. */
int spm_serve_cancel(spm_batch* b, uint64_t ticket);
int spm_serve_pending(spm_batch* b, int* queued, int* in_flight);
/* Replays an arrival schedule (arrival_s[i] seconds after the start, non-
 * decreasing) through submit/progress/poll on the calling thread; fills
 * out[i] and latency_s[i] (arrival to completion) and the stream summary. */
int spm_serve_replay(spm_batch* b, int count, const spm_request* reqs, const double* arrival_s, spm_result* out,
                     double* latency_s, spm_serve_stats* stats);
/* Engine-lifetime totals of late_publish and retried (spm_serve_stats); both
 * stay 0 unless a result reached host memory after its batch's completion
 * or failed a check.  No reference counterpart (a health check for callers
 * of the batch engine). */
int spm_batch_counters(spm_batch* b, uint64_t* late_publish, uint64_t* retried);
/* The next `launches` batch launches fail on the host without touching the
 * device; their requests complete with SPM_EFAULTINJ. */
int spm_batch_inject_fault(spm_batch* b, int launches);
/* The next `results` results handed out count as failing their checks (no
 * reference counterpart; for tests of the re-run path): each such request
 * runs again, up to twice, then completes with SPM_EHIP. */
int spm_batch_inject_bad_results(spm_batch* b, int results);
/* Which HIP call failed last (empty if none); valid until the next call on
 * the same thread. */
const char* spm_batch_last_error(spm_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* SUPERMAN_AMD_H */
