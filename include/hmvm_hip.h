/*
 * hmvm_hip.h -- C ABI of the MI355X (gfx950) H-matrix-vector multiply.
 *
 * Drop-in part: the functions mirror the reference HMVM entry points one for
 * one (same argument lists and meaning), so a caller of hmvm_seq / hmvm_omp /
 * hmvm_cuda can switch by renaming the call:
 *
 *   reference                                   here
 *   hmvm_seq(mat, mat2, b, dump)  hacapk.h:36   hmvm_hip(mat, mat2, b, dump)
 *   hmvm_seq_bench(mat, mat2, b)  hacapk.h:37   hmvm_hip_bench(mat, mat2, b)
 *   hmvm_seq_1 (v, mat,  b)       hmvm_seq.c:9    hmvm_hip_1 (v, mat,  b)
 *   hmvm_seq_1t(v, mat,  b)       hmvm_seq.c:61   hmvm_hip_1t(v, mat,  b)
 *   hmvm_seq_2 (v, mat2, b)       hmvm_seq.c:113  hmvm_hip_2 (v, mat2, b)
 *   hmvm_seq_2t(v, mat2, b)       hmvm_seq.c:169  hmvm_hip_2t(v, mat2, b)
 *   hmvm_cuda1<T>(mat2, b, kernel, dump, nbench)  hmvm_cuda.h:5
 *                                               hmvm_hip<T>(...) in hmvm_hip.hpp
 *
 * Semantics (as the sequential reference): hmvm_hip_X ACCUMULATES, v += A*b;
 * the caller zeroes v for a plain product.  All pointers are host pointers
 * owned by the caller.  The first call on a matrix uploads it to the GPU and
 * keeps the device copy, keyed by the struct address; call hmvm_hip_release()
 * after changing or freeing the matrix.  Calls block until the result is back
 * in host memory.  Errors never exit the process: the function returns, the
 * reason is printed once to stderr and hmvm_hip_last_error() reports it.
 * Not re-entrant on one matrix; one host thread per device.
 *
 * Device-resident part (no reference counterpart; used by Krylov loops,
 * benchmarks and batched callers): an explicit plan object that holds the
 * repacked matrix in HBM and applies y (+)= A x to device vectors on a stream.
 */
#ifndef HACAPK_AMD_HMVM_HIP_H
#define HACAPK_AMD_HMVM_HIP_H

#include <stddef.h>

#include "hacapk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
enum {
  HMVM_OK = 0,
  HMVM_ERR_ARG = -101,      /* NULL pointer or invalid option               */
  HMVM_ERR_MATRIX = -102,   /* leaf descriptors inconsistent with nd / len  */
  HMVM_ERR_HIP = -103,      /* HIP runtime error (message has the details)  */
  HMVM_ERR_NOMEM = -104,    /* host allocation failed                       */
  HMVM_ERR_NODEVICE = -105, /* no HIP device is visible                     */
  HMVM_ERR_UNSUPPORTED = -106,
  HMVM_ERR_BREAKDOWN = -107, /* Krylov iteration met a zero divisor or a     */
                             /* non-finite residual (u is the last finite    */
                             /* iterate)                                     */
  HMVM_ERR_WAIT = -108       /* a bounded wait of the persistent launch gave */
                             /* up: that launch's results are invalid        */
};

/* ---- drop-in entry points ------------------------------------------------ */

/* Runs variants 1, 1t (when mat != NULL) and 2, 2t (when mat2 != NULL) on a
 * zeroed v and, when dump_result != 0, writes result_hip_<variant>_d.txt with
 * one "%.3E" value per line (reference src_c/hmvm_seq.c:226-286). */
void hmvm_hip(const matrix *mat, const matrix2 *mat2, const double *b, int dump_result);

/* Times each available variant: 10 runs, statistics over the last 5, printed
 * as "TIME 10 hmvm_hip_<variant> min <s> max <s> avg <s>" (reference
 * src_c/hmvm_seq.c:289-390).  The timed region is the GPU kernels only, with
 * b and v already in device memory, as the reference GPU proxies time it
 * (src_cpp/hmvm_cuda1.cu:236-302). */
void hmvm_hip_bench(const matrix *mat, const matrix2 *mat2, const double *b);

/* v += A*b from the per-leaf arrays; a2 (1) or a2t (1t) for the low-rank
 * second stage (reference src_c/hmvm_seq.c:9, :61). */
void hmvm_hip_1(double *v, const matrix *mat, const double *b);
void hmvm_hip_1t(double *v, const matrix *mat, const double *b);
/* v += A*b from matrix2.rowmat (2) or matrix2.rowmat_t (2t)
 * (reference src_c/hmvm_seq.c:113, :169). */
void hmvm_hip_2(double *v, const matrix2 *mat2, const double *b);
void hmvm_hip_2t(double *v, const matrix2 *mat2, const double *b);

/* Status of the last call on this thread (HMVM_OK or an error code), and its
 * message.  No reference counterpart: the reference calls exit(-1) instead
 * (src_cpp/hmvm_cuda0.cu:11). */
int hmvm_hip_last_error(void);
const char *hmvm_hip_last_error_message(void);

/* Drops the cached device copy of `mat` (a matrix* or matrix2*).  No
 * reference counterpart (the reference re-uploads and leaks per call). */
void hmvm_hip_release(const void *mat);

/* ---- device-resident plan API --------------------------------------------- */
enum { HMVM_STORAGE_F64 = 0, HMVM_STORAGE_F32 = 1, HMVM_STORAGE_BF16 = 2 };
enum { HMVM_A2_ROWMAJOR = 0, HMVM_A2_TRANSPOSED = 1 };

typedef struct hmvm_options {
  int storage;           /* HMVM_STORAGE_*: element type kept in HBM            */
  int w1;                /* target elements per stage-1 tile (0: by storage,    */
                         /* capped at max(1024, len / 4096) for small matrices) */
  int w2;                /* target elements per dense tile (and per stage-2     */
                         /* tile unless w3 is set; 0: ditto)                    */
  int block_threads;     /* threads per workgroup, 64..256 step 64 (0: by       */
                         /* storage, per launch)                                */
  int kernel;            /* tile-kernel variant 0..7 (-1: by storage, per       */
                         /* launch): bit0 vector entries of wide tiles read     */
                         /* 64 columns at a time and broadcast (readlane),      */
                         /* bit1 software pipelining of wide tiles, bit2        */
                         /* non-temporal matrix loads                           */
  int order;             /* tile order in a launch: 0 memory, 1 largest first   */
                         /* (launch A: stage-1 and dense tiles in one order)    */
  int mode;              /* 0: launches on one stream; 1: one persistent launch */
                         /* of as many workgroups as the device holds, its four */
                         /* stages separated by grid barriers, tiles dealt to   */
                         /* the waves on the host (any t_mode); 2: launches,    */
                         /* dense tiles on a second stream concurrent with      */
                         /* stage 1/2                                           */
  int t_mode;            /* stage-2 vector entries: 0 sum the stage-1 partial   */
                         /* slots on the fly, 1 read a materialised t (one more */
                         /* launch; with several right-hand sides the stage-1   */
                         /* tiles of a row group of one column chunk write      */
                         /* their t rows themselves and the t sum adds the      */
                         /* rest), 2 choose per leaf (likewise),                */
                         /* 3 t written directly by the stage-1 tiles (bands of */
                         /* rank rows over all of a column cluster's columns;   */
                         /* no extra launch), -1 default (1)                    */
  int max_rows;          /* rows per tile (0: by storage); at most 192*V-V+1,  */
                         /* V = 16 bytes / element size; when set, also for the */
                         /* stage-2 tiles unless max_rows_b is set              */
  int max_rows_b;        /* rows per stage-2 tile of launch B (0: max_rows if   */
                         /* set, else by storage)                               */
  int w3;                /* target elements per stage-2 tile (0: w2 if set,     */
                         /* else by storage; capped like w2)                    */
} hmvm_options;

typedef struct hmvm_plan_info {
  int nd, nlf, nlowrank, ndense;
  long long len;             /* reference element count                         */
  long long device_elems;    /* elements in HBM incl. alignment padding         */
  int storage, elem_bytes;
  int ntiles_a, ntiles_b;    /* tiles of launch A (stage 1 + dense), B (stage 2) */
  int npanels;
  int ncol_groups, nrow_groups;
  int t_fused;               /* 1: stage 2 sums stage-1 partials itself         */
  long long n_t, n_tslots, n_zslots;
  long long elems_a, elems_b;                     /* elements per launch        */
  long long elems_stage1, elems_dense, elems_stage2;
  long long vec_reads_a, vec_reads_b;
  long long algorithmic_bytes; /* elem_bytes*len + 16*nd (x read, y written)    */
  long long device_bytes;      /* HBM held by the plan                          */
  int mode;                    /* execution mode an apply runs (hmvm_options)   */
  int t_direct;                /* 1: stage-1 tiles write t (t_mode 3)           */
} hmvm_plan_info;

typedef struct hmvm_timing {
  int iters;
  double ms_avg, ms_min, ms_median; /* whole HMVM (all launches)                 */
  double ms_a, ms_tsum, ms_b, ms_ysum; /* per-launch averages (A, t sum, B, y sum)   */
  double ms_stream; /* iters applies enqueued back to back, total / iters       */
} hmvm_timing;

typedef struct hmvm_plan hmvm_plan;

void hmvm_options_default(hmvm_options *opt);

/* Builds the schedule from the descriptors, uploads the values and repacks
 * them on the GPU into the plan's layout.  a2_source selects rowmat
 * (HMVM_A2_ROWMAJOR) or rowmat_t (HMVM_A2_TRANSPOSED) of mat2. */
int hmvm_plan_create(const matrix2 *mat2, int a2_source, const hmvm_options *opt, hmvm_plan **out);
/* Same from the array-of-structs form: a2 (ROWMAJOR) or a2t (TRANSPOSED). */
int hmvm_plan_create_aos(const matrix *mat, int a2_source, const hmvm_options *opt, hmvm_plan **out);
void hmvm_plan_destroy(hmvm_plan *plan);
int hmvm_plan_get_info(const hmvm_plan *plan, hmvm_plan_info *info);

/* d_y (+)= A d_x on `stream` (a hipStream_t; NULL = the plan's own stream).
 * d_x and d_y are device pointers of nd doubles and must not overlap.
 * Asynchronous; no host synchronisation and no allocation. */
int hmvm_plan_apply(hmvm_plan *plan, const double *d_x, double *d_y, int accumulate, void *stream);
/* Host-pointer form: copies x in, applies, copies y back (blocking). */
int hmvm_plan_apply_host(hmvm_plan *plan, const double *x, double *y, int accumulate);
/* Copies x in once, then times `iters` back-to-back applies on the device
 * (kernels only) and returns the seconds per apply (the C++ hmvm_hip<T>
 * benchmark loop). */
int hmvm_plan_bench_apply(hmvm_plan *plan, const double *x, int iters, double *seconds);
/* Times applies on the plan's internal device vectors after `warmup` untimed
 * ones: ms_avg/min/median over `iters` single applies (an event pair around
 * each), the per-launch averages over `iters` more with an event after every
 * launch, ms_stream over `iters` applies enqueued back to back. */
int hmvm_plan_time(hmvm_plan *plan, int warmup, int iters, hmvm_timing *timing);
/* Times `iters` single applies (after one untimed), each preceded by a kernel
 * that writes flush_bytes of scratch memory so that no factor is still in L2
 * or the Infinity Cache: ms_avg/min/median per apply (ms_stream = ms_avg; the
 * per-launch fields are 0).  For matrices smaller than twice the 256 MB
 * Infinity Cache, where back-to-back applies can hit in it.  No reference
 * counterpart (the reference times without flushing, hmvm_cuda1.cu:236-302). */
int hmvm_plan_time_flushed(hmvm_plan *plan, long long flush_bytes, int iters, hmvm_timing *timing);

/* Switches the tile-kernel variant and workgroup size of an existing plan
 * (used by the autotuner; results are identical across variants up to the
 * summation order inside a tile, which every variant keeps). */
int hmvm_plan_set_launch(hmvm_plan *plan, int kernel, int block_threads);
/* The same for one launch only: which = 0 (A: stage 1 + dense) or 1 (B). */
int hmvm_plan_set_phase_launch(hmvm_plan *plan, int which, int kernel, int block_threads);
/* Back to the measured per-storage defaults of both launches. */
int hmvm_plan_reset_launch(hmvm_plan *plan);

/* Several right-hand sides on the matrix cores: Y (+)= A X with X, Y of
 * nd x nrhs doubles, row-major (X[i*nrhs + c]); nrhs = 16, 32 or 64.  FP64
 * storage runs v_mfma_f64_16x16x4, FP32 v_mfma_f32_16x16x4 (FP32 accumulate
 * in a tile, FP64 across tiles), BF16 v_mfma_f32_16x16x16_bf16 with the
 * vector entries rounded to BF16 (FP32 and BF16 storage read an FP32 copy
 * of X that the apply makes on the device).  hmvm_plan_reserve_multi
 * allocates the per-plan work buffers for up to nrhs columns and lays the
 * tiles out for the 8 XCDs (workgroups over the same columns of X share an
 * L2); the apply itself allocates nothing unless the plan's workgroup size
 * changed since (it then lays the tiles out again, a synchronous upload:
 * call hmvm_plan_reserve_multi again before capturing the apply in a graph).
 * The host form reserves by itself.  d_X and d_Y need 8-byte alignment; a
 * 16-byte-aligned d_Y gets 16-byte stores, and an FP64 plan copies a d_X
 * that is not 16-byte aligned to its own buffer first (same values). */
int hmvm_plan_reserve_multi(hmvm_plan *plan, int nrhs);
int hmvm_plan_apply_multi(hmvm_plan *plan, const double *d_X, double *d_Y, int nrhs, int accumulate, void *stream);
int hmvm_plan_apply_multi_host(hmvm_plan *plan, const double *X, double *Y, int nrhs, int accumulate);
int hmvm_plan_time_multi(hmvm_plan *plan, int nrhs, int warmup, int iters, hmvm_timing *timing);

/* Selects the execution mode of an existing plan (see hmvm_options.mode).
 * Mode 1 deals the tiles to the persistent launch's waves here, and again
 * when hmvm_plan_set_launch / _set_phase_launch / _reset_launch /
 * _set_tile_flags / _set_tile_order change them (a synchronous upload: make
 * such changes before capturing applies in a graph). */
int hmvm_plan_set_mode(hmvm_plan *plan, int mode);
/* 1 when a bounded wait inside a persistent launch gave up since the last
 * report (the results of that launch are invalid), 0 otherwise; synchronises
 * the plan's stream and clears the flag.  The blocking entry points
 * (hmvm_plan_apply_host, the timing functions) check it themselves and return
 * HMVM_ERR_WAIT. */
int hmvm_plan_fault(hmvm_plan *plan);

/* Achievable HBM bandwidth on this device: a read-only stream and a copy of
 * `bytes` (each repeated `iters` times); GB/s counts bytes read (+ written). */
int hmvm_bandwidth_probe(size_t bytes, int iters, double *read_gbs, double *copy_gbs);

/* Host-only check of the schedule a plan would use for this matrix (no GPU
 * access): every element streamed once, every access in bounds. */
int hmvm_plan_check(const matrix2 *mat2, const hmvm_options *opt, hmvm_plan_info *info);

/* ---- int64 form ---------------------------------------------------------------
 * Matrices of more than 2^31-1 stored elements (matrix2_64, hacapk.h); the
 * reference has no counterpart, its int32 len overflows there.  Semantics as
 * hmvm_plan_create and hmvm_hip_2 / hmvm_hip_2t (accumulate, cached upload,
 * hmvm_hip_release drops the device copy). */
int hmvm_plan_create64(const matrix2_64 *mat2, int a2_source, const hmvm_options *opt, hmvm_plan **out);
void hmvm_hip_2_64(double *v, const matrix2_64 *mat2, const double *b);
void hmvm_hip_2t_64(double *v, const matrix2_64 *mat2, const double *b);

/* Plan straight from a format 1, 2 or 2L file: the leaf descriptors are read
 * first, then the values stream through two pinned host buffers into HBM and
 * are repacked there window by window; the whole matrix is never held in host
 * memory (replaces loadHmatrix/loadHmatrix2 + upload, src_c/loadmatrix.c:8-232). */
int hmvm_plan_create_file(const char *fname, const hmvm_options *opt, hmvm_plan **out);

/* One plan over `count` independent matrices of any sizes: they are placed on
 * the diagonal of one block-diagonal matrix (matrix q owns rows and columns
 * row_offsets[q] .. row_offsets[q+1]-1 of x and y), so a single apply (one
 * set of launches, or one graph replay) multiplies all of them.
 * row_offsets (count+1 ints) may be NULL. */
int hmvm_plan_create_batch(const matrix2 *const *mats, int count, int a2_source, const hmvm_options *opt,
                           hmvm_plan **out, int *row_offsets);

/* ---- hipGraph form -------------------------------------------------------------
 * `count` applies d_y (+)= A d_x captured once and replayed by
 * hmvm_graph_launch on any stream; d_x/d_y are fixed at capture. */
typedef struct hmvm_graph hmvm_graph;
int hmvm_plan_graph_create(hmvm_plan *plan, const double *d_x, double *d_y, int accumulate, int count,
                           hmvm_graph **out);
int hmvm_graph_launch(hmvm_graph *graph, void *stream);
void hmvm_graph_destroy(hmvm_graph *graph);
/* Times graph replays of `count` applies on the plan's internal vectors;
 * results per apply (ms_avg/min/median per replay / count; ms_stream with
 * replays back to back). */
int hmvm_plan_time_graph(hmvm_plan *plan, int count, int warmup, int iters, hmvm_timing *timing);

/* ---- tile-level measurement and scheduling (autotuner, cost model) ------------
 * No reference counterpart: the reference chooses its kernel variant by a
 * human reading timing CSVs (src_cpp/hmvm1.cpp:82-96).
 *
 * A tile is the work of one wavefront.  hmvm_plan_tile_features writes
 * HMVM_NFEAT floats per tile of launch A (which = 0) or B (which = 1):
 *   [0..2] one-hot kind: stage-1 / dense / stage-2   [3] log2 rows
 *   [4] log2(1+columns)  [5] log2 panels  [6] log2 lanes per column
 *   [7] log2(1+elements read incl. lane padding)  [8] log2(1+mean partial
 *   slots summed per vector entry)  [9] log2 bytes per element
 *   [10..12] kernel variant bits  [13] log2(block_threads/64)  [14] log2 nrhs
 *   [15] log2 min(tiles in the launch, W)  [16] log2 W, where W is the
 *   number of wavefronts of the launch's variant resident at once on the
 *   device (hmvm_tile_occupancy): how saturated the launch is, bounded so
 *   that no launch size lies outside what a model has seen
 *   [17] 1 when the tile runs the lane-per-row path (more than 32 lanes per
 *   column, or HMVM_TILE_ROW_PER_LANE set), 0 for the column-split path
 * hmvm_plan_profile_tiles runs the launch with a clock-stamping copy of the
 * tile kernel (same results) and returns each tile's mean duration and start
 * offset in microseconds, plus the mean launch time.  Before each timed
 * launch it overwrites 512 MB of scratch (twice the last-level cache), so
 * the launch reads its matrix values from HBM even when the matrix would fit
 * in cache: small matrices then give the durations of the streaming regime
 * that large ones run in. 
 * hmvm_plan_set_tile_order permutes the dispatch order of a launch (results
 * are unchanged: tiles write disjoint slots).  Modes 1 and 2 need launch A as
 * "stage-1 tiles, then dense": selecting them regroups launch A so, keeping
 * each part's order. */
enum { HMVM_NFEAT = 18 };
int hmvm_plan_tile_features(const hmvm_plan *plan, int which, int nrhs, float *feat);
/* Per-tile kernel path of launch A (which = 0) or B (1), in the launch's
 * current dispatch order: flags[i] = HMVM_TILE_ROW_PER_LANE makes tile i run
 * the lane-per-row path (each lane owns row groups and reads whole columns)
 * instead of the column-split path (lane groups take different columns, an
 * LDS reduction adds them); 0 restores the default.  Only tiles of at most
 * 32 lanes per column have the choice; the flag is ignored on the others.
 * Results differ from the default only in the summation order inside a
 * tile.  hmvm_plan_get_tile_flags reads them back. */
enum { HMVM_TILE_ROW_PER_LANE = 1 };
int hmvm_plan_set_tile_flags(hmvm_plan *plan, int which, const int *flags);
int hmvm_plan_get_tile_flags(const hmvm_plan *plan, int which, int *flags);
int hmvm_plan_profile_tiles(hmvm_plan *plan, int which, int iters, double *tile_us, double *start_us,
                            double *launch_ms);
int hmvm_plan_set_tile_order(hmvm_plan *plan, int which, const int *perm);
/* Wavefronts of the tile kernel (storage, variant, workgroup size) that can be
 * resident on the current device at once. */
int hmvm_tile_occupancy(int storage, int kernel, int block_threads, int *waves);

/* ---- the learned cost model: a small MLP on the GPU ---------------------------
 * nin -> nh -> nh -> 1, ReLU, FP32; inputs are standardised with the mean and
 * standard deviation stored in the model.  Forward, backward and Adam steps
 * are HIP kernels; the host only moves data.  Parameters are laid out as
 * W1 (nh x nin, row-major), b1 (nh), W2 (nh x nh), b2 (nh), w3 (nh), b3 (1). */
typedef struct hmvm_mlp hmvm_mlp;
int hmvm_mlp_create(int nin, int nh, unsigned long long seed, hmvm_mlp **out);
void hmvm_mlp_destroy(hmvm_mlp *m);
int hmvm_mlp_num_params(const hmvm_mlp *m);
int hmvm_mlp_get_params(const hmvm_mlp *m, float *params, float *mean, float *stdev);
int hmvm_mlp_set_params(hmvm_mlp *m, const float *params, const float *mean, const float *stdev);
/* Sets mean/stdev from the columns of X (n x nin). */
int hmvm_mlp_fit_scaler(hmvm_mlp *m, const float *X, int n);
/* y[i] = f(X[i]) for n rows (host arrays). */
int hmvm_mlp_predict(hmvm_mlp *m, const float *X, int n, float *y);
/* Loss 0.5/n * sum (f(X_i) - y_i)^2 and its gradient wrt the parameters
 * (computed on the GPU, deterministic order). */
int hmvm_mlp_loss_grad(hmvm_mlp *m, const float *X, const float *y, int n, float *loss, float *grad);
/* Adam (lr, beta1 0.9, beta2 0.999) over shuffled minibatches; returns the
 * final full-data loss. */
int hmvm_mlp_train(hmvm_mlp *m, const float *X, const float *y, int n, int epochs, int batch, float lr,
                   unsigned long long seed, float *final_loss);

/* ---- Krylov driver -----------------------------------------------------------
 * BiCGSTAB for A u = b with this plan as A, iterating as the reference's
 * HACApK_bicgstab_lfmtx_hyp (src_hdump/m_HACApK_solve.f90:306-380): two HMVMs
 * per iteration, all scalars on the device.  use_graph != 0 captures one
 * iteration in a hipGraph and replays it.  As in the reference, the solve
 * stops at the first iteration whose ||r||/||b|| is below tol (tol <= 0: run
 * max_iter iterations): the device records that iteration and turns the
 * updates of later ones into no-ops, so u and iters are exactly those of the
 * reference loop even though the host only reads the state back every
 * check_every iterations (0: only at the end) and launches whole chunks.  A
 * zero divisor or a non-finite scalar stops the updates the same way and the
 * call returns HMVM_ERR_BREAKDOWN with u the last finite iterate.  u is the
 * initial guess on entry and the iterate on return (host pointers). */
typedef struct hmvm_krylov_result {
  int iters;          /* iterations that updated u (the reference's nstp)      */
  int hmvms;          /* HMVMs the solve needed: 2 per iteration + 1           */
  double relres;      /* ||r||/||b|| of the recurrence at the end              */
  double ms_loop;     /* device time of the iteration loop (0 when not timed)  */
  double ms_per_iter;
  int graph_used;
} hmvm_krylov_result;

int hmvm_bicgstab(hmvm_plan *plan, const double *b, double *u, int max_iter, double tol, int use_graph,
                  int check_every, hmvm_krylov_result *result);

/* hmvm_bicgstab without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight solver
 * loops, harnesses that forbid hipMalloc).  b and u are DEVICE vectors of nd
 * doubles (u: initial guess on entry, iterate on return); the work vectors
 * live in a fixed array of the library, one per device, so nd is at most
 * HMVM_BICGSTAB_STATIC_MAX_N (larger plans get HMVM_ERR_ARG).  Runs on
 * `stream` (NULL: the null stream) with the plan's device current, eagerly
 * (no graph), and returns when the solve has stopped; ms_loop is not timed
 * (0).  Iterates, stopping rule and results are those of hmvm_bicgstab with
 * use_graph = 0.  Every call on a device shares that device's work array:
 * calls that overlap on one device race. */
#define HMVM_BICGSTAB_STATIC_MAX_N 1048576
int hmvm_bicgstab_static(hmvm_plan *plan, const double *d_b, double *d_u, int max_iter, double tol,
                         int check_every, void *stream, hmvm_krylov_result *result);

/* ---- schedule introspection (host only) -------------------------------------
 * The schedule a plan runs, as plain arrays: the autotuner's features, and
 * the CPU tests that replay it.  Element (r, j) of a tile's panel p is
 * values[val_off + r + j*ld]; row r of a tile goes to slot out_off + r, or to
 * t[tmap[map_off + r]] when map_off >= 0 (in the multi-RHS launches: when
 * map_off_multi >= 0, and their t sum leaves the entries so written alone);
 * the panel's vector entry j is
 * x[vec_off + j] (vec_sel 0), t[vec_off + j] (1) or
 * sum_{c<nch} slots[vec_off + c*stride + j] (2, vec_chunks = nch<<16 | stride).  Repack record: values[dst_off + r +
 * c*dst_cs] = source[src_off + r*src_rs + c*src_cs].  Segment sums:
 * out[i] (+)= sum_{e in ptr[seg[i]]..ptr[seg[i]+1]} slots[list[e] + i]. */
typedef struct hmvm_panel {
  long long val_off;
  int ld, ncols, vec_off, vec_sel, factor, vec_chunks;
} hmvm_panel;
/* first: a copy of panel panel0; ncols: the columns of all its panels;
 * map_off_multi: the row map the multi-RHS launches use in place of map_off */
typedef struct hmvm_tile {
  int nrows, panel0, npanel, out_off, map_off, flags, ncols, map_off_multi;
  hmvm_panel first;
} hmvm_tile;
typedef struct hmvm_repack {
  long long src_off, dst_off;
  int rows, cols, src_rs, src_cs, dst_cs, factor;
} hmvm_repack;
typedef struct hmvm_segsum {
  int n, nseg;
  int *seg, *ptr;
  long long *list;
} hmvm_segsum;
typedef struct hmvm_schedule {
  int ntiles_a, ntiles_b, npanels, nrepack, t_fused;
  long long nvals, n_t, n_tslots, n_zslots;
  hmvm_tile *tiles_a, *tiles_b;
  hmvm_panel *panels;
  hmvm_repack *repack;
  hmvm_segsum tsum, ysum;
  int t_direct;               /* stage-1 tiles with map_off >= 0 write t[tmap[map_off + r]] */
  long long ntmap;
  int *tmap;
} hmvm_schedule;

int hmvm_schedule_export(const matrix2 *mat2, int a2_source, const hmvm_options *opt, hmvm_schedule *out);
/* The schedule of hmvm_plan_create64 for the same matrix. */
int hmvm_schedule_export64(const matrix2_64 *mat2, int a2_source, const hmvm_options *opt, hmvm_schedule *out);
/* The schedule of hmvm_plan_create_batch for the same matrices. */
int hmvm_schedule_export_batch(const matrix2 *const *mats, int count, int a2_source, const hmvm_options *opt,
                               hmvm_schedule *out);
void hmvm_schedule_release(hmvm_schedule *s);

/* ---- small device helpers for callers without a HIP runtime of their own -- */
int hmvm_device_count(void);
int hmvm_device_alloc(void **ptr, size_t bytes);
int hmvm_device_free(void *ptr);
int hmvm_memcpy_h2d(void *dst, const void *src, size_t bytes);
int hmvm_memcpy_d2h(void *dst, const void *src, size_t bytes);
int hmvm_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* HACAPK_AMD_HMVM_HIP_H */
