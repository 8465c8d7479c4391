/*
 * rt_spmspm.h -- C ABI of the MI355X-native SpGEMM (C = A * B, CSR in, CSR out).
 *
 * The reference (escalab/RTSpMSpM) has no library API: its only entry point is
 * the `optixSpMSpM` command line (optixSpMSpM/src/optixSpMSpM/optixSpMSpM.cpp:496-660),
 * which parses two Matrix Market files, traces one ray per nonzero of A against
 * one sphere per nonzero of B (optixSpMSpM.cu:113-201,220-281) and writes C plus
 * a latency.  This header is the thin C boundary that replaces the work done
 * between the file parse and the result dump of that program:
 *
 *   mat1ToGPU / storeSphereData   (optixSpMSpM.cpp:73-162)   -> rt_csr inputs
 *   buildGAS .. optixLaunch       (optixSpMSpM.cpp:585-611)  -> rt_spmspm / rt_spmspm_ex
 *   resultBufferSetup, printResult(optixSpMSpM.cpp:164-171,452-494) -> rt_csr output
 *
 * Every function returns 0 on success and a negative RT_E* code on failure;
 * no C++ exception crosses this boundary (the reference turns CUDA_CHECK /
 * OPTIX_CHECK exceptions into `return 1` in main, optixSpMSpM.cpp:654-658).
 * rt_last_error() gives the message of the last failure on the calling thread.
 *
 * CSR convention for every rt_csr: 0-based, int64 row pointers of length
 * rows+1, int32 column indices sorted ascending inside each row, no duplicate
 * columns, fp32 values.  C has dims (A.rows x B.cols) and its structure is the
 * union pattern of the products: an entry produced by cancellation (a sum that
 * is exactly 0.0f) is KEPT as a stored zero.
 */
#ifndef RT_SPMSPM_H
#define RT_SPMSPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define RT_OK            0
#define RT_E_INVALID    -1   /* bad argument, malformed CSR, NULL pointer      */
#define RT_E_SHAPE      -2   /* A.cols != B.rows                               */
#define RT_E_NOMEM      -3   /* device or host allocation failed               */
#define RT_E_HIP        -4   /* a HIP runtime call failed                      */
#define RT_E_IO         -5   /* file could not be opened / parsed              */
#define RT_E_LIMIT      -6   /* size beyond what the kernels support           */
#define RT_E_NODEVICE   -7   /* no GPU visible                                 */

/* ---- memory spaces of an rt_csr ---------------------------------------- */
#define RT_MEM_HOST      0
#define RT_MEM_DEVICE    1

typedef struct rt_csr {
  int64_t  rows, cols, nnz;
  int64_t* row_ptr;   /* rows+1 entries                                    */
  int32_t* col;       /* nnz entries                                       */
  float*   val;       /* nnz entries                                       */
  int32_t  mem;       /* RT_MEM_HOST or RT_MEM_DEVICE                      */
  int32_t  owner;     /* 0: caller's arrays; 1: allocated by this library  */
} rt_csr;

/* ---- algorithm selection ------------------------------------------------ */
#define RT_MODE_PLAIN    0  /* exact two-phase: symbolic hash, scan, numeric  */
#define RT_MODE_RAY      1  /* faithful one-work-item-per-A-nonzero launch    */
#define RT_MODE_LEARNED  2  /* MLP-predicted row sizing, exact rerun on overflow */

/* Ablations of the faithful ray kernel.  They mirror the reference's compile
 * time variants (optixSpMSpM.cu:106-111 NOTHING, :151-189 NOINT,
 * :222-227 ARCHSUP, :228-229 NOMEM; src/CMakeLists.txt:348-376) as run-time
 * switches and exist for timing.  Only RT_ABL_FULL produces a product.  With
 * the dense output (RT_RAY_DENSE, and RT_RAY_AUTO where it picks dense) the
 * other ablations touch no output entry, so C comes back with the right
 * shape and no entries, as the reference's ablation builds leave their dense
 * result all zero.  With RT_RAY_SLOT they return C's exact structure (from
 * the symbolic pass) with unspecified values.                               */
#define RT_ABL_FULL        0  /* ATOMIC: one global atomic add per product    */
#define RT_ABL_NO_ATOMIC   1  /* ARCHSUP: non-atomic +1 into a 256-entry buffer */
#define RT_ABL_TRAVERSE    2  /* NOMEM: walk B's row, accumulate nothing      */
#define RT_ABL_NO_HIT      3  /* NOINT: no B access, dummy loop of nnz/5 steps */
#define RT_ABL_EMPTY       4  /* NOTHING: launch only                         */

/* Output buffer of the faithful kernel.                                     */
#define RT_RAY_AUTO        0  /* dense when rows*cols*5 bytes <= 4 GiB        */
#define RT_RAY_DENSE       1  /* dense fp32 C + touch map, as the reference   */
#define RT_RAY_SLOT        2  /* CSR slots located by the exact symbolic pass */

/* Long rows (more than the last hash bin holds) in the learned method.     */
#define RT_LONG_DEFAULT    0  /* the tuned choice (exact path untuned)        */
#define RT_LONG_LEARNED    1  /* sized by the predictor, learned tile kernel   */
#define RT_LONG_EXACT      2  /* symbolic bitmap + segmented numeric kernels   */

/* rt_opts.flags.                                                           */
#define RT_FLAG_TRUSTED    1  /* device operands are known-valid CSR: skip the
                                 on-device structure check (k_validate).  A
                                 malformed operand then reads out of bounds. */

typedef struct rt_opts {
  int32_t mode;        /* RT_MODE_*                                         */
  int32_t ablation;    /* RT_ABL_* (RT_MODE_RAY only)                       */
  int32_t ray_output;  /* RT_RAY_* (RT_MODE_RAY only)                       */
  int32_t profile;     /* 1: time every kernel launch with hipEvents        */
  int32_t use_tuned;   /* 1: use/record autotuned bin thresholds            */
  int32_t graph;       /* 0 = default (on for batched calls), 1 = capture
                          each stage in a hipGraph, 2 = direct launches    */
  int32_t learned_long;/* RT_MODE_LEARNED: RT_LONG_* handling of long rows */
  int32_t flags;       /* RT_FLAG_* bits                                    */
  /* outputs */
  double   ms_total;       /* hipEvent time, inputs resident -> C complete  */
  int64_t  products;       /* intermediate products P                       */
  int64_t  overflow_rows;  /* learned mode: rows rerun through exact path   */
  int64_t  predicted_rows; /* learned mode: rows sized by the predictor     */
} rt_opts;

/* ---- core entry points --------------------------------------------------- */

/* C = A * B with the default exact two-phase method.  A and B are borrowed
 * and must live in the same memory space; C's arrays are allocated by the
 * library in that same space (owner = 1) and released with rt_csr_free.    */
int rt_spmspm(const rt_csr* A, const rt_csr* B, rt_csr* C);

/* As rt_spmspm with an explicit method and an explicit HIP stream
 * (hipStream_t passed as void*, NULL = the library's own stream).  The call
 * is synchronous: when it returns C is complete.  opts may be NULL.         */
int rt_spmspm_ex(const rt_csr* A, const rt_csr* B, rt_csr* C,
                 rt_opts* opts, void* stream);

/* Many independent products in one pass.  A_stack holds the A_b matrices
 * stacked row-wise (rows = sum of a_rows[b]; col indices local to each pair,
 * i.e. in [0, b_rows[b])); B_stack holds the B_b stacked row-wise (rows = sum
 * of b_rows[b]; col indices local, in [0, b_cols[b])).  C_stack receives the
 * C_b stacked row-wise with local column indices; rows of pair b start at
 * sum_{c<b} a_rows[c].  Device memory only.  mode RT_MODE_PLAIN or
 * RT_MODE_LEARNED.                                                           */
int rt_spmspm_batched(int32_t count, const int64_t* a_rows, const int64_t* b_rows,
                      const int64_t* b_cols, const rt_csr* A_stack,
                      const rt_csr* B_stack, rt_csr* C_stack,
                      rt_opts* opts, void* stream);

/* ---- product without device allocation ----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the exact two-phase product of device operands A and B
 * written into the caller's arrays, with every scratch buffer in one
 * file-scope device array of RT_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends.  No hipMalloc, hipFree or allocating library call is made.
 *   c_row_ptr  device array of A.rows + 1 entries (always written when the
 *              product fits the scratch)
 *   c_col, c_val  device arrays of c_capacity entries (may be NULL when
 *              c_capacity is 0)
 *   *c_nnz     nnz(C); -1 when the product needs more scratch than there is
 * Returns RT_E_LIMIT when C needs more than c_capacity entries (c_row_ptr
 * then holds C's row pointer and c_col / c_val are untouched: call again
 * with arrays of *c_nnz entries), and RT_E_LIMIT with a message when the
 * operands need more scratch than RT_NOALLOC_SCRATCH_BYTES (rows of A and B
 * up to a few hundred thousand fit).  flags: RT_FLAG_TRUSTED skips the
 * device check of the operands.  Calls on one GPU share its scratch: within
 * a process they are serialised; calls that overlap on one GPU from several
 * processes race.  Synchronous on `stream` (NULL: the library's stream).    */
#define RT_NOALLOC_SCRATCH_BYTES (64ll << 20)
int rt_spmspm_noalloc(const rt_csr* A, const rt_csr* B, int64_t* c_row_ptr, int32_t* c_col,
                      float* c_val, int64_t c_capacity, int64_t* c_nnz, int32_t flags,
                      void* stream);

/* ---- numeric-only re-run on a fixed structure ---------------------------
 * Counterpart of cuSPARSE's SpGEMMreuse path used by the reference's
 * comparator when built with REUSE (cuSparse/src/main.cpp:98-363, switch at
 * :648-658): the structure of C (symbolic pass, row binning, long-row tile
 * counts) is computed once by rt_plan_create, which also returns C; each
 * rt_plan_execute recomputes C's values from new values of A and B with the
 * same structure (same row_ptr / col of A and B).  Device operands only;
 * A's and B's row_ptr/col arrays are borrowed by the plan until it is
 * destroyed.                                                                 */
typedef struct rt_plan rt_plan;
int rt_plan_create(const rt_csr* A, const rt_csr* B, rt_plan** plan, rt_csr* C,
                   void* stream);
/* a_val / b_val: device arrays of at least A.nnz / B.nnz new values (the
 * nnz of the operands given to rt_plan_create; shorter arrays are read past
 * their end); C must be the rt_csr returned by rt_plan_create (its val array
 * is overwritten).  ms (may be NULL) receives the device time.              */
int rt_plan_execute(rt_plan* plan, const float* a_val, const float* b_val, rt_csr* C,
                    double* ms);
void rt_plan_destroy(rt_plan* plan);

/* Release arrays that the library allocated (owner == 1); zeroes *C.       */
void rt_csr_free(rt_csr* C);

/* Message of the last failure on this thread ("" if none).                 */
const char* rt_last_error(void);

/* ---- memory helpers ------------------------------------------------------ */
int rt_csr_to_device(const rt_csr* host, rt_csr* dev);   /* dev: owner = 1  */
int rt_csr_to_host(const rt_csr* dev, rt_csr* host);     /* host: owner = 1 */
/* Plain device -> host copy of `bytes` bytes (e.g. only C.row_ptr).        */
int rt_memcpy_d2h(void* dst, const void* src, int64_t bytes);
/* Device -> device copy with the library's streaming copy kernel
 * ("device_copy"; 16-byte vector loads/stores).  Also the reference point
 * for reading hardware byte counters.  *ms: kernel time (may be NULL).     */
int rt_device_copy(void* dst, const void* src, int64_t bytes, double* ms);
/* Device buffer from the library pool (and back), for callers without HIP. */
int rt_device_alloc(void** ptr, int64_t bytes);
int rt_device_free(void* ptr);
int rt_device_synchronize(void);
/* Return cached device blocks of the library's pool to the driver.         */
int rt_pool_trim(void);
/* Bytes and blocks handed out by the pool and not yet returned, and bytes
 * cached for reuse (any pointer may be NULL).                               */
int rt_pool_stats(int64_t* live_bytes, int64_t* live_blocks, int64_t* cached_bytes);
/* Test hook for the out-of-memory paths: the n-th pool allocation from now
 * (n >= 1) fails as a failed hipMalloc does (RT_E_NOMEM), once; n = 0
 * disarms it.  Lets tests check that every entry point then returns
 * RT_E_NOMEM and releases what it had taken (rt_pool_stats).                */
int rt_debug_fail_alloc(int64_t n);
/* A HIP stream of the library's device at the default priority (returned as
 * void*), for callers without HIP that run products concurrently from several
 * host threads, one stream each.                                            */
int rt_stream_create(void** stream);
int rt_stream_destroy(void* stream);
int rt_device_info(char* name, int64_t name_len, int64_t* total_mem, int32_t* cus);

/* ---- autotuner ------------------------------------------------------------
 * Times a set of candidate row-bin settings (bin thresholds, long-row cut,
 * lanes per row of each hash bin and the last bin's table size, learned
 * handling of rows) for `mode` (RT_MODE_PLAIN or RT_MODE_LEARNED) on the
 * given device operands, `reps` timed runs each, and records the fastest
 * for the operands' matrix class in the tuning cache (RTSPMSPM_TUNE_CACHE or
 * ~/.cache/rtspmspm/tune.txt).  Later calls with rt_opts.use_tuned = 1 use
 * it.  desc receives a one-line summary: the class, each candidate's time,
 * the chosen bins as slots/max nnz/lanes, and "thresholds sym <5> num <5>
 * learned_heavy <0|1> shapes <5> learned_bins <mask>".
 * No counterpart in the reference (its pipeline is fixed at build time).   */
int rt_autotune(const rt_csr* A, const rt_csr* B, int32_t mode, int32_t reps, char* desc,
                int64_t desc_len);
/* As rt_autotune for the batched call (rt_spmspm_batched with the same
 * arguments): the batched class is tuned and cached separately.             */
int rt_autotune_batched(int32_t count, const int64_t* a_rows, const int64_t* b_rows,
                        const int64_t* b_cols, const rt_csr* A_stack, const rt_csr* B_stack,
                        int32_t mode, int32_t reps, char* desc, int64_t desc_len);
/* Reads a tuning-cache file as the library does and counts the lines it
 * would use (*valid) and those it would drop as malformed, out of range or
 * of another format version (*rejected).  Host only.                        */
int rt_tune_cache_check(const char* path, int32_t* valid, int32_t* rejected);

/* ---- per-kernel timing (rt_opts.profile = 1) -----------------------------
 * Kernel launch durations measured with hipEvents recorded on the stream
 * each kernel is launched on, accumulated since the last reset.             */
void rt_prof_reset(void);
int32_t rt_prof_count(void);
/* i-th entry: name (copied into name[name_len]), launches, total ms.       */
int rt_prof_get(int32_t i, char* name, int64_t name_len, int64_t* launches,
                double* total_ms);

/* ---- learned row-nnz predictor ------------------------------------------ */
#define RT_NFEAT 8
/* Per-row features and exact row nnz of C = A*B (device computation; the
 * arrays are host arrays of A.rows*RT_NFEAT floats and A.rows int64).       */
int rt_row_features(const rt_csr* A, const rt_csr* B, float* feat, int64_t* row_nnz);
/* Number of fp32 parameters of the MLP (8 -> 32 -> 32 -> 1).               */
int64_t rt_mlp_param_count(void);
/* Load / fetch the predictor weights used by RT_MODE_LEARNED (host arrays). */
int rt_mlp_set_weights(const float* params, int64_t n, float margin);
int rt_mlp_get_weights(float* params, int64_t n, float* margin);
/* Train on the device with the library's forward/backward/Adam kernels.
 * feat: n x RT_NFEAT, target: n (log2(1+nnz)); host arrays.  Returns the
 * final mean-squared loss in *loss.                                         */
int rt_mlp_train(const float* feat, const float* target, int64_t n,
                 int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 float* loss);
/* One forward+backward on the device for a given batch: loss and gradient
 * (host arrays; grad has rt_mlp_param_count() entries).                     */
int rt_mlp_loss_grad(const float* params, const float* feat, const float* target,
                     int64_t n, float* loss, float* grad);
/* Device inference: pred[n] = network(feat) (log2(1+nnz) units).            */
int rt_mlp_predict(const float* params, const float* feat, int64_t n, float* pred);

/* ---- seeded synthetic generators (host CSR, owner = 1) ------------------- */
/* n x n, exactly k distinct uniform columns per row, values U(-1,1).       */
int rt_gen_uniform_k(int64_t n, int64_t k, uint64_t seed, rt_csr* out);
/* R-MAT 2^scale x 2^scale, edge_factor*2^scale edges, quadrant probs a,b,c
 * (d = 1-a-b-c), duplicates merged, self-loops kept, values U(-1,1).        */
int rt_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                uint64_t seed, rt_csr* out);
/* 27-point stencil on an n^3 grid: diag 26, off-diagonals -1 + U(-.01,.01). */
int rt_gen_stencil27(int64_t n, uint64_t seed, rt_csr* out);
/* 2x2x2 aggregation of an n^3 grid: P (n^3 x (n/2)^3) and R = P^T.           */
int rt_gen_aggregation(int64_t n, rt_csr* P, rt_csr* R);
/* rows x cols, row lengths clip(round(lognormal(mu,sigma)),lo,hi), uniform
 * columns, values N(0,1).                                                   */
int rt_gen_lognormal_rows(int64_t rows, int64_t cols, double mu, double sigma,
                          int64_t lo, int64_t hi, uint64_t seed, uint64_t stream,
                          rt_csr* out);
/* rows x cols, exactly k uniform distinct columns per row, values N(0,1).   */
int rt_gen_fixed_normal(int64_t rows, int64_t cols, int64_t k, uint64_t seed,
                        uint64_t stream, rt_csr* out);
/* n x n with Poisson(lambda) columns per row (clipped to n), U(-1,1).       */
int rt_gen_poisson(int64_t n, double lambda, uint64_t seed, uint64_t stream,
                   rt_csr* out);
/* `count` independent square pairs, n_b = round(exp(U(ln nmin, ln nmax))),
 * Poisson(lambda_b) columns per row with lambda_b ~ U(lmin, lmax), values
 * U(-1,1); written stacked row-wise in the layout rt_spmspm_batched takes.
 * a_rows/b_rows/b_cols: caller arrays of `count` entries (all = n_b).       */
int rt_gen_batch_poisson(int32_t count, int64_t nmin, int64_t nmax, double lmin,
                         double lmax, uint64_t seed, int64_t* a_rows, int64_t* b_rows,
                         int64_t* b_cols, rt_csr* A_stack, rt_csr* B_stack);

/* ---- Matrix Market I/O (host) -------------------------------------------- */
/* Coordinate real/integer/pattern general|symmetric; 1-based; any entry
 * order; duplicates summed; result sorted CSR (owner = 1).                  */
int rt_mtx_read(const char* path, rt_csr* out);
/* ref_compat = 0: every stored entry, row-major, 9 significant digits.
 * ref_compat = 1: the reference's printResult format (optixSpMSpM.cpp:452-494):
 *                 exact zeros dropped, 6 significant digits, true count.
 * ref_compat = 2: as 1, byte for byte what printResult leaves on disk: the
 *                 size line is written as "rows cols " and later overwritten
 *                 in place by "rows cols 0" (its entry map is never filled,
 *                 :465,488-491), which is one byte longer and so replaces
 *                 the first byte of the first entry line.                   */
int rt_mtx_write(const char* path, const rt_csr* C, int32_t ref_compat);

#ifdef __cplusplus
}
#endif
#endif /* RT_SPMSPM_H */
