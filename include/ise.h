/*
 * ise.h -- C ABI of the MI355X batched canonicity test.
 *
 * Plain C: pointers, sizes and status codes only.  One context per host
 * thread (the reference keeps one CGPU_CanonChecker per enumerator thread,
 * Source/ThreadEnumerator.h:96, each with its own stream,
 * Source/MatrixCanonChecker.h:264); calls on one context are not thread-safe.
 *
 * Wire format of a batch of BIBD-type candidates (all little-endian):
 *   rows  [n][stride] uint64  bit c of word i = entry (row i, column c)
 *   cells [n][stride] uint64  column-orbit start mask of row i, i.e. bit c set
 *                             when c == 0 or columns c-1 and c differ in some
 *                             row j < i.  May be NULL: it is then derived from
 *                             the rows (the candidate's columns must be
 *                             grouped so that orbits are contiguous with ones
 *                             left-justified, as MakeRow builds them,
 *                             Source/MatrixCanonChecker.cpp:14-93).
 *   meta  [n] uint32          depth | ncols << 8 | lenStab << 16 | outInfo << 24
 * with 1 <= depth <= 32, 1 <= ncols <= 64, depth <= stride.
 *
 * Results (structure of arrays; any pointer except verdict may be NULL):
 *   verdict        [n] uint8   1 canonical, 0 not canonical, 2 search limit hit
 *   fail_row       [n] uint8   row at which the permuted matrix was larger (255)
 *   fail_col       [n] uint8   first column of the deciding column orbit (255)
 *   row_to_change  [n] uint8   CCanonicityChecker::rowToChange of fail_row (255)
 *   group_order    [n] uint64  |Aut| (row action) when outInfo has
 *                              ISE_SAVE_ROW_PERMUTATIONS and canonical, else
 *                              1 (canonical) / 0 (not canonical)
 *   perm_row       [n][32] uint8  row permutation at exit (witness if verdict 0)
 *   perm_col       [n][64] uint8  column permutation at exit
 *   orbits         [n][32] uint8  row orbits under the automorphisms found
 *   n_gens         [n] uint16  automorphisms found
 *   gens           [n][gen_cap][64] uint8  per automorphism: the permutation
 *                              it induces on the column orbits of the next row
 *                              (partial check) or the row permutation
 *                              (ISE_SAVE_ROW_PERMUTATIONS); unused bytes 255
 *   counters       [n][4] uint32  permutations tried, rows compared, column
 *                              orbits compared, columns examined
 *   group_order_hi [n] uint64  bits 64..127 of |Aut|: the reference keeps
 *                              group orders in a 128-bit UInt
 *                              (k-Factorization/TripleSys/Include/Global.h:197-211),
 *                              and the row group of a 32-row matrix can
 *                              exceed 2^64 (32! > 2^117); 0 whenever
 *                              group_order is not |Aut|
 */
#ifndef ISE_H
#define ISE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISE_MAX_ROWS 32
#define ISE_MAX_COLS 64

/* outInfo flags: same bits as t_canonOutInfo (Source/CanonicityChecker.h:27-32) */
#define ISE_SAVE_ROW_PERMUTATIONS 1u
#define ISE_SAVE_COL_PERMUTATIONS 2u
#define ISE_SAVE_ROW_TO_CHANGE    4u

#define ISE_VERDICT_NONCANONICAL 0
#define ISE_VERDICT_CANONICAL    1
#define ISE_VERDICT_LIMIT        2
#define ISE_VERDICT_INVALID      3   /* record failed validation on the device */

/* status codes */
#define ISE_OK            0
#define ISE_ERR_ARG      -1   /* invalid argument or candidate            */
#define ISE_ERR_HIP      -2   /* a HIP runtime call failed                */
#define ISE_ERR_NODEV    -3   /* no usable gfx950 device                  */
#define ISE_ERR_CAPACITY -4   /* more candidates than the context holds   */
#define ISE_ERR_STATE    -5   /* call out of order (collect before submit) */

typedef struct ise_ctx ise_ctx;

typedef struct {
    uint8_t  *verdict;
    uint8_t  *fail_row;
    uint8_t  *fail_col;
    uint8_t  *row_to_change;
    uint64_t *group_order;
    uint8_t  *perm_row;
    uint8_t  *perm_col;
    uint8_t  *orbits;
    uint16_t *n_gens;
    uint8_t  *gens;
    uint32_t *counters;
    uint32_t  gen_cap;
    uint64_t *group_order_hi;
} ise_results;

/* Permutation budgets of the first and second pass by default (ise_set_defer). */
#define ISE_DEFAULT_DEFER_PERMS  0u
#define ISE_DEFAULT_DEFER_PERMS2 0u

/* Launch configuration of the canonicity kernel. */
typedef struct {
    uint32_t group_lanes;     /* lanes per candidate: 16, 32, 64; 0 = smallest that fits */
    uint32_t waves_per_block; /* 1..16 */
    uint32_t blocks;          /* persistent grid size; 0 = from occupancy */
    uint32_t max_perms;       /* search limit per candidate (verdict 2 beyond it) */
} ise_launch_cfg;

/* Search order of the canonicity test (the verdict never depends on it):
 *  ISE_ORDER_PLAIN    the reference's search alone (its first witness);
 *  ISE_ORDER_HUNT     first a witness hunt of at most `hunt_budget` row
 *                     comparisons, depth-first over row prefixes in increasing
 *                     row order without automorphism pruning, then the plain
 *                     search for the candidates it leaves open;
 *  ISE_ORDER_LEARNED  the same with the rows of each position tried in the
 *                     order given per candidate by `d_order`
 *                     ([n][order_dim][order_dim] bytes, row i of a candidate
 *                     lists the rows to try at position i), e.g. ranked by the
 *                     learned scorer (ise_mlp_order).
 * A hunt witness is reported in place of the plain one (verdict, fail_row,
 * rowToChange and the witness permutation refer to it); canonical verdicts,
 * |Aut| and generators always come from the plain search.
 *  ISE_ORDER_LEAD     (ise_mlp_use_lead) the learned order of the first row:
 *                     what ise_mlp_check_lead_device does, on every launch of
 *                     the context that reports no column permutation and
 *                     takes no cell masks (those run the plain search, as
 *                     does the canonical-form loop, ise_canonize_*); the
 *                     witness is the lead order's, every canonical
 *                     candidate's results are the plain search's.
 */
#define ISE_ORDER_PLAIN   0
#define ISE_ORDER_HUNT    1
#define ISE_ORDER_LEARNED 2
#define ISE_ORDER_LEAD    3   /* set by ise_mlp_use_lead, not by ise_set_order */
int ise_set_order(ise_ctx *ctx, int mode, uint32_t hunt_budget, const uint8_t *d_order,
                  uint32_t order_dim);

/* Open a context on HIP device `device` able to hold `capacity` candidates of
 * up to `stride` rows per submit.  Replaces the construction of
 * CGPU_CanonChecker (Source/MatrixCanonChecker.h:204-290) and of the
 * per-thread CCanonicityChecker.  (The k-Sys C entry points createCanonizer /
 * runCanonizer, k-Factorization/Utils/Include/CDTools.h:9-11, are
 * ise_cd_create_canonizer / ise_cd_run_canonizer below.) */
int ise_open(ise_ctx **out, int device, uint32_t capacity, uint32_t stride);

/* Release the context (CGPU_CanonChecker dtor). */
int ise_close(ise_ctx *ctx);

int ise_set_launch(ise_ctx *ctx, const ise_launch_cfg *cfg);
int ise_get_launch(const ise_ctx *ctx, ise_launch_cfg *cfg);

/* Queue one matrix for the next batch (CGPU_CanonChecker::assignChecker,
 * Source/MatrixCanonChecker.h:248-256).  rows: `depth` packed words. */
int ise_assign(ise_ctx *ctx, const uint64_t *rows, uint32_t depth, uint32_t ncols,
               uint32_t len_stab, uint32_t out_info);
/* Number of matrices queued since the last launch. */
int ise_num_assigned(const ise_ctx *ctx, uint32_t *n);

/* Submit a whole batch from host memory (copied to the context's device
 * buffers) and launch asynchronously on the context stream
 * (CGPU_CanonChecker::LaunchCanonicityTesting, MatrixCanonChecker.h:303-329).
 * n == 0 with rows == NULL launches the matrices queued by ise_assign. */
int ise_submit(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rows,
               const uint64_t *cells, const uint32_t *meta, uint32_t gen_cap);

/* Wait for the last submit and copy its results into host arrays. */
int ise_collect(ise_ctx *ctx, ise_results *out_host);

/* Synchronous single-matrix test: CCanonicityChecker::TestCanonicity
 * (Source/CanonicityChecker.h:58) on a row-major uint8 0/1 matrix. */
int ise_test_canonicity(ise_ctx *ctx, const uint8_t *mat, uint32_t nrows, uint32_t ncols,
                        uint32_t len_stab, uint32_t out_info, uint32_t gen_cap,
                        ise_results *out_host);

/* ---- canonical form (CMatrixCanonChecker::CanonizeMatrix with k = 0,
 * Source/MatrixCanonChecker.cpp:420-564) ------------------------------------
 * The test with row permutations saved is repeated on the matrix rearranged
 * by each witness (rows by the witness, then the columns put back in orbit
 * form: decreasing column order with row 0 most significant, the order
 * GenerateBinaryColumnOrbits, :186-263, leaves) until the test finds none.
 * The result is the largest column-sorted form over the row permutations that
 * fix the first lenStab rows -- for lenStab = 0 the canonical form, the same
 * for every row and column relabelling of the input.  Per matrix: status 1
 * (canonical form reached), 2 (max_iters tests did not reach it; the last
 * matrix is returned), 3 (invalid record); |Aut| of the canonical form (row
 * action; 0 unless status 1); the tests run.  Matrices use the candidate wire
 * format (meta = depth | ncols << 8 | lenStab << 16; outInfo ignored). */
#define ISE_CANONIZE_MAX_ITERS 4096u

/* Device-resident batch (every pointer device memory; d_aut and d_iters may be
 * NULL), n up to the context capacity and stride up to the context's;
 * synchronous on the context stream (it reads the number of open matrices
 * every few rounds and returns when none is left).  Uses context buffers
 * (allocated on the first call). */
int ise_canonize_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                              const uint32_t *d_meta, uint64_t *d_canon, uint8_t *d_status, uint64_t *d_aut,
                              uint32_t *d_iters, uint32_t max_iters);

/* The same under an initial column partition `part` (bit c set: column c
 * starts a block; 0 or 1: one block, the call above): CanonizeMatrix with
 * k > 0, whose first row of labels makes every label's columns a column orbit
 * (MatrixCanonChecker.cpp:437-449, the k-Sys matrix of runCanonizer); the
 * label row itself is not part of the rows.  Columns stay inside their block;
 * the result is the largest form over the row permutations and the column
 * permutations inside every block. */
int ise_canonize_batch_device_part(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                   const uint32_t *d_meta, uint64_t part, uint64_t *d_canon, uint8_t *d_status,
                                   uint64_t *d_aut, uint32_t *d_iters, uint32_t max_iters);

/* The same on host arrays (canon_out [n][stride]; aut_out, iters_out optional). */
int ise_canonize_batch(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                       uint64_t *canon_out, uint8_t *status_out, uint64_t *aut_out, uint32_t *iters_out,
                       uint32_t max_iters);

/* One row-major uint8 0/1 matrix (nrows x ncols), CanonizeMatrix's own
 * interface: out receives the canonical matrix (same layout), *aut its |Aut|,
 * *iters the tests run (both optional).  ISE_ERR_STATE if ISE_CANONIZE_MAX_ITERS
 * tests did not reach it. */
int ise_canonize_matrix(ise_ctx *ctx, const uint8_t *mat, uint32_t nrows, uint32_t ncols, uint32_t len_stab,
                        uint8_t *out, uint64_t *aut, uint32_t *iters);

/* Words of device memory behind the work-queue argument of the device entry
 * points: one queue head per XCD, each on its own 128-byte line. */
#define ISE_QUEUE_WORDS 256

/* Device-resident batch: every pointer is device memory, results are written
 * in place; enqueued on `stream` (hipStream_t, NULL = the context stream).
 * `d_counter` is ISE_QUEUE_WORDS uint32 of device memory used as the work
 * queue (cleared by the call).  No allocation happens in this call. */
int ise_check_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride,
                           const uint64_t *d_rows, const uint64_t *d_cells,
                           const uint32_t *d_meta, const ise_results *d_out,
                           uint32_t *d_counter, void *stream);

/* Allocation-free variant of ise_check_batch_device.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * results and launch settings as ise_check_batch_device, but the work-queue
 * words and the schedule list live in a device array of the library
 * (ISE_QUEUE_WORDS + ISE_NOALLOC_MAX_CANDIDATES + 128 uint32 words, about
 * 4 MiB, held by every GPU the library runs on until the process ends), so no
 * d_counter is passed, nothing is allocated or freed, and the launch runs in
 * one pass whatever ise_set_defer says (results are the same either way).
 * n above ISE_NOALLOC_MAX_CANDIDATES gives ISE_ERR_CAPACITY.  Every call on one
 * GPU shares that array: calls that overlap in time on one GPU (other
 * streams, other contexts or threads) race and must not be issued.
 * Enqueued on `stream` (NULL = the context stream); asynchronous. */
#define ISE_NOALLOC_MAX_CANDIDATES (1u << 20)
int ise_check_batch_device_noalloc(ise_ctx *ctx, uint32_t n, uint32_t stride,
                                   const uint64_t *d_rows, const uint64_t *d_cells,
                                   const uint32_t *d_meta, const ise_results *d_out,
                                   void *stream);

/* Device-resident batch launched `repeats` times back to back on the context
 * stream; ms_out[i] receives the time of launch i from HIP events recorded on
 * that stream around it.  Same arguments and results as
 * ise_check_batch_device.  Synchronous. */
int ise_check_batch_device_timed(ise_ctx *ctx, uint32_t n, uint32_t stride,
                                 const uint64_t *d_rows, const uint64_t *d_cells,
                                 const uint32_t *d_meta, const ise_results *d_out,
                                 uint32_t *d_counter, uint32_t repeats, float *ms_out);

/* Device memory owned by the caller through the context's device: allocate
 * (zero-filled), free, and synchronous copies on the context stream. */
int ise_device_alloc(ise_ctx *ctx, void **d_ptr, uint64_t bytes);
int ise_device_free(ise_ctx *ctx, void *d_ptr);
int ise_copy_to_device(ise_ctx *ctx, void *d_dst, const void *src, uint64_t bytes);
int ise_copy_to_host(ise_ctx *ctx, void *dst, const void *d_src, uint64_t bytes);
int ise_copy_on_device(ise_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes);

/* Capture every launch of ise_submit as a hipGraph (queue reset, schedule
 * and check kernels), one executable graph per batch shape (stride, group
 * size, inputs and outputs, launch and order settings), instantiated once and
 * replayed afterwards.  A batch of another size updates that graph in place
 * (hipGraphExecUpdate), so an enumeration replays one graph for every level
 * and tail chunk.  Cached graphs are dropped whenever a context buffer they
 * address is reallocated (a larger gen_cap, a device batch above capacity)
 * and by ise_set_defer / ise_set_schedule.  on = 0 launches directly. */
int ise_use_graphs(ise_ctx *ctx, int on);

/* Whether ise_submit reports the column permutation (perm_col; the
 * reference's t_saveColPermutations output, CanonicityChecker.h:27-32).
 * Default on.  With it off the kernel runs the column-set form of the same
 * search (every other result identical) and ise_collect refuses a perm_col
 * array with ISE_ERR_STATE.  The device-resident entry points decide the same
 * way from whether d_out->perm_col is null. */
int ise_report_columns(ise_ctx *ctx, int on);

/* Long searches together or in queue order.  With first_pass_perms > 0 a
 * batch runs in passes: every candidate with that permutation budget; then,
 * from the start, the candidates that ran past it, with budget
 * second_pass_perms (0: no limit but the context's max_perms); then, if there
 * was a second budget, those that ran past it, with no limit.  0 (the
 * default, ISE_DEFAULT_DEFER_PERMS) runs one launch.  Results are identical
 * either way.  The lists of deferred candidates are a context buffer: batches
 * on one context must not run concurrently, and the first device-resident
 * batch larger than the context capacity grows it (hipMalloc).
 * second_pass_perms must exceed first_pass_perms when set. */
int ise_set_defer(ise_ctx *ctx, uint32_t first_pass_perms, uint32_t second_pass_perms);

/* Order in which the candidates of a batch are started (results do not
 * depend on it): ISE_SCHEDULE_QUEUE in index order, ISE_SCHEDULE_DEEPEST_FIRST
 * (default) by decreasing depth, so that the long searches of the deepest
 * partial matrices do not start at the end of the batch.  The schedule list
 * is a context buffer, like the deferred lists of ise_set_defer. */
#define ISE_SCHEDULE_QUEUE          0
#define ISE_SCHEDULE_DEEPEST_FIRST  1
int ise_set_schedule(ise_ctx *ctx, int schedule);
/* Number of executable graphs instantiated so far and replays of them. */
int ise_graph_stats(const ise_ctx *ctx, uint32_t *graphs, uint64_t *replays);
/* In-place updates of cached graphs for batches of another size so far. */
int ise_graph_updates(const ise_ctx *ctx, uint64_t *updates);

/* Autotuner: times the device-resident batch (same arguments as
 * ise_check_batch_device) for every lane-group size that holds the batch
 * (16/32/64) and every workgroup size of 1, 2, 4, 8 waves (which sets the
 * LDS tile of orbit geometry per workgroup), `repeats` launches each, keeps
 * the fastest in the context's launch configuration and writes the
 * measurements (ms per launch, -1 if not run) to table[3][4] (rows: 16, 32,
 * 64 lanes; columns: 1, 2, 4, 8 waves).  max_depth/max_cols describe the
 * batch (the caller knows them; device data is not inspected). */
int ise_autotune(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                 const uint64_t *d_cells, const uint32_t *d_meta, const ise_results *d_out,
                 uint32_t *d_counter, uint32_t max_depth, uint32_t max_cols, uint32_t repeats,
                 float *table, ise_launch_cfg *best);

/* Wait for all work on the context stream. */
int ise_sync(ise_ctx *ctx);

/* Kernel time of the last collected ise_submit in milliseconds (HIP events
 * around the kernel launches on the context stream). */
float ise_last_kernel_ms(const ise_ctx *ctx);

/* Last HIP error string of the context ("" if none). */
const char *ise_last_error(const ise_ctx *ctx);

/* ---- learned row order ---------------------------------------------------
 * A small MLP scores (candidate, position i, row j) pairs from 16 features of
 * row j placed at position i under the identity prefix (the search's own
 * orbit-by-orbit comparison of row j with row i, intersections, positions);
 * the rows of each position are then tried by decreasing score in the
 * witness hunt (ISE_ORDER_LEARNED).  Network: 16 -> 128 -> 128 -> 1 with ReLU,
 * bf16 weights and FP32 accumulation on MFMA; trained on the GPU (binary
 * cross-entropy, Adam) by the library's own forward and backward kernels.
 * Parameters are exchanged as ISE_MLP_PARAMS floats packed as W1[16][128],
 * b1[128], W2[128][128], b2[128], W3[128][16] (column 0 used), b3[16]. */
#define ISE_MLP_FEATURES 16
#define ISE_MLP_PARAMS 20752

typedef struct ise_mlp ise_mlp;

/* Model on the context's device with workspace for up to max_rows pair rows
 * per call (rounded up to 16); weights initialised from `seed`. */
int ise_mlp_create(ise_ctx *ctx, ise_mlp **out, uint32_t max_rows, uint64_t seed);
int ise_mlp_destroy(ise_mlp *m);
int ise_mlp_get_params(const ise_mlp *m, float *host_params);
int ise_mlp_set_params(ise_mlp *m, const float *host_params);

/* Pair features of a device batch: d_X[(c*od + i)*od + j][16] bf16, zero for
 * i or j >= depth.  od in 1..32, every depth <= od. */
int ise_mlp_features(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                     const uint32_t *d_meta, uint32_t od, uint16_t *d_X);

/* Scores (logits) of M pair rows (M a multiple of 16, <= max_rows):
 * d_score[m] = network(d_X[m]). */
int ise_mlp_forward(ise_mlp *m, const uint16_t *d_X, uint32_t M, float *d_score);

/* Gradient of the weighted mean binary cross-entropy of the logits against
 * labels d_y with weights d_w (all device, M rows): writes the gradient
 * (ISE_MLP_PARAMS floats, packed like the parameters) to host_grad and the
 * loss to *loss.  Does not change the parameters. */
int ise_mlp_gradient(ise_mlp *m, const uint16_t *d_X, const float *d_y, const float *d_w, uint32_t M,
                     float *host_grad, float *loss);

/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the same loss;
 * *loss receives the loss before the step. */
int ise_mlp_train_step(ise_mlp *m, const uint16_t *d_X, const float *d_y, const float *d_w, uint32_t M,
                       float lr, float *loss);

/* ise_mlp_train_step repeated `steps` times over consecutive batches of
 * `batch` rows (wrapping around M_total; batches with zero total weight are
 * skipped); loss_curve[s] (optional, host) receives each step's loss. */
int ise_mlp_fit(ise_mlp *m, const uint16_t *d_X, const float *d_y, const float *d_w, uint64_t M_total,
                uint32_t batch, uint32_t steps, float lr, float *loss_curve);

/* Row order for a device batch, features, network and sort fused in one
 * kernel: d_order[c][i][0..d-1] lists the rows of candidate c by decreasing
 * score for position i (ties by increasing row); entries past the depth, and
 * positions i >= depth, hold the identity.  This is the input of
 * ISE_ORDER_LEARNED.  d_score (optional) receives the scores [c][i][j]. */
int ise_mlp_order(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint32_t *d_meta,
                  uint32_t od, uint8_t *d_order, float *d_score);

/* Timed learned-order checking on the context's stream; ms_out[r] = the
 * whole repeat, order_ms_out[r] (optional) = up to the end of the ordering.
 * With a gate (ise_mlp_set_gate, default ISE_DEFAULT_LEARNED_GATE) a repeat
 * is: the plain search on every candidate with a budget of `gate`
 * permutations (the verdict of those it decides is final), the ordering
 * kernel on the candidates it leaves open only, then on those the witness
 * hunt in the learned order (`hunt_budget` comparisons) and the plain search
 * for what the hunt leaves open.  Without a gate: the ordering kernel on every
 * candidate, then the canonicity kernel in ISE_ORDER_LEARNED mode.  The
 * context's previous order mode is restored. */
int ise_mlp_check_batch_device_timed(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                     const uint32_t *d_meta, uint32_t od, uint8_t *d_order, uint32_t hunt_budget,
                                     const ise_results *d_out, uint32_t *d_counter, uint32_t repeats, float *ms_out,
                                     float *order_ms_out);

/* Features of every row of a device batch as the row of position 0 (the
 * lead order's network input): d_X [n][16][16] bf16 bits, row j's 16
 * features at [c][j], zero past the depth and for depths outside 2..16. */
int ise_mlp_lead_features(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                          const uint32_t *d_meta, uint16_t *d_X);

/* Learned-order test of a device batch in one launch (the lead order): each
 * candidate in orbit form with depth 2..16, lenStab 0 and at most 56 columns
 * runs the plain search with a budget of `gate` permutations
 * (ise_mlp_set_gate); when that leaves it open without an automorphism, the
 * network scores its rows as the row of position 0 and a depth-first search
 * tries them in that order (every later position in increasing row order).
 * It stops at a witness of that order (verdict 0, valid, not necessarily the
 * plain order's first); when it ends without witness and automorphism the
 * candidate is canonical with a trivial group and gets exactly the plain
 * search's results; when it meets an automorphism the plain search decides
 * from the start.  Verdicts and, for canonical candidates, every field equal
 * the plain search's.  Other candidates run the plain search.  Counters (if
 * d_out->counters): the gate's work, the lead search's, and the plain
 * search's when it ran.  No column permutation (d_out->perm_col must be
 * null).  d_lead_in (optional [n][16]) replaces the network's order of the
 * rows of position 0 (rows >= depth and repeats skipped, unlisted rows after
 * the listed ones in increasing order); d_lead_out / d_score_out (optional
 * [n][16]) receive the order used (0xff past the depth) and the scores, for
 * the candidates the lead search ran on.  Replaces, for the learned order,
 * the permutation walk of CCanonicityChecker::TestCanonicity
 * (Source/CanonicityChecker.h:255-586; GroupOrder.h:136-248). */
int ise_mlp_check_lead_device(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                              const uint32_t *d_meta, const ise_results *d_out, uint32_t *d_counter,
                              const uint8_t *d_lead_in, uint8_t *d_lead_out, float *d_score_out);
/* The same, `repeats` launches on the context's stream, ms_out[r] each. */
int ise_mlp_check_lead_device_timed(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                    const uint32_t *d_meta, const ise_results *d_out, uint32_t *d_counter,
                                    uint32_t repeats, float *ms_out);

/* The lead order (ISE_ORDER_LEAD) with this network on every launch of its
 * context (on != 0), or back to the plain order (on == 0, if the context uses
 * this network).  The context keeps the network's device buffers: changed
 * parameters and ise_mlp_set_gate apply to its next launches, and
 * ise_mlp_destroy returns the context to the plain order.  Cached graphs of
 * the context are dropped.  Lets the enumerators (ise_enumerate_bibd*) and
 * ise_submit / ise_collect run the learned order. */
int ise_mlp_use_lead(ise_mlp *m, int on);

/* Permutations of plain search after which a candidate of a learned-order
 * batch is ordered by the network (0: every candidate is). */
#define ISE_DEFAULT_LEARNED_GATE 16u
int ise_mlp_set_gate(ise_mlp *m, uint32_t gate_perms);

/* Witness-log labels (host arrays): for every non-canonical candidate and
 * every position i <= fail_row, pair (i, j), j < depth, gets label 1 when the
 * plain search's witness put row j at position i, else 0, and weight 1; all
 * other pairs weight 0.  y and w are [n][od][od]. */
int ise_mlp_witness_labels(uint32_t n, uint32_t od, const uint32_t *meta, const uint8_t *verdict,
                           const uint8_t *fail_row, const uint8_t *perm_row, float *y, float *w);

/* ---- 1-factorizations of K_n (groups of two players) --------------------
 * Wire format: rounds[cand][stride] uint64, nibble p of word i = player at
 * position p of round i (the n/2 pairs of the round in order); meta[cand] =
 * nrounds | n << 8, with n even, 4 <= n <= 16, 1 <= nrounds <= min(n-1, 15).
 * The rounds must be edge-disjoint perfect matchings.  Canonical = no
 * relabelling of the players (with any round moved first) gives a smaller
 * normal form (pairs sorted, pairs by first player, rounds by the partner of
 * player 0), compared round by round -- alldata::canonizator / cnvCheckKm1 /
 * kmProcessMatrix (k-Factorization/TripleSys/Source/cnv.cpp:65-351,
 * kmSupport.cpp:275-313). */
/* Records are checked on the device too (the device entry points get no host
 * check): a meta outside the sizes above, a round that is not a perfect
 * matching of players 0..n-1, or a pair in two rounds gives verdict 3; the
 * host entry points reject such records (ISE_ERR_ARG) instead. */
typedef struct {
    uint8_t  *verdict;  /* [n] 1 canonical, 0 not, 3 invalid record            */
    uint64_t *witness;  /* [n] plain-order index of the first smaller relabelling (~0 if none) */
    uint64_t *n_auts;   /* [n] relabellings that reproduce the matrix, counted up to the
                           decision (|Aut| when canonical)                       */
    uint64_t *n_tried;  /* [n] relabellings evaluated up to the decision          */
    uint16_t *player_index; /* [n] optional (needs witness): the k-Sys backjump position
                           round * players + slot (alldata::setPlayerIndex,
                           kmSupport.cpp:303-356): set at the first smaller row of the
                           witness's normal form unless an earlier row of it came from
                           a later round; nrounds * players - 3 (cnv.cpp:213) when
                           canonical or kept; 0xffff for an invalid record.  Every
                           matrix agreeing on the rounds before that round and on its
                           slots up to that one is not canonical either.          */
} ise_kf_results;

/* Host arrays in and out, synchronous (cnvCheckNew of one matrix per record,
 * cnv.cpp:211-270).  Up to the context capacity per call. */
int ise_kf_check(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rounds,
                 const uint32_t *meta, ise_kf_results *out_host);

/* Device-resident batch on `stream` (NULL = context stream); `d_queue` is
 * ISE_QUEUE_WORDS words of device memory.  With the split launch on (the
 * default, ise_kf_set_split) the call uses context buffers sized for the
 * context capacity: a larger batch grows them on its first call (hipMalloc),
 * and calls on one context must not overlap in time.  With the split off
 * nothing is allocated and concurrent calls need only different d_queue. */
int ise_kf_check_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                              const uint32_t *d_meta, const ise_kf_results *d_out,
                              uint32_t *d_queue, void *stream);
/* Same, launched `repeats` times on the context stream; ms_out[i] = time of
 * launch i from HIP events.  Synchronous. */
int ise_kf_check_batch_device_timed(ise_ctx *ctx, uint32_t n, uint32_t stride,
                                    const uint64_t *d_rounds, const uint32_t *d_meta,
                                    const ise_kf_results *d_out, uint32_t *d_queue,
                                    uint32_t repeats, float *ms_out);

/* Two-row canonizer of perfect 1-factorizations (every pair of rounds one
 * Hamiltonian cycle): alldata::cnvCheck2P1F with create2P1FTr and
 * kmProcessMatrix2p1f (k-Factorization/TripleSys/Source/cnvSupport.cpp:3-95,
 * p1fSupport.cpp:3-38, kmSupport.cpp:519-598), the check the reference selects
 * for P1F runs (allData.cpp:123-145).  Same records and result fields as
 * ise_kf_check; the relabellings tried are those sending an ordered pair of
 * rounds onto rounds 0, 1, in the reference's order: index
 * ((pair * 2 + swap) * n + start player), pair = last * (last - 1) / 2 + ind
 * for rounds ind < last, swap choosing (ind, last) or (last, ind) as the
 * rounds sent to 0 and 1.  Rounds sent to 0, 1 are not compared; the round
 * whose relabelled partner of player 0 is v (v = 3..nrounds) is compared with
 * round v - 1 as given.  Rounds must be in normal form (pairs (a, b) with
 * a < b, pairs by first player) and every pair of them one Hamiltonian cycle:
 * host arrays are rejected with ISE_ERR_ARG otherwise, device records get
 * ISE_VERDICT_INVALID.  Host path: synchronous, up to the context capacity;
 * device path: no allocation. */
int ise_kf_check_2p1f(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rounds,
                      const uint32_t *meta, ise_kf_results *out_host);
int ise_kf_check_2p1f_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const ise_kf_results *d_out, void *stream);
int ise_kf_check_2p1f_batch_device_timed(ise_ctx *ctx, uint32_t n, uint32_t stride,
                                         const uint64_t *d_rounds, const uint32_t *d_meta,
                                         const ise_kf_results *d_out, uint32_t repeats, float *ms_out);

/* Two-row canonizer of uniform 1-factorizations: alldata::cnvCheck2U1F
 * (k-Factorization/TripleSys/Source/cnvSupport2.cpp:3-142) with the cycle
 * tables of getCyclesAndPathFromNeighbors (p1fCheck.cpp:153-228), the mapping
 * order of CycleMapping (cycles.cpp:111-211) and createU1FTr
 * (TrCycles.cpp:3-50); selected for U1F runs at allData.cpp:123-145.  Every
 * ordered pair of rounds whose cycles have the lengths of rounds 0, 1's is
 * sent onto rounds 0, 1 by every mapping of its cycles onto theirs (cycles of
 * one length permuted, a start offset and a direction per cycle); the
 * relabelled matrix is normalised and compared in full.  Index: ordered pair *
 * T + ((permutation rank * V + offsets) * 2^cycles + directions), T the
 * relabellings per pair, V the product of half cycle lengths (directions: last
 * cycle lowest bit; offsets: first cycle lowest digit; ranks: lexicographic,
 * first group of equal lengths lowest).  A pair of rounds with other cycle
 * lengths makes the matrix not canonical with witness ~0 and n_tried the
 * relabellings before it.  Rounds must be in normal form (ISE_ERR_ARG on host
 * arrays, ISE_VERDICT_INVALID on the device).  Device path: no allocation. */
int ise_kf_check_2u1f(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rounds,
                      const uint32_t *meta, ise_kf_results *out_host);
int ise_kf_check_2u1f_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const ise_kf_results *d_out, void *stream);
int ise_kf_check_2u1f_batch_device_timed(ise_ctx *ctx, uint32_t n, uint32_t stride,
                                         const uint64_t *d_rounds, const uint32_t *d_meta,
                                         const ise_kf_results *d_out, uint32_t repeats, float *ms_out);

/* alldata::cnvCheck3U1F (k-Factorization/TripleSys/Source/cnvSupport.cpp:94-339)
 * for groups of 2 on complete graphs, the 1-factorizations this library takes,
 * with the reference's defaults there (AutSaveTestedTrs 0, undefined cycles not
 * allowed, no second-row database): the check the reference selects for
 * groups of 2 with UseFastCanonizerForG2, for canonized input and for the
 * second-row generation (allData.cpp:123-143).  The run's cycle types are
 * u1f_cycles[n_types][ISE_KF_MAX_U1F_CYCLES] (the u1fCycles parameter: cycle
 * lengths in players, ascending, zero after the last); n_types 0 means one
 * cycle through every player.  Rounds 0, 1 of a type not listed make the
 * matrix not canonical before any relabelling (cyclesFor2Rows collects no
 * cycle set: witness ~0, n_tried 0).  Otherwise the relabellings are those of
 * ise_kf_check_2u1f, each ordered pair's tried 32 times over: the loop over
 * MAX_CYCLE_SETS (Global.h:12) cycle sets (cnvSupport.cpp:208-216) recomputes
 * the same set for groups of 2 (collectOneCyclesSet, :341-352), and the
 * repeated relabellings add nothing to the group (n_auts counts distinct
 * automorphisms).  Index: ((ordered pair * 32) + repetition) * T + position,
 * T and position as for ise_kf_check_2u1f; n_tried counts every relabelling
 * tried.  Groups of 3 and graphs that are not complete are outside the
 * matrices this library takes.  player_index is not returned (ISE_ERR_ARG).
 * Device path: no allocation. */
#define ISE_KF_MAX_U1F_TYPES  4
#define ISE_KF_MAX_U1F_CYCLES 4
int ise_kf_check_3u1f(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *rounds,
                      const uint32_t *meta, const uint8_t *u1f_cycles, uint32_t n_types,
                      ise_kf_results *out_host);
int ise_kf_check_3u1f_batch_device(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const uint8_t *u1f_cycles, uint32_t n_types,
                                   const ise_kf_results *d_out, void *stream);

/* Seeded relabelled prefixes of perfect 1-factorizations (a candidate stream
 * for the two-row canonizer): for candidate i a source matrix, an ordered pair
 * of its rounds and a start player are drawn, the pair is sent onto the
 * source's rounds 0, 1 as create2P1FTr does, the result is put in normal form
 * and its first nrounds rounds are kept (nrounds uniform in [rmin, rounds of
 * the source]).  Sources: nsrc records of src_stride words with src_meta as
 * for ise_kf_check.  Deterministic per seed. */
int ise_kf_relabel_p1f(const uint64_t *src, const uint32_t *src_meta, uint32_t nsrc, uint32_t src_stride,
                       uint64_t count, int rmin, uint64_t seed, uint32_t stride, uint64_t *rounds_out,
                       uint32_t *meta_out);

/* The same for uniform 1-factorizations: the drawn pair of rounds goes onto
 * the source's rounds 0, 1 cycle by cycle (cycles of one length in a random
 * assignment, a random start and direction per cycle).  Every pair of rounds
 * of a source must have the cycle lengths of its rounds 0, 1. */
int ise_kf_relabel_u1f(const uint64_t *src, const uint32_t *src_meta, uint32_t nsrc, uint32_t src_stride,
                       uint64_t count, int rmin, uint64_t seed, uint32_t stride, uint64_t *rounds_out,
                       uint32_t *meta_out);

/* Split launch of the 1-factorization test.  A canonical candidate must try
 * every relabelling (46080 x rounds for K_12), which on one wave would be the
 * batch's tail.  With first_chunks > 0 every candidate first gets a look at
 * its first 4 relabellings (sixteen candidates per wave), those still open run
 * up to 64 x first_chunks relabellings, and the rest of each candidate still
 * undecided is cut into up to max_ranges contiguous ranges run by independent
 * waves and combined in plain order (first witness, automorphisms before it).  Results
 * are identical to first_chunks = 0, one wave per candidate to its decision.
 * No counterpart in the reference, whose test is one sequential loop. */
#define ISE_DEFAULT_KF_FIRST_CHUNKS 4u
#define ISE_DEFAULT_KF_MAX_RANGES 1024u
int ise_kf_set_split(ise_ctx *ctx, uint32_t first_chunks, uint32_t max_ranges);

/* Seeded random partial 1-factorizations of K_players with nrounds uniform in
 * [rmin, rmax]: each round is a uniformly drawn perfect matching avoiding the
 * edges already used (the smallest free player takes a uniform free partner;
 * a stuck round is redrawn, a stuck matrix restarted), then relabelled so that
 * round 0 is (0 1)(2 3)... and put in normal form.  Deterministic per seed. */
int ise_kf_random(int players, int rmin, int rmax, uint64_t count, uint64_t seed,
                  uint32_t stride, uint64_t *rounds_out, uint32_t *meta_out);

/* ---- k-Sys binary canonizer (createCanonizer / runCanonizer /
 * releaseCanonizer, k-Factorization/Utils/Include/CDTools.h:9-11,
 * Source/CDTools.cpp:5-51) ---------------------------------------------------
 * A k-system is `days` rows of v player bytes, each row the v/k groups of k
 * players of one day (the k-Sys result matrix).  runCanonizer makes it the
 * binary matrix of C_InSys::convertToBinaryMatrix (Source/matrix.h:198-210):
 * player p's row has a 1 in column j * v/k + g when p is in group g of day j,
 * under a first row of day labels that keeps every day's columns together
 * (day 0 first).  CanonizeMatrix (MatrixCanonChecker.cpp:420-564) returns the
 * largest such matrix over the player orders and the column orders inside
 * each day -- and, when every one of the (v-1)/(k-1) days is present, over
 * the day orders as well (its walk over the day permutations keeps the
 * largest matrix and starts again whenever it grows).  The result is the v x b
 * player rows, b = days * v/k: equal for two k-systems exactly when one is a
 * relabelling of the other (for full days, with the days reordered too).
 * Here every day order is canonized on the GPU at once (days! matrices, days
 * <= ISE_CD_MAX_WALK_DAYS) and the largest kept.  Limits: v <= 32 players,
 * b <= 64 columns. */
#define ISE_CD_MAX_WALK_DAYS 10u
typedef struct ise_cd_canonizer ise_cd_canonizer;

/* createCanonizer(v, k) on HIP device `device`.  ISE_ERR_ARG unless k >= 2,
 * k divides v, k - 1 divides v - 1 and v <= 32.  A canonizer owns its own
 * context and buffers; like the reference's, it serves one caller at a time. */
int ise_cd_create_canonizer(ise_cd_canonizer **out, int device, uint32_t v, uint32_t k);

/* releaseCanonizer. */
int ise_cd_release_canonizer(ise_cd_canonizer *h);

/* runCanonizer(h, matrix, k, day): `day` days of `matrix` (day <= 0: all of
 * them); *out receives the canonical v x b player rows (one byte per entry),
 * valid until the next call on h, *ncols = b.  k must be the canonizer's. */
int ise_cd_run_canonizer(ise_cd_canonizer *h, const uint8_t *matrix, uint32_t k, int day, const uint8_t **out,
                         uint32_t *ncols);

/* The group orders of the last ise_cd_run_canonizer call: day_order = the
 * order of the day-permutation group of the result, i.e. the day orders whose
 * canonical matrix is the result (the class group order CanonizeMatrix keeps
 * in extraGroupOrder, MatrixCanonChecker.cpp:548-551; 1 when fewer than all
 * days are given, where no day permutation is tried), fixed_order = |Aut| of
 * the result with every day fixed (player permutations that, with column
 * permutations inside the days, keep the matrix).  For a full k-system
 * day_order x fixed_order = |Aut(M)|.  Either pointer may be NULL. */
int ise_cd_group_order(const ise_cd_canonizer *h, uint64_t *day_order, uint64_t *fixed_order);

/* A batch of n k-systems of `day` days each (matrices [n][days][v]); out
 * [n][v][b] bytes; day_orders / fixed_orders [n] (optional, NULL: not
 * written) as ise_cd_group_order per k-system. */
int ise_cd_run_canonizer_batch(ise_cd_canonizer *h, uint32_t n, const uint8_t *matrices, int day, uint8_t *out,
                               uint64_t *day_orders, uint64_t *fixed_orders);

/* The same on device memory: d_matrices [n][days][v] bytes, d_canon [n][v]
 * packed rows (bit c = column c), d_day_orders / d_fixed_orders [n] device
 * words or NULL; synchronous. */
int ise_cd_run_canonizer_device(ise_cd_canonizer *h, uint32_t n, const uint8_t *d_matrices, int day,
                                uint64_t *d_canon, uint64_t *d_day_orders, uint64_t *d_fixed_orders);

/* Last error message of h. */
const char *ise_cd_error(const ise_cd_canonizer *h);

/* ---- the automorphism group after the test ------------------------------
 * Generators returned by the test (ise_results.gens: column-orbit
 * permutations of the next row, or row permutations with
 * ISE_SAVE_ROW_PERMUTATIONS) as byte permutations of length len. */

/* Closure of the generators (CPermutStorage::constructGroup,
 * Source/PermutStorage.h:183-225): every group element, identity first, into
 * elements[cap][len]; *order_out = |G|.  ISE_ERR_CAPACITY (with *order_out
 * >= cap + 1) when the group has more than cap elements. */
int ise_group_closure(const uint8_t *gens, uint32_t n_gens, uint32_t len, uint8_t *elements, uint64_t cap,
                      uint64_t *order_out);

/* Generators restricted to the column orbits idx[0..len_idx-1]
 * (CPermutStorage::adjustGenerators, Source/PermutStorage.h:155-181):
 * out[g][j] = position in idx of gens[g][idx[j]].  ISE_ERR_ARG if the subset
 * is not invariant. */
int ise_adjust_generators(const uint8_t *gens, uint32_t n_gens, uint32_t len, const uint32_t *idx,
                          uint32_t len_idx, uint8_t *out);

/* Orbits of row solutions under a group (CPermutStorage::findSolutionIndex,
 * Source/PermutStorage.h:281-325), on the GPU: canon_out[s] = the
 * lexicographically largest image of solution s (image[i] = sol[g[i]] over
 * every element g); rep_out[s] (optional) = the first listed solution equal
 * to that image, or 0xffffffff if none is listed.  len <= 64. */
int ise_solution_orbits(ise_ctx *ctx, const uint8_t *group, uint64_t order, uint32_t len, const uint8_t *sols,
                        uint32_t n_sol, uint8_t *canon_out, uint32_t *rep_out);

/* The improved solution of CCanonicityChecker::reconstructSolution
 * (Source/CanonicityChecker.cpp:227-277), the input the reference passes to
 * CRowSolution::moveNoncanonicalSolutions (Source/RowSolution.cpp:158-281)
 * when the last row of a partial matrix makes it non-canonical (fail_row ==
 * depth - 1): per column orbit of the last row (cells[depth-1], or derived
 * from the rows when cells is NULL), the ones of row perm_row[depth-1] among
 * the columns perm_col[] places in that orbit.  perm_row, perm_col: the
 * witness of the test as reported with the column permutation
 * (ise_report_columns on, or a non-null perm_col on the device paths).
 * Writes *n_orbits entries to solution_out (<= 64).  Host only; ISE_ERR_ARG
 * when fail_row is not the last row or the witness is not a permutation. */
int ise_improved_solution(const uint64_t *rows, uint32_t depth, uint32_t ncols, const uint64_t *cells,
                          const uint8_t *perm_row, const uint8_t *perm_col, uint32_t fail_row,
                          uint8_t *solution_out, uint32_t *n_orbits);

/* ---- text output of results (host) --------------------------------------
 * Both write a NUL-terminated string into buf (cap bytes); *len_out receives
 * the length needed (without the NUL), and ISE_ERR_CAPACITY is returned when
 * it does not fit. */

/* One matrix as the reference prints it (CMatrixData::printOut,
 * Source/Matrix.cpp:43-155): "\nMatrix # %3llu[.sub][  *** Non-canonical ***]
 * |Aut(M)| = %6llu" then one line of 0/1 per row; number 0 prints the rows
 * only. */
int ise_format_matrix(const uint64_t *rows, int nrows, int ncols, uint64_t number, uint64_t sub, int canon,
                      uint64_t aut, char *buf, uint64_t cap, uint64_t *len_out);

/* The enumeration summary: "BIBD(v, b, r, k, lambda)" and the |Aut(D)| table
 * (CGroupsInfo::printGroupInfo, Source/GroupsInfo.cpp:96-150): per group
 * order the numbers of designs, of simple designs (no replicated blocks), of
 * designs transitive on the rows and of both, then the totals.  designs
 * [n][v] packed rows, auts [n], orbits [n][32] row orbits from the test
 * (transitive = every row in orbit 0). */
int ise_format_aut_table(int v, int b, int r, int k, int lambda, uint32_t n, const uint64_t *designs,
                         const uint64_t *auts, const uint8_t *orbits, char *buf, uint64_t cap, uint64_t *len_out);

/* One 1-factorization as the k-Sys result files print it (K_OrderedMatrices.txt,
 * TableAut::name / Table::printTable / outMatrix, k-Factorization/TripleSys/
 * Include/Table.h:5-41, 93-115, 333-361): "\n\n%5llu: |Aut(M)| =  <aut>,
 * Cycles:<sets>" with the cycle types of every pair of rounds counted
 * (matrixStatOutput / collectCyclesAndPath, TripleSys/Source/p1fCheck.cpp:
 * 77-118, 499-517), then one quoted line per round and an empty line.  rounds
 * in the ise_kf_check wire format (nibble p of word r = player at position p).
 * Same buffer convention as the two formatters above. */
int ise_kf_format_matrix(const uint64_t *rounds, int nrounds, int players, uint64_t number, uint64_t aut,
                         char *buf, uint64_t cap, uint64_t *len_out);

/* ---- host-side candidate producers (stay on the CPU) -------------------- */

/* Level-synchronous enumeration of 2-(v,k,lambda) designs: every level's
 * candidates are checked as one batch on the GPU (the reference checks them
 * one by one after every row, Source/Enumerator.cpp:435-501, 350-433).
 * Canonical designs go to designs[] (v packed words each, in the order of
 * discovery, up to max_designs) and |Aut| to auts[].
 * stats[0..3] = candidates tested, canonical partial matrices, fully
 * constructed matrices, designs; stats[4] = GPU kernel time in ns.
 * level_counts (may be NULL) receives per depth 1..v the candidates tested
 * ([v+1] entries).  Returns designs found (>= 0) or a status code. */
long ise_enumerate_bibd(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs,
                        uint64_t *auts, long max_designs, uint64_t *stats,
                        uint64_t *level_counts);

/* The same enumeration with the solutions of every canonical matrix pruned by
 * its automorphism group (CRowSolution::sortSolutionsByGroup,
 * Source/RowSolution.h:325-383, called from CBIBD_Enumerator::checkSolutions,
 * Source/BIBD_Enumerator.h:90-114): the test of each level also returns the
 * generators (up to 32 per matrix; fewer prune less, never wrongly), and of
 * each orbit of next-row solutions under them only the lexicographically
 * largest is extended -- the children of the others are isomorphic to it with
 * the same first rows and cannot be canonical.  Same designs, |Aut| and
 * outputs as ise_enumerate_bibd; fewer candidates tested. */
long ise_enumerate_bibd_pruned(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs,
                               uint64_t *auts, long max_designs, uint64_t *stats,
                               uint64_t *level_counts);

/* Same enumeration, but every candidate at depth in [dmin,dmax] is also
 * copied out (stride words each, up to cap; *n_out receives the count):
 * the "enumerator output" part of the candidate streams. */
long ise_enumerate_bibd_candidates(ise_ctx *ctx, int v, int k, int lambda,
                                   int dmin, int dmax, uint32_t stride,
                                   uint64_t *rows_out, uint32_t *meta_out,
                                   long cap, long *n_out);

/* Row extension on the GPU (the reference's CInSysSolver / CRowSolution,
 * Source/InSysSolver.h:92-286): for n parents of `depth` rows (device,
 * [n][stride] words), every next row whose column-orbit counts sum to r,
 * meet every earlier row in lambda columns and leave every column able to
 * reach k ones; ones left-justified inside each orbit, counts tried from the
 * largest down (the order of the host solver).  d_counts/d_offsets [n]
 * (device) receive the solutions per parent and their exclusive prefix sum;
 * the children (parent rows + the new row, zero-padded to stride) go to
 * d_children [children_cap][stride] in parent order.  *n_children = total;
 * ISE_ERR_CAPACITY (with the total) when it exceeds children_cap. */
int ise_solve_rows_device(ise_ctx *ctx, int v, int k, int lambda, uint32_t n, uint32_t stride,
                          const uint64_t *d_parents, int depth, uint32_t *d_counts, uint32_t *d_offsets,
                          uint64_t *d_children, uint64_t children_cap, uint64_t *n_children);

/* ise_enumerate_bibd with every step on the device: row extension
 * (ise_solve_rows_device kernels), canonicity, and an order-preserving
 * compaction of the canonical children per level; the host reads only the
 * level sizes and the final designs.  Same outputs as ise_enumerate_bibd. */
long ise_enumerate_bibd_device(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs,
                               uint64_t *auts, long max_designs, uint64_t *stats,
                               uint64_t *level_counts);

/* Beam-sampled enumeration for parameter sets too large to enumerate in
 * full: level by level as above, but when more than `beam` (>= 1) candidates
 * of a level are canonical only `beam` of them are extended, chosen by a
 * partial Fisher-Yates shuffle of their indices with splitmix64 from `seed`
 * (t = 0..beam-1: swap idx[t] with idx[t + next() % (count - t)]) and kept in
 * level order; the generator state carries over between levels.  Candidates
 * at depth in [dmin,dmax] are copied out as above; level_counts (optional,
 * [v+1]) receives the candidates tested per depth.  Returns the number of
 * candidates written or a status code. */
long ise_enumerate_bibd_sampled(ise_ctx *ctx, int v, int k, int lambda, uint32_t beam, uint64_t seed,
                                int dmin, int dmax, uint32_t stride, uint64_t *rows_out,
                                uint32_t *meta_out, long cap, long *n_out, uint64_t *level_counts);

/* Seeded random valid extensions: each new row takes r columns uniformly
 * among those whose weight is below k and that keep every pairwise row
 * intersection <= lambda (randomized backtracking, 256 draws); columns are then sorted in
 * decreasing lexicographic order so that orbits are contiguous.  Depth is
 * uniform in [dmin, dmax].  Writes n candidates (stride words each) and
 * meta.  Deterministic for a given seed.  Returns ISE_OK or a status. */
int ise_random_extensions(int v, int k, int lambda, uint64_t n, int dmin, int dmax,
                          uint64_t seed, uint32_t stride, uint64_t *rows_out,
                          uint32_t *meta_out);

/* The same with rejection sampling of the depth: a walk that falls short of
 * its candidate's depth starts again from the empty matrix, up to `attempts`
 * walks (1 = ise_random_extensions; near-complete depths, where the rows are
 * all but forced, are reached by few walks), and the deepest matrix met is
 * kept.  A candidate whose first walk reaches its depth is the same for every
 * `attempts`.  With ctx the batch is generated on ctx's device, one thread
 * per candidate, identical to the host one (host output buffers); with NULL,
 * on the host threads.  No counterpart in the reference (stress batches of
 * the benchmark).  Returns ISE_OK or a status. */
#define ISE_RANDOM_MAX_ATTEMPTS 4096u
int ise_random_extensions_ex(ise_ctx *ctx, int v, int k, int lambda, uint64_t n, int dmin, int dmax,
                             uint64_t seed, uint32_t attempts, uint32_t stride, uint64_t *rows_out,
                             uint32_t *meta_out);

#ifdef __cplusplus
}
#endif
#endif
