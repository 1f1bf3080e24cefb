/*
 * ennemi_hip.h -- C ABI of the MI355X (gfx950) nearest-neighbour estimator
 * library libennemi_hip.so.
 *
 * The reference (polsys/ennemi 1.5.0) has no FFI: its hot path sits behind
 * four private Python estimators and a thread-pool fan-out.  This ABI
 * replaces, in one batched call,
 *
 *   _map_maybe_parallel(_lagged_mi, params, ...)   ennemi/_driver.py:492,715,736-785
 *
 * whose tasks each end in one of
 *
 *   _estimate_single_mi(x, y, k)                    ennemi/_entropy_estimators.py:69-113
 *   _estimate_conditional_mi(x, y, cond, k)         ennemi/_entropy_estimators.py:116-156
 *   _estimate_semidiscrete_mi(x, y, k)              ennemi/_entropy_estimators.py:159-200
 *   _estimate_conditional_semidiscrete_mi(...)      ennemi/_entropy_estimators.py:203-247
 *   _estimate_single_entropy(x, k)                  ennemi/_entropy_estimators.py:21-42
 *
 * and the SciPy calls inside them, cKDTree.query(k=[k+1], p=inf) and
 * cKDTree.query_ball_point(r, p=inf, return_length=True).  Each instance in a
 * batch is one such estimator call on already validated, already rescaled
 * float64 data (the host keeps ennemi's validation and preprocessing).
 *
 * Conventions: plain C types only.  Pointers named d_* are device pointers
 * (HBM of the context's GPU); everything else is host memory, read during the
 * call only.  Every function returns 0 on success and a negative ENM_E_* code
 * on failure; nothing throws or aborts across the ABI, and the message of the
 * last failure is kept in the context (enm_last_error).  A context is not
 * thread-safe: use one per host thread or serialise calls.
 */
#ifndef ENNEMI_HIP_H
#define ENNEMI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ENM_ABI_VERSION 1

/* Estimator kinds (enm_instance.kind). */
#define ENM_KIND_KSG   0 /* _estimate_single_mi                    :69-113  */
#define ENM_KIND_CMI   1 /* _estimate_conditional_mi               :116-156 */
#define ENM_KIND_ROSS  2 /* _estimate_semidiscrete_mi              :159-200 */
#define ENM_KIND_CROSS 3 /* _estimate_conditional_semidiscrete_mi  :203-247 */
#define ENM_KIND_KL    4 /* _estimate_single_entropy               :21-42   */

/* Maximum number of float64 columns (dimensions) of one instance. */
#define ENM_MAX_COLS 16

/* Flags for enm_run_batch. */
#define ENM_F_FP64          0x1u  /* plain FP64 distances (no FP32 filter)            */
#define ENM_F_BRUTE         0x2u  /* scan every candidate (no sorted-window pruning)  */
#define ENM_F_GRAPH         0x4u  /* capture the batch in a hipGraph, replay on reuse */
#define ENM_F_PER_INSTANCE  0x8u  /* one launch sequence per instance (no batching)   */
#define ENM_F_TIMING       0x10u  /* record HIP events between stages (not in graphs) */

/* Stages timed with ENM_F_TIMING (enm_stats.stage_ms). */
#define ENM_STAGE_FILL    0
#define ENM_STAGE_SORT    1
#define ENM_STAGE_GATHER  2
#define ENM_STAGE_KNN     3
#define ENM_STAGE_COUNT   4
#define ENM_STAGE_COMBINE 5
#define ENM_N_STAGES      6

/* Error codes. */
#define ENM_OK            0
#define ENM_E_ARG        -1  /* invalid argument (message says which)       */
#define ENM_E_HIP        -2  /* a HIP runtime call failed                   */
#define ENM_E_NOMEM      -3  /* device workspace allocation failed          */
#define ENM_E_UNSUPPORTED -4 /* shape outside what the kernels support      */

typedef struct enm_ctx enm_ctx;

/*
 * One estimator call.  The instance's points are numbered 0..n-1; column c
 * holds coordinate c of every point at d_arena[col[c] .. col[c]+n).
 * Columns are ordered x (dx of them), then y (dy), then z (dz):
 *
 *   KSG   : x, y        -> psi(n) + psi(k) - mean(psi(n_x) + psi(n_y))
 *   CMI   : x, y, z     -> psi(k) - mean(psi(n_xz) + psi(n_yz) - psi(n_z))
 *   ROSS  : x           -> psi(n) + psi(k) - mean(psi(n_x))
 *                          - sum_c psi(n_c) n_c / n
 *   CROSS : x, z        -> psi(k) - mean(psi(n_xz) + psi(n_yz) - psi(n_z))
 *   KL    : x           -> psi(n) - psi(k) + dx (mean(log eps) + log 2)
 *
 * ROSS and CROSS need the points grouped by discrete class: n_segs classes
 * of sizes seg_len[seg_index .. seg_index+n_segs) (sum = n) in class order,
 * as np.unique orders them.  Other kinds use n_segs = 1.
 * Counts use the radius eps - 1e-12 with an inclusive test, as the reference
 * does (ennemi/_entropy_estimators.py:104-110); psi is the reference's
 * truncated series with psi(0) = +inf poisoning the mean (:327-350).
 */
typedef struct enm_instance {
    int32_t kind;
    int32_t k;
    int64_t n;
    int32_t dx, dy, dz;
    int32_t n_segs;
    int64_t seg_index;
    int64_t col[ENM_MAX_COLS];
} enm_instance;

/* Work counters of the last enm_run_batch (device-side, exact). */
typedef struct enm_stats {
    uint64_t knn_pairs;       /* (query, candidate) distance evaluations in kNN     */
    uint64_t knn_rechecks;    /* FP64 re-evaluations triggered by the FP32 filter   */
    uint64_t count_pairs;     /* pair evaluations in multi-dimensional ball counts  */
    uint64_t count_rechecks;  /* FP64 re-evaluations in the count guard band        */
    uint64_t count_queries_1d;/* binary-search (1-D) ball counts                    */
    uint64_t sort_keys;       /* keys sorted                                        */
    uint32_t n_launches;      /* kernel launches (0 when a graph was replayed)      */
    uint32_t graph_replayed;  /* 1 if the batch ran as a hipGraph replay            */
    float stage_ms[ENM_N_STAGES]; /* per-stage device time with ENM_F_TIMING, else 0  */
    float total_ms;           /* first to last event with ENM_F_TIMING, else 0     */
} enm_stats;

/* Create a context on HIP device `device`. */
int enm_init(int device, enm_ctx** out_ctx);

/* Release the context, its workspace and cached graphs. */
void enm_shutdown(enm_ctx* ctx);

/* Message of the last failure on ctx (or of enm_init when ctx is NULL). */
const char* enm_last_error(const enm_ctx* ctx);

/*
 * Run n_inst estimator instances and write one float64 value per instance to
 * d_out[0..n_inst).  Optional exports (either pointer may be NULL), in each
 * instance's own point numbering, instance after instance (point offset of
 * instance i = sum of n over instances before i):
 *   d_eps    : the k-th-neighbour radius of every point (float64);
 *   d_counts : 3 int32 ball counts per point, laid out [space][point] per
 *              instance (spaces x,y for KSG; xz,yz,z for CMI/CROSS; x for
 *              ROSS; unused slots are 0).
 * `stream` is a hipStream_t (NULL = the legacy default stream); all work is
 * enqueued on it and the call returns without synchronising.
 */
int enm_run_batch(enm_ctx* ctx,
                  const enm_instance* inst, int32_t n_inst,
                  const int64_t* seg_len,
                  const double* d_arena,
                  double* d_out,
                  double* d_eps,
                  int32_t* d_counts,
                  uint32_t flags,
                  void* stream);

/*
 * enm_run_batch_hinted: enm_run_batch with sort hints, for batches whose
 * columns come from a few series (lag scans, pairwise scans), whose orders
 * one sort of each series yields (enm_prep_hints).  Replaces the same
 * reference code as enm_run_batch; the hints stand in for the kd-tree builds
 * (ennemi/_entropy_estimators.py:36,100-102,135-139).
 *
 * d_hint (device, int32, may be NULL) is parallel to d_arena: for an
 * instance column at d_arena[col .. col+n) that a sort is keyed on,
 * d_hint[col .. col+n) may hold the column's ascending order -- the point
 * indices of its values from smallest to largest, equal values by index --
 * or d_hint[col] < 0 for none.  Every hinted order is checked in full
 * (indices in range, each (value, index) pair above the one before, which
 * makes it the sorted permutation), and columns whose hint is absent or
 * wrong are sorted as enm_run_batch sorts them: the results are
 * enm_run_batch's whatever d_hint holds.  Hints are read when the
 * batch's longest sorted column has 2049 to 2^20 points (the sample-sort
 * path), for columns outside class segments.
 */
int enm_run_batch_hinted(enm_ctx* ctx,
                         const enm_instance* inst, int32_t n_inst,
                         const int64_t* seg_len,
                         const double* d_arena,
                         const int32_t* d_hint,
                         double* d_out,
                         double* d_eps,
                         int32_t* d_counts,
                         uint32_t flags,
                         void* stream);
/*
 * enm_run_batch_static: enm_run_batch without any device allocation.
 *
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and exports as
 * enm_run_batch, with these differences:
 *   - the whole workspace (plan tables, sort buffers, orders, radii, counts)
 *     lives in one static device array of ENM_STATIC_WORKSPACE_BYTES, one
 *     copy per GPU, held from the first enm_init on that GPU until the
 *     process ends; a batch that needs more returns ENM_E_NOMEM and says how
 *     much it needs (about 140 bytes per point for KSG, so about 1.9 million
 *     points per call);
 *   - nothing is cached: every call plans the batch, copies its tables from
 *     pageable host memory and synchronises the stream once before
 *     enqueueing the kernels (the call still returns before they finish);
 *   - ENM_F_GRAPH and ENM_F_TIMING are rejected (ENM_E_ARG).
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (from two contexts, or two streams
 * of one context) race, so serialise them.  enm_get_stats reports the last
 * call, as for enm_run_batch.
 */
#define ENM_STATIC_WORKSPACE_BYTES (256ull << 20)
int enm_run_batch_static(enm_ctx* ctx,
                         const enm_instance* inst, int32_t n_inst,
                         const int64_t* seg_len,
                         const double* d_arena,
                         double* d_out,
                         double* d_eps,
                         int32_t* d_counts,
                         uint32_t flags,
                         void* stream);

/*
 * Device-side preparation of estimate_mi tasks, replacing per task of
 *
 *   _lagged_mi windows        ennemi/_driver.py:795-806
 *   _apply_masks              ennemi/_driver.py:834-854
 *   _rescale_data             ennemi/_driver.py:871-902
 *
 * for continuous x, y and condition.  The raw series live in one float64
 * device array d_raw; a task names its x and y series and condition columns
 * by offset and the first row of each window (x0 = max_lag - lag,
 * y0 = max_lag, z0[c] = max_lag - cond_lag[c]; all windows have w rows, the
 * mask (one byte per row of the full series, NULL for none) is read at
 * y0 .. y0+w).  Two calls:
 *
 *   enm_prep_count   kept rows n[t] of every task (mask, then, with
 *                    ENM_P_DROP_NAN, rows with a NaN in x, y or any
 *                    condition column dropped) and h_nan[t] = 1 when a NaN
 *                    is left among them; synchronous (host outputs are
 *                    written before it returns).  The host validates with
 *                    these (k < n, no NaN left) before going on.
 *   enm_prep_write   the kept rows of task t, in row order, to
 *                    d_arena[arena_off[t] ...]: x (n values), y (n), then
 *                    each condition column (n each); with ENM_P_RESCALE each
 *                    variable standardised bit for bit as NumPy does in
 *                    _rescale_data, plus the noise d_noise[...] (the first
 *                    values of default_rng(2_718281828).normal(0, 1e-10, .),
 *                    at least max_t n[t] * (2 + dz[t]) of them).
 *                    d_const[t] (device) receives bit 0/1/2 when x/y/the
 *                    condition was constant (std < 1e-20: left unscaled, no
 *                    noise drawn), for the caller's warnings.  Must follow
 *                    the enm_prep_count of the same tasks, d_raw, d_mask and
 *                    ENM_P_DROP_NAN on the same context.  Asynchronous.
 */
typedef struct enm_prep_task {
    int64_t x_col, y_col;               /* d_raw offsets of the x and y series   */
    int64_t x0, y0;                     /* first row of the x and y windows      */
    int64_t w;                          /* window rows                           */
    int32_t dz;                         /* condition columns (0 .. 14)           */
    int32_t pad;
    int64_t z_col[ENM_MAX_COLS - 2];    /* d_raw offsets of the condition series */
    int64_t z0[ENM_MAX_COLS - 2];       /* first row of each condition window    */
} enm_prep_task;

#define ENM_P_DROP_NAN 0x1u
#define ENM_P_RESCALE  0x2u
#define ENM_P_HINTS    0x4u  /* enm_prep_write keeps what enm_prep_hints needs */

int enm_prep_count(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks,
                   const double* d_raw, const uint8_t* d_mask, uint32_t flags,
                   int64_t* h_n, uint32_t* h_nan, void* stream);
int enm_prep_write(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks,
                   const double* d_raw, const uint8_t* d_mask,
                   const double* d_noise, int64_t noise_len, uint32_t flags,
                   const int64_t* arena_off, double* d_arena, uint32_t* d_const,
                   void* stream);
/*
 * enm_prep_hints: sort hints for enm_run_batch_hinted, for the columns the
 * preceding enm_prep_write (with ENM_P_HINTS; same tasks, inputs, noise,
 * flags, arena_off and d_arena) made.  d_hint (device, int32) is parallel
 * to d_arena: column c of task t (x, y, then the condition columns) gets at
 * d_hint[arena_off[t] + c * n[t] ...] its ascending order (task positions of
 * its values, smallest first, equal values by position), or
 * d_hint[arena_off[t] + c * n[t]] = -1 where it was not derived.
 *
 * Each raw series is sorted once over the rows that any window reads; a
 * column's order is its series' sorted rows filtered to the task's kept
 * rows, with the runs of near-equal values that the noise may have
 * reordered sorted again (a run longer than 32 values, ties in the raw
 * data, leaves the column without a hint).  Asynchronous.
 */
int enm_prep_hints(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks,
                   const double* d_raw, const uint8_t* d_mask,
                   const double* d_noise, int64_t noise_len, uint32_t flags,
                   const int64_t* arena_off, const double* d_arena, int32_t* d_hint,
                   void* stream);

/*
 * enm_unique_rows: the distinct rows of row sets and their counts, as
 * np.unique(rows, axis=0, return_inverse=True, return_counts=True) gives
 * them, for the discrete plug-in estimators
 *
 *   _estimate_discrete_entropy             ennemi/_entropy_estimators.py:44-56
 *   _estimate_discrete_mi                  ennemi/_entropy_estimators.py:257-289
 *   _estimate_conditional_discrete_mi      ennemi/_entropy_estimators.py:291-320
 *
 * A row set is ncol float64 columns of n values in d_values (numeric data;
 * -0.0 and 0.0 are one value; no NaN).  Its rows are numbered by their
 * place among the distinct rows in lexicographic order (column 0 first):
 * for row set s, placed at o_s = n_0 + ... + n_{s-1},
 *   d_codes[o_s + p]   the number of point p's distinct row (int32),
 *   d_counts[o_s + r]  the points of distinct row r (int64, r < h_distinct[s]).
 * Synchronous.  ENM_E_UNSUPPORTED when the product of the distinct counts
 * of a multi-column row set's columns exceeds 2^53 (nothing is written).
 */
typedef struct enm_rowset {
    int64_t n;                     /* rows                                   */
    int32_t ncol;                  /* columns (1 .. ENM_MAX_COLS)            */
    int32_t pad;
    int64_t col[ENM_MAX_COLS];     /* d_values offset of each column         */
} enm_rowset;

int enm_unique_rows(enm_ctx* ctx, const enm_rowset* sets, int32_t n_sets,
                    const double* d_values, int32_t* d_codes, int64_t* d_counts,
                    int64_t* h_distinct, void* stream);

/* Counters of the last enm_run_batch; synchronises the stream it used. */
int enm_get_stats(enm_ctx* ctx, enm_stats* out);

/*
 * Tile shapes.  They change how the work is cut, never a result: every
 * radius, count and estimate is the same bits under every shape.
 *   strip_points      the kNN runs on a grid order: the key-sorted points
 *                     cut into strips of this many consecutive points, each
 *                     strip sorted by the second column (0 = plain
 *                     key-sorted order); a multiple of 64 up to 65536;
 *   waves_per_cu      sizes the persistent grids (1..64);
 *   list_slots        register-list slots of the kNN (4, 8, 16 or 32: at
 *                     least k+1 are used; 0 = the fewest that hold k+1);
 *                     lists of k+1 > 32 are heaps whatever this says;
 *   items_per_ticket  most work items a wave takes per pop of its queue
 *                     (1..64; 0 = 8); fewer when the launch is small, so
 *                     that every wave of the grid gets work.
 * The planner takes each instance's shape from the context: the shape
 * enm_autotune kept for its (joint dimension class, k+1 class,
 * floor(log2 n)) if there is one, else the default; until enm_set_tuning(_ex)
 * is first called, the default gives instances of three or more joint
 * columns strips of 4096 points.  enm_autotune sweeps
 * strips {0, 256, ..., 8192}, then waves {8, 16, 32}, then list slots
 * {0, 8, 16}, then items per ticket {0, 1, 2, 4}, each with the best of the
 * ones before (two enm_run_batch calls per shape, the second timed), and
 * keeps the winner for each shape of the batch (the default is unchanged).
 * enm_set_tuning(_ex) sets the default and forgets the kept shapes.
 */
typedef struct enm_tuning {
    int32_t strip_points;
    int32_t waves_per_cu;
    int32_t list_slots;
    int32_t items_per_ticket;
} enm_tuning;

int enm_autotune(enm_ctx* ctx,
                 const enm_instance* inst, int32_t n_inst,
                 const int64_t* seg_len, const double* d_arena,
                 double* d_out, uint32_t flags, void* stream);
int enm_set_tuning(enm_ctx* ctx, int32_t strip_points, int32_t waves_per_cu);
int enm_set_tuning_ex(enm_ctx* ctx, const enm_tuning* tuning);
int enm_get_tuning(const enm_ctx* ctx, int32_t* strip_points, int32_t* waves_per_cu);
/* The shape an instance of `dims` joint columns, k and n would run with. */
int enm_get_tuning_for(const enm_ctx* ctx, int32_t dims, int32_t k, int64_t n, enm_tuning* out);

#ifdef __cplusplus
}
#endif

#endif /* ENNEMI_HIP_H */
