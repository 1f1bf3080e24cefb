// Host-side launchers of the kernels in enm_*.hip (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "enm_internal.h"
#include "../../include/ennemi_hip.h"

hipError_t enm_launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st);
hipError_t enm_launch_fill_i32(int* p, int64_t n, int v, hipStream_t st);

hipError_t enm_launch_sort(const SortSeg* d_segs, int nsegs, int ntiles, int64_t maxlen, int tile,
                           const double* arena, const int* iarena, double* k0, int* i0, double* k1, int* i1,
                           int* split, int* final_buf, int* n_launch, hipStream_t st,
                           const int* skip = nullptr);

// Sample sort of the first-stage segments (all at most ENM_SS_MAXLEN keys),
// laid out straight into the orders that consume them (the first-stage
// gather); (k1, i1) hold the scattered buckets on the way, (k0, i0) the
// result of a run too large for LDS.
struct SampleSortArgs {
    const SortSeg* segs;
    int nsegs, ntiles;
    const SampleSeg* ss;
    const SortSeg* samp_segs;   // the segments' sample ranges (those with samples)
    int n_samp_segs, samp_tiles;
    int64_t samp_maxlen;
    double *samp_k0, *samp_k1;
    int *samp_i0, *samp_i1, *samp_split;
    unsigned short* bid;        // bucket of every key, by sort-buffer position
    int* counts;                // (chunk, bucket) counts, then scatter offsets
    SsTile* tiles;              // run queues (queue q from qbase[q]; at most its buckets)
    const int* qbase;           // ENM_NQ + 1 queue offsets (device)
    unsigned* qctl;             // 2 * ENM_NQ + 1 counters, ENM_CSTRIDE apart: appends, pops, long runs
    SsTile* long_runs;          // runs of one bucket larger than an LDS run (at most its buckets)
    int long_grid;
    int tile_grid;
    const SsCons* cons;         // consumers of each segment (SampleSeg::cons0, ncons)
    const OrderDev* orders;
    double* col64;
    float* col32;
    int* perm;
    unsigned long long* maxabs;
    int with32;
    const int* hint;            // caller's sort hints, parallel to the arena (NULL: none)
    int* hint_bad;              // per segment: 1 when its hint was absent or wrong; then one
                                // more: 1 when any was (all nsegs + 1 zeroed before)
    int all_hinted;             // every segment has a hint: nothing else is sorted unless one fails
};
hipError_t enm_launch_sample_sort(const SampleSortArgs& a, const double* arena, double* k0, int* i0,
                                  double* k1, int* i1, int* n_launch, hipStream_t st);
hipError_t enm_launch_strip_sort(const GatherItem* d_items, int nitems, const SortSeg* d_segs,
                                 const OrderDev* d_orders, int tile, double* col64, float* col32, int* perm,
                                 int with32, int* n_launch, hipStream_t st);
hipError_t enm_launch_gather(const GatherItem* d_items, int nitems, const SortSeg* d_segs,
                             const OrderDev* d_orders, const double* arena,
                             const double* skey, const int* sidx,
                             double* col64, float* col32, int* perm,
                             unsigned long long* maxabs, int with32, int* n_launch, hipStream_t st);

int enm_knn_dim_class(int d);
int enm_knn_list_class(int k1);
hipError_t enm_launch_knn(const Item* items, const int* qoff, int n_items, unsigned* counter,
                          const KnnDev* descs, const OrderDev* orders,
                          const unsigned long long* maxabs, const int* perm,
                          const double* col64, const float* col32, double* res64,
                          unsigned long long* stats, int dclass, int pclass, bool f32, bool brute,
                          int walk, int grid, int k1max, double* heap, int ticket, int smax, char* far,
                          unsigned* far_count, int64_t far_bytes, hipStream_t st);
// Grid of a heap-list kNN launch (ENM_LIST_HEAP_GLOBAL) whose heaps fit
// ENM_HEAP_GLOBAL_BYTES: at most `grid` blocks of 64 x k1max doubles.
int enm_knn_heap_grid(int grid, int k1max);

int enm_count_dim_class(int d);
hipError_t enm_launch_count(const Item* items, const int* qoff, int n_items, unsigned* counter,
                            const CountDev* descs, const OrderDev* orders,
                            const unsigned long long* maxabs, const int* perm, const double* psi_tab,
                            const double* col64, const float* col32, double* res64, int* res32,
                            unsigned long long* stats, int dclass, bool f32, bool brute,
                            int grid, unsigned* deferred, unsigned* n_deferred, unsigned seg_cap,
                            int ticket, hipStream_t st);
hipError_t enm_launch_psi_table(double* tab, int64_t n, hipStream_t st);
hipError_t enm_launch_psi_tiles(const Item* items, int n_items, const CountDev* descs, const double* psi_tab,
                                const int* counts, double* res64, int* res32, hipStream_t st);

hipError_t enm_launch_combine(const InstDev* insts, const double* cls, int first, int count,
                              const double* res64, const int* res32, double* out, hipStream_t st);
hipError_t enm_launch_reduce_pt(const RedItem* items, int n_items, const RedDev* descs, double* res64,
                                hipStream_t st);
hipError_t enm_launch_export_eps(const double* src, const int* perm, int64_t n, double* dst, hipStream_t st);
hipError_t enm_launch_export_cnt(const int* src, const int* perm, int64_t n, int* dst, hipStream_t st);

hipError_t enm_launch_prep_count(const PrepTaskDev* tasks, int64_t stride, int n_tasks,
                                 const double* raw, const unsigned char* mask, int drop_nan, int* blk_count,
                                 int64_t* blk_off, unsigned* nan_flag, int64_t* n_out, hipStream_t st);
hipError_t enm_launch_prep_write(PrepTaskDev* tasks, int64_t stride, int n_tasks,
                                 const double* raw, const unsigned char* mask, int drop_nan, int rescale,
                                 const int64_t* blk_off, const PrepSumDev* sums, int n_sums,
                                 const int* chunk_sum_of, const int* chunk_index, const int* chunk_list,
                                 int n_chunks, int n_full, double* chunk_val, double* stats, int64_t sstride,
                                 const double* noise, unsigned* const_flag, double* arena, KeptWord* kept,
                                 hipStream_t st);
// Sort hints (enm_hint.hip).
hipError_t enm_launch_hint_keys(const HintSeries* series, int n_series, int64_t maxlen, const double* raw,
                                double* keys, hipStream_t st);
// Per (column, segment of ENM_HINT_SEG sorted rows) of the hint filter.
struct HintSegs {
    int* cnt;        // kept rows of the segment
    double* last;    // value before the noise of its last kept row
    int64_t* off;    // its first output position
    double* cin;     // value before the noise of the last kept row before it
    int* chas;       // whether there is one
};
// The hint of every column: noise maximum, segmented filter (count, scan,
// filter), runs re-sorted, columns without a hint marked.
// Filter segments per column for n_cols columns of at most max_len sorted
// rows and a GPU holding target_blocks workgroups at once, and the rows of
// a segment (whole chunks).
int hint_segments(int n_cols, int64_t max_len, int target_blocks);
int64_t hint_segment_rows(int64_t max_len, int segs_per_col);
hipError_t enm_launch_hint_order(const HintCol* cols, int n_cols, int64_t max_w, int64_t max_n, int64_t max_len,
                                 int target_blocks,
                                 const int* sidx, const double* skey, const KeptWord* kept,
                                 const PrepTaskDev* tasks, const double* stats, int rescale, const double* noise,
                                 int64_t noise_len, double* noise_max, unsigned* cont, int64_t cont_words,
                                 const HintSegs& sg, const double* arena, int* hint, int* bad, hipStream_t st);

hipError_t enm_launch_uq_codes(const UqSeg* segs, int nsegs, const double* skeys, const int* sidx, int* codes,
                               int64_t* starts, int64_t* counts, int64_t* distinct, hipStream_t st);
hipError_t enm_launch_uq_keys(const UqRowKey* rk, int n_rk, int64_t max_n, const int* ccodes, double* keys,
                              hipStream_t st);
