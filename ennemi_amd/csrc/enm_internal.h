// Internal layout shared by the HIP kernels and the host planner of
// libennemi_hip.so.  Nothing here is part of the public ABI.
//
// HBM layout of one batch.  All offsets are element offsets into five
// workspace buffers, so each kernel reads candidates through one read-only
// base pointer per element type (which lets the compiler keep wave-uniform
// candidate reads on the scalar path) and writes through others:
//
//   col64 (double) : per order, its columns in sorted order [col][pos]
//                    (column 0 is the sort key).
//   col32 (float)  : per order, the FP32 rounding of col64's columns.
//   perm  (int32)  : per order, the permutation pos -> instance point index.
//   res64 (double) : radii eps_A (kNN order) and eps_orig (instance
//                    numbering), per-tile partial sums.
//   res32 (int32)  : per-space ball counts, per-tile zero flags.
//
// An "order" is one sorted view of an instance's points: the points of each
// scope range (one class segment, or all points) sorted by the key column,
// with the columns the stage needs gathered into that order.
#pragma once
#include <stdint.h>

#define ENM_TILE 64            // queries per work item (one per lane of a wave64)
#define ENM_NQ 8               // work queues per persistent launch (one per XCD)
#define ENM_C1_THREADS 512     // threads of a 1-D count workgroup
#define ENM_C1_Q 4             // queries per thread of a 1-D count item
#define ENM_C1_ITEM (ENM_C1_THREADS * ENM_C1_Q)   // positions per 1-D count item (2048)
#define ENM_C1_HALO 1528       // keys staged on each side of a 1-D count item
#define ENM_C1_CAP (ENM_C1_ITEM + 2 * ENM_C1_HALO)  // LDS window (doubles, 40 KB: four workgroups per CU)
#define ENM_C1_OFF_BITS 11     // position bits of a deferred 1-D query code
#define ENM_NDEF 64            // segments (and counters) of the deferred 1-D query list
#define ENM_CSTRIDE 32         // unsigned slots between two atomic counters (one per 128 B line)
#define ENM_KCHUNK 8           // kNN / multi-D count items taken per ticket
#define ENM_PSI_TILE 512       // positions per psi partial sum of a 1-D space
#define ENM_SORT_TILE 2048     // elements per block-sort tile and merge chunk
#define ENM_SORT_THREADS 256
#define ENM_MAX_SPACES 3
// kNN list classes beyond the register lists of 4..32 slots: a per-lane
// max-heap of k+1 values in LDS (k+1 <= ENM_LIST_LDS_MAX) or in a global
// buffer (any larger k+1).
#define ENM_LIST_HEAP_LDS 0
#define ENM_LIST_HEAP_GLOBAL -1
#define ENM_LIST_LDS_MAX 128
#define ENM_HEAP_GLOBAL_BYTES (256ll << 20)  // most global heap storage of one launch
// Two-column kNN: strips up to this size are walked a strip per workgroup
// with the strip and its two neighbours staged in LDS (3 x 1024 x 16 B).
#define ENM_STRIP_LDS_MAX 1024
// Default strip of instances of three or more joint columns.
#define ENM_STRIP_WIDE 4096

// One contiguous key range to sort: keys at arena[src + t], t in [0, len),
// carried with their instance point index p0 + t.
struct SortSeg {
    int64_t src;    // arena offset of the first key
    int64_t sbase;  // base position in the sort buffers
    int32_t len;
    int32_t first_tile;  // index of its first ENM_SORT_TILE chunk
    int32_t p0;          // instance point index of the first key
    int32_t hint;        // 1: the caller's hint array may hold this segment's order at src
};

// Sample sort of the first-stage segments (segments of up to ENM_SS_MAXLEN
// keys): each segment is split into nb buckets by nb - 1 splitters drawn from
// ENM_SS_OVER * nb sorted samples, its keys are scattered bucket by bucket,
// and runs of whole buckets of at most ENM_SS_TILE keys are sorted in LDS.
#define ENM_SS_BUCKET 768      // target keys per bucket
#define ENM_SS_OVER 16         // samples per bucket
#define ENM_SS_MAXB 1024       // most buckets of one segment
#define ENM_SS_MAXLEN (ENM_SS_MAXB * 1024)
#define ENM_SS_TILE 4096       // keys per LDS-sorted run of buckets
#define ENM_SS_THREADS 512

// Per first-stage segment (same index as its SortSeg).
struct SampleSeg {
    int64_t cbase;  // its (chunk, bucket) count table: chunk-major, nb per chunk
    int32_t samp;   // offset of its samples in the sample buffers
    int32_t ns;     // number of samples: ENM_SS_OVER * nb, or 0 when nb == 1
    int32_t nb;     // number of buckets
    int32_t q;      // run queue (one per XCD, by instance)
    int32_t cons0;  // its consumers: SsCons entries [cons0, cons0 + ncons)
    int32_t ncons;
};

// An order laid out from a sorted segment: the segment's sorted element t
// goes to order position pos0 + t.
struct SsCons {
    int64_t pos0;
    int32_t order;
    int32_t pad;
};

// One run of whole buckets to sort: sort-buffer positions [pos, pos + len)
// of first-stage segment seg.
struct SsTile {
    int64_t pos;
    int32_t len;
    int32_t seg;
};

// Gather work: one ENM_SORT_TILE chunk of one sort segment into one order.
struct GatherItem {
    int32_t order;
    int32_t sseg;
    int32_t off;   // chunk offset within the sort segment
    int32_t pos0;  // order position of the sort segment's first element
};

struct OrderDev {
    int64_t c64;        // col64 offset of column 0; column c at c64 + c*n
    int64_t c32;        // col32 offset of column 0
    int64_t perm;       // perm offset
    int64_t n;          // points of the instance
    int32_t ncol;
    int32_t inst;
    int64_t src[16];    // arena offsets of the instance columns gathered (src[0] = key)
    // Grid orders are sorted in a second stage, within strips of a base
    // order: their sort payload is a base-order position; the point and its
    // columns are read from the base order (perm offset base_perm, col64 /
    // col32 offsets base_c64 / base_c32), inside the same strip.  -1: none.
    int64_t base_perm;
    int64_t base_c64;
    int64_t base_c32;
    // >= 0 (even): col64 offset of an interleaved copy of columns 0 and 1,
    // (c0, c1) of position p at xy + 2p, for the two-column kNN walk (one
    // 16-byte load per candidate).  -1: none.
    int64_t xy;
    // 1: nothing reads the column-major columns (only the strip kNN, which
    // reads xy), so the layout writes xy and the permutation only.
    int32_t xy_only;
    int32_t pad2;
};

// kNN over one order (its scope ranges are the class segments, or all points).
// With strip > 0 the order is a grid order: positions [lo + s*strip,
// lo + (s+1)*strip) of each scope range [lo, hi) form strip s, whose points
// are consecutive in the base order (sorted by column 0, whose sorted values
// start at col64 offset base_key) and are sorted by column 1 inside the strip.
struct KnnDev {
    int32_t order;
    int32_t inst;
    int32_t k;
    int32_t kl;         // 1: an entropy instance (radii are summed as logs)
    int64_t eps;        // res64 offset: radius per order position (or -1)
    int64_t eps_orig;   // res64 offset: radius per instance point (or -1)
    int64_t part;       // unused (-1)
    int64_t base_key;   // col64 offset of the base order's sorted column 0
    int32_t strip;      // points per strip (0: plain key-sorted order)
    int32_t pad;
};

// Ball count in one space (a subset of an order's columns, key first).
struct CountDev {
    int32_t order;
    int32_t inst;
    int32_t ncol;
    int32_t same_order;  // 1: radius read at the same order position (eps_A)
    int32_t cols[16];    // indices into the order's columns; cols[0] == 0
    int64_t eps;         // res64 offset of the radius source
    int64_t eps_perm;    // perm offset of the order's permutation (radius via eps_orig)
    int64_t counts;      // res32 offset: count per order position
    int64_t part;        // res64 offset: per-tile sum of psi(count)
    int64_t zero;        // res32 offset: per-tile "some count is 0" flag
    // With strip > 0 the order is a grid order (as for KnnDev) and the space
    // starts with the order's columns 0 and 1.
    int64_t base_key;
    int32_t strip;
    int32_t pad;
    // >= 0: res64 offset of a point-order array receiving psi(count) (+inf
    // for a count of 0) instead of per-tile partial sums; a reduction in
    // point order then makes the estimate independent of this order.
    int64_t psi_pt;
};

// Fixed-chunk sum of a point-order array (psi values, or log radii).
#define ENM_RED_CHUNK 4096
struct RedDev {
    int64_t src;    // res64 offset of n values in point order
    int64_t n;
    int64_t part;   // res64 offset of one partial per chunk
    int32_t log;    // 1: sum log(value)
    int32_t pad;
};
struct RedItem {
    int32_t desc;
    int32_t chunk;
};

// One work item: a tile of consecutive order positions [t0, t1) inside the
// scope range [lo, hi) (64 for kNN and multi-D counts, ENM_C1_ITEM for 1-D
// counts, ENM_PSI_TILE for 1-D psi sums).  A persistent launch takes its
// items from ENM_NQ queues, queue q being items [qoff[q], qoff[q+1]).
struct Item {
    int32_t desc;
    int32_t tile;   // tile index within its descriptor (partials slot)
    int32_t lo, hi;
    int32_t t0, t1;
};

// Per-instance data for the final combination.
struct InstDev {
    int32_t kind;
    int32_t k;
    int64_t n;
    int32_t dx;
    int32_t nspace;
    int32_t n_segs;
    int32_t pad0;
    int64_t seg_len;     // offset of the class sizes (as doubles) in the class table, or -1
    int64_t part[ENM_MAX_SPACES];  // res64 offsets of per-tile partials
    int64_t zero[ENM_MAX_SPACES];  // res32 offsets of per-tile zero flags
    int32_t ntile[ENM_MAX_SPACES];
    int32_t pad1;
};

// Device counters (enm_stats order).
enum { ST_KNN_PAIRS = 0, ST_KNN_RECHECK, ST_CNT_PAIRS, ST_CNT_RECHECK, ST_CNT_1D, ST_SORT_KEYS, ST_N };

// Device-side preparation of estimate_mi tasks (enm_prep_count / enm_prep_write).
#define ENM_PREP_BLOCK 2048     // window rows per compaction block
#define ENM_PREP_MAXZ 14        // condition columns of one task (ENM_MAX_COLS - 2)

struct PrepTaskDev {
    int64_t x_col, y_col;       // raw offsets of the x and y series
    int64_t x0, y0;             // first row of each window (the mask is aligned with y)
    int64_t w;                  // window rows
    int64_t z_col[ENM_PREP_MAXZ], z0[ENM_PREP_MAXZ];
    int32_t dz, nblk;           // condition columns; compaction blocks
    int64_t arena_off;          // arena offset of its x column (then y, then z columns, n each)
    int64_t n;                  // kept rows
    int64_t stat;               // its first stats slot: mean, std of x, y, then each z column
    int64_t noise_x, noise_y, noise_z;   // noise offsets (set on the device)
    uint32_t flags, pad;        // bit 0/1/2: x/y/z constant (not rescaled)
    int64_t pmap;               // first word of its kept-row map (sort hints): ceil(w / 32) words
};

// One word of a task's kept-row map: bit b is window row 32 * i + b (1:
// kept), prefix the kept rows before row 32 * i.  A kept row r is at task
// position prefix + popcount(bits below r).
struct KeptWord {
    uint32_t bits;
    int32_t prefix;
};

// Sort hints (enm_prep_hints).  A series: the rows [lo, lo + len) of a raw
// column that some task windows, sorted once at key-buffer position kbase.
struct HintSeries {
    int64_t col, lo, len, kbase;
};

// One task column: its series' sorted rows filtered to the task's kept
// rows give the column's order before the noise.
struct HintCol {
    int64_t sb;      // key-buffer position of the series' first sorted row
    int64_t ws;      // the window's first row, relative to the series range
    int64_t w;       // window rows
    int64_t pm;      // the task's kept-row map (first word)
    int64_t out;     // arena (and hint) offset of the column
    int64_t n;       // kept rows
    int64_t cbits;   // first word of its continuation bits (one per position)
    int32_t len;     // rows of the series range
    int32_t task;    // the task (its statistics and constant flags)
    int32_t slot;    // the column: 0 x, 1 y, 2 + c condition column c
    int32_t pad;
};
#define ENM_HINT_BLOCK 1024   // sorted series rows per filter chunk (256 threads x 4)
#define ENM_HINT_RUN 32       // longest run of near-equal values re-sorted (longer: no hint)

// One array whose mean and standard deviation NumPy's way are needed.
struct PrepSumDev {
    int64_t src;        // arena offset
    int64_t n;
    int32_t stat;       // stats slot of the mean (the std follows)
    int32_t sequential; // 1: accumulated row by row (a column of a multi-column condition)
    int32_t chunk0;     // pairwise arrays: first of its 8192-value chunks
    int32_t n_chunks;
};

// Distinct rows (enm_unique_rows): one sorted segment of keys.
struct UqSeg {
    int64_t sbase;      // sort-buffer position of its first key
    int64_t n;
    int64_t code_off;   // codes of its points, by point index
    int64_t start_off;  // run starts (n + 1 slots)
    int64_t count_off;  // points per code
    int32_t slot;       // its distinct-count slot
    int32_t pad;
};

// The mixed-radix keys of a multi-column row set.
struct UqRowKey {
    int64_t n;
    int64_t key_off;            // where its keys go
    int32_t ncol, pad;
    int64_t code_off[16];       // each column's codes
    int64_t radix[16];          // each column's distinct count
};
