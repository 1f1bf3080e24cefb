// Segmented sort of float64 keys with their point index, and the gather that
// lays an instance's columns out in sorted order (FP64 and FP32 copies).
//
// Replaces the kd-tree builds of the reference
// (ennemi/_entropy_estimators.py:36,100-102,135-139,184-185,222-228): a sorted
// key column is all the brute-force window kernels and the 1-D binary-search
// counts need.
//
// Algorithm: every segment is cut into ENM_SORT_TILE chunks; a 256-thread
// block sorts one chunk in LDS with a bitonic network, then merge passes of
// doubling run width merge neighbouring runs inside each segment.  Per pass,
// one thread per chunk finds the merge-path split of the chunk's first output
// (co_rank_kernel), then a merge block owns ENM_SORT_TILE outputs: it stages
// its two input slices in LDS and each thread merges 8 outputs.  Keys compare as
// (value, index) pairs, so all keys are distinct and the order is total.
// Traffic per pass is 12 B read + 12 B written per key (HBM-bound).
//
// First-stage segments of up to ENM_SS_MAXLEN keys are sample-sorted
// instead: nb - 1 splitters from ENM_SS_OVER * nb sorted samples cut a
// segment into nb buckets (ss_count_kernel counts every 2048-key chunk's keys
// per bucket, ss_scan_kernel turns the counts into scatter offsets and packs
// consecutive buckets into runs of at most ENM_SS_TILE keys,
// ss_scatter_kernel moves the keys bucket by bucket), and ss_tile_kernel
// sorts each run in LDS.  56 B of traffic per key instead of 24 B per merge
// pass.  A run of one bucket larger than ENM_SS_TILE (sampling bad luck) is
// merge-sorted by one block through global memory, so every input sorts.
//
// Sort hints: a caller may pass, parallel to the arena, the ascending order
// of a key column (hint[src + i] = point index of the i-th smallest key,
// ties by index; hint[src] < 0 for none).  ss_hint_kernel checks a hinted
// segment's order element by element against its neighbour -- every index
// inside the segment and every (key, index) pair strictly above the one
// before, which makes it the sorted permutation -- and lays it out into the
// orders directly; the sample sort then passes over the segments whose hint
// held and sorts the others (overwriting whatever the check laid out).
#include "enm_device.h"
#include "enm_launch.h"

namespace {

__device__ __forceinline__ bool kless(double ka, int ia, double kb, int ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Index of the segment owning chunk `tile` (segments sorted by first_tile).
__device__ __forceinline__ int seg_of_tile(const SortSeg* __restrict__ segs, int nsegs, int tile) {
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].first_tile <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Compare-exchange of two (key, index) pairs.
__device__ __forceinline__ void cex_pair(double& ka, int& ia, double& kb, int& ib) {
    const bool sw = kless(kb, ib, ka, ia);
    const double k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const int i0 = sw ? ib : ia, i1 = sw ? ia : ib;
    ka = k0; ia = i0; kb = k1; ib = i1;
}

// Sort PER (key, index) pairs held in registers (PER = 2, 4, 8 or 16;
// Batcher's odd-even merge sorting networks: 1, 5, 19 and 63 comparators).
template <int PER>
__device__ __forceinline__ void sort_regs(double (&k)[PER], int (&x)[PER]) {
#define ENM_CX(a, b) cex_pair(k[a], x[a], k[b], x[b])
    if constexpr (PER == 2) {
        ENM_CX(0, 1);
    } else if constexpr (PER == 4) {
        ENM_CX(0, 1); ENM_CX(2, 3); ENM_CX(0, 2); ENM_CX(1, 3); ENM_CX(1, 2);
    } else if constexpr (PER == 8) {
        ENM_CX(0, 1); ENM_CX(2, 3); ENM_CX(4, 5); ENM_CX(6, 7);
        ENM_CX(0, 2); ENM_CX(1, 3); ENM_CX(4, 6); ENM_CX(5, 7);
        ENM_CX(1, 2); ENM_CX(5, 6);
        ENM_CX(0, 4); ENM_CX(1, 5); ENM_CX(2, 6); ENM_CX(3, 7);
        ENM_CX(2, 4); ENM_CX(3, 5);
        ENM_CX(1, 2); ENM_CX(3, 4); ENM_CX(5, 6);
    } else {
        static_assert(PER == 16, "2, 4, 8 or 16 per thread");
        // Two sorted halves of 8 (the 19-comparator network on each) ...
        ENM_CX(0, 1); ENM_CX(2, 3); ENM_CX(0, 2); ENM_CX(1, 3); ENM_CX(1, 2);
        ENM_CX(4, 5); ENM_CX(6, 7); ENM_CX(4, 6); ENM_CX(5, 7); ENM_CX(5, 6);
        ENM_CX(0, 4); ENM_CX(2, 6); ENM_CX(2, 4); ENM_CX(1, 5); ENM_CX(3, 7); ENM_CX(3, 5);
        ENM_CX(1, 2); ENM_CX(3, 4); ENM_CX(5, 6);
        ENM_CX(8, 9); ENM_CX(10, 11); ENM_CX(8, 10); ENM_CX(9, 11); ENM_CX(9, 10);
        ENM_CX(12, 13); ENM_CX(14, 15); ENM_CX(12, 14); ENM_CX(13, 15); ENM_CX(13, 14);
        ENM_CX(8, 12); ENM_CX(10, 14); ENM_CX(10, 12); ENM_CX(9, 13); ENM_CX(11, 15); ENM_CX(11, 13);
        ENM_CX(9, 10); ENM_CX(11, 12); ENM_CX(13, 14);
        // ... and their odd-even merge (25 comparators).
        ENM_CX(0, 8); ENM_CX(4, 12); ENM_CX(4, 8); ENM_CX(2, 10); ENM_CX(6, 14); ENM_CX(6, 10);
        ENM_CX(2, 4); ENM_CX(6, 8); ENM_CX(10, 12); ENM_CX(1, 9); ENM_CX(5, 13); ENM_CX(5, 9);
        ENM_CX(3, 11); ENM_CX(7, 15); ENM_CX(7, 11); ENM_CX(3, 5); ENM_CX(7, 9); ENM_CX(11, 13);
        ENM_CX(1, 2); ENM_CX(3, 4); ENM_CX(5, 6); ENM_CX(7, 8); ENM_CX(9, 10); ENM_CX(11, 12);
        ENM_CX(13, 14);
    }
#undef ENM_CX
}

// Number of A elements among the first d outputs of merge(A, B).  Keys
// are compared first; the indices are read only where some lane's keys tie.
__device__ int co_rank(int d, const double* ka, const int* ia, int na,
                       const double* kb, const int* ib, int nb) {
    int lo = max(0, d - nb), hi = min(d, na);
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int j = d - i - 1;
        const double x = ka[i], y = kb[j];
        bool lt = x < y;
        if (x == y) lt = ia[i] < ib[j];
        if (lt) lo = i + 1; else hi = i;
    }
    return lo;
}

// LDS tiles of (key, index) pairs keep one unused slot after every 32
// elements (element i at lp(i)).  A thread's PER consecutive elements then
// start at a different bank from its neighbour's: without the slot, the 64
// lanes' 8-element runs (64 B apart) fall on the same few banks and every
// register-phase read and write serializes 16 ways.
#define ENM_LDS_PAD(n) ((n) + ((n) >> 5))
__device__ __forceinline__ int lp(int i) { return i + (i >> 5); }

// co_rank over an LDS tile: A at elements [a0, a0 + na), B at [b0, b0 + nb).
__device__ __forceinline__ int co_rank_lds(int d, const double* sk, const int* si, int a0, int na,
                                           int b0, int nb) {
    int lo = max(0, d - nb), hi = min(d, na);
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int pa = lp(a0 + i), pb = lp(b0 + d - i - 1);
        const double x = sk[pa], y = sk[pb];
        bool lt = x < y;
        if (x == y) lt = si[pa] < si[pb];
        if (lt) lo = i + 1; else hi = i;
    }
    return lo;
}

// The PER outputs of merge(A, B) of an LDS tile (A at [a0, a0 + na), B at
// [b0, b0 + nb), na, nb >= 1) from merge-path split (a, b): the two heads
// stay in registers and each step reloads only the side it consumed.
// Clamped reads stay inside the slices; the flags decide.
template <int PER>
__device__ __forceinline__ void seq_merge(const double* sk, const int* si, int a0, int na, int b0, int nb,
                                          int a, int b, int ne, double (&k)[PER], int (&x)[PER]) {
    double ha = sk[lp(a0 + min(a, na - 1))], hb = sk[lp(b0 + min(b, nb - 1))];
    int xa = si[lp(a0 + min(a, na - 1))], xb = si[lp(b0 + min(b, nb - 1))];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        if (e < ne) {
            const bool take_a = b >= nb || (a < na && kless(ha, xa, hb, xb));
            k[e] = take_a ? ha : hb;
            x[e] = take_a ? xa : xb;
            if (e + 1 < PER) {
                a += take_a ? 1 : 0;
                b += take_a ? 0 : 1;
                const int src = lp(take_a ? a0 + min(a, na - 1) : b0 + min(b, nb - 1));
                const double v = sk[src];
                const int vi = si[src];
                ha = take_a ? v : ha;
                xa = take_a ? vi : xa;
                hb = take_a ? hb : v;
                xb = take_a ? xb : vi;
            }
        }
    }
}

// Sort the TILE (key, index) pairs of an LDS tile (ENM_LDS_PAD(TILE) slots)
// with an NT-thread block: each thread sorts TILE/NT consecutive pairs in
// registers, then merge levels of doubling run width, each thread producing
// TILE/NT consecutive outputs of its pair of runs from a merge-path split
// (outputs kept in registers across a barrier, so the merge works in place).
// Padding keys (+inf, INT_MAX) sort last.  Starts and ends with a barrier.
template <int TILE, int NT = ENM_SORT_THREADS>
__device__ __forceinline__ void block_sort_lds(double* sk, int* si) {
    constexpr int PER = TILE / NT;
    __syncthreads();
    const int e0 = threadIdx.x * PER;
    double k[PER];
    int x[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) { k[e] = sk[lp(e0 + e)]; x[e] = si[lp(e0 + e)]; }
    sort_regs<PER>(k, x);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) { sk[lp(e0 + e)] = k[e]; si[lp(e0 + e)] = x[e]; }
    __syncthreads();
    for (int run = PER; run < TILE; run <<= 1) {
        const int pb = (e0 / (2 * run)) * (2 * run);
        const int d = e0 - pb;
        const int a = co_rank_lds(d, sk, si, pb, run, pb + run, run);
        seq_merge<PER>(sk, si, pb, run, pb + run, run, a, d - a, PER, k, x);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PER; ++e) { sk[lp(e0 + e)] = k[e]; si[lp(e0 + e)] = x[e]; }
        __syncthreads();
    }
}

// Merge the staged sorted slices [0, la) and [la, la + lb) of an LDS tile
// (la + lb <= NT * PER) in place: each thread takes PER consecutive outputs
// into registers from its merge-path split, and they go back after a
// barrier.  The caller synchronizes after staging; ends with a barrier.
template <int NT, int PER>
__device__ __forceinline__ void lds_merge(double* sk, int* si, int la, int lb) {
    if (la == 0 || lb == 0) {   // one slice: already in order (block-uniform)
        __syncthreads();
        return;
    }
    const int e0 = threadIdx.x * PER;
    const int ne = min(PER, la + lb - e0);
    double ok[PER];
    int oi[PER];
    if (ne > 0) {
        const int a = co_rank_lds(e0, sk, si, 0, la, la, lb);
        seq_merge<PER>(sk, si, 0, la, la, lb, a, e0 - a, ne, ok, oi);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        if (e < ne) {
            sk[lp(e0 + e)] = ok[e];
            si[lp(e0 + e)] = oi[e];
        }
    }
    __syncthreads();
}

// Sort one TILE chunk of a segment: keys arena[src + off + t] with their
// point indices p0 + off + t (or iarena[src + off + t] when given), into the
// sort buffers.
template <int TILE>
__global__ __launch_bounds__(ENM_SORT_THREADS)
void block_sort_kernel(const SortSeg* __restrict__ segs, int nsegs,
                       const double* __restrict__ arena, const int* __restrict__ iarena,
                       double* __restrict__ kout, int* __restrict__ iout, const int* __restrict__ skip) {
    __shared__ double sk[ENM_LDS_PAD(TILE)];
    __shared__ int si[ENM_LDS_PAD(TILE)];
    if (skip != nullptr && *skip == 0) return;   // nothing to sort this run
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    const int off = (tile - sg.first_tile) * TILE;
    const int len = min(TILE, sg.len - off);
    for (int t = threadIdx.x; t < TILE; t += ENM_SORT_THREADS) {
        if (t < len) {
            sk[lp(t)] = arena[sg.src + off + t];
            si[lp(t)] = iarena ? iarena[sg.src + off + t] : sg.p0 + off + t;
        } else {
            sk[lp(t)] = INFINITY;
            si[lp(t)] = 0x7fffffff;
        }
    }
    block_sort_lds<TILE>(sk, si);
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        kout[sg.sbase + off + t] = sk[lp(t)];
        iout[sg.sbase + off + t] = si[lp(t)];
    }
}

// Build a grid order strip by strip: one block per strip (at most TILE
// points) sorts the strip's points by the base order's column 1 in LDS and
// copies each point's columns and instance index from the base order into
// the grid order directly (strip-local reads, coalesced writes), with no
// pass through the sort buffers.  Item: the strip's sort segment (src =
// column 1 of the base order at the strip, p0 = the strip's first position,
// which is the same in both orders).
template <int TILE>
__global__ __launch_bounds__(ENM_SORT_THREADS)
void strip_sort_kernel(const GatherItem* __restrict__ items, const SortSeg* __restrict__ segs,
                       const OrderDev* __restrict__ orders, double* col64, float* col32, int* perm,
                       int with32) {
    __shared__ double sk[ENM_LDS_PAD(TILE)];
    __shared__ int si[ENM_LDS_PAD(TILE)];
    const GatherItem it = items[blockIdx.x];
    const SortSeg sg = segs[it.sseg];
    const OrderDev& od = orders[it.order];
    const int len = sg.len;
    // A thread's PER elements: every load of a column goes out before any
    // store (a store between two loads would hold the second one back).
    constexpr int PER = TILE / ENM_SORT_THREADS;
    {
        double kv[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = k * ENM_SORT_THREADS + (int)threadIdx.x;
            kv[k] = t < len ? col64[sg.src + t] : INFINITY;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = k * ENM_SORT_THREADS + (int)threadIdx.x;
            sk[lp(t)] = kv[k];
            si[lp(t)] = t < len ? sg.p0 + t : 0x7fffffff;
        }
    }
    block_sort_lds<TILE>(sk, si);
    const int64_t n = od.n;
    const int ncol = od.ncol;
    int idx[PER];
    auto in = [&](int k) { return k * ENM_SORT_THREADS + (int)threadIdx.x < len; };
    auto pos = [&](int k) { return it.pos0 + k * ENM_SORT_THREADS + (int)threadIdx.x; };
#pragma unroll
    for (int k = 0; k < PER; ++k) idx[k] = in(k) ? si[lp(k * ENM_SORT_THREADS + (int)threadIdx.x)] : 0;
    {
        int pv[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) pv[k] = in(k) ? perm[od.base_perm + idx[k]] : 0;
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (in(k)) perm[od.perm + pos(k)] = pv[k];
    }
    for (int c = 0; c < (od.xy_only ? 0 : ncol); ++c) {
        double v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) v[k] = in(k) ? col64[od.base_c64 + c * n + idx[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (in(k)) col64[od.c64 + c * n + pos(k)] = v[k];
        if (with32) {
            float f[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) f[k] = in(k) ? col32[od.base_c32 + c * n + idx[k]] : 0.0f;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (in(k)) col32[od.c32 + c * n + pos(k)] = f[k];
        }
    }
    if (od.xy >= 0) {
        // Column 1 is the sort key: its sorted values are in LDS already.
        double2 v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            v[k].x = in(k) ? col64[od.base_c64 + idx[k]] : 0.0;
            v[k].y = in(k) ? sk[lp(k * ENM_SORT_THREADS + (int)threadIdx.x)] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (in(k)) reinterpret_cast<double2*>(col64 + od.xy)[pos(k)] = v[k];
    }
}

// Geometry of chunk `tile` in a merge pass of run width `width`: its pair of
// runs A = [base, base+na), B = [base+width, base+width+nb) in the sort
// buffer and its output range [d0, d1) within the pair.
struct MergeChunk {
    int64_t base;
    int na, nb, d0, d1;
};

__device__ __forceinline__ MergeChunk merge_chunk(const SortSeg& sg, int tile, int64_t width) {
    MergeChunk m;
    const int64_t off = (int64_t)(tile - sg.first_tile) * ENM_SORT_TILE;
    const int64_t len = sg.len;
    const int64_t pb = (off / (2 * width)) * (2 * width);
    m.na = (int)min(width, len - pb);
    m.nb = (int)max((int64_t)0, min(2 * width, len - pb) - width);
    m.d0 = (int)(off - pb);
    m.d1 = min(m.d0 + ENM_SORT_TILE, m.na + m.nb);
    m.base = sg.sbase + pb;
    return m;
}

// Merge-path split of every chunk's first output, one thread per chunk, all
// in parallel (each a binary search of ~17 dependent loads).
__global__ __launch_bounds__(ENM_SORT_THREADS)
void co_rank_kernel(const SortSeg* __restrict__ segs, int nsegs, int ntiles, int64_t width,
                    const double* __restrict__ kin, const int* __restrict__ iin, int* __restrict__ split,
                    const int* __restrict__ skip) {
    const int tile = blockIdx.x * ENM_SORT_THREADS + threadIdx.x;
    if (tile >= ntiles || (skip != nullptr && *skip == 0)) return;
    const SortSeg sg = segs[seg_of_tile(segs, nsegs, tile)];
    const MergeChunk m = merge_chunk(sg, tile, width);
    split[tile] = m.nb == 0 ? m.d0
                            : co_rank(m.d0, kin + m.base, iin + m.base, m.na, kin + m.base + width,
                                      iin + m.base + width, m.nb);
}

__global__ __launch_bounds__(ENM_SORT_THREADS)
void merge_pass_kernel(const SortSeg* __restrict__ segs, int nsegs, int64_t width,
                       const double* __restrict__ kin, const int* __restrict__ iin,
                       const int* __restrict__ split,
                       double* __restrict__ kout, int* __restrict__ iout, const int* __restrict__ skip) {
    __shared__ double sk[ENM_LDS_PAD(ENM_SORT_TILE)];
    __shared__ int si[ENM_LDS_PAD(ENM_SORT_TILE)];
    if (skip != nullptr && *skip == 0) return;
    const int tile = blockIdx.x;
    const SortSeg sg = segs[seg_of_tile(segs, nsegs, tile)];
    const MergeChunk m = merge_chunk(sg, tile, width);
    const int64_t base = m.base;
    const int na = m.na, nb = m.nb, d0 = m.d0, d1 = m.d1;
    if (nb == 0) {
        for (int t = d0 + threadIdx.x; t < d1; t += ENM_SORT_THREADS) {
            kout[base + t] = kin[base + t];
            iout[base + t] = iin[base + t];
        }
        return;
    }
    const double* ka = kin + base;
    const int* ia = iin + base;
    const double* kb = kin + base + width;
    const int* ib = iin + base + width;
    // The next chunk's split is this chunk's end split, unless the pair ends.
    const int i0 = split[tile];
    const int i1 = d1 == na + nb ? na : split[tile + 1];
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int la = i1 - i0, lb = j1 - j0;
    for (int t = threadIdx.x; t < la + lb; t += ENM_SORT_THREADS) {
        if (t < la) { sk[lp(t)] = ka[i0 + t]; si[lp(t)] = ia[i0 + t]; }
        else        { sk[lp(t)] = kb[j0 + t - la]; si[lp(t)] = ib[j0 + t - la]; }
    }
    __syncthreads();
    // Merged in LDS, then out with coalesced stores.
    lds_merge<ENM_SORT_THREADS, ENM_SORT_TILE / ENM_SORT_THREADS>(sk, si, la, lb);
    for (int t = threadIdx.x; t < la + lb; t += ENM_SORT_THREADS) {
        kout[base + d0 + t] = sk[lp(t)];
        iout[base + d0 + t] = si[lp(t)];
    }
}

// Lay one sorted chunk out into an order: permutation, key column and the
// other gathered columns (FP64, and FP32 copies when a kernel will read
// them: with32), and fold |value| into the
// instance's maximum magnitude (for the FP32 guard band).  For a grid order
// the sort payload is a base-order position in the same strip: the point and
// every column are copied from the base order (strip-local reads; the base
// order's magnitudes are already folded in).
__global__ __launch_bounds__(ENM_SORT_THREADS)
void gather_kernel(const GatherItem* __restrict__ items,
                   const SortSeg* __restrict__ segs,
                   const OrderDev* __restrict__ orders,
                   const double* __restrict__ arena,
                   const double* __restrict__ skey, const int* __restrict__ sidx,
                   double* col64, float* col32,
                   int* perm, unsigned long long* __restrict__ maxabs, int with32) {
    // A thread's PER elements: every load of a column goes out before any
    // store (a store between two loads would hold the second one back).
    constexpr int PER = ENM_SORT_TILE / ENM_SORT_THREADS;
    const GatherItem it = items[blockIdx.x];
    const SortSeg sg = segs[it.sseg];
    // Fields read one by one: a local copy of the whole descriptor, indexed
    // by a loop counter, would live in scratch memory.
    const OrderDev& od = orders[it.order];
    const int len = min(ENM_SORT_TILE, sg.len - it.off);
    const int64_t n = od.n;
    const int64_t sp0 = sg.sbase + it.off, p0 = it.pos0 + it.off;
    int idx[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int t = k * ENM_SORT_THREADS + (int)threadIdx.x;
        idx[k] = t < len ? sidx[sp0 + t] : 0;
    }
    auto in = [&](int k) { return k * ENM_SORT_THREADS + (int)threadIdx.x < len; };
    auto pos = [&](int k) { return p0 + k * ENM_SORT_THREADS + (int)threadIdx.x; };
    if (od.base_perm >= 0) {
        // idx is a base-order position in the same strip: read locally (the
        // base order's magnitudes are already folded in).
        {
            int pv[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) pv[k] = in(k) ? perm[od.base_perm + idx[k]] : 0;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (in(k)) perm[od.perm + pos(k)] = pv[k];
        }
        for (int c = 0; c < (od.xy_only ? 0 : od.ncol); ++c) {
            double v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) v[k] = in(k) ? col64[od.base_c64 + c * n + idx[k]] : 0.0;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (in(k)) col64[od.c64 + c * n + pos(k)] = v[k];
            if (with32) {
                float f[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k) f[k] = in(k) ? col32[od.base_c32 + c * n + idx[k]] : 0.0f;
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (in(k)) col32[od.c32 + c * n + pos(k)] = f[k];
            }
        }
        if (od.xy >= 0) {
            double2 v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                v[k].x = in(k) ? col64[od.base_c64 + idx[k]] : 0.0;
                v[k].y = in(k) ? col64[od.base_c64 + n + idx[k]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (in(k)) reinterpret_cast<double2*>(col64 + od.xy)[pos(k)] = v[k];
        }
        return;
    }
    double m = 0.0;
    {
        double key[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) key[k] = in(k) ? skey[sp0 + k * ENM_SORT_THREADS + threadIdx.x] : 0.0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (in(k)) {
                perm[od.perm + pos(k)] = idx[k];
                col64[od.c64 + pos(k)] = key[k];
                if (with32) col32[od.c32 + pos(k)] = (float)key[k];
                m = fmax(m, fabs(key[k]));
            }
        }
    }
    for (int c = 1; c < od.ncol; ++c) {
        const double* src = arena + od.src[c];
        double v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) v[k] = in(k) ? src[idx[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (in(k)) {
                col64[od.c64 + c * n + pos(k)] = v[k];
                if (with32) col32[od.c32 + c * n + pos(k)] = (float)v[k];
                m = fmax(m, fabs(v[k]));
            }
        }
    }
    // One atomic per block (atomics from every wave on one address queue
    // behind each other); |v| >= 0, so the IEEE bit patterns order like
    // the values.
    __shared__ double red[ENM_SORT_THREADS / ENM_WAVE];
    m = wave_max(m);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) red[threadIdx.x / ENM_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = red[0];
        for (int w = 1; w < ENM_SORT_THREADS / ENM_WAVE; ++w) mm = fmax(mm, red[w]);
        atomicMax(maxabs + od.inst, (unsigned long long)__double_as_longlong(mm));
    }
}

// ---------------------------------------------------------------------------
// Sample sort of the first-stage segments.

// Exclusive prefix sums of v[0, n) in place (n <= 4 * NT), `part` being
// NT / 64 ints of LDS scratch.  Call after a barrier; ends with one.
template <int NT>
__device__ __forceinline__ void block_exscan(int* v, int n, int* part) {
    constexpr int EPT = 4;
    const int t = threadIdx.x;
    int loc[EPT];
    int s = 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int i = t * EPT + e;
        loc[e] = i < n ? v[i] : 0;
        s += loc[e];
    }
    const int lane = t & (ENM_WAVE - 1);
    int inc = s;
#pragma unroll
    for (int off = 1; off < ENM_WAVE; off <<= 1) {
        const int y = __shfl_up(inc, off, ENM_WAVE);
        if (lane >= off) inc += y;
    }
    if (lane == ENM_WAVE - 1) part[t / ENM_WAVE] = inc;
    __syncthreads();
    int run = inc - s;
    for (int w = 0; w < t / ENM_WAVE; ++w) run += part[w];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int i = t * EPT + e;
        if (i < n) v[i] = run;
        run += loc[e];
    }
    __syncthreads();
}

// Position of sample i of ns in a segment of len keys: one position drawn
// from each of ns equal strides by an integer hash of i, so periodic inputs
// do not bias the splitters.  The positions are distinct.
__device__ __forceinline__ int64_t ss_sample_pos(int i, int ns, int64_t len) {
    uint32_t h = (uint32_t)i * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    const int64_t step = len / ns;
    return ((int64_t)i * len) / ns + (int64_t)(h % (uint32_t)step);
}

// One block per segment: copy its samples (key, point index) out.  Block 0
// also clears the run-queue counters (appends, then pops: 2 * ENM_NQ, then
// the long runs' count: each
// on its own line) that ss_scan_kernel and ss_tile_kernel use.
__global__ __launch_bounds__(ENM_SORT_THREADS)
void ss_sample_kernel(const SortSeg* __restrict__ segs, const SampleSeg* __restrict__ ss,
                      const double* __restrict__ arena, double* __restrict__ sk, int* __restrict__ si,
                      unsigned* __restrict__ qctl, const int* __restrict__ skip) {
    if (skip != nullptr && *skip == 0) return;   // every hint held: nothing to sort this run
    if (blockIdx.x == 0 && threadIdx.x <= 2 * ENM_NQ) qctl[threadIdx.x * ENM_CSTRIDE] = 0u;
    const SortSeg sg = segs[blockIdx.x];
    const SampleSeg q = ss[blockIdx.x];
    for (int i = threadIdx.x; i < q.ns; i += ENM_SORT_THREADS) {
        const int64_t p = ss_sample_pos(i, q.ns, sg.len);
        sk[q.samp + i] = arena[sg.src + p];
        si[q.samp + i] = sg.p0 + (int)p;
    }
}

// One block per 2048-key chunk: the bucket of each key (the number of
// splitters not above it in (value, index) order), kept for the scatter, and
// the chunk's keys per bucket.  Splitter j (j = 1 .. nb-1) is sorted sample
// ENM_SS_OVER * j.
__global__ __launch_bounds__(ENM_SORT_THREADS)
void ss_count_kernel(const SortSeg* __restrict__ segs, int nsegs, const SampleSeg* __restrict__ ss,
                     const double* __restrict__ arena, const double* __restrict__ samp_k,
                     const int* __restrict__ samp_i, unsigned short* __restrict__ bid, int* __restrict__ counts,
                     const int* __restrict__ hint_bad, const int* __restrict__ skip) {
    __shared__ double spk[ENM_SS_MAXB];
    __shared__ int spi[ENM_SS_MAXB];
    __shared__ int hist[ENM_SS_MAXB];
    if (skip != nullptr && *skip == 0) return;
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    if (sg.hint && hint_bad != nullptr && hint_bad[s] == 0) return;   // laid out from its hint
    const SampleSeg q = ss[s];
    const int nb = q.nb, nsp = nb - 1;
    for (int j = threadIdx.x; j < nsp; j += ENM_SORT_THREADS) {
        spk[j] = samp_k[q.samp + ENM_SS_OVER * (j + 1)];
        spi[j] = samp_i[q.samp + ENM_SS_OVER * (j + 1)];
    }
    for (int b = threadIdx.x; b < nb; b += ENM_SORT_THREADS) hist[b] = 0;
    __syncthreads();
    const int off = (tile - sg.first_tile) * ENM_SORT_TILE;
    const int len = min(ENM_SORT_TILE, sg.len - off);
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const double key = arena[sg.src + off + t];
        const int idx = sg.p0 + off + t;
        int lo = 0, hi = nsp;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const double sv = spk[mid];
            bool lt = key < sv;
            if (key == sv) lt = idx < spi[mid];   // indices read only on ties
            if (lt) hi = mid; else lo = mid + 1;
        }
        bid[sg.sbase + off + t] = (unsigned short)lo;
        atomicAdd(&hist[lo], 1);
    }
    __syncthreads();
    int* row = counts + q.cbase + (int64_t)(tile - sg.first_tile) * nb;
    for (int b = threadIdx.x; b < nb; b += ENM_SORT_THREADS) row[b] = hist[b];
}

// One block per segment: turn the (chunk, bucket) counts into segment
// positions (bucket-major, chunk-minor: the first position of chunk c's keys
// of bucket b), and pack consecutive buckets greedily into runs of at most
// ENM_SS_TILE keys (a larger bucket is a run of its own), appended to the
// segment's run queue (queue q's runs start at tiles[qbase[q]]).  A thread
// per bucket walks its column of the count table SCAN_BATCH chunks at a
// time, all loads of a batch issued before its additions and stores (a
// store between loads would keep the next load from issuing early: a
// 2^20-key segment has 512 chunks, one wait each).
constexpr int SCAN_THREADS = ENM_SS_MAXB;
constexpr int SCAN_BATCH = 16;
__global__ __launch_bounds__(SCAN_THREADS)
void ss_scan_kernel(const SortSeg* __restrict__ segs, const SampleSeg* __restrict__ ss,
                    int* __restrict__ counts, SsTile* __restrict__ tiles, const int* __restrict__ qbase,
                    unsigned* __restrict__ qctl, SsTile* __restrict__ long_runs, const int* __restrict__ hint_bad,
                    const int* __restrict__ skip) {
    __shared__ int tot[ENM_SS_MAXB];
    __shared__ int start[ENM_SS_MAXB];
    __shared__ int part[SCAN_THREADS / ENM_WAVE];
    if (skip != nullptr && *skip == 0) return;
    const SortSeg sg = segs[blockIdx.x];
    if (sg.hint && hint_bad != nullptr && hint_bad[blockIdx.x] == 0) return;   // no runs to sort
    const SampleSeg q = ss[blockIdx.x];
    const int nb = q.nb;
    const int nch = (sg.len + ENM_SORT_TILE - 1) / ENM_SORT_TILE;
    int* cnt = counts + q.cbase;
    const int b = threadIdx.x;   // nb <= ENM_SS_MAXB = SCAN_THREADS
    if (b < nb) {
        int sum = 0;
        int c = 0;
        for (; c + SCAN_BATCH <= nch; c += SCAN_BATCH) {
            int v[SCAN_BATCH];
#pragma unroll
            for (int u = 0; u < SCAN_BATCH; ++u) v[u] = cnt[(int64_t)(c + u) * nb + b];
#pragma unroll
            for (int u = 0; u < SCAN_BATCH; ++u) sum += v[u];
        }
        for (; c < nch; ++c) sum += cnt[(int64_t)c * nb + b];
        tot[b] = sum;
        start[b] = sum;
    }
    __syncthreads();
    block_exscan<SCAN_THREADS>(start, nb, part);
    if (b < nb) {
        int run = start[b];
        int c = 0;
        for (; c + SCAN_BATCH <= nch; c += SCAN_BATCH) {
            int v[SCAN_BATCH];
#pragma unroll
            for (int u = 0; u < SCAN_BATCH; ++u) v[u] = cnt[(int64_t)(c + u) * nb + b];
#pragma unroll
            for (int u = 0; u < SCAN_BATCH; ++u) {
                cnt[(int64_t)(c + u) * nb + b] = run;
                run += v[u];
            }
        }
        for (; c < nch; ++c) {
            const int v = cnt[(int64_t)c * nb + b];
            cnt[(int64_t)c * nb + b] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0) {
        // Runs of one bucket larger than an LDS run (sampling bad luck) go
        // to the long-run list for ss_long_kernel; the others to the queue.
        int nt = 0, cur = 0;
        for (int b = 0; b < nb; ++b) {
            const int z = tot[b];
            if (cur > 0 && cur + z > ENM_SS_TILE) { nt += cur <= ENM_SS_TILE; cur = 0; }
            cur += z;
        }
        if (cur > 0) nt += cur <= ENM_SS_TILE;
        SsTile* out = tiles + qbase[q.q] + atomicAdd(qctl + q.q * ENM_CSTRIDE, (unsigned)nt);
        int k = 0, t0 = 0;
        cur = 0;
        auto emit = [&](int len) {
            const SsTile tl{sg.sbase + t0, len, (int)blockIdx.x};
            if (len <= ENM_SS_TILE) out[k++] = tl;
            else long_runs[atomicAdd(qctl + 2 * ENM_NQ * ENM_CSTRIDE, 1u)] = tl;
        };
        for (int b = 0; b < nb; ++b) {
            const int z = tot[b];
            if (cur > 0 && cur + z > ENM_SS_TILE) {
                emit(cur);
                t0 += cur;
                cur = 0;
            }
            cur += z;
        }
        if (cur > 0) emit(cur);
    }
}

// One block per 2048-key chunk: order the chunk's keys by bucket in LDS (a
// counting sort; the order inside a bucket does not matter, as the runs are
// sorted whole later) and write each bucket's keys to their positions with
// coalesced stores.
__global__ __launch_bounds__(ENM_SORT_THREADS)
void ss_scatter_kernel(const SortSeg* __restrict__ segs, int nsegs, const SampleSeg* __restrict__ ss,
                       const double* __restrict__ arena, const unsigned short* __restrict__ bid,
                       const int* __restrict__ offs, double* __restrict__ kout, int* __restrict__ iout,
                       const int* __restrict__ hint_bad, const int* __restrict__ skip) {
    constexpr int PER = ENM_SORT_TILE / ENM_SORT_THREADS;
    __shared__ double sk[ENM_SORT_TILE];
    __shared__ int si[ENM_SORT_TILE];
    __shared__ unsigned short sb[ENM_SORT_TILE];
    __shared__ int hist[ENM_SS_MAXB];
    __shared__ int goff[ENM_SS_MAXB];
    __shared__ int part[ENM_SORT_THREADS / ENM_WAVE];
    if (skip != nullptr && *skip == 0) return;
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    if (sg.hint && hint_bad != nullptr && hint_bad[s] == 0) return;   // laid out from its hint
    const SampleSeg q = ss[s];
    const int nb = q.nb;
    const int off = (tile - sg.first_tile) * ENM_SORT_TILE;
    const int len = min(ENM_SORT_TILE, sg.len - off);
    const int* row = offs + q.cbase + (int64_t)(tile - sg.first_tile) * nb;
    for (int b = threadIdx.x; b < nb; b += ENM_SORT_THREADS) {
        hist[b] = 0;
        goff[b] = row[b];
    }
    __syncthreads();
    double key[PER];
    int bk[PER], rk[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int t = threadIdx.x + e * ENM_SORT_THREADS;
        if (t < len) {
            key[e] = arena[sg.src + off + t];
            bk[e] = bid[sg.sbase + off + t];
            rk[e] = atomicAdd(&hist[bk[e]], 1);
        }
    }
    __syncthreads();
    block_exscan<ENM_SORT_THREADS>(hist, nb, part);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int t = threadIdx.x + e * ENM_SORT_THREADS;
        if (t < len) {
            const int p = hist[bk[e]] + rk[e];
            sk[p] = key[e];
            si[p] = sg.p0 + off + t;
            sb[p] = (unsigned short)bk[e];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const int b = sb[t];
        const int64_t dst = sg.sbase + goff[b] + (t - hist[b]);
        kout[dst] = sk[t];
        iout[dst] = si[t];
    }
}

// Merge sort of one run longer than ENM_SS_TILE by the whole block: sort
// ENM_SS_TILE chunks in LDS from the scattered buffers (k1, i1) into the
// output buffers (k0, i0), then merge passes of doubling width between the
// two buffers' ranges of the run (which no other block touches), each
// ENM_SS_TILE outputs merged in LDS from a merge-path split; the result ends
// in (k0, i0).
__device__ void ss_sort_long(double* sk, int* si, int* sp, int64_t pos, int len,
                             double* k1, int* i1, double* k0, int* i0) {
    constexpr int T = ENM_SS_TILE, NT = ENM_SS_THREADS;
    for (int c0 = 0; c0 < len; c0 += T) {
        const int cl = min(T, len - c0);
        for (int t = threadIdx.x; t < T; t += NT) {
            if (t < cl) {
                sk[lp(t)] = k1[pos + c0 + t];
                si[lp(t)] = i1[pos + c0 + t];
            } else {
                sk[lp(t)] = INFINITY;
                si[lp(t)] = 0x7fffffff;
            }
        }
        block_sort_lds<T, NT>(sk, si);
        for (int t = threadIdx.x; t < cl; t += NT) {
            k0[pos + c0 + t] = sk[lp(t)];
            i0[pos + c0 + t] = si[lp(t)];
        }
        __syncthreads();
    }
    double* ks = k0;
    int* is = i0;
    double* kd = k1;
    int* id = i1;
    for (int w = T; w < len; w *= 2) {
        for (int pb = 0; pb < len; pb += 2 * w) {
            const int na = min(w, len - pb), nb = max(0, min(2 * w, len - pb) - w);
            const double* ka = ks + pos + pb;
            const int* ia = is + pos + pb;
            const double* kb = ka + w;
            const int* ib = ia + w;
            for (int d0 = 0; d0 < na + nb; d0 += T) {
                const int d1 = min(d0 + T, na + nb);
                if (threadIdx.x < 2) {
                    const int d = threadIdx.x ? d1 : d0;
                    sp[threadIdx.x] = nb == 0 ? d : co_rank(d, ka, ia, na, kb, ib, nb);
                }
                __syncthreads();
                const int a0 = sp[0], a1 = sp[1];
                const int b0 = d0 - a0;
                const int la = a1 - a0, lb = (d1 - a1) - b0;
                for (int t = threadIdx.x; t < la + lb; t += NT) {
                    if (t < la) { sk[lp(t)] = ka[a0 + t]; si[lp(t)] = ia[a0 + t]; }
                    else        { sk[lp(t)] = kb[b0 + t - la]; si[lp(t)] = ib[b0 + t - la]; }
                }
                __syncthreads();
                lds_merge<NT, T / NT>(sk, si, la, lb);
                for (int t = threadIdx.x; t < la + lb; t += NT) {
                    kd[pos + pb + d0 + t] = sk[lp(t)];
                    id[pos + pb + d0 + t] = si[lp(t)];
                }
                __syncthreads();
            }
        }
        double* tk = ks; ks = kd; kd = tk;
        int* ti = is; is = id; id = ti;
    }
    if (ks != k0) {
        for (int t = threadIdx.x; t < len; t += NT) {
            k0[pos + t] = ks[pos + t];
            i0[pos + t] = is[pos + t];
        }
        __syncthreads();
    }
}

// Lay a sorted run out into every order that consumes its segment: the
// point index, the key column and the order's other columns (read from the
// arena at the point, FP32 copies when with32), and fold the magnitudes into
// the instance's maximum (for the FP32 guard band).  Sorted element t of the
// run is at LDS slot lp(t) (LDS) or at kg[t], ig[t] (global).
struct SsLayout {
    const SortSeg* segs;
    const SampleSeg* ss;
    const SsCons* cons;
    const OrderDev* orders;
    const double* arena;
    double* col64;
    float* col32;
    int* perm;
    unsigned long long* maxabs;
    int with32;
};

// Lays out elements [0, len) of a sorted run into the order `od` from
// position p0 on: element t has key key_at(t) and point index idx_at(t),
// and the order's other columns are gathered from the arena.  A thread
// takes elements threadIdx.x + NT * k, PER at a time, and issues all its
// gathers of a column before it stores any of them (a store between two
// gathers would hold the second one back: the gathers may alias the
// stores as far as the compiler knows).  Returns the largest |value|.
template <int NT, int PER, class K, class I>
__device__ __forceinline__ double layout_elems(const SsLayout& L, const OrderDev& od, int64_t p0, int len,
                                               K key_at, I idx_at) {
    const int64_t n = od.n;
    const int ncol = od.ncol;
    double m = 0.0;
    for (int base = 0; base < len; base += NT * PER) {
        int idx[PER];
        double key[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = base + k * NT + (int)threadIdx.x;
            const bool in = t < len;
            idx[k] = in ? idx_at(t) : 0;
            key[k] = in ? key_at(t) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = base + k * NT + (int)threadIdx.x;
            if (t < len) {
                const int64_t p = p0 + t;
                L.perm[od.perm + p] = idx[k];
                L.col64[od.c64 + p] = key[k];
                if (L.with32) L.col32[od.c32 + p] = (float)key[k];
                m = fmax(m, fabs(key[k]));
            }
        }
        for (int cc = 1; cc < ncol; ++cc) {
            const double* src = L.arena + od.src[cc];
            double v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int t = base + k * NT + (int)threadIdx.x;
                v[k] = t < len ? src[idx[k]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int t = base + k * NT + (int)threadIdx.x;
                if (t < len) {
                    const int64_t p = p0 + t;
                    L.col64[od.c64 + cc * n + p] = v[k];
                    if (L.with32) L.col32[od.c32 + cc * n + p] = (float)v[k];
                    m = fmax(m, fabs(v[k]));
                }
            }
        }
    }
    return m;
}

template <bool LDS>
__device__ void ss_layout_run(const SsLayout& L, const SsTile& tl, const double* kk, const int* ii,
                              double* red) {
    const SortSeg& sg = L.segs[tl.seg];
    const SampleSeg& q = L.ss[tl.seg];
    const int64_t r0 = tl.pos - sg.sbase;
    for (int c = 0; c < q.ncons; ++c) {
        const SsCons cn = L.cons[q.cons0 + c];
        // Fields read one by one: a local copy of the descriptor, indexed by
        // a loop counter, would live in scratch memory.
        const OrderDev& od = L.orders[cn.order];
        const int64_t n = od.n;
        const int ncol = od.ncol;
        double m = 0.0;
        for (int t = threadIdx.x; t < tl.len; t += ENM_SS_THREADS) {
            const double key = LDS ? kk[lp(t)] : kk[t];
            const int idx = LDS ? ii[lp(t)] : ii[t];
            const int64_t p = cn.pos0 + r0 + t;
            L.perm[od.perm + p] = idx;
            L.col64[od.c64 + p] = key;
            if (L.with32) L.col32[od.c32 + p] = (float)key;
            m = fmax(m, fabs(key));
            for (int cc = 1; cc < ncol; ++cc) {
                const double v = L.arena[od.src[cc] + idx];
                L.col64[od.c64 + cc * n + p] = v;
                if (L.with32) L.col32[od.c32 + cc * n + p] = (float)v;
                m = fmax(m, fabs(v));
            }
        }
        m = wave_max(m);
        if ((threadIdx.x & (ENM_WAVE - 1)) == 0) red[threadIdx.x / ENM_WAVE] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mm = red[0];
            for (int w = 1; w < ENM_SS_THREADS / ENM_WAVE; ++w) mm = fmax(mm, red[w]);
            // |v| >= 0, so the IEEE bit patterns order like the values.
            atomicMax(L.maxabs + od.inst, (unsigned long long)__double_as_longlong(mm));
        }
        __syncthreads();
    }
}

// One block per ENM_SORT_TILE chunk of a hinted first-stage segment: check
// the hint's order over the chunk (each element against the one before it,
// so the chunks together cover the whole segment) and lay the chunk out
// into every order of the segment, as ss_tile_kernel would.  The chunk's
// indices and keys (and the element before it) are gathered once into LDS.
// A failed check marks the segment for the sample sort.
__global__ __launch_bounds__(ENM_SORT_THREADS)
void ss_hint_kernel(const SortSeg* __restrict__ segs, int nsegs, const int* __restrict__ hint,
                    int* __restrict__ hint_bad, SsLayout L) {
    __shared__ double s_key[ENM_SORT_TILE + 1];
    __shared__ int s_idx[ENM_SORT_TILE + 1];
    __shared__ int s_bad;
    __shared__ double red[ENM_SORT_THREADS / ENM_WAVE];
    // A segment's chunks on one XCD: its columns' gathers hit one L2.
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    if (!sg.hint) return;
    const SampleSeg q = L.ss[s];
    const int off = (tile - sg.first_tile) * ENM_SORT_TILE;
    const int len = min(ENM_SORT_TILE, sg.len - off);
    const int* h = hint + sg.src;
    if (threadIdx.x == 0) s_bad = h[0] < 0 ? 1 : 0;
    // Slot t holds the element at position off - 1 + t (-1: none or out of
    // range).  All of a thread's hint loads go out first, then all its key
    // gathers.
    {
        constexpr int PER = ENM_SORT_TILE / ENM_SORT_THREADS + 1;
        int hi[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = k * ENM_SORT_THREADS + (int)threadIdx.x;
            const int p = off - 1 + t;
            hi[k] = (t <= len && p >= 0) ? h[p] : -1;
        }
        double hk[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const bool in = hi[k] >= 0 && hi[k] < sg.len;
            hk[k] = in ? L.arena[sg.src + hi[k]] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int t = k * ENM_SORT_THREADS + (int)threadIdx.x;
            if (t <= len) {
                const bool in = hi[k] >= 0 && hi[k] < sg.len;
                s_idx[t] = in ? hi[k] : -1;
                s_key[t] = hk[k];
            }
        }
    }
    __syncthreads();
    bool bad = false;
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const int i = s_idx[t + 1], j = s_idx[t];
        const double v = s_key[t + 1], u = s_key[t];
        // NaN keys fail both comparisons.
        const bool ok = i >= 0 && (off + t == 0 || (j >= 0 && (u < v || (u == v && j < i))));
        bad = bad || !ok;
    }
    if (wave_any(bad) && (threadIdx.x & (ENM_WAVE - 1)) == 0) s_bad = 1;
    __syncthreads();
    if (s_bad) {
        if (threadIdx.x == 0) {
            hint_bad[s] = 1;
            hint_bad[nsegs] = 1;   // some segment is sorted this run
        }
        return;
    }
    for (int c = 0; c < q.ncons; ++c) {
        const SsCons cn = L.cons[q.cons0 + c];
        const OrderDev& od = L.orders[cn.order];
        const int p0 = sg.p0;
        double m = layout_elems<ENM_SORT_THREADS, ENM_SORT_TILE / ENM_SORT_THREADS>(
            L, od, cn.pos0 + off, len, [&](int t) { return s_key[t + 1]; },
            [&](int t) { return p0 + s_idx[t + 1]; });
        // One atomic per block: atomics from every wave on one address
        // queue behind each other.
        m = wave_max(m);
        if ((threadIdx.x & (ENM_WAVE - 1)) == 0) red[threadIdx.x / ENM_WAVE] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mm = red[0];
            for (int w = 1; w < ENM_SORT_THREADS / ENM_WAVE; ++w) mm = fmax(mm, red[w]);
            atomicMax(L.maxabs + od.inst, (unsigned long long)__double_as_longlong(mm));
        }
        __syncthreads();
    }
}

// Persistent blocks sort the runs in LDS, from the scattered buffers (k1,
// i1), and lay them out into the orders.  Runs sit in ENM_NQ queues by
// instance; block b starts on queue b mod ENM_NQ (its XCD, so one
// instance's random column reads stay in one L2) and takes other queues'
// runs when its own is empty.  Three blocks fit a CU's LDS; registers are
// held to six waves per SIMD to match.
__global__ __launch_bounds__(ENM_SS_THREADS, 6)
void ss_tile_kernel(const SsTile* __restrict__ tiles, const int* __restrict__ qbase, unsigned* qctl,
                    double* k1, int* i1, double* k0, int* i0, SsLayout L, const int* __restrict__ skip) {
    __shared__ double sk[ENM_LDS_PAD(ENM_SS_TILE)];
    __shared__ int si[ENM_LDS_PAD(ENM_SS_TILE)];
    __shared__ unsigned s_ticket;
    __shared__ double red[ENM_SS_THREADS / ENM_WAVE];
    if (skip != nullptr && *skip == 0) return;   // the queues were neither cleared nor filled
    const int home = (int)(blockIdx.x % ENM_NQ);
    for (int r = 0; r < ENM_NQ; ++r) {
        const int qq = (home + r) % ENM_NQ;
        // Appends finished with the previous kernel: the count is final.
        const unsigned qn = qctl[qq * ENM_CSTRIDE];
        for (;;) {
            if (threadIdx.x == 0) s_ticket = atomicAdd(qctl + (ENM_NQ + qq) * ENM_CSTRIDE, 1u);
            __syncthreads();
            const unsigned t = (unsigned)__builtin_amdgcn_readfirstlane((int)s_ticket);
            __syncthreads();
            if (t >= qn) break;
            const SsTile tl = tiles[qbase[qq] + (int)t];
            for (int u = threadIdx.x; u < ENM_SS_TILE; u += ENM_SS_THREADS) {
                if (u < tl.len) {
                    sk[lp(u)] = k1[tl.pos + u];
                    si[lp(u)] = i1[tl.pos + u];
                } else {
                    sk[lp(u)] = INFINITY;
                    si[lp(u)] = 0x7fffffff;
                }
            }
            block_sort_lds<ENM_SS_TILE, ENM_SS_THREADS>(sk, si);
            ss_layout_run<true>(L, tl, sk, si, red);
        }
    }
}

// Runs of one bucket larger than an LDS run (rare: sampling bad luck),
// listed by ss_scan_kernel: each merge-sorted by one block through global
// memory (ss_sort_long) and laid out into the orders.  Kept out of
// ss_tile_kernel, whose registers then hold its LDS runs without spilling.
__global__ __launch_bounds__(ENM_SS_THREADS)
void ss_long_kernel(const SsTile* __restrict__ runs, const unsigned* __restrict__ qctl, double* k1, int* i1,
                    double* k0, int* i0, SsLayout L, const int* __restrict__ skip) {
    __shared__ double sk[ENM_LDS_PAD(ENM_SS_TILE)];
    __shared__ int si[ENM_LDS_PAD(ENM_SS_TILE)];
    __shared__ int sp[2];
    __shared__ double red[ENM_SS_THREADS / ENM_WAVE];
    if (skip != nullptr && *skip == 0) return;
    const unsigned n = qctl[2 * ENM_NQ * ENM_CSTRIDE];   // final: ss_scan_kernel has finished
    for (unsigned t = blockIdx.x; t < n; t += gridDim.x) {
        const SsTile tl = runs[t];
        ss_sort_long(sk, si, sp, tl.pos, tl.len, k1, i1, k0, i0);
        ss_layout_run<false>(L, tl, k0 + tl.pos, i0 + tl.pos, red);
    }
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

hipError_t enm_launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int threads = 256;
    const int64_t blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(threads), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t enm_launch_sort(const SortSeg* d_segs, int nsegs, int ntiles, int64_t maxlen, int tile,
                           const double* arena, const int* iarena, double* k0, int* i0, double* k1, int* i1,
                           int* split, int* final_buf, int* n_launch, hipStream_t st, const int* skip) {
    *final_buf = 0;
    if (ntiles <= 0) return hipSuccess;
    // Segments no longer than the tile are sorted by the block sort alone;
    // longer ones (always with 2048-key tiles) continue with merge passes.
    if (tile == 512)
        hipLaunchKernelGGL(block_sort_kernel<512>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, iarena, k0, i0, skip);
    else if (tile == 1024)
        hipLaunchKernelGGL(block_sort_kernel<1024>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, iarena, k0, i0, skip);
    else if (tile == ENM_SORT_TILE)
        hipLaunchKernelGGL(block_sort_kernel<ENM_SORT_TILE>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, iarena, k0, i0, skip);
    else
        return hipErrorInvalidValue;
    if (tile != ENM_SORT_TILE && maxlen > tile) return hipErrorInvalidValue;
    ++*n_launch;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int cur = 0;
    for (int64_t w = ENM_SORT_TILE; w < maxlen; w *= 2) {
        const double* kin = cur ? k1 : k0;
        const int* iin = cur ? i1 : i0;
        double* kout = cur ? k0 : k1;
        int* iout = cur ? i0 : i1;
        hipLaunchKernelGGL(co_rank_kernel, dim3((ntiles + ENM_SORT_THREADS - 1) / ENM_SORT_THREADS),
                           dim3(ENM_SORT_THREADS), 0, st, d_segs, nsegs, ntiles, w, kin, iin, split, skip);
        hipLaunchKernelGGL(merge_pass_kernel, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, w, kin, iin, split, kout, iout, skip);
        *n_launch += 2;
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur ^= 1;
    }
    *final_buf = cur;
    return hipSuccess;
}

hipError_t enm_launch_sample_sort(const SampleSortArgs& a, const double* arena, double* k0, int* i0,
                                  double* k1, int* i1, int* n_launch, hipStream_t st) {
    if (a.nsegs <= 0) return hipSuccess;
    if (a.qctl == nullptr || a.qbase == nullptr) return hipErrorInvalidValue;
    SsLayout L;
    L.segs = a.segs;
    L.ss = a.ss;
    L.cons = a.cons;
    L.orders = a.orders;
    L.arena = arena;
    L.col64 = a.col64;
    L.col32 = a.col32;
    L.perm = a.perm;
    L.maxabs = a.maxabs;
    L.with32 = a.with32;
    // Hinted segments first: when every segment is hinted, the rest of the
    // sample sort runs only if some hint failed (hint_bad[nsegs], set by
    // ss_hint_kernel), and otherwise every kernel returns at once.
    const int* hint_bad = nullptr;
    const int* skip = nullptr;
    if (a.hint != nullptr) {
        if (a.hint_bad == nullptr) return hipErrorInvalidValue;
        hipLaunchKernelGGL(ss_hint_kernel, dim3(a.ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           a.segs, a.nsegs, a.hint, a.hint_bad, L);
        ++*n_launch;
        hint_bad = a.hint_bad;
        if (a.all_hinted) skip = a.hint_bad + a.nsegs;
    }
    // Samples go to the second sample buffers, sorted into the first (or
    // back, as the merge passes end).
    hipLaunchKernelGGL(ss_sample_kernel, dim3(a.nsegs), dim3(ENM_SORT_THREADS), 0, st,
                       a.segs, a.ss, arena, a.samp_k1, a.samp_i1, a.qctl, skip);
    ++*n_launch;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int fb = 0;
    if (a.n_samp_segs > 0) {
        e = enm_launch_sort(a.samp_segs, a.n_samp_segs, a.samp_tiles, a.samp_maxlen, ENM_SORT_TILE,
                            a.samp_k1, a.samp_i1, a.samp_k0, a.samp_i0, a.samp_k1, a.samp_i1, a.samp_split,
                            &fb, n_launch, st, skip);
        if (e != hipSuccess) return e;
    }
    const double* spk = fb ? a.samp_k1 : a.samp_k0;
    const int* spi = fb ? a.samp_i1 : a.samp_i0;
    hipLaunchKernelGGL(ss_count_kernel, dim3(a.ntiles), dim3(ENM_SORT_THREADS), 0, st,
                       a.segs, a.nsegs, a.ss, arena, spk, spi, a.bid, a.counts, hint_bad, skip);
    hipLaunchKernelGGL(ss_scan_kernel, dim3(a.nsegs), dim3(SCAN_THREADS), 0, st,
                       a.segs, a.ss, a.counts, a.tiles, a.qbase, a.qctl, a.long_runs, hint_bad, skip);
    hipLaunchKernelGGL(ss_scatter_kernel, dim3(a.ntiles), dim3(ENM_SORT_THREADS), 0, st,
                       a.segs, a.nsegs, a.ss, arena, a.bid, a.counts, k1, i1, hint_bad, skip);
    hipLaunchKernelGGL(ss_tile_kernel, dim3(a.tile_grid), dim3(ENM_SS_THREADS), 0, st,
                       a.tiles, a.qbase, a.qctl, k1, i1, k0, i0, L, skip);
    hipLaunchKernelGGL(ss_long_kernel, dim3(a.long_grid), dim3(ENM_SS_THREADS), 0, st,
                       a.long_runs, a.qctl, k1, i1, k0, i0, L, skip);
    ++*n_launch;
    *n_launch += 4;
    return hipGetLastError();
}

hipError_t enm_launch_strip_sort(const GatherItem* d_items, int nitems, const SortSeg* d_segs,
                                 const OrderDev* d_orders, int tile, double* col64, float* col32, int* perm,
                                 int with32, int* n_launch, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    if (tile == 512)
        hipLaunchKernelGGL(strip_sort_kernel<512>, dim3(nitems), dim3(ENM_SORT_THREADS), 0, st,
                           d_items, d_segs, d_orders, col64, col32, perm, with32);
    else if (tile == 1024)
        hipLaunchKernelGGL(strip_sort_kernel<1024>, dim3(nitems), dim3(ENM_SORT_THREADS), 0, st,
                           d_items, d_segs, d_orders, col64, col32, perm, with32);
    else if (tile == ENM_SORT_TILE)
        hipLaunchKernelGGL(strip_sort_kernel<ENM_SORT_TILE>, dim3(nitems), dim3(ENM_SORT_THREADS), 0, st,
                           d_items, d_segs, d_orders, col64, col32, perm, with32);
    else
        return hipErrorInvalidValue;
    ++*n_launch;
    return hipGetLastError();
}

hipError_t enm_launch_gather(const GatherItem* d_items, int nitems, const SortSeg* d_segs,
                             const OrderDev* d_orders, const double* arena,
                             const double* skey, const int* sidx,
                             double* col64, float* col32, int* perm,
                             unsigned long long* maxabs, int with32, int* n_launch, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(nitems), dim3(ENM_SORT_THREADS), 0, st,
                       d_items, d_segs, d_orders, arena, skey, sidx, col64, col32, perm, maxabs, with32);
    ++*n_launch;
    return hipGetLastError();
}
