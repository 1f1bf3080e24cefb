// Chebyshev (max-norm) k-th-neighbour radius, replacing
// cKDTree.query(pts, k=[k+1], p=np.inf) of the reference
// (ennemi/_entropy_estimators.py:39,108,142,194,240).
//
// eps_i is the (k+1)-th smallest D(i,j) = max_d fl64(|a_d[i] - a_d[j]|) over
// all j of the query's scope range, i itself included (distance 0); fewer
// than k+1 points leave it +inf, as cKDTree does.
//
// Grid orders: the key-sorted points cut into strips, each strip sorted by
// the second column.  A work item is one wave and 64 consecutive positions
// of one strip (one query per lane).  With two columns each lane walks the
// strips around its own point by itself (knn_walk): its own strip from its
// own position, then the strips outward that its current radius still
// reaches on the key axis, each from the lane's y position in both
// directions until the y distance reaches the radius.  Candidates arrive in
// about increasing distance, so each lane evaluates about ten candidates.
// Each lane keeps an ascending FP64 list of its K1 = k+1 smallest distances
// in registers (P slots, the lowest P-K1 pinned at -inf).
//
// Grid orders of three or more columns, where the two gridded columns prune
// less and a lane's walk would be long, share candidates instead: a wave
// takes a tile of 64 consecutive positions of a strip and walks candidates
// in wave-uniform order (grid_sweep: its own strip, then the strips outward,
// each within the union of the lanes' y intervals).  Other orders (one
// column, or ENM_F_BRUTE) do the same over the key-sorted order, 8
// candidates at a time through the scalar cache (candidate coordinates are
// SGPR operands); with one column, from the tile outward in key order, alternating
// left and right in steps of 64, a side stopping once every lane's key
// distance to the last visited candidate is >= its current K1-th distance T.
// With ENM_F_BRUTE the whole scope range is scanned.
//
// Precision.  Every distance that enters a list is exact FP64.  The FP32
// filter (brute force in FP32-filter mode) computes the bulk distance from
// FP32 copies, and a chunk of 8 candidates is settled in FP64 only when some
// lane's FP32 distance is <= fl32_up(T + g), with the per-instance guard g
// bounding |d32 - d64|.  T only decreases, so one pass is exact: radii are
// bit-identical to the FP64 mode.  While most chunks improve some lane's list
// the wave skips the filter and evaluates chunks in FP64 directly.  List
// updates for up to 8 slots are branch-free: the 8 exact distances are
// sorted by a 19-comparator network and merged into the list.
//
// Work distribution.  A persistent grid pops items from ENM_NQ queues, one
// per XCD (for_each_wave_item), in the order the host laid them out.
#pragma once
#include "enm_device.h"
#include "enm_launch.h"

#include <algorithm>


// Arguments of every kNN kernel (shared by the per-dimension translation
// units, which instantiate the kernels of one joint dimension each).
struct KnnArgs {
    const Item* items;
    const int* qoff;    // ENM_NQ + 1 queue offsets into items
    unsigned* counter;  // ENM_NQ ticket counters
    int chunk;          // items per ticket
    const KnnDev* descs;
    const OrderDev* orders;
    const unsigned long long* maxabs;
    const int* perm;
    int brute;
    double* heap;       // global heap lists (list class ENM_LIST_HEAP_GLOBAL): per block
    int64_t heap_stride;  // doubles per block in `heap` (64 lanes x the group's largest k+1)
    int smax;           // strip kernel: the group's largest strip (LDS slot size, points)
    // Strip kernel: queries whose radius still reaches past the staged
    // neighbours are listed here (KnnFar records, as many as far_bytes
    // hold; the others are finished in place) and finished by
    // knn_far_kernel; NULL: all finished in place.
    char* far;
    unsigned* far_count;  // zeroed before the launch
    int64_t far_bytes;
};

// A strip-kNN query whose radius reaches past its staged neighbours: its
// item, grid-order position, the sides still open (bit 0 left, bit 1
// right) and its list so far; knn_far_kernel continues its walk.
template <int P>
struct KnnFar {
    int32_t item;
    int32_t i;
    int32_t open;
    int32_t pad;
    double L[P];
};

// Launch of the kNN kernels of one joint dimension D (one translation unit
// per dimension class, so the kernels compile in parallel).
#define ENM_KNN_DIMS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(12) X(16)
#define ENM_KNN_DECLARE(D) \
    hipError_t enm_knn_launch_d##D(const KnnArgs& a, int P, bool f32, int walk, int grid, \
                                   const double* col64, const float* col32, double* res64, \
                                   unsigned long long* stats, hipStream_t st);
ENM_KNN_DIMS(ENM_KNN_DECLARE)

#ifdef ENM_KNN_KERNELS
namespace {

constexpr int CH = 8;     // candidates per filter chunk
constexpr int SWEEP = 64; // candidates per side per sweep step



// The K1 = k+1 smallest distances of a lane.  Register lists (P = 4, 8,
// 16, 32 slots) keep them sorted ascending, the lowest P-K1 slots pinned at
// -inf, so the last slot is the K1-th smallest.  For K1 above 32 a max-heap
// of exactly K1 values per lane (initially +inf) keeps the same multiset:
// a value below the root replaces it and sinks, so the root is the K1-th
// smallest.  The heap lives in LDS (P = ENM_LIST_HEAP_LDS, up to
// ENM_LIST_LDS_MAX values per lane) or in a global buffer
// (P = ENM_LIST_HEAP_GLOBAL), lane-interleaved: slot t of a lane at
// h[t * 64].  Either way top() is the current radius.
template <int P>
struct RegList {
    double L[P];
    __device__ __forceinline__ void init(int K1, double*) {
#pragma unroll
        for (int t = 0; t < P; ++t) L[t] = (t < P - K1) ? -INFINITY : INFINITY;
    }
    __device__ __forceinline__ void insert(double v) { list_insert<P>(L, v); }
    __device__ __forceinline__ double top() const { return L[P - 1]; }
};

struct HeapList {
    double* h;
    int K1;
    double root;
    __device__ __forceinline__ void init(int k1, double* base) {
        h = base;
        K1 = k1;
        for (int t = 0; t < K1; ++t) h[t * ENM_WAVE] = INFINITY;
        root = INFINITY;
    }
    __device__ __forceinline__ void insert(double v) {
        if (!(v < root)) return;
        int i = 0;
        for (;;) {
            int c = 2 * i + 1;
            if (c >= K1) break;
            double cv = h[c * ENM_WAVE];
            if (c + 1 < K1) {
                const double c2 = h[(c + 1) * ENM_WAVE];
                if (c2 > cv) { ++c; cv = c2; }
            }
            if (!(cv > v)) break;
            h[i * ENM_WAVE] = cv;
            i = c;
        }
        h[i * ENM_WAVE] = v;
        root = h[0];
    }
    __device__ __forceinline__ double top() const { return root; }
};

template <int P> struct ListOf { using type = RegList<P>; };
template <> struct ListOf<ENM_LIST_HEAP_LDS> { using type = HeapList; };
template <> struct ListOf<ENM_LIST_HEAP_GLOBAL> { using type = HeapList; };

// Dynamic LDS of the heap kernels (64 lanes x K1 doubles).
extern __shared__ double enm_heap_lds[];

// This lane's heap storage (list classes with a heap), else nullptr.
template <int P>
__device__ __forceinline__ double* heap_base(const KnnArgs& A) {
    if constexpr (P == ENM_LIST_HEAP_LDS) return enm_heap_lds + threadIdx.x;
    else if constexpr (P == ENM_LIST_HEAP_GLOBAL) return A.heap + blockIdx.x * A.heap_stride + threadIdx.x;
    else return nullptr;
}

template <int D, int P, bool F32>
struct KnnTile {
    const double* c64[D];
    const float* c32[D];
    double q64[D];
    float q32[D];
    typename ListOf<P>::type L;
    double T;
    float thr;
    double guard;
    unsigned rechecks;
    bool dense;   // FP32 mode: evaluating chunks directly in FP64
    double* stage;   // LDS, 64 candidates x D doubles (three or more columns)

    __device__ __forceinline__ void set_thr() {
        if (F32) thr = f32_up(T + guard);
    }

    __device__ __forceinline__ void recheck(int j) {
        double e = fabs(c64[0][j] - q64[0]);
#pragma unroll
        for (int d = 1; d < D; ++d) e = fmax(e, fabs(c64[d][j] - q64[d]));
        ++rechecks;
        if (e < T) {
            L.insert(e);
            T = L.top();
            set_thr();
        }
    }

    // Candidate coordinates of chunk j (wave-uniform: scalar loads).
    __device__ __forceinline__ void load64(double (&cx)[D][CH], int j) const {
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int c = 0; c < CH; ++c) cx[d][c] = c64[d][j + c];
    }

    __device__ __forceinline__ bool chunk64(int j) {
        double cx[D][CH];
        load64(cx, j);
        return chunk64v(cx);
    }

    // FP64 chunk: exact distances; merge when some lane improves.  Returns
    // whether any lane could improve its list.
    __device__ __forceinline__ bool chunk64v(const double (&cx)[D][CH]) {
        double dd[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            double m = fabs(cx[0][c] - q64[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmax(m, fabs(cx[d][c] - q64[d]));
            dd[c] = m;
        }
        return take(dd);
    }

    // List update with the exact distances of one chunk.
    __device__ __forceinline__ bool take(double (&dd)[CH]) {
        double mn = dd[0];
#pragma unroll
        for (int c = 1; c < CH; ++c) mn = fmin(mn, dd[c]);
        if (!wave_any(mn < T)) return false;
        bool merged = false;
        if constexpr (P > 0 && P <= 8) {
            // After a chunk that improved some list (dense regime) most of
            // the 8 candidates matter to some lane: one network merge.
            if (dense) {
                sort8(dd);
                list_merge8<P>(L.L, dd);
                T = L.top();
                merged = true;
            }
        }
        if (!merged) {
            // Sparse regime: only the few candidates some lane needs.
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (wave_any(dd[c] < T) && dd[c] < T) {
                    L.insert(dd[c]);
                    T = L.top();
                }
            }
        }
        if (F32) set_thr();
        return true;
    }

    // FP32-filter chunk: FP32 distances; if some lane's minimum passes the
    // guard-widened threshold, the chunk is settled exactly in FP64.
    // Returns whether that happened.
    __device__ __forceinline__ bool chunk32(int j) {
        float dd[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            float m = fabsf(c32[0][j + c] - q32[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmaxf(m, fabsf(c32[d][j + c] - q32[d]));
            dd[c] = m;
        }
        float mn = dd[0];
#pragma unroll
        for (int c = 1; c < CH; ++c) mn = fminf(mn, dd[c]);
        if (!wave_any(mn <= thr)) return false;
#pragma unroll
        for (int c = 0; c < CH; ++c) rechecks += dd[c] <= thr;
        // A candidate the filter rejected has an FP64 distance > T, so
        // settling the whole chunk in FP64 gives the same list.
        (void)chunk64(j);
        return true;
    }

    // One chunk of 8 candidates.  The wave adapts to how often chunks
    // improve some lane's list: after such a chunk ("dense") it evaluates the
    // next in FP64 directly and updates lists with the merge network; after a
    // chunk that improved none it filters in FP32 (FP32-filter mode) and
    // inserts the few passing candidates one by one.
    __device__ __forceinline__ void chunk(int j) {
        if (F32 && !dense) dense = chunk32(j);
        else dense = chunk64(j);
    }

    __device__ __forceinline__ void single(int j) {
        if (F32) {
            float m = fabsf(c32[0][j] - q32[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmaxf(m, fabsf(c32[d][j] - q32[d]));
            if (wave_any(m <= thr) && m <= thr) recheck(j);
        } else {
            double m = fabs(c64[0][j] - q64[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j] - q64[d]));
            if (m < T) {
                L.insert(m);
                T = L.top();
            }
        }
    }

    // Process the wave-uniform candidate range [a, b).  In FP64 mode with
    // up to two columns the next chunk's coordinates are loaded before the
    // current chunk is evaluated (software pipelining: the scalar loads of a
    // chunk often miss the scalar cache, and their L2 latency then overlaps
    // a chunk of arithmetic instead of stalling the wave); both buffers fit
    // in SGPRs (2 x 2 x 8 doubles).
    // Three or more columns: candidates come through LDS.  The wave loads
    // 64 candidates at a time, one per lane (coalesced per column, the next
    // 64 in flight while these are evaluated), writes them to `stage`, and
    // every lane reads them back with broadcast LDS reads.  (Scalar loads
    // of D columns would not fit two chunks in SGPRs, so each chunk would
    // wait for its loads.)  Distances are exact FP64 in both modes.
    __device__ __forceinline__ void range_staged(int a, int b) {
        const int lane = (int)(threadIdx.x & (ENM_WAVE - 1));
        double nx[D];
        auto fetch = [&](int j0) {
            const int p = j0 + lane;
            const bool ok = p < b;
#pragma unroll
            for (int d = 0; d < D; ++d) nx[d] = ok ? c64[d][p] : 0.0;
        };
        fetch(a);
        for (int j = a; j < b; j += ENM_WAVE) {
            __syncthreads();   // every lane has read the previous block
#pragma unroll
            for (int d = 0; d < D; ++d) stage[d * ENM_WAVE + lane] = nx[d];
            __syncthreads();
            const int m = min(ENM_WAVE, b - j);
            if (j + ENM_WAVE < b) fetch(j + ENM_WAVE);
            int c = 0;
            for (; c + CH <= m; c += CH) {
                double dd[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    double e = fabs(stage[c + u] - q64[0]);
#pragma unroll
                    for (int d = 1; d < D; ++d) e = fmax(e, fabs(stage[d * ENM_WAVE + c + u] - q64[d]));
                    dd[u] = e;
                }
                dense = take(dd);
            }
            for (; c < m; ++c) {
                double e = fabs(stage[c] - q64[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) e = fmax(e, fabs(stage[d * ENM_WAVE + c] - q64[d]));
                if (e < T) {
                    L.insert(e);
                    T = L.top();
                }
            }
        }
        if (F32) set_thr();
    }

    __device__ __forceinline__ void range(int a, int b) {
        if constexpr (D >= 3) {
            if (b > a) range_staged(a, b);
            return;
        }
        int j = a;
        if constexpr (!F32 && D <= 2) {
            // FP64 with up to two columns: the next chunk's coordinates are
            // loaded before the current chunk is evaluated (software
            // pipelining of the scalar loads; both buffers fit in SGPRs).
            if (j + CH <= b) {
                double cur[D][CH];
                load64(cur, j);
                for (; j + CH <= b; j += CH) {
                    double nxt[D][CH];
                    // The last chunk reloads itself (in bounds, unused).
                    load64(nxt, j + 2 * CH <= b ? j + CH : j);
                    dense = chunk64v(cur);
#pragma unroll
                    for (int d = 0; d < D; ++d)
#pragma unroll
                        for (int c = 0; c < CH; ++c) cur[d][c] = nxt[d][c];
                }
            }
        } else {
            for (; j + CH <= b; j += CH) chunk(j);
        }
        for (; j < b; ++j) single(j);
    }
};

// First position in [ks, ke) whose key is >= y, one search per lane (the
// trip count depends only on ke - ks, up to one step).
__device__ __forceinline__ int lane_lower_bound(const double* __restrict__ v, int ks, int ke, double y) {
    int first = ks, len = ke - ks;
    while (len > 0) {
        const int half = len >> 1;
        const bool go = v[first + half] < y;
        first = go ? first + half + 1 : first;
        len = go ? len - half - 1 : half;
    }
    return first;
}

// Every live lane's lower bound lane_lower_bound(v, ks, ke, y), found by
// the whole wave: the bounds [a, b) of the live lanes' y range in the strip
// (wave_bounds, two rounds of independent probes for a 1024-point strip),
// then, when that window holds at most 64 keys (the usual case: a wave's
// 64 queries are consecutive in y, and so are their entry points), each
// lane counts the window keys below its y by a binary search over the
// window held one key per lane (shuffles, no memory round trips).  Wider
// windows fall back to a per-lane search inside the window.  The result is
// the exact lower bound of every live lane; other lanes get ks.
__device__ __forceinline__ int wave_lower_bounds(const double* __restrict__ v, int ks, int ke, double y, bool live) {
    const double lo = uni_f64(wave_min(live ? y : INFINITY));
    const double hi = uni_f64(wave_max(live ? y : -INFINITY));
    if (!(lo <= hi)) return ks;
    int a, b;
    wave_bounds(v, ks, ke, lo, hi, a, b);
    const int L = b - a;
    const int lane = (int)(threadIdx.x & (ENM_WAVE - 1));
    if (L <= ENM_WAVE) {
        const double key = lane < L ? v[a + lane] : INFINITY;
        int pos = 0;
#pragma unroll
        for (int step = ENM_WAVE; step > 0; step >>= 1) {
            const int idx = pos + step - 1;
            const double kv = __shfl(key, idx < ENM_WAVE ? idx : ENM_WAVE - 1, ENM_WAVE);
            if (idx < L && kv < y) pos += step;
        }
        return live ? a + pos : ks;
    }
    return live ? lane_lower_bound(v, a, b, y) : ks;
}

// Walk of one strip [ks, ke) (sorted by column 1) for every live lane, from
// its own start: upward from u and downward from dn, one candidate per
// direction and iteration, so candidates come in about increasing y
// distance.  A direction stops at the first candidate whose y distance is
// >= the lane's current radius T (the strip is sorted, so every later one
// is further still, and T only decreases); the others are evaluated and
// offered to the list (an insertion at or above T leaves it unchanged).
// Candidates are (x, y) pairs of the interleaved copy, one 16-byte load
// each (the lanes of a wave walk different positions, so every load
// instruction touches up to 64 cache lines: one per candidate, not two);
// the next candidate is loaded before the current one is evaluated.  Adds
// the lane's evaluated candidates to `visited`.
template <int P, bool F32>
__device__ __forceinline__ void walk_strip(KnnTile<2, P, F32>& w, const double2* __restrict__ xy, int ks, int ke,
                                           int u, int dn, bool live, unsigned& visited) {
    const double qx = w.q64[0], qy = w.q64[1];
    bool up = live && u < ke;
    bool down = live && dn >= ks;
    double2 cu = xy[up ? u : ks], cd = xy[down ? dn : ks];
    while (wave_any(up || down)) {
        const double2 nu = xy[up && u + 1 < ke ? u + 1 : ks];
        const double2 nd = xy[down && dn - 1 >= ks ? dn - 1 : ks];
        {
            const double dy = cu.y - qy;
            up = up && dy < w.T;
            const double m = fmax(dy, fabs(cu.x - qx));
            w.L.insert(up ? m : INFINITY);
            w.T = w.L.top();
            visited += up;
        }
        {
            const double dy = qy - cd.y;
            down = down && dy < w.T;
            const double m = fmax(dy, fabs(cd.x - qx));
            w.L.insert(down ? m : INFINITY);
            w.T = w.L.top();
            visited += down;
        }
        ++u;
        --dn;
        up = up && u < ke;
        down = down && dn >= ks;
        cu = nu;
        cd = nd;
    }
}

// kNN on a grid order by per-lane walks: each lane walks its own strip
// from its own position, then the strips outward on each side that its
// radius still reaches on the key axis (x gap to the strip < T), entering
// each at its own y position (per-lane binary search).  A side stops for
// the wave once no lane reaches the next strip (strips further out are
// further in x).  Every candidate that could enter a lane's list is
// evaluated, in about increasing distance, so lists settle early; a lane
// evaluates a few dozen candidates instead of the whole tile's window.
// Returns the lane's number of evaluated candidates (the query itself
// included).
template <int P, bool F32>
__device__ unsigned knn_walk(KnnTile<2, P, F32>& w, const double2* __restrict__ xy, const KnnDev& kd,
                             const double* __restrict__ col64, int lo, int hi, int t0, int i, bool active) {
    const int S = kd.strip;
    const double* xk = col64 + kd.base_key;  // base order: column 0 sorted over [lo, hi)
    const int s = (t0 - lo) / S;
    const int nstrip = (hi - lo + S - 1) / S;
    unsigned visited = active ? 1u : 0u;
    if (active) {
        w.L.insert(0.0);  // the query itself
        w.T = w.L.top();
    }
    {
        const int ss = lo + s * S, se = min(hi, ss + S);
        walk_strip(w, xy, ss, se, i + 1, i - 1, active, visited);
    }
    bool ldone = s == 0, rdone = s == nstrip - 1;
    for (int d = 1; !(ldone && rdone); ++d) {
        if (!ldone) {
            const int sl = s - d;
            if (sl < 0) {
                ldone = true;
            } else {
                const int ks = lo + sl * S, ke = ks + S;  // a full strip (left of s)
                const bool need = active && w.q64[0] - xk[ke - 1] < w.T;
                ldone = !wave_any(need);
                if (!ldone) {
                    const int p = wave_lower_bounds(w.c64[1], ks, ke, w.q64[1], need);
                    walk_strip(w, xy, ks, ke, p, p - 1, need, visited);
                }
            }
        }
        if (!rdone) {
            const int sr = s + d;
            if (sr >= nstrip) {
                rdone = true;
            } else {
                const int ks = lo + sr * S, ke = min(hi, ks + S);
                const bool need = active && xk[ks] - w.q64[0] < w.T;
                rdone = !wave_any(need);
                if (!rdone) {
                    const int p = wave_lower_bounds(w.c64[1], ks, ke, w.q64[1], need);
                    walk_strip(w, xy, ks, ke, p, p - 1, need, visited);
                }
            }
        }
    }
    return visited;
}

// Grid sweep of one tile (already seeded with its own 64 points): the rest
// of the tile's strip near the tile in y order, then whole strips outward on
// both sides.  A strip is visited only within the y window of the lanes'
// current radii; a side stops at the first strip whose x bound is at least
// every lane's radius away (all strips beyond it are further in x).  Each
// candidate is visited at most once, and every candidate that could still
// enter a lane's list is visited.  Returns the number of candidates visited.
template <int D, int P, bool F32>
__device__ int grid_sweep(KnnTile<D, P, F32>& w, const KnnDev& kd, const double* __restrict__ col64,
                          int lo, int hi, int t0, int t1, bool active) {
    constexpr int SEED = 64;
    const int S = kd.strip;
    const double* xk = col64 + kd.base_key;  // base order: column 0 sorted over [lo, hi)
    const double* yc = w.c64[1];             // column 1, sorted inside each strip
    const int s = (t0 - lo) / S;
    const int nstrip = (hi - lo + S - 1) / S;
    const int ss = lo + s * S, se = min(hi, ss + S);
    const double yq = w.q64[1];
    int visited = 0;
    // Seed: up to SEED neighbours in y on each side of the tile.
    const int a0 = max(ss, t0 - SEED), b0 = min(se, t1 + SEED);
    w.range(a0, t0);
    w.range(t1, b0);
    visited += (t0 - a0) + (b0 - t1);
    {
        int a, b;
        strip_window(yc, ss, se, yq, w.T, active, a, b);
        if (a < a0) { w.range(a, a0); visited += a0 - a; }
        if (b > b0) { w.range(b0, b); visited += b - b0; }
    }
    bool ldone = s == 0, rdone = s == nstrip - 1;
    for (int d = 1; !(ldone && rdone); ++d) {
        if (!ldone) {
            const int sl = s - d;
            if (sl < 0) {
                ldone = true;
            } else {
                const int le = lo + (sl + 1) * S;  // strip sl is [le - S, le)
                const double xb = xk[le - 1];      // largest x of strip sl
                // Lanes whose radius reaches this strip on the key axis.
                const bool need = w.q64[0] - xb < w.T;
                ldone = !wave_any(need);
                if (!ldone) {
                    int a, b;
                    strip_window(yc, le - S, le, yq, w.T, need, a, b);
                    w.range(a, b);
                    visited += b - a;
                }
            }
        }
        if (!rdone) {
            const int sr = s + d;
            if (sr >= nstrip) {
                rdone = true;
            } else {
                const int rs = lo + sr * S;
                const double xb = xk[rs];          // smallest x of strip sr
                const bool need = xb - w.q64[0] < w.T;
                rdone = !wave_any(need);
                if (!rdone) {
                    int a, b;
                    strip_window(yc, rs, min(hi, rs + S), yq, w.T, need, a, b);
                    w.range(a, b);
                    visited += b - a;
                }
            }
        }
    }
    return visited;
}

template <int D, int P, bool F32>
__device__ void knn_item(const KnnArgs& A, const Item& it, const double* __restrict__ col64,
                         const float* __restrict__ col32, double* __restrict__ res64,
                         unsigned long long& pairs, unsigned long long& rechecks, double* stage) {
    const KnnDev kd = A.descs[it.desc];
    const OrderDev od = A.orders[kd.order];
    const int lane = threadIdx.x;
    const int64_t n = od.n;
    const int i = it.t0 + lane;
    const bool active = i < it.t1;
    const int ncol = od.ncol;

    KnnTile<D, P, F32> w;
    bool f32ok = true;
    w.guard = enm_guard(A.maxabs[od.inst], &f32ok);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        // Dimensions beyond the order's columns repeat column 0, which
        // leaves the max-norm unchanged.
        const int c = d < ncol ? d : 0;
        w.c64[d] = col64 + od.c64 + c * n;
        w.c32[d] = col32 + od.c32 + c * n;
        w.q64[d] = active ? w.c64[d][i] : 0.0;
        w.q32[d] = (float)w.q64[d];
    }
    const int K1 = kd.k + 1;
    w.L.init(K1, heap_base<P>(A));
    w.T = active ? INFINITY : -INFINITY;
    w.thr = active ? INFINITY : -INFINITY;
    w.rechecks = 0;
    w.dense = true;   // a tile starts with its nearest candidates
    w.stage = stage;

    const int lo = it.lo, hi = it.hi, t0 = it.t0, t1 = it.t1;
    int visited = t1 - t0;
    w.range(t0, t1);
    bool swept = false;
    if constexpr (D >= 2) {
        if (kd.strip > 0 && !A.brute) {
            visited += grid_sweep(w, kd, col64, lo, hi, t0, t1, active);
            swept = true;
        }
    }
    if (swept) {
    } else if (A.brute) {
        w.range(lo, t0);
        w.range(t1, hi);
        visited = hi - lo;
    } else {
        int left = t0, right = t1;
        bool ldone = left <= lo, rdone = right >= hi;
        while (!(ldone && rdone)) {
            if (!ldone) {
                const int a = max(lo, left - SWEEP);
                w.range(a, left);
                visited += left - a;
                left = a;
                const double ck = w.c64[0][left];
                ldone = left <= lo || wave_all(!(w.q64[0] - ck < w.T));
            }
            if (!rdone) {
                const int b = min(hi, right + SWEEP);
                w.range(right, b);
                visited += b - right;
                right = b;
                const double ck = w.c64[0][right - 1];
                rdone = right >= hi || wave_all(!(ck - w.q64[0] < w.T));
            }
        }
    }
    const double eps = w.L.top();
    if (active) {
        if (kd.eps >= 0) res64[kd.eps + i] = eps;
        if (kd.eps_orig >= 0) res64[kd.eps_orig + A.perm[od.perm + i]] = eps;
    }
    pairs += (unsigned long long)(t1 - t0) * (unsigned long long)visited;
    if (F32) rechecks += wave_sum_u64(w.rechecks);
}

// One grid-order item of a two-column order by per-lane walks (knn_walk).
template <int P>
__device__ void walk_item(const KnnArgs& A, const Item& it, const double* __restrict__ col64,
                          double* __restrict__ res64, unsigned long long& pairs) {
    const KnnDev kd = A.descs[it.desc];
    const OrderDev od = A.orders[kd.order];
    const int64_t n = od.n;
    const int i = it.t0 + (int)threadIdx.x;
    const bool active = i < it.t1;
    KnnTile<2, P, false> w;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        w.c64[d] = col64 + od.c64 + d * n;
        w.q64[d] = active ? w.c64[d][i] : 0.0;
    }
    const double2* xy = reinterpret_cast<const double2*>(col64 + od.xy);
    const int K1 = kd.k + 1;
    w.L.init(K1, heap_base<P>(A));
    w.T = active ? INFINITY : -INFINITY;
    const unsigned v = knn_walk(w, xy, kd, col64, it.lo, it.hi, it.t0, i, active);
    const double eps = w.L.top();
    if (active) {
        if (kd.eps >= 0) res64[kd.eps + i] = eps;
        if (kd.eps_orig >= 0) res64[kd.eps_orig + A.perm[od.perm + i]] = eps;
    }
    pairs += wave_sum_u64(v);
}

template <int P>
__global__ __launch_bounds__(ENM_WAVE)
void knn_walk_kernel(KnnArgs A, const double* __restrict__ col64, double* __restrict__ res64,
                     unsigned long long* __restrict__ stats) {
    unsigned long long pairs = 0;
    for_each_wave_item(A.qoff, A.counter, A.chunk, [&](int idx) {
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        walk_item<P>(A, it, col64, res64, pairs);
    });
    if (threadIdx.x == 0) atomicAdd(stats + ST_KNN_PAIRS, pairs);
}

// ---------------------------------------------------------------------------
// kNN of two-column grid orders, a strip per workgroup (knn_strip_kernel).
//
// A workgroup takes one strip s of a grid order and stages the (x, y) pairs
// of strips s-1, s and s+1 in LDS (coalesced loads, three slots of `smax`
// pairs).  Each thread then finds the k+1 nearest of its queries one after
// the other: it walks its own strip from its own position in LDS, up and
// down, a direction stopping at the first candidate whose y distance is at
// least its current radius T; then the strips outward that T still reaches
// on the x axis, entering each at its y (binary search) and walking the
// same way -- in LDS for s-1 and s+1, in global memory further out.  Every
// candidate that could enter the list is visited, so the radii are those
// of the per-lane walk and of brute force.  Candidate reads are LDS reads
// instead of one cache line per lane per load.

constexpr int SK_THREADS_MAX = 512;

// First position in [0, len) (len >= 1) of a y-sorted run whose y is >= qy,
// i.e. the number of entries below qy, by a four-way search (for runs in
// global memory, where each round trip is long): the last entry is tested
// on its own first, then each round reads three entries a quarter of the
// remaining range apart, together (y < qy holds on a prefix, so the number
// of hits is how many quarters to skip).  L rounds cover 4^L positions: a
// strip of 1024 takes five rounds where a binary search takes ten.
__device__ __forceinline__ int slot_lower_bound(const double2* __restrict__ w, int len, double qy) {
    const int n = len - 1;
    double yl = w[n].y;
    int st = 1;   // 4 * st - 1 >= n: the rounds reach any position
    while (4 * st < len) st *= 4;
    int p = 0;
    for (; st > 0; st >>= 2) {
        double v[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) v[m] = w[max(min(p + (m + 1) * st, n), 1) - 1].y;
#pragma unroll
        for (int m = 0; m < 3; ++m) pin(v[m]);
        pin(yl);   // read with the first round
        int c = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m) c += (p + (m + 1) * st <= n && v[m] < qy) ? 1 : 0;
        p += c * st;
    }
    return yl < qy ? len : p;
}

// The same for two queries at once (fixed trip count, the two chains of
// LDS reads interleaved): the number of entries of w[0, len) below each y.
__device__ __forceinline__ void slot_lower_bound2(const double2* __restrict__ w, int len, double y0, double y1,
                                                  int& p0, int& p1) {
    int top = 1;
    while (top < len) top <<= 1;
    p0 = 0;
    p1 = 0;
    for (int step = top; step > 0; step >>= 1) {
        const int j0 = p0 + step, j1 = p1 + step;
        const double v0 = w[min(j0, len) - 1].y, v1 = w[min(j1, len) - 1].y;
        if (j0 <= len && v0 < y0) p0 = j0;
        if (j1 <= len && v1 < y1) p1 = j1;
    }
}

// The walk of one y-sorted run of candidates w[0, len) from positions u
// (upward) and dn (downward): a direction stops at the first candidate
// whose y distance is at least the current radius T.
template <class L>
__device__ __forceinline__ void run_walk(L& list, double& T, const double2* __restrict__ w, int len, int u, int dn,
                                         double qx, double qy, unsigned& visited) {
    bool up = u < len, down = dn >= 0;
    while (up || down) {
        // Both next candidates are read together (a finished direction
        // rereads the other's), so their round trips overlap.
        double2 cu = w[up ? u : dn], cd = w[down ? dn : u];
        pin(cu.x); pin(cu.y); pin(cd.x); pin(cd.y);
        if (up) {
            const double dy = cu.y - qy;
            if (dy < T) {
                list.insert(vmax_abs(dy, cu.x - qx));
                T = list.top();
                ++visited;
                up = ++u < len;
            } else {
                up = false;
            }
        }
        if (down) {
            const double dy = qy - cd.y;
            if (dy < T) {
                list.insert(vmax_abs(dy, cd.x - qx));
                T = list.top();
                ++visited;
                down = --dn >= 0;
            } else {
                down = false;
            }
        }
    }
}

// One query's walk state in run_walk2.
struct WalkQ {
    double2 q;
    double T;
    int u, dn;
    bool up, down;
};

// Two queries walking the same run in lockstep: the four next candidates
// (up and down of each) are read together, so their LDS latencies overlap;
// each query still takes its candidates in its own order, up then down.
template <class L>
__device__ __forceinline__ void run_walk2(L& l0, WalkQ& a, L& l1, WalkQ& b, const double2* __restrict__ w, int len,
                                          unsigned& visited) {
    a.up = a.up && a.u < len;
    a.down = a.down && a.dn >= 0;
    b.up = b.up && b.u < len;
    b.down = b.down && b.dn >= 0;
    while (a.up || a.down || b.up || b.down) {
        double2 ca = w[a.up ? a.u : 0], da = w[a.down ? a.dn : 0];
        double2 cb = w[b.up ? b.u : 0], db = w[b.down ? b.dn : 0];
        pin(ca.x); pin(ca.y); pin(da.x); pin(da.y);
        pin(cb.x); pin(cb.y); pin(db.x); pin(db.y);
        if (a.up) {
            const double dy = ca.y - a.q.y;
            if (dy < a.T) {
                l0.insert(vmax_abs(dy, ca.x - a.q.x));
                a.T = l0.top();
                ++visited;
                a.up = ++a.u < len;
            } else {
                a.up = false;
            }
        }
        if (a.down) {
            const double dy = a.q.y - da.y;
            if (dy < a.T) {
                l0.insert(vmax_abs(dy, da.x - a.q.x));
                a.T = l0.top();
                ++visited;
                a.down = --a.dn >= 0;
            } else {
                a.down = false;
            }
        }
        if (b.up) {
            const double dy = cb.y - b.q.y;
            if (dy < b.T) {
                l1.insert(vmax_abs(dy, cb.x - b.q.x));
                b.T = l1.top();
                ++visited;
                b.up = ++b.u < len;
            } else {
                b.up = false;
            }
        }
        if (b.down) {
            const double dy = b.q.y - db.y;
            if (dy < b.T) {
                l1.insert(vmax_abs(dy, db.x - b.q.x));
                b.T = l1.top();
                ++visited;
                b.down = --b.dn >= 0;
            } else {
                b.down = false;
            }
        }
    }
}

// Strips further out than the staged neighbours, for one query: in global
// memory, from distance d = 2 on each side that is not done.
template <class L>
__device__ __forceinline__ void far_strips(L& list, double& T, const double2* __restrict__ xy,
                                           const double* __restrict__ xk, int lo, int hi, int S, int s, int nstrip,
                                           double2 q, bool ldone, bool rdone, unsigned& visited) {
    for (int d = 2; !(ldone && rdone); ++d) {
        if (!ldone) {
            const int sl = s - d;
            if (sl < 0 || !(q.x - xk[lo + (sl + 1) * S - 1] < T)) {
                ldone = true;
            } else {
                const double2* w = xy + lo + sl * S;
                const int p = slot_lower_bound(w, S, q.y);
                run_walk(list, T, w, S, p, p - 1, q.x, q.y, visited);
            }
        }
        if (!rdone) {
            const int sr = s + d;
            if (sr >= nstrip || !(xk[lo + sr * S] - q.x < T)) {
                rdone = true;
            } else {
                const int a = lo + sr * S, len = min(hi, a + S) - a;
                const double2* w = xy + a;
                const int p = slot_lower_bound(w, len, q.y);
                run_walk(list, T, w, len, p, p - 1, q.x, q.y, visited);
            }
        }
    }
}

// Four-slot lists are held to 80 VGPRs, six waves per SIMD (a few bytes
// spill: the far-list bookkeeping would otherwise cost a wave per SIMD).
template <int P>
__global__ __launch_bounds__(SK_THREADS_MAX, P == 4 ? 6 : 1)
void knn_strip_kernel(KnnArgs A, const double* __restrict__ col64, double* __restrict__ res64,
                      unsigned long long* __restrict__ stats) {
    extern __shared__ double2 sxy[];   // three slots of A.smax pairs: strips s-1, s, s+1
    __shared__ unsigned s_ticket;
    const int NT = (int)blockDim.x;
    unsigned pairs = 0;   // this thread's candidates (a few thousand): one register
    for_each_block_item(A.qoff, A.counter, &s_ticket, [&](int idx) {
        const Item it = A.items[idx];
        const KnnDev kd = A.descs[it.desc];
        const OrderDev od = A.orders[kd.order];
        const int S = kd.strip;
        const double2* __restrict__ xy = reinterpret_cast<const double2*>(col64 + od.xy);
        const double* __restrict__ xk = col64 + kd.base_key;   // column 0 sorted over the scope range
        const int lo = it.lo, hi = it.hi;
        const int s = (it.t0 - lo) / S;
        const int nstrip = (hi - lo + S - 1) / S;
        int slo[3], slen[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ss = s - 1 + k;
            slo[k] = lo + ss * S;
            slen[k] = (ss < 0 || ss >= nstrip) ? 0 : min(hi, slo[k] + S) - slo[k];
        }
        // Staged a pair per strip per round, the three loads of a round
        // issued before its LDS writes (two per strip would spill).
        const int most = max(slen[0], max(slen[1], slen[2]));
        for (int t = (int)threadIdx.x; t < most; t += NT) {
            double2 v[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = t < slen[k] ? xy[slo[k] + t] : make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (t < slen[k]) sxy[k * A.smax + t] = v[k];
        }
        // The x bounds of the neighbours: largest x left, smallest x right;
        // and of the strips beyond them.
        const double xl = uni_f64(slen[0] > 0 ? xk[slo[1] - 1] : 0.0);
        const double xr = uni_f64(slen[2] > 0 ? xk[slo[2]] : 0.0);
        const double xl2 = uni_f64(s >= 2 ? xk[slo[0] - 1] : 0.0);
        const double xr2 = uni_f64(s + 2 < nstrip ? xk[lo + (s + 2) * S] : 0.0);
        __syncthreads();
        const double2* own = sxy + A.smax;
        // Two queries per thread, NT apart, walked in lockstep.
        for (int i0 = it.t0 + (int)threadIdx.x; i0 < it.t1; i0 += 2 * NT) {
            const int i1 = i0 + NT;
            const bool act1 = i1 < it.t1;
            RegList<P> l0, l1;
            l0.init(kd.k + 1, nullptr);
            l1.init(kd.k + 1, nullptr);
            WalkQ a, b;
            const int p0 = i0 - slo[1], p1 = act1 ? i1 - slo[1] : p0;
            a.q = own[p0];
            b.q = own[p1];
            l0.insert(0.0);   // the queries themselves
            if (act1) l1.insert(0.0);
            a.T = l0.top();
            b.T = l1.top();
            unsigned visited = 1 + (act1 ? 1u : 0u);
            a.u = p0 + 1; a.dn = p0 - 1; a.up = a.down = true;
            b.u = p1 + 1; b.dn = p1 - 1; b.up = b.down = act1;
            run_walk2(l0, a, l1, b, own, slen[1], visited);
            // The staged neighbours, where the radius reaches them on x.
            const bool la = slen[0] > 0 && a.q.x - xl < a.T, lb = act1 && slen[0] > 0 && b.q.x - xl < b.T;
            if (la || lb) {
                int e0, e1;
                slot_lower_bound2(sxy, slen[0], a.q.y, b.q.y, e0, e1);
                a.u = e0; a.dn = e0 - 1; a.up = a.down = la;
                b.u = e1; b.dn = e1 - 1; b.up = b.down = lb;
                run_walk2(l0, a, l1, b, sxy, slen[0], visited);
            }
            const bool ra = slen[2] > 0 && xr - a.q.x < a.T, rb = act1 && slen[2] > 0 && xr - b.q.x < b.T;
            if (ra || rb) {
                const double2* w = sxy + 2 * A.smax;
                int e0, e1;
                slot_lower_bound2(w, slen[2], a.q.y, b.q.y, e0, e1);
                a.u = e0; a.dn = e0 - 1; a.up = a.down = ra;
                b.u = e1; b.dn = e1 - 1; b.up = b.down = rb;
                run_walk2(l0, a, l1, b, w, slen[2], visited);
            }
            // Further strips in global memory (a side is done once its
            // nearer strip was out of reach).  A query that needs them goes
            // to the far list, so that its long walk through global memory
            // holds up neither its wave nor, at the next barrier, its
            // workgroup; knn_far_kernel finishes it with every lane busy.
            const bool fa = (la && s >= 2 && a.q.x - xl2 < a.T) || (ra && s + 2 < nstrip && xr2 - a.q.x < a.T);
            const bool fb = act1 && ((lb && s >= 2 && b.q.x - xl2 < b.T) ||
                                     (rb && s + 2 < nstrip && xr2 - b.q.x < b.T));
            bool da = false, db = false;
            if (A.far != nullptr) {
                const unsigned long long ma = __builtin_amdgcn_ballot_w64(fa);
                const unsigned long long mb = __builtin_amdgcn_ballot_w64(fb);
                const unsigned tot = (unsigned)(__popcll(ma) + __popcll(mb));
                if (tot != 0) {
                    const int lane = (int)(threadIdx.x & (ENM_WAVE - 1));
                    const int leader = __ffsll((long long)__builtin_amdgcn_read_exec()) - 1;
                    unsigned base = 0;
                    if (lane == leader) base = atomicAdd(A.far_count, tot);
                    base = (unsigned)__shfl((int)base, leader, ENM_WAVE);
                    const unsigned ra_ = __builtin_amdgcn_mbcnt_hi((unsigned)(ma >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)ma, 0u));
                    const unsigned rb_ = __builtin_amdgcn_mbcnt_hi((unsigned)(mb >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)mb, 0u));
                    const unsigned sa = base + ra_, sb = base + (unsigned)__popcll(ma) + rb_;
                    KnnFar<P>* far = reinterpret_cast<KnnFar<P>*>(A.far);
                    const int64_t cap = A.far_bytes / (int64_t)sizeof(KnnFar<P>);
                    da = fa && (int64_t)sa < cap;
                    db = fb && (int64_t)sb < cap;
                    if (da) {
                        KnnFar<P>& f = far[sa];
                        f.item = idx;
                        f.i = i0;
                        f.open = (la ? 1 : 0) | (ra ? 2 : 0);
#pragma unroll
                        for (int t = 0; t < P; ++t) f.L[t] = l0.L[t];
                    }
                    if (db) {
                        KnnFar<P>& f = far[sb];
                        f.item = idx;
                        f.i = i1;
                        f.open = (lb ? 1 : 0) | (rb ? 2 : 0);
#pragma unroll
                        for (int t = 0; t < P; ++t) f.L[t] = l1.L[t];
                    }
                }
            }
            if (fa && !da) far_strips(l0, a.T, xy, xk, lo, hi, S, s, nstrip, a.q, !la, !ra, visited);
            if (fb && !db) far_strips(l1, b.T, xy, xk, lo, hi, S, s, nstrip, b.q, !lb, !rb, visited);
            const double e0 = l0.top(), e1 = l1.top();
            const bool w0 = !da, w1 = act1 && !db;
            if (kd.eps >= 0) {
                if (w0) res64[kd.eps + i0] = e0;
                if (w1) res64[kd.eps + i1] = e1;
            }
            if (kd.eps_orig >= 0) {
                // Both permutation entries are read before either store.
                const int o0 = A.perm[od.perm + i0], o1 = A.perm[od.perm + (act1 ? i1 : i0)];
                if (w0) res64[kd.eps_orig + o0] = e0;
                if (w1) res64[kd.eps_orig + o1] = e1;
            }
            pairs += visited;
        }
        __syncthreads();   // the slots are restaged by the next item
    });
    const unsigned long long total = wave_sum_u64((unsigned long long)pairs);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) atomicAdd(stats + ST_KNN_PAIRS, total);
}

// The far list of a knn_strip_kernel launch, one query per thread: its
// walk continues through the strips beyond its staged neighbours
// (far_strips, from its list so far), and its radius is written as the
// strip kernel would have.
template <int P>
__global__ __launch_bounds__(256)
void knn_far_kernel(KnnArgs A, const double* __restrict__ col64, double* __restrict__ res64,
                    unsigned long long* __restrict__ stats) {
    const int64_t cap = A.far_bytes / (int64_t)sizeof(KnnFar<P>);
    const unsigned n = (unsigned)min((int64_t)*A.far_count, cap);
    const KnnFar<P>* far = reinterpret_cast<const KnnFar<P>*>(A.far);
    unsigned long long pairs = 0;
    for (unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const KnnFar<P>& f = far[r];
        const Item it = A.items[f.item];
        const KnnDev& kd = A.descs[it.desc];
        const OrderDev& od = A.orders[kd.order];
        const int S = kd.strip;
        const double2* __restrict__ xy = reinterpret_cast<const double2*>(col64 + od.xy);
        const double* __restrict__ xk = col64 + kd.base_key;
        const int lo = it.lo, hi = it.hi;
        const int s = (it.t0 - lo) / S;
        const int nstrip = (hi - lo + S - 1) / S;
        RegList<P> L;
#pragma unroll
        for (int t = 0; t < P; ++t) L.L[t] = f.L[t];
        double T = L.top();
        const int i = f.i;
        unsigned visited = 0;
        far_strips(L, T, xy, xk, lo, hi, S, s, nstrip, xy[i], !(f.open & 1), !(f.open & 2), visited);
        const double e = L.top();
        if (kd.eps >= 0) res64[kd.eps + i] = e;
        if (kd.eps_orig >= 0) res64[kd.eps_orig + A.perm[od.perm + i]] = e;
        pairs += visited;
    }
    pairs = wave_sum_u64(pairs);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0 && pairs) atomicAdd(stats + ST_KNN_PAIRS, pairs);
}

template <int D, int P, bool F32>
__global__ __launch_bounds__(ENM_WAVE)
void knn_kernel(KnnArgs A, const double* __restrict__ col64, const float* __restrict__ col32,
                double* __restrict__ res64, unsigned long long* __restrict__ stats) {
    // Counters accumulate per wave (the same value in every lane) and go
    // out once at the end: one atomic per item on one address would queue
    // a million atomics on a single L2 channel.
    unsigned long long pairs = 0, rechecks = 0;
    __shared__ double stage[D >= 3 ? D * ENM_WAVE : 1];   // candidates of three or more columns
    for_each_wave_item(A.qoff, A.counter, A.chunk, [&](int idx) {
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        if (F32) {
            const KnnDev& kd = A.descs[it.desc];
            bool ok = true;
            (void)enm_guard(A.maxabs[A.orders[kd.order].inst], &ok);
            if (ok) knn_item<D, P, true>(A, it, col64, col32, res64, pairs, rechecks, stage);
            else    knn_item<D, P, false>(A, it, col64, col32, res64, pairs, rechecks, stage);
        } else {
            knn_item<D, P, false>(A, it, col64, col32, res64, pairs, rechecks, stage);
        }
    });
    if (threadIdx.x == 0) {
        atomicAdd(stats + ST_KNN_PAIRS, pairs);
        if (F32) atomicAdd(stats + ST_KNN_RECHECK, rechecks);
    }
}

template <int D, int P>
hipError_t launch_dp(const KnnArgs& a, bool f32, int walk, int grid, const double* col64, const float* col32,
                     double* res64, unsigned long long* stats, hipStream_t st) {
    // LDS heaps: 64 lanes x the group's largest K1 doubles per workgroup.
    const unsigned shmem = P == ENM_LIST_HEAP_LDS ? (unsigned)(a.heap_stride * sizeof(double)) : 0u;
    // Per-lane walks and strip workgroups exist for two-column grid orders
    // only; strip workgroups for register lists (their LDS holds strips).
    if (walk) {
        if constexpr (D == 2) {
            if constexpr (P > 0) {
                if (walk == 2) {
                    const int nt = a.smax >= SK_THREADS_MAX ? SK_THREADS_MAX : ((a.smax + 63) / 64) * 64;
                    const unsigned lds = (unsigned)(3 * a.smax * sizeof(double2));
                    hipLaunchKernelGGL((knn_strip_kernel<P>), dim3(grid), dim3(nt), lds, st, a, col64, res64, stats);
                    // The far list: eight 4-wave blocks per CU of the strip grid's
                    // three (latency-bound walks need every wave slot).
                    if (a.far != nullptr)
                        hipLaunchKernelGGL((knn_far_kernel<P>), dim3(std::max(1, grid) * 8 / 3 + 1), dim3(256), 0, st,
                                           a, col64, res64, stats);
                    return hipGetLastError();
                }
            }
            if (walk == 1) {
                hipLaunchKernelGGL((knn_walk_kernel<P>), dim3(grid), dim3(ENM_WAVE), shmem, st, a, col64, res64,
                                   stats);
                return hipGetLastError();
            }
        }
        return hipErrorInvalidValue;
    }
    if (f32)
        hipLaunchKernelGGL((knn_kernel<D, P, true>), dim3(grid), dim3(ENM_WAVE), shmem, st,
                           a, col64, col32, res64, stats);
    else
        hipLaunchKernelGGL((knn_kernel<D, P, false>), dim3(grid), dim3(ENM_WAVE), shmem, st,
                           a, col64, col32, res64, stats);
    return hipGetLastError();
}

template <int D>
hipError_t launch_d(const KnnArgs& a, int P, bool f32, int walk, int grid, const double* col64,
                    const float* col32, double* res64, unsigned long long* stats, hipStream_t st) {
    switch (P) {
        case 4: return launch_dp<D, 4>(a, f32, walk, grid, col64, col32, res64, stats, st);
        case 8: return launch_dp<D, 8>(a, f32, walk, grid, col64, col32, res64, stats, st);
        case 16: return launch_dp<D, 16>(a, f32, walk, grid, col64, col32, res64, stats, st);
        case 32: return launch_dp<D, 32>(a, f32, walk, grid, col64, col32, res64, stats, st);
        case ENM_LIST_HEAP_LDS: return launch_dp<D, ENM_LIST_HEAP_LDS>(a, f32, walk, grid, col64, col32, res64,
                                                                      stats, st);
        case ENM_LIST_HEAP_GLOBAL: return launch_dp<D, ENM_LIST_HEAP_GLOBAL>(a, f32, walk, grid, col64, col32,
                                                                            res64, stats, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

#define ENM_KNN_DEFINE(D) \
    hipError_t enm_knn_launch_d##D(const KnnArgs& a, int P, bool f32, int walk, int grid, \
                                   const double* col64, const float* col32, double* res64, \
                                   unsigned long long* stats, hipStream_t st) { \
        return launch_d<D>(a, P, f32, walk, grid, col64, col32, res64, stats, st); \
    }
#endif
