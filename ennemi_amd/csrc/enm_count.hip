// Inclusive max-norm ball counts, replacing
// cKDTree.query_ball_point(pts, eps - 1e-12, p=np.inf, return_length=True)
// of the reference (ennemi/_entropy_estimators.py:109-110,152-154,196,243-245):
//   n_i = #{ j in scope : max_d fl64(|a_d[i] - a_d[j]|) <= fl64(eps_i - 1e-12) },
// which is 0 when the shifted radius is negative.  Queries and candidates are
// the same key-sorted order, so a query's own position is known.
//
// 1-D spaces (the marginals of KSG, Ross and a 1-D condition): the ball is a
// contiguous run of the sorted key column around the query, because
// fl64(v_j - v_i) is monotone in v_j on each side.  Each query binary-searches
// both boundaries in an LDS copy of the keys around its item, with the same
// FP64 predicate the reference evaluates, so counts are exact; the few balls
// wider than the staged keys are searched afterwards in global memory.
// Algorithmic traffic: 8 B radius + 8 B key + 4 B count per query.  In every
// mode (including ENM_F_BRUTE, which only removes window pruning) 1-D counts
// use this search.
//
// Multi-D spaces: on a grid order (the space's first two columns are the
// order's columns 0 and 1) a wave takes 64 queries of one strip and scans its
// own strip and the strips outward, each within the y window of the tile
// widened by the wave's largest radius, in FP64 (count_grid); with two
// columns, a strip whose whole key range lies within a lane's radius is
// counted for that lane by binary searches of its y column.  Otherwise the
// wave bounds the union of its 64 balls on the key axis and scans that window
// of the key-sorted order with wave-uniform candidate loads; in FP32-filter
// mode each lane counts candidates whose FP32 distance is <= the radius
// rounded down minus the guard (certainly inside) and <= the radius rounded
// up plus the guard (possibly inside), and if any lane saw a candidate in
// between the wave rescans its window and settles those in FP64.  (On this
// ISA FP32 and FP64 compares issue at the same rate, so the grid path counts
// in FP64 directly: the band bookkeeping would cost more than it saves.)
//
// psi(n_i) comes from a table filled once per launch (psi_tab[c], +inf for
// c = 0: psi(0) = +inf in the reference, :338-339).  Multi-D spaces write it
// in point order for a fixed-order reduction; 1-D spaces are summed per
// ENM_PSI_TILE positions of their sorted order with a flag for n_i == 0.
#include "enm_device.h"
#include "enm_launch.h"

#include <algorithm>

namespace {

constexpr int CH = 8;

struct CountArgs {
    const Item* items;
    const int* qoff;    // ENM_NQ + 1 queue offsets into items
    unsigned* counter;  // ENM_NQ ticket counters
    int chunk;          // items per ticket (multi-D counts)
    const CountDev* descs;
    const OrderDev* orders;
    const unsigned long long* maxabs;
    const int* perm;
    const double* psi_tab;  // psi_tab[c] = psi(c), psi_tab[0] = +inf
    int brute;
};

__device__ __forceinline__ double query_radius(const CountDev& cd, const CountArgs& A,
                                               const double* __restrict__ res64, int i) {
    const double eps = cd.same_order ? res64[cd.eps + i] : res64[cd.eps + A.perm[cd.eps_perm + i]];
    return eps - 1e-12;
}

__device__ __forceinline__ void finish_item(const CountDev& cd, const Item& it, bool active,
                                            int cnt, int* __restrict__ res32,
                                            double* __restrict__ res64, int i, const int* perm,
                                            const double* __restrict__ psi_tab) {
    if (active) res32[cd.counts + i] = cnt;
    if (cd.psi_pt >= 0) {
        // Point order; a zero count is +inf (psi_tab[0]), as psi(0) poisons
        // the mean.
        if (active) res64[cd.psi_pt + perm[cd.eps_perm + i]] = psi_tab[cnt];
        return;
    }
    const double p = (active && cnt > 0) ? psi_tab[cnt] : 0.0;
    const double s = wave_sum(p);
    const int z = wave_or(active && cnt == 0);
    if (threadIdx.x == 0) {
        res64[cd.part + it.tile] = s;
        res32[cd.zero + it.tile] = z;
    }
}

// Boundary searches over the whole scope range [lo, hi) in global memory:
// last position in [i, hi) with fl(v[j] - x) <= r (right) or first in
// [lo, i] with fl(x - v[j]) <= r (left); position i holds x itself.
// fl(v - x) is monotone in v, so the predicate holds on a prefix (right)
// or suffix (left).
__device__ __forceinline__ int global_right(const double* __restrict__ v, int i, int hi, double x, double r) {
    int good = i, bad = hi;
    while (bad - good > 1) {
        const int mid = (good + bad) >> 1;
        if (v[mid] - x <= r) good = mid; else bad = mid;
    }
    return good;
}
__device__ __forceinline__ int global_left(const double* __restrict__ v, int lo, int i, double x, double r) {
    int good = i, bad = lo - 1;
    while (good - bad > 1) {
        const int mid = (good + bad) >> 1;
        if (x - v[mid] <= r) good = mid; else bad = mid;
    }
    return good;
}

// 1-D counts.  An item is ENM_C1_ITEM consecutive positions of a key-sorted
// order, ENM_C1_Q per thread of a 512-thread workgroup.  The workgroup
// copies the keys of the item and ENM_C1_HALO positions on each side into
// LDS with coalesced loads (40 KB, shared by its eight waves: four
// workgroups fill a CU's 32 wave slots; 256 threads with eight queries
// each left half of them empty and took 18 % longer), and each
// query binary-searches both boundaries there.  A query whose ball reaches
// past the staged keys on either side (its predicate still holds at the
// window's edge, about 1 % of queries on Gaussian-like data, but in almost
// every item) is deferred: its code (item << ENM_C1_OFF_BITS | offset) goes
// to a list that count1d_deferred_kernel then searches over the whole scope
// range with every lane busy, instead of stalling the rest of its wave.  The
// list has ENM_NDEF segments with a counter each (workgroup b appends to
// segment b % ENM_NDEF, one atomic per wave and round), so appends do not
// queue on one address; a full segment makes the query search in place.
// Counts only: the psi sums of 1-D spaces are taken afterwards per
// ENM_PSI_TILE positions (psi_tile_kernel), in a fixed order.
__global__ __launch_bounds__(ENM_C1_THREADS, 8)
void count1d_kernel(CountArgs A, const double* __restrict__ col64, const double* __restrict__ res64,
                    int* __restrict__ res32, unsigned long long* __restrict__ stats,
                    unsigned* __restrict__ deferred, unsigned* __restrict__ n_deferred, unsigned seg_cap) {
    // Slot t + 1 holds key w0 + t; slots 0 and nw + 1 hold NaN, which fails
    // both boundary predicates, so the searches need no range checks.
    __shared__ double win[ENM_C1_CAP + 2];
    __shared__ unsigned s_ticket;
    const int tid = threadIdx.x;
    const int seg = (int)(blockIdx.x % ENM_NDEF);
    unsigned* dcount = n_deferred + seg * ENM_CSTRIDE;
    unsigned* dlist = deferred + (size_t)seg * seg_cap;
    unsigned long long queries = 0;
    for_each_block_item(A.qoff, A.counter, &s_ticket, [&](int idx) {
        Item it = A.items[idx];
        it.desc = uni(it.desc);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        const CountDev& cd = A.descs[it.desc];
        const double* v = col64 + A.orders[cd.order].c64;  // sorted key column
        const int64_t counts = cd.counts;
        const int w0 = max(it.lo, it.t0 - ENM_C1_HALO);
        const int w1 = min(it.hi, it.t1 + ENM_C1_HALO);
        const int nw = w1 - w0;
        {
            // All of a thread's window loads go out before its LDS writes.
            constexpr int WPER = (ENM_C1_CAP + ENM_C1_THREADS - 1) / ENM_C1_THREADS;
            double wv[WPER];
#pragma unroll
            for (int k = 0; k < WPER; ++k) {
                const int t = tid + k * ENM_C1_THREADS;
                wv[k] = t < nw ? v[w0 + t] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < WPER; ++k) {
                const int t = tid + k * ENM_C1_THREADS;
                if (t < nw) win[t + 1] = wv[k];
            }
        }
        if (tid == 0) {
            win[0] = __builtin_nan("");
            win[nw + 1] = __builtin_nan("");
        }
        __syncthreads();
        const double wlo = win[1], whi = win[nw];
        // All radii first, each step of their dependent gathers for every
        // query at once (the permutation entries, then the radii), so the
        // round trips overlap instead of stalling every query in turn.
        const bool same = uni((int)cd.same_order) != 0;
        const int64_t eps_at = cd.eps, eps_perm = cd.eps_perm;
        int src[ENM_C1_Q];
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) {
            const int i = min(it.t0 + tid + ENM_C1_THREADS * q, it.t1 - 1);
            src[q] = same ? i : A.perm[eps_perm + i];
        }
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) pin(src[q]);
        double rq[ENM_C1_Q];
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) rq[q] = res64[eps_at + src[q]];
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) {
            pin(rq[q]);
            const int i = it.t0 + tid + ENM_C1_THREADS * q;
            rq[q] = i < it.t1 ? rq[q] - 1e-12 : -1.0;
        }
        // Both boundaries of all ENM_C1_Q queries are searched in lockstep
        // (halving power-of-two steps from the query's own position, which
        // together span the whole window): 2 * ENM_C1_Q independent LDS
        // reads per step instead of one dependent chain per search.
        double xq[ENM_C1_Q];
        int pr[ENM_C1_Q], pl[ENM_C1_Q];
        bool live[ENM_C1_Q], dq[ENM_C1_Q];
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) {
            const int i = it.t0 + tid + ENM_C1_THREADS * q;
            const double r = rq[q];
            const int p = i < it.t1 ? i - w0 : 0;
            const double x = win[p + 1];
            xq[q] = x;
            pr[q] = p;
            pl[q] = p;
            live[q] = i < it.t1 && !(r < 0.0);
            dq[q] = live[q] && ((w1 < it.hi && whi - x <= r) || (w0 > it.lo && x - wlo <= r));
        }
        // A query sits at most ENM_C1_ITEM + ENM_C1_HALO - 1 slots from
        // either end of the window, and steps top, top / 2, ..., 1 reach
        // any offset up to 2 * top - 1.
        int top = 1;
        while (2 * top - 1 < ENM_C1_ITEM + ENM_C1_HALO - 1) top *= 2;
        for (int step = top; step >= 1; step >>= 1) {
#pragma unroll
            for (int q = 0; q < ENM_C1_Q; ++q) {
                const int jr = pr[q] + step, jl = pl[q] - step;
                // Past either end of the window the read is a NaN sentinel.
                const double vr = win[min(jr, nw) + 1], vl = win[max(jl, -1) + 1];
                if (vr - xq[q] <= rq[q]) pr[q] = jr;
                if (xq[q] - vl <= rq[q]) pl[q] = jl;
            }
        }
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) {
            const int i = it.t0 + tid + ENM_C1_THREADS * q;
            const bool active = i < it.t1;
            const double r = rq[q];
            bool defer = dq[q];
            const int cnt0 = live[q] && !defer ? pr[q] - pl[q] + 1 : 0;
            int cnt = cnt0;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(defer);
            if (m != 0) {
                // One atomic per wave: every lane adds, lane 0 the wave's
                // total, the others 0; lane 0's old value is the base.
                const unsigned tot = (unsigned)__popcll(m);
                const unsigned base = (unsigned)__builtin_amdgcn_readfirstlane(
                    (int)atomicAdd(dcount, (tid & (ENM_WAVE - 1)) == 0 ? tot : 0u));
                const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if (defer) {
                    const unsigned slot = base + rank;
                    if (slot < seg_cap) {
                        dlist[slot] = ((unsigned)idx << ENM_C1_OFF_BITS) | (unsigned)(i - it.t0);
                    } else {
                        const double x = v[i];
                        cnt = global_right(v, i, it.hi, x, r) - global_left(v, it.lo, i, x, r) + 1;
                        defer = false;
                    }
                }
            }
            if (active && !defer) res32[counts + i] = cnt;
        }
        queries += (unsigned long long)(it.t1 - it.t0);
        __syncthreads();  // the window is reloaded by the next item
    });
    if (tid == 0) atomicAdd(stats + ST_CNT_1D, queries);
}

// Deferred 1-D queries: one per thread over all segments at once (entry e
// of the concatenated segments found through their prefix sums in LDS),
// both boundaries by binary search over the whole scope range.
__global__ __launch_bounds__(256)
void count1d_deferred_kernel(CountArgs A, const double* __restrict__ col64, const double* __restrict__ res64,
                             int* __restrict__ res32, const unsigned* __restrict__ deferred,
                             const unsigned* __restrict__ n_deferred, unsigned seg_cap) {
    __shared__ unsigned pre[ENM_NDEF + 1];
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int s = 0; s < ENM_NDEF; ++s) {
            pre[s] = acc;
            acc += min(n_deferred[s * ENM_CSTRIDE], seg_cap);
        }
        pre[ENM_NDEF] = acc;
    }
    __syncthreads();
    const unsigned total = pre[ENM_NDEF];
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        int lo = 0, hi = ENM_NDEF - 1;   // last segment s with pre[s] <= e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const unsigned code = deferred[(size_t)lo * seg_cap + (e - pre[lo])];
        const Item it = A.items[code >> ENM_C1_OFF_BITS];
        const CountDev& cd = A.descs[it.desc];
        const double* v = col64 + A.orders[cd.order].c64;
        const int i = it.t0 + (int)(code & ((1u << ENM_C1_OFF_BITS) - 1));
        const double r = query_radius(cd, A, res64, i);
        const double x = v[i];
        // Both boundaries in lockstep by halving steps from i (the steps add
        // up to at least the farther end of the scope range): two
        // independent reads per step.  A deferred query has r >= 0, so the
        // predicate holds at i itself.
        const int span = max(it.hi - 1 - i, i - it.lo);
        int step = 1;
        while (2 * step <= span) step *= 2;
        int pr = i, pl = i;
        for (; step >= 1 && span > 0; step >>= 1) {
            const int jr = pr + step, jl = pl - step;
            const double vr = v[min(jr, it.hi - 1)], vl = v[max(jl, it.lo)];
            if (jr < it.hi && vr - x <= r) pr = jr;
            if (jl >= it.lo && x - vl <= r) pl = jl;
        }
        res32[cd.counts + i] = pr - pl + 1;
    }
}

// psi sums of 1-D spaces: one wave per ENM_PSI_TILE positions of a space's
// key-sorted order, 8 per lane in lane-major order, then a butterfly sum;
// psi values come from the table psi_tab[c] = psi(c).  A zero count sets the
// tile's flag (psi(0) poisons the reference's mean).
__global__ __launch_bounds__(ENM_WAVE)
void psi_tile_kernel(const Item* __restrict__ items, int n_items, const CountDev* __restrict__ descs,
                     const double* __restrict__ psi_tab, const int* __restrict__ counts_in,
                     double* __restrict__ res64, int* __restrict__ res32) {
    const int item = blockIdx.x;
    if (item >= n_items) return;
    const Item it = items[item];
    const CountDev& cd = descs[it.desc];
    const int lane = threadIdx.x;
    constexpr int Q = ENM_PSI_TILE / ENM_WAVE;
    // Every count is read (positions past the tile reread its last one),
    // then every table entry, each batch issued together; the sum then takes
    // the tile's positions in order.
    int c[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) c[q] = counts_in[cd.counts + min(it.t0 + lane + ENM_WAVE * q, it.t1 - 1)];
#pragma unroll
    for (int q = 0; q < Q; ++q) pin(c[q]);
    double pv[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) pv[q] = psi_tab[c[q]];
#pragma unroll
    for (int q = 0; q < Q; ++q) pin(pv[q]);
    double ps = 0.0;
    bool zero = false;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int i = it.t0 + lane + ENM_WAVE * q;
        if (i < it.t1) {
            if (c[q] > 0) ps += pv[q];
            else zero = true;
        }
    }
    const double s = wave_sum(ps);
    const int z = wave_or(zero);
    if (lane == 0) {
        res64[cd.part + it.tile] = s;
        res32[cd.zero + it.tile] = z;
    }
}

// psi_tab[c] = psi(c) for c in [1, n], psi_tab[0] = +inf.
__global__ void psi_table_kernel(double* __restrict__ tab, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= n) tab[c] = c == 0 ? INFINITY : enm_psi((double)c);
}

// Count of candidates j in [a, b) with max_d fl64(|c_d[j] - q_d|) <= r, for
// one lane (wave-uniform candidate range).
template <int D>
__device__ __forceinline__ int count_range64(const double* const (&c64)[D], const double (&q64)[D],
                                             double r, int a, int b) {
    int cnt = 0;
    int j = a;
    for (; j + CH <= b; j += CH) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            double m = fabs(c64[0][j + c] - q64[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j + c] - q64[d]));
            cnt += (m <= r);
        }
    }
    for (; j < b; ++j) {
        double m = fabs(c64[0][j] - q64[0]);
#pragma unroll
        for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j] - q64[d]));
        cnt += (m <= r);
    }
    return cnt;
}

// count_range64 for three or more columns, candidates through LDS: the wave
// loads 64 candidates at a time, one per lane (coalesced per column, the
// next 64 in flight meanwhile), into `stage` (D x 64 doubles), and every
// lane reads them back with broadcast LDS reads.  [a, b) is wave-uniform.
template <int D>
__device__ __forceinline__ int count_range_staged(const double* const (&c64)[D], const double (&q64)[D],
                                                  double r, int a, int b, double* __restrict__ stage) {
    const int lane = (int)(threadIdx.x & (ENM_WAVE - 1));
    int cnt = 0;
    double nx[D];
    auto fetch = [&](int j0) {
        const int p = j0 + lane;
        const bool ok = p < b;
#pragma unroll
        for (int d = 0; d < D; ++d) nx[d] = ok ? c64[d][p] : 0.0;
    };
    if (a < b) fetch(a);
    for (int j = a; j < b; j += ENM_WAVE) {
        __syncthreads();   // every lane has read the previous block
#pragma unroll
        for (int d = 0; d < D; ++d) stage[d * ENM_WAVE + lane] = nx[d];
        __syncthreads();
        const int m = min(ENM_WAVE, b - j);
        if (j + ENM_WAVE < b) fetch(j + ENM_WAVE);
        int c = 0;
        for (; c + CH <= m; c += CH) {
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                double e = fabs(stage[c + u] - q64[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) e = fmax(e, fabs(stage[d * ENM_WAVE + c + u] - q64[d]));
                cnt += (e <= r);
            }
        }
        for (; c < m; ++c) {
            double e = fabs(stage[c] - q64[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) e = fmax(e, fabs(stage[d * ENM_WAVE + c] - q64[d]));
            cnt += (e <= r);
        }
    }
    return cnt;
}

// Number of positions j in [a, b) of an ascending column with
// fl64(|y_j - q|) <= r (r >= 0), for one lane: fl64(y - q) is monotone in y,
// so they are the run from the first j with fl64(y_j - q) >= -r to the first
// with fl64(y_j - q) > r, found by two binary searches run together (their
// loads are independent).  [a, b) must hold every such j.
__device__ __forceinline__ int count_sorted(const double* __restrict__ y, int a, int b, double q, double r) {
    int l0 = a, n0 = b - a;   // first j with fl(y_j - q) >= -r
    int l1 = a, n1 = b - a;   // first j with fl(y_j - q) > r
    while (n0 > 0 || n1 > 0) {
        const int h0 = n0 >> 1, h1 = n1 >> 1;
        // Both reads every round (a finished search rereads a position of
        // the range), issued together ahead of the branches that use them.
        double v0 = y[min(l0 + h0, b - 1)], v1 = y[min(l1 + h1, b - 1)];
        pin(v0);
        pin(v1);
        if (n0 > 0) {
            if (v0 - q >= -r) n0 = h0;
            else { l0 += h0 + 1; n0 -= h0 + 1; }
        }
        if (n1 > 0) {
            if (v1 - q > r) n1 = h1;
            else { l1 += h1 + 1; n1 -= h1 + 1; }
        }
    }
    return l1 - l0;
}

// Grid count: the tile's own strip and then strips outward, each within the
// union of the y intervals [y_i - r_i, y_i + r_i] of the lanes that can
// reach it (strip_window).  A side stops at the first strip whose column-0
// bound is more than every live lane's radius away (strictly: a count
// includes distance == r), and all strips beyond it are further still.
// With two columns, a lane whose ball covers a strip's whole column-0 range
// (both end keys within r: fl64(x_j - x_i) is monotone in x_j) counts that
// strip as a run of its y column, by two binary searches inside the window
// of such lanes, instead of scanning; the other lanes scan theirs.
// Returns candidates scanned.
template <int D>
__device__ int count_grid(const CountDev& cd, const double* __restrict__ col64,
                          const double* const (&c64)[D], const double (&q64)[D], double r, bool live,
                          int lo, int hi, int t0, int t1, int& cnt, double* stage) {
    const int S = cd.strip;
    const double* xk = col64 + cd.base_key;
    const double* yc = c64[1];
    const int s = (t0 - lo) / S;
    const int nstrip = (hi - lo + S - 1) / S;
    if (!wave_any(live)) return 0;
    int visited = 0;
    // Strip [ks, ke) for the lanes in `need`.
    auto visit = [&](int ks, int ke, bool need) {
        int a, b;
        if constexpr (D == 2) {
            const bool inner = need && fabs(xk[ks] - q64[0]) <= r && fabs(xk[ke - 1] - q64[0]) <= r;
            const bool edge = need && !inner;
            if (wave_any(inner)) {
                strip_window(yc, ks, ke, q64[1], r, inner, a, b);
                if (inner) cnt += count_sorted(yc, a, b, q64[1], r);
            }
            if (wave_any(edge)) {
                strip_window(yc, ks, ke, q64[1], r, edge, a, b);
                const int c = count_range64<D>(c64, q64, r, a, b);
                if (!inner) cnt += c;   // lanes outside `need` find no candidate here
                visited += b - a;
            }
        } else {
            strip_window(yc, ks, ke, q64[1], r, need, a, b);
            cnt += count_range_staged<D>(c64, q64, r, a, b, stage);
            visited += b - a;
        }
    };
    visit(lo + s * S, min(hi, lo + s * S + S), live);
    bool ldone = s == 0, rdone = s == nstrip - 1;
    for (int d = 1; !(ldone && rdone); ++d) {
        if (!ldone) {
            const int sl = s - d;
            if (sl < 0) {
                ldone = true;
            } else {
                const int le = lo + (sl + 1) * S;
                const double xb = xk[le - 1];
                const bool need = live && !(q64[0] - xb > r);
                ldone = !wave_any(need);
                if (!ldone) visit(le - S, le, need);
            }
        }
        if (!rdone) {
            const int sr = s + d;
            if (sr >= nstrip) {
                rdone = true;
            } else {
                const int rs = lo + sr * S;
                const double xb = xk[rs];
                const bool need = live && !(xb - q64[0] > r);
                rdone = !wave_any(need);
                if (!rdone) visit(rs, min(hi, rs + S), need);
            }
        }
    }
    return visited;
}

template <int D, bool F32>
__device__ void countnd_item(const CountArgs& A, const Item& it, const CountDev& cd,
                             const double* __restrict__ col64, const float* __restrict__ col32,
                             double* __restrict__ res64, int* __restrict__ res32,
                             unsigned long long& pairs, unsigned long long& rechecks, double* stage) {
    const OrderDev od = A.orders[cd.order];
    const int64_t n = od.n;
    const int i = it.t0 + (int)threadIdx.x;
    const bool active = i < it.t1;
    bool f32ok = true;
    const double g = enm_guard(A.maxabs[od.inst], &f32ok);

    const double* c64[D];
    const float* c32[D];
    double q64[D];
    float q32[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int c = d < cd.ncol ? cd.cols[d] : cd.cols[0];
        c64[d] = col64 + od.c64 + c * n;
        c32[d] = col32 + od.c32 + c * n;
        q64[d] = active ? c64[d][i] : 0.0;
        q32[d] = (float)q64[d];
    }
    double r = active ? query_radius(cd, A, res64, i) : -1.0;
    const bool live = active && !(r < 0.0);
    if (!live) r = -INFINITY;

    if constexpr (D >= 2) {
        if (cd.strip > 0) {
            int cnt = 0;
            const int visited = count_grid<D>(cd, col64, c64, q64, r, live, it.lo, it.hi, it.t0, it.t1, cnt, stage);
            finish_item(cd, it, active, cnt, res32, res64, i, A.perm, A.psi_tab);
            pairs += (unsigned long long)(it.t1 - it.t0) * (unsigned long long)visited;
            return;
        }
    }

    // Window on the key axis (a superset of every live lane's ball).
    int w0 = it.lo, w1 = it.hi;
    if (!A.brute) {
        const double key = q64[0];
        const double slack = 0x1p-50 * (fabs(key) + fabs(r));
        double lb = live ? (key - r * (1.0 + 0x1p-50)) - slack : INFINITY;
        double ub = live ? (key + r * (1.0 + 0x1p-50)) + slack : -INFINITY;
        lb = wave_min(lb);
        ub = wave_max(ub);
        if (lb > ub) { w0 = w1 = it.t0; }
        else {
            wave_bounds(c64[0], it.lo, it.hi, lb, ub, w0, w1);
        }
    }

    int cnt = 0;
    if (F32) {
        const float lo32 = live ? f32_down(r - g) : -INFINITY;
        const float hi32 = live ? f32_up(r + g) : -INFINITY;
        int chi = 0;
        int j = w0;
        for (; j + CH <= w1; j += CH) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                float m = fabsf(c32[0][j + c] - q32[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) m = fmaxf(m, fabsf(c32[d][j + c] - q32[d]));
                cnt += (m <= lo32);
                chi += (m <= hi32);
            }
        }
        for (; j < w1; ++j) {
            float m = fabsf(c32[0][j] - q32[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmaxf(m, fabsf(c32[d][j] - q32[d]));
            cnt += (m <= lo32);
            chi += (m <= hi32);
        }
        unsigned rc = 0;
        if (wave_any(chi != cnt)) {
            const bool band = chi != cnt;
            for (j = w0; j < w1; ++j) {
                float m = fabsf(c32[0][j] - q32[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) m = fmaxf(m, fabsf(c32[d][j] - q32[d]));
                if (band && m > lo32 && m <= hi32) {
                    double e = fabs(c64[0][j] - q64[0]);
#pragma unroll
                    for (int d = 1; d < D; ++d) e = fmax(e, fabs(c64[d][j] - q64[d]));
                    cnt += (e <= r);
                    ++rc;
                }
            }
        }
        rechecks += wave_sum_u64(rc);
    } else {
        int j = w0;
        for (; j + CH <= w1; j += CH) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                double m = fabs(c64[0][j + c] - q64[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j + c] - q64[d]));
                cnt += (m <= r);
            }
        }
        for (; j < w1; ++j) {
            double m = fabs(c64[0][j] - q64[0]);
#pragma unroll
            for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j] - q64[d]));
            cnt += (m <= r);
        }
    }
    finish_item(cd, it, active, cnt, res32, res64, i, A.perm, A.psi_tab);
    pairs += (unsigned long long)(it.t1 - it.t0) * (unsigned long long)(w1 - w0);
}

template <int D, bool F32>
__global__ __launch_bounds__(ENM_WAVE)
void countnd_kernel(CountArgs A, const double* __restrict__ col64, const float* __restrict__ col32,
                    double* __restrict__ res64, int* __restrict__ res32,
                    unsigned long long* __restrict__ stats) {
    // Counters accumulate per wave and go out once (see knn_kernel).
    unsigned long long pairs = 0, rechecks = 0;
    __shared__ double stage[D >= 3 ? D * ENM_WAVE : 1];   // grid candidates of three or more columns
    for_each_wave_item(A.qoff, A.counter, A.chunk, [&](int idx) {
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        const CountDev cd = A.descs[it.desc];
        if (F32) {
            bool ok = true;
            (void)enm_guard(A.maxabs[A.orders[cd.order].inst], &ok);
            if (ok) countnd_item<D, true>(A, it, cd, col64, col32, res64, res32, pairs, rechecks, stage);
            else    countnd_item<D, false>(A, it, cd, col64, col32, res64, res32, pairs, rechecks, stage);
        } else {
            countnd_item<D, false>(A, it, cd, col64, col32, res64, res32, pairs, rechecks, stage);
        }
    });
    if (threadIdx.x == 0) {
        atomicAdd(stats + ST_CNT_PAIRS, pairs);
        if (F32) atomicAdd(stats + ST_CNT_RECHECK, rechecks);
    }
}

template <int D>
hipError_t launch_nd(const CountArgs& a, bool f32, int grid, const double* col64, const float* col32,
                     double* res64, int* res32, unsigned long long* stats, hipStream_t st) {
    if (f32)
        hipLaunchKernelGGL((countnd_kernel<D, true>), dim3(grid), dim3(ENM_WAVE), 0, st,
                           a, col64, col32, res64, res32, stats);
    else
        hipLaunchKernelGGL((countnd_kernel<D, false>), dim3(grid), dim3(ENM_WAVE), 0, st,
                           a, col64, col32, res64, res32, stats);
    return hipGetLastError();
}

}  // namespace

int enm_count_dim_class(int d) {
    if (d <= 6) return d;
    if (d <= 8) return 8;
    if (d <= 12) return 12;
    if (d <= 16) return 16;
    return -1;
}

hipError_t enm_launch_count(const Item* items, const int* qoff, int n_items, unsigned* counter,
                            const CountDev* descs, const OrderDev* orders,
                            const unsigned long long* maxabs, const int* perm, const double* psi_tab,
                            const double* col64, const float* col32, double* res64, int* res32,
                            unsigned long long* stats, int dclass, bool f32, bool brute,
                            int grid, unsigned* deferred, unsigned* n_deferred, unsigned seg_cap,
                            int ticket, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    // Tickets of up to ENM_KCHUNK items, at least four per wave of the grid.
    // Tickets of up to ENM_KCHUNK items (or the tuned items per ticket), but
    // at least four per wave of the grid, so small launches still fill it.
    const int chunk = std::max(1, std::min(ticket > 0 ? ticket : ENM_KCHUNK, n_items / (4 * std::max(1, grid))));
    CountArgs a{items, qoff, counter, chunk, descs, orders, maxabs, perm, psi_tab, brute ? 1 : 0};
    if (dclass == 1) {
        // `grid` counts waves; a 1-D workgroup has ENM_C1_THREADS / 64.
        const int blocks = std::max(1, std::min(n_items, grid / (ENM_C1_THREADS / ENM_WAVE)));
        hipLaunchKernelGGL(count1d_kernel, dim3(blocks), dim3(ENM_C1_THREADS), 0, st, a, col64, res64, res32,
                           stats, deferred, n_deferred, seg_cap);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(count1d_deferred_kernel, dim3(std::max(1, grid / 4)), dim3(256), 0, st, a, col64,
                           res64, res32, deferred, n_deferred, seg_cap);
        return hipGetLastError();
    }
    switch (dclass) {
        case 2: return launch_nd<2>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 3: return launch_nd<3>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 4: return launch_nd<4>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 5: return launch_nd<5>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 6: return launch_nd<6>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 8: return launch_nd<8>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 12: return launch_nd<12>(a, f32, grid, col64, col32, res64, res32, stats, st);
        case 16: return launch_nd<16>(a, f32, grid, col64, col32, res64, res32, stats, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t enm_launch_psi_table(double* tab, int64_t n, hipStream_t st) {
    const int threads = 256;
    const int64_t blocks = (n + threads) / threads;
    hipLaunchKernelGGL(psi_table_kernel, dim3((unsigned)blocks), dim3(threads), 0, st, tab, n);
    return hipGetLastError();
}

hipError_t enm_launch_psi_tiles(const Item* items, int n_items, const CountDev* descs, const double* psi_tab,
                                const int* counts, double* res64, int* res32, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(psi_tile_kernel, dim3(n_items), dim3(ENM_WAVE), 0, st, items, n_items, descs, psi_tab,
                       counts, res64, res32);
    return hipGetLastError();
}
