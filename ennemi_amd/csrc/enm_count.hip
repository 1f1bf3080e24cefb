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
// FP64 predicate the reference evaluates, so counts are exact; a boundary
// beyond the staged keys is finished in global memory with 8 independent
// probes per round.  Algorithmic traffic: 8 B radius + 8 B key + 4 B count
// per query.  In every mode (including ENM_F_BRUTE, which only removes window
// pruning) 1-D counts use this search.
//
// Multi-D spaces: on a grid order (the space's first two columns are the
// order's columns 0 and 1) a wave takes 64 queries of one strip and scans its
// own strip and the strips outward, each within the y window of the tile
// widened by the wave's largest radius, in FP64 (count_grid).  Otherwise the
// wave bounds the union of its 64 balls on the key axis and scans that window
// of the key-sorted order with wave-uniform candidate loads; in FP32-filter
// mode each lane counts candidates whose FP32 distance is <= the radius
// rounded down minus the guard (certainly inside) and <= the radius rounded
// up plus the guard (possibly inside), and if any lane saw a candidate in
// between the wave rescans its window and settles those in FP64.  (On this
// ISA FP32 and FP64 compares issue at the same rate, so the grid path counts
// in FP64 directly: the band bookkeeping would cost more than it saves.)
//
// Each item also writes the sum over its 64 queries of psi(n_i) and a flag
// for n_i == 0 (psi(0) = +inf in the reference, :338-339).
#include "enm_device.h"
#include "enm_launch.h"

namespace {

constexpr int CH = 8;

struct CountArgs {
    const Item* items;
    int n_items;
    unsigned* counter;
    const CountDev* descs;
    const OrderDev* orders;
    const unsigned long long* maxabs;
    const int* perm;
    int brute;
};

__device__ __forceinline__ double query_radius(const CountDev& cd, const CountArgs& A,
                                               const double* __restrict__ res64, int i) {
    const double eps = cd.same_order ? res64[cd.eps + i] : res64[cd.eps + A.perm[cd.eps_perm + i]];
    return eps - 1e-12;
}

__device__ __forceinline__ void finish_item(const CountDev& cd, const Item& it, bool active,
                                            int cnt, int* __restrict__ res32,
                                            double* __restrict__ res64, int i, const int* perm) {
    if (active) res32[cd.counts + i] = cnt;
    if (cd.psi_pt >= 0) {
        // Point order; a zero count is +inf, as psi(0) poisons the mean.
        if (active) res64[cd.psi_pt + perm[cd.eps_perm + i]] = cnt > 0 ? enm_psi((double)cnt) : INFINITY;
        return;
    }
    const double p = (active && cnt > 0) ? enm_psi((double)cnt) : 0.0;
    const double s = wave_sum(p);
    const int z = wave_or(active && cnt == 0);
    if (threadIdx.x == 0) {
        res64[cd.part + it.tile] = s;
        res32[cd.zero + it.tile] = z;
    }
}

// Search state of one side of a 1-D ball: P(good) holds, P(bad) fails (or
// bad is the sentinel just outside the scope range), where for the right
// side P(j) = fl(v[j] - x) <= r and for the left side P(j) = fl(x - v[j]) <= r.
// Each round probes 8 positions whose loads are independent, so a lane waits
// on about 5 rounds of memory latency instead of ~20 dependent probes:
// first an exponential phase (offsets step*2^m, m = 0..7, step growing by
// 256 while every probe holds), then an 8-ary refinement of (good, bad).
struct Side {
    int good, bad;   // positions; direction is +1 (right) or -1 (left)
    long long step;
    bool expo;
};

template <int DIR>
__device__ __forceinline__ bool side_pred(const double* __restrict__ v, int j, double x, double r) {
    return DIR > 0 ? (v[j] - x <= r) : (x - v[j] <= r);
}

template <int DIR>
__device__ __forceinline__ void side_round(Side& S, const double* __restrict__ v, int lo, int hi,
                                           double x, double r) {
    if (S.expo) {
        int pos[8];
        bool ok[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const long long p = (long long)S.good + DIR * S.step * (1ll << m);
            const bool in = DIR > 0 ? (p < hi) : (p >= lo);
            pos[m] = in ? (int)p : S.bad;
            ok[m] = in && side_pred<DIR>(v, pos[m], x, r);
        }
        int first_fail = 8;
#pragma unroll
        for (int m = 7; m >= 0; --m) if (!ok[m]) first_fail = m;
        if (first_fail == 8) {
            S.good = pos[7];
            S.step <<= 8;
        } else {
            if (first_fail > 0) S.good = pos[first_fail - 1];
            S.bad = pos[first_fail];
            S.expo = false;
        }
        return;
    }
    const int L = DIR > 0 ? S.bad - S.good : S.good - S.bad;
    if (L <= 1) return;
    int last_ok = S.good, first_fail = S.bad;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        // Offsets ceil((m+1) L / 9), strictly inside (good, bad).
        const int off = min(L - 1, (int)(((long long)(m + 1) * L + 8) / 9));
        const int q = S.good + DIR * off;
        const bool okm = side_pred<DIR>(v, q, x, r);
        if (okm && (DIR > 0 ? q > last_ok : q < last_ok)) last_ok = q;
        if (!okm && (DIR > 0 ? q < first_fail : q > first_fail)) first_fail = q;
    }
    S.good = last_ok;
    S.bad = first_fail;
}

// Boundary search inside an LDS copy w[0, nw) of the window: last position
// in [i, nw) with fl(w[j] - x) <= r (right) or first in [0, i] with
// fl(x - w[j]) <= r (left); position i holds x itself.
__device__ __forceinline__ int lds_right(const double* w, int i, int nw, double x, double r) {
    int good = i, bad = nw;
    while (bad - good > 1) {
        const int mid = (good + bad) >> 1;
        if (w[mid] - x <= r) good = mid; else bad = mid;
    }
    return good;
}
__device__ __forceinline__ int lds_left(const double* w, int i, double x, double r) {
    int good = i, bad = -1;
    while (good - bad > 1) {
        const int mid = (good + bad) >> 1;
        if (x - w[mid] <= r) good = mid; else bad = mid;
    }
    return good;
}

// Finish a boundary search in global memory from a known state (the 8-probe
// rounds of side_round).
template <int DIR>
__device__ __forceinline__ int global_side(const double* __restrict__ v, int good, int bad, int lo,
                                           int hi, double x, double r) {
    Side S{good, bad, 1, true};
    for (;;) {
        const bool done = !S.expo && (DIR > 0 ? S.bad - S.good : S.good - S.bad) <= 1;
        if (done) break;
        side_round<DIR>(S, v, lo, hi, x, r);
    }
    return S.good;
}

// 1-D counts.  An item is ENM_C1_ITEM consecutive positions of a key-sorted
// order, ENM_C1_Q per lane.  The wave copies the keys of the item and
// ENM_C1_HALO positions on each side into LDS with coalesced loads, and each
// query binary-searches both boundaries there.  A boundary that lies beyond
// the staged keys (the query's predicate still holds at the window's edge)
// is finished from that edge in global memory, so a few wide balls do not
// widen the window of the whole item.
__global__ __launch_bounds__(ENM_WAVE)
void count1d_kernel(CountArgs A, const double* __restrict__ col64, double* __restrict__ res64,
                    int* __restrict__ res32, unsigned long long* __restrict__ stats) {
    __shared__ double win[ENM_C1_CAP];
    const int lane = threadIdx.x;
    for (;;) {
        const unsigned idx = pop_ticket(A.counter);
        if (idx >= (unsigned)A.n_items) break;
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        const CountDev cd = A.descs[it.desc];
        const OrderDev od = A.orders[cd.order];
        const double* v = col64 + od.c64;  // sorted key column
        const int w0 = max(it.lo, it.t0 - ENM_C1_HALO);
        const int w1 = min(it.hi, it.t1 + ENM_C1_HALO);
        const int nw = w1 - w0;
        for (int t = lane; t < nw; t += ENM_WAVE) win[t] = v[w0 + t];
        __syncthreads();
        double ps = 0.0;
        bool zero = false;
#pragma unroll
        for (int q = 0; q < ENM_C1_Q; ++q) {
            const int i = it.t0 + lane + ENM_WAVE * q;
            if (i < it.t1) {
                const double r = query_radius(cd, A, res64, i);
                int cnt = 0;
                if (!(r < 0.0)) {
                    const int p = i - w0;
                    const double x = win[p];
                    int right, left;
                    if (w1 < it.hi && win[nw - 1] - x <= r)
                        right = global_side<1>(v, w1 - 1, it.hi, it.lo, it.hi, x, r);
                    else
                        right = w0 + lds_right(win, p, nw, x, r);
                    if (w0 > it.lo && x - win[0] <= r)
                        left = global_side<-1>(v, w0, it.lo - 1, it.lo, it.hi, x, r);
                    else
                        left = w0 + lds_left(win, p, x, r);
                    cnt = right - left + 1;
                }
                res32[cd.counts + i] = cnt;
                if (cnt > 0) ps += enm_psi((double)cnt);
                else zero = true;
            }
        }
        __syncthreads();
        const double s = wave_sum(ps);
        const int z = wave_or(zero);
        if (lane == 0) {
            res64[cd.part + it.tile] = s;
            res32[cd.zero + it.tile] = z;
            atomicAdd(stats + ST_CNT_1D, (unsigned long long)(it.t1 - it.t0));
        }
    }
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

// Grid count: the tile's own strip and then strips outward, each within the
// y window of the tile's column-1 range widened by the wave's largest radius
// R.  A side stops at the first strip whose column-0 bound is more than every
// live lane's radius away (strictly: a count includes distance == r), and
// all strips beyond it are further still.  Returns candidates visited.
template <int D>
__device__ int count_grid(const CountDev& cd, const double* __restrict__ col64,
                          const double* const (&c64)[D], const double (&q64)[D], double r, bool live,
                          int lo, int hi, int t0, int t1, int& cnt) {
    const int S = cd.strip;
    const double* xk = col64 + cd.base_key;
    const double* yc = c64[1];
    const int s = (t0 - lo) / S;
    const int nstrip = (hi - lo + S - 1) / S;
    const int ss = lo + s * S, se = min(hi, ss + S);
    const double R = uni_f64(wave_max(live ? r : -INFINITY));
    if (!(R >= 0.0)) return 0;  // no live lane
    const double ylo = yc[t0], yhi = yc[t1 - 1];
    int visited = 0;
    int a, b;
    {
        const double slack = 0x1p-50 * (fmax(fabs(ylo), fabs(yhi)) + R);
        const double lb = uni_f64((ylo - R * (1.0 + 0x1p-50)) - slack);
        const double ub = uni_f64((yhi + R * (1.0 + 0x1p-50)) + slack);
        a = uni(lower_bound_u(yc, ss, se, lb));
        b = uni(upper_bound_u(yc, a, se, ub));
        cnt += count_range64<D>(c64, q64, r, a, b);
        visited += b - a;
        bool ldone = s == 0, rdone = s == nstrip - 1;
        for (int d = 1; !(ldone && rdone); ++d) {
            if (!ldone) {
                const int sl = s - d;
                if (sl < 0) {
                    ldone = true;
                } else {
                    const int le = lo + (sl + 1) * S;
                    const double xb = xk[le - 1];
                    ldone = wave_all(!live || (q64[0] - xb > r));
                    if (!ldone) {
                        a = uni(lower_bound_u(yc, le - S, le, lb));
                        b = uni(upper_bound_u(yc, a, le, ub));
                        cnt += count_range64<D>(c64, q64, r, a, b);
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
                    const int re = min(hi, rs + S);
                    const double xb = xk[rs];
                    rdone = wave_all(!live || (xb - q64[0] > r));
                    if (!rdone) {
                        a = uni(lower_bound_u(yc, rs, re, lb));
                        b = uni(upper_bound_u(yc, a, re, ub));
                        cnt += count_range64<D>(c64, q64, r, a, b);
                        visited += b - a;
                    }
                }
            }
        }
    }
    return visited;
}

template <int D, bool F32>
__device__ void countnd_item(const CountArgs& A, const Item& it, const CountDev& cd,
                             const double* __restrict__ col64, const float* __restrict__ col32,
                             double* __restrict__ res64, int* __restrict__ res32,
                             unsigned long long* __restrict__ stats) {
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
            const int visited = count_grid<D>(cd, col64, c64, q64, r, live, it.lo, it.hi, it.t0, it.t1, cnt);
            finish_item(cd, it, active, cnt, res32, res64, i, A.perm);
            if (threadIdx.x == 0)
                atomicAdd(stats + ST_CNT_PAIRS, (unsigned long long)(it.t1 - it.t0) * (unsigned long long)visited);
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
            w0 = uni(lower_bound_u(c64[0], it.lo, it.hi, lb));
            w1 = uni(upper_bound_u(c64[0], w0, it.hi, ub));
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
        const unsigned long long rcs = wave_sum_u64(rc);
        if (threadIdx.x == 0) atomicAdd(stats + ST_CNT_RECHECK, rcs);
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
    finish_item(cd, it, active, cnt, res32, res64, i, A.perm);
    if (threadIdx.x == 0)
        atomicAdd(stats + ST_CNT_PAIRS, (unsigned long long)(it.t1 - it.t0) * (unsigned long long)(w1 - w0));
}

template <int D, bool F32>
__global__ __launch_bounds__(ENM_WAVE)
void countnd_kernel(CountArgs A, const double* __restrict__ col64, const float* __restrict__ col32,
                    double* __restrict__ res64, int* __restrict__ res32,
                    unsigned long long* __restrict__ stats) {
    for (;;) {
        const unsigned idx = pop_ticket(A.counter);
        if (idx >= (unsigned)A.n_items) break;
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        const CountDev cd = A.descs[it.desc];
        if (F32) {
            bool ok = true;
            (void)enm_guard(A.maxabs[A.orders[cd.order].inst], &ok);
            if (ok) countnd_item<D, true>(A, it, cd, col64, col32, res64, res32, stats);
            else    countnd_item<D, false>(A, it, cd, col64, col32, res64, res32, stats);
        } else {
            countnd_item<D, false>(A, it, cd, col64, col32, res64, res32, stats);
        }
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

hipError_t enm_launch_count(const Item* items, int n_items, unsigned* counter,
                            const CountDev* descs, const OrderDev* orders,
                            const unsigned long long* maxabs, const int* perm,
                            const double* col64, const float* col32, double* res64, int* res32,
                            unsigned long long* stats, int dclass, bool f32, bool brute,
                            int grid, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    CountArgs a{items, n_items, counter, descs, orders, maxabs, perm, brute ? 1 : 0};
    if (dclass == 1) {
        hipLaunchKernelGGL(count1d_kernel, dim3(grid), dim3(ENM_WAVE), 0, st, a, col64, res64, res32, stats);
        return hipGetLastError();
    }
    switch (dclass) {
        case 1: return launch_nd<1>(a, f32, grid, col64, col32, res64, res32, stats, st);
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
