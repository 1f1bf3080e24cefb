// Chebyshev (max-norm) k-th-neighbour radius, replacing
// cKDTree.query(pts, k=[k+1], p=np.inf) of the reference
// (ennemi/_entropy_estimators.py:39,108,142,194,240).
//
// eps_i is the (k+1)-th smallest D(i,j) = max_d fl64(|a_d[i] - a_d[j]|) over
// all j of the query's scope range, i itself included (distance 0); fewer
// than k+1 points leave it +inf, as cKDTree does.
//
// Shape.  A work item is one wave and a tile of 64 consecutive positions of
// the key-sorted order (one query per lane, coordinates in VGPRs).  The wave
// walks candidates in wave-uniform order, so candidate coordinates are
// wave-uniform loads served by the scalar cache and used as SGPR operands;
// no LDS is needed.  Each lane keeps an ascending FP64 list of its K1 = k+1
// smallest distances in registers (P slots, the lowest P-K1 pinned at -inf).
//
// Pruning.  Candidates are visited from the tile outward in key order,
// alternating left and right in steps of 64.  A side stops once every lane's
// key distance to the last visited candidate is >= its current K1-th
// distance T: every candidate further out on that side has a key distance,
// and so a Chebyshev distance, >= T and cannot enter the list.  With
// ENM_F_BRUTE the whole scope range is scanned instead.
//
// Precision.  In FP64 mode every distance is exact FP64.  In FP32-filter mode
// the bulk distance is computed from FP32 copies, and a candidate is
// re-evaluated in FP64 only when its FP32 distance is <= fl32_up(T + g), with
// the per-instance guard g bounding |d32 - d64|.  T only decreases, so one
// pass is exact: radii are bit-identical to the FP64 mode.
//
// Work distribution.  A persistent grid pops items from an atomic counter in
// the (largest-first) order the host laid them out.
#include "enm_device.h"
#include "enm_launch.h"

namespace {

constexpr int CH = 8;     // candidates per filter chunk
constexpr int SWEEP = 64; // candidates per side per sweep step

struct KnnArgs {
    const Item* items;
    int n_items;
    unsigned* counter;
    const KnnDev* descs;
    const OrderDev* orders;
    const unsigned long long* maxabs;
    const int* perm;
    int brute;
};

template <int D, int P, bool F32>
struct KnnTile {
    const double* c64[D];
    const float* c32[D];
    double q64[D];
    float q32[D];
    double L[P];
    double T;
    float thr;
    double guard;
    unsigned rechecks;

    __device__ __forceinline__ void set_thr() {
        if (F32) thr = f32_up(T + guard);
    }

    __device__ __forceinline__ void recheck(int j) {
        double e = fabs(c64[0][j] - q64[0]);
#pragma unroll
        for (int d = 1; d < D; ++d) e = fmax(e, fabs(c64[d][j] - q64[d]));
        ++rechecks;
        if (e < T) {
            list_insert<P>(L, e);
            T = L[P - 1];
            set_thr();
        }
    }

    __device__ __forceinline__ void chunk(int j) {
        if (F32) {
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
            if (wave_any(mn <= thr)) {
#pragma unroll
                for (int c = 0; c < CH; ++c)
                    if (dd[c] <= thr) recheck(j + c);
            }
        } else {
            double dd[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                double m = fabs(c64[0][j + c] - q64[0]);
#pragma unroll
                for (int d = 1; d < D; ++d) m = fmax(m, fabs(c64[d][j + c] - q64[d]));
                dd[c] = m;
            }
            double mn = dd[0];
#pragma unroll
            for (int c = 1; c < CH; ++c) mn = fmin(mn, dd[c]);
            if (wave_any(mn < T)) {
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    if (dd[c] < T) {
                        list_insert<P>(L, dd[c]);
                        T = L[P - 1];
                    }
                }
            }
        }
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
                list_insert<P>(L, m);
                T = L[P - 1];
            }
        }
    }

    // Process the wave-uniform candidate range [a, b).
    __device__ __forceinline__ void range(int a, int b) {
        int j = a;
        for (; j + CH <= b; j += CH) chunk(j);
        for (; j < b; ++j) single(j);
    }
};

template <int D, int P, bool F32>
__device__ void knn_item(const KnnArgs& A, const Item& it, const double* __restrict__ col64,
                         const float* __restrict__ col32, double* __restrict__ res64,
                         unsigned long long* __restrict__ stats) {
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
#pragma unroll
    for (int t = 0; t < P; ++t) w.L[t] = (t < P - K1) ? -INFINITY : INFINITY;
    w.T = active ? INFINITY : -INFINITY;
    w.thr = active ? INFINITY : -INFINITY;
    w.rechecks = 0;

    const int lo = it.lo, hi = it.hi, t0 = it.t0, t1 = it.t1;
    int visited = t1 - t0;
    w.range(t0, t1);
    if (A.brute) {
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
    const double eps = w.L[P - 1];
    if (active) {
        res64[kd.eps + i] = eps;
        if (kd.eps_orig >= 0) res64[kd.eps_orig + A.perm[od.perm + i]] = eps;
    }
    if (kd.kl) {
        const double s = wave_sum(active ? log(eps) : 0.0);
        if (lane == 0) res64[kd.part + it.tile] = s;
    }
    const unsigned long long nact = (unsigned long long)(t1 - t0);
    const unsigned long long rc = wave_sum_u64(w.rechecks);
    if (lane == 0) {
        atomicAdd(stats + ST_KNN_PAIRS, nact * (unsigned long long)visited);
        if (F32) atomicAdd(stats + ST_KNN_RECHECK, rc);
    }
}

template <int D, int P, bool F32>
__global__ __launch_bounds__(ENM_WAVE)
void knn_kernel(KnnArgs A, const double* __restrict__ col64, const float* __restrict__ col32,
                double* __restrict__ res64, unsigned long long* __restrict__ stats) {
    for (;;) {
        const unsigned idx = pop_ticket(A.counter);
        if (idx >= (unsigned)A.n_items) break;
        Item it = A.items[idx];
        it.desc = uni(it.desc); it.tile = uni(it.tile);
        it.lo = uni(it.lo); it.hi = uni(it.hi); it.t0 = uni(it.t0); it.t1 = uni(it.t1);
        if (F32) {
            const KnnDev& kd = A.descs[it.desc];
            bool ok = true;
            (void)enm_guard(A.maxabs[A.orders[kd.order].inst], &ok);
            if (ok) knn_item<D, P, true>(A, it, col64, col32, res64, stats);
            else    knn_item<D, P, false>(A, it, col64, col32, res64, stats);
        } else {
            knn_item<D, P, false>(A, it, col64, col32, res64, stats);
        }
    }
}

template <int D, int P>
hipError_t launch_dp(const KnnArgs& a, bool f32, int grid, const double* col64, const float* col32,
                     double* res64, unsigned long long* stats, hipStream_t st) {
    if (f32)
        hipLaunchKernelGGL((knn_kernel<D, P, true>), dim3(grid), dim3(ENM_WAVE), 0, st,
                           a, col64, col32, res64, stats);
    else
        hipLaunchKernelGGL((knn_kernel<D, P, false>), dim3(grid), dim3(ENM_WAVE), 0, st,
                           a, col64, col32, res64, stats);
    return hipGetLastError();
}

template <int D>
hipError_t launch_d(const KnnArgs& a, int P, bool f32, int grid, const double* col64,
                    const float* col32, double* res64, unsigned long long* stats, hipStream_t st) {
    switch (P) {
        case 4: return launch_dp<D, 4>(a, f32, grid, col64, col32, res64, stats, st);
        case 8: return launch_dp<D, 8>(a, f32, grid, col64, col32, res64, stats, st);
        case 16: return launch_dp<D, 16>(a, f32, grid, col64, col32, res64, stats, st);
        case 32: return launch_dp<D, 32>(a, f32, grid, col64, col32, res64, stats, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int enm_knn_dim_class(int d) {
    if (d <= 6) return d;
    if (d <= 8) return 8;
    if (d <= 12) return 12;
    if (d <= 16) return 16;
    return -1;
}

int enm_knn_list_class(int k1) {
    if (k1 <= 4) return 4;
    if (k1 <= 8) return 8;
    if (k1 <= 16) return 16;
    if (k1 <= 32) return 32;
    return -1;
}

hipError_t enm_launch_knn(const Item* items, int n_items, unsigned* counter,
                          const KnnDev* descs, const OrderDev* orders,
                          const unsigned long long* maxabs, const int* perm,
                          const double* col64, const float* col32, double* res64,
                          unsigned long long* stats, int dclass, int pclass, bool f32, bool brute,
                          int grid, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    KnnArgs a{items, n_items, counter, descs, orders, maxabs, perm, brute ? 1 : 0};
    switch (dclass) {
        case 1: return launch_d<1>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 2: return launch_d<2>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 3: return launch_d<3>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 4: return launch_d<4>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 5: return launch_d<5>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 6: return launch_d<6>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 8: return launch_d<8>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 12: return launch_d<12>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        case 16: return launch_d<16>(a, pclass, f32, grid, col64, col32, res64, stats, st);
        default: return hipErrorInvalidValue;
    }
}
