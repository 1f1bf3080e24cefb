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
// doubling run width merge neighbouring runs inside each segment.  A merge
// block owns ENM_SORT_TILE outputs: two lanes find the merge-path split of the
// block's first and last output by binary search, the block stages the two
// input slices in LDS, and each thread merges 8 outputs.  Keys compare as
// (value, index) pairs, so all keys are distinct and the order is total.
// Traffic per pass is 12 B read + 12 B written per key (HBM-bound).
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

__global__ __launch_bounds__(ENM_SORT_THREADS)
void block_sort_kernel(const SortSeg* __restrict__ segs, int nsegs,
                       const double* __restrict__ arena,
                       double* __restrict__ kout, int* __restrict__ iout) {
    __shared__ double sk[ENM_SORT_TILE];
    __shared__ int si[ENM_SORT_TILE];
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    const int off = (tile - sg.first_tile) * ENM_SORT_TILE;
    const int len = min(ENM_SORT_TILE, sg.len - off);
    for (int t = threadIdx.x; t < ENM_SORT_TILE; t += ENM_SORT_THREADS) {
        if (t < len) {
            sk[t] = arena[sg.src + off + t];
            si[t] = sg.p0 + off + t;
        } else {
            sk[t] = INFINITY;
            si[t] = 0x7fffffff;
        }
    }
    __syncthreads();
    for (int size = 2; size <= ENM_SORT_TILE; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = threadIdx.x; q < ENM_SORT_TILE / 2; q += ENM_SORT_THREADS) {
                const int i = 2 * q - (q & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const double ki = sk[i], kj = sk[j];
                const int ii = si[i], ij = si[j];
                if (kless(kj, ij, ki, ii) == asc) {
                    sk[i] = kj; sk[j] = ki;
                    si[i] = ij; si[j] = ii;
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        kout[sg.sbase + off + t] = sk[t];
        iout[sg.sbase + off + t] = si[t];
    }
}

// Number of A elements among the first d outputs of merge(A, B).
__device__ int co_rank(int d, const double* ka, const int* ia, int na,
                       const double* kb, const int* ib, int nb) {
    int lo = max(0, d - nb), hi = min(d, na);
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int j = d - i - 1;
        if (kless(ka[i], ia[i], kb[j], ib[j])) lo = i + 1; else hi = i;
    }
    return lo;
}

__global__ __launch_bounds__(ENM_SORT_THREADS)
void merge_pass_kernel(const SortSeg* __restrict__ segs, int nsegs, int64_t width,
                       const double* __restrict__ kin, const int* __restrict__ iin,
                       double* __restrict__ kout, int* __restrict__ iout) {
    __shared__ double sk[ENM_SORT_TILE];
    __shared__ int si[ENM_SORT_TILE];
    __shared__ int split[2];
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    const int64_t off = (int64_t)(tile - sg.first_tile) * ENM_SORT_TILE;
    const int64_t len = sg.len;
    const int64_t pb = (off / (2 * width)) * (2 * width);
    const int na = (int)min(width, len - pb);
    const int nb = (int)max((int64_t)0, min(2 * width, len - pb) - width);
    const int d0 = (int)(off - pb);
    const int d1 = min(d0 + ENM_SORT_TILE, na + nb);
    const int64_t base = sg.sbase + pb;
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
    if (threadIdx.x < 2) {
        const int d = threadIdx.x == 0 ? d0 : d1;
        split[threadIdx.x] = co_rank(d, ka, ia, na, kb, ib, nb);
    }
    __syncthreads();
    const int i0 = split[0], i1 = split[1];
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int la = i1 - i0, lb = j1 - j0;
    for (int t = threadIdx.x; t < la + lb; t += ENM_SORT_THREADS) {
        if (t < la) { sk[t] = ka[i0 + t]; si[t] = ia[i0 + t]; }
        else        { sk[t] = kb[j0 + t - la]; si[t] = ib[j0 + t - la]; }
    }
    __syncthreads();
    const int per = ENM_SORT_TILE / ENM_SORT_THREADS;
    const int e0 = threadIdx.x * per;
    const int e1 = min(e0 + per, la + lb);
    if (e0 < e1) {
        int a = co_rank(e0, sk, si, la, sk + la, si + la, lb);
        int b = e0 - a;
        for (int e = e0; e < e1; ++e) {
            bool take_a;
            if (a >= la) take_a = false;
            else if (b >= lb) take_a = true;
            else take_a = kless(sk[a], si[a], sk[la + b], si[la + b]);
            const int src = take_a ? a : la + b;
            kout[base + d0 + e] = sk[src];
            iout[base + d0 + e] = si[src];
            if (take_a) ++a; else ++b;
        }
    }
}

// Lay one sorted chunk out into an order: permutation, key column and the
// other gathered columns (FP64 and FP32), and fold |value| into the
// instance's maximum magnitude (for the FP32 guard band).  For a grid order
// the sort payload is a base-order position: the point is read through the
// base permutation and every column, key included, is gathered from the
// arena.
__global__ __launch_bounds__(ENM_SORT_THREADS)
void gather_kernel(const GatherItem* __restrict__ items,
                   const SortSeg* __restrict__ segs,
                   const OrderDev* __restrict__ orders,
                   const double* __restrict__ arena,
                   const double* __restrict__ skey, const int* __restrict__ sidx,
                   double* __restrict__ col64, float* __restrict__ col32,
                   int* perm, unsigned long long* __restrict__ maxabs) {
    __shared__ double red[ENM_SORT_THREADS / ENM_WAVE];
    const GatherItem it = items[blockIdx.x];
    const SortSeg sg = segs[it.sseg];
    const OrderDev od = orders[it.order];
    const int len = min(ENM_SORT_TILE, sg.len - it.off);
    const int64_t n = od.n;
    const bool indirect = od.base_perm >= 0;
    double m = 0.0;
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const int64_t sp = sg.sbase + it.off + t;
        const int64_t p = it.pos0 + it.off + t;
        const int idx = sidx[sp];
        const int pt = indirect ? perm[od.base_perm + idx] : idx;
        perm[od.perm + p] = pt;
        int c0 = 0;
        if (!indirect) {
            const double key = skey[sp];
            col64[od.c64 + p] = key;
            col32[od.c32 + p] = (float)key;
            m = fmax(m, fabs(key));
            c0 = 1;
        }
        for (int c = c0; c < od.ncol; ++c) {
            const double v = arena[od.src[c] + pt];
            col64[od.c64 + c * n + p] = v;
            col32[od.c32 + c * n + p] = (float)v;
            m = fmax(m, fabs(v));
        }
    }
    m = wave_max(m);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) red[threadIdx.x / ENM_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = red[0];
        for (int w = 1; w < ENM_SORT_THREADS / ENM_WAVE; ++w) mm = fmax(mm, red[w]);
        // |v| >= 0, so the IEEE bit patterns order like the values.
        atomicMax(maxabs + od.inst, (unsigned long long)__double_as_longlong(mm));
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

hipError_t enm_launch_sort(const SortSeg* d_segs, int nsegs, int ntiles, int64_t maxlen,
                           const double* arena, double* k0, int* i0, double* k1, int* i1,
                           int* final_buf, int* n_launch, hipStream_t st) {
    *final_buf = 0;
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_sort_kernel, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                       d_segs, nsegs, arena, k0, i0);
    ++*n_launch;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int cur = 0;
    for (int64_t w = ENM_SORT_TILE; w < maxlen; w *= 2) {
        const double* kin = cur ? k1 : k0;
        const int* iin = cur ? i1 : i0;
        double* kout = cur ? k0 : k1;
        int* iout = cur ? i0 : i1;
        hipLaunchKernelGGL(merge_pass_kernel, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, w, kin, iin, kout, iout);
        ++*n_launch;
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur ^= 1;
    }
    *final_buf = cur;
    return hipSuccess;
}

hipError_t enm_launch_gather(const GatherItem* d_items, int nitems, const SortSeg* d_segs,
                             const OrderDev* d_orders, const double* arena,
                             const double* skey, const int* sidx,
                             double* col64, float* col32, int* perm,
                             unsigned long long* maxabs, int* n_launch, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(nitems), dim3(ENM_SORT_THREADS), 0, st,
                       d_items, d_segs, d_orders, arena, skey, sidx, col64, col32, perm, maxabs);
    ++*n_launch;
    return hipGetLastError();
}
