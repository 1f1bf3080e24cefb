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

// Sort PER (key, index) pairs held in registers (PER = 2, 4 or 8; Batcher's
// odd-even merge networks).
template <int PER>
__device__ __forceinline__ void sort_regs(double (&k)[PER], int (&x)[PER]) {
#define ENM_CX(a, b) cex_pair(k[a], x[a], k[b], x[b])
    if constexpr (PER == 2) {
        ENM_CX(0, 1);
    } else if constexpr (PER == 4) {
        ENM_CX(0, 1); ENM_CX(2, 3); ENM_CX(0, 2); ENM_CX(1, 3); ENM_CX(1, 2);
    } else {
        static_assert(PER == 8, "2, 4 or 8 per thread");
        ENM_CX(0, 1); ENM_CX(2, 3); ENM_CX(4, 5); ENM_CX(6, 7);
        ENM_CX(0, 2); ENM_CX(1, 3); ENM_CX(4, 6); ENM_CX(5, 7);
        ENM_CX(1, 2); ENM_CX(5, 6);
        ENM_CX(0, 4); ENM_CX(1, 5); ENM_CX(2, 6); ENM_CX(3, 7);
        ENM_CX(2, 4); ENM_CX(3, 5);
        ENM_CX(1, 2); ENM_CX(3, 4); ENM_CX(5, 6);
    }
#undef ENM_CX
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

// Sort the TILE (key, index) pairs of sk/si in LDS with a 256-thread block:
// each thread sorts TILE/256 consecutive pairs in registers, then merge
// levels of doubling run width, each thread producing TILE/256 consecutive
// outputs of its pair of runs from a merge-path split (outputs kept in
// registers across a barrier, so the merge works in place).  Padding keys
// (+inf, INT_MAX) sort last.  Starts and ends with a barrier.
template <int TILE>
__device__ __forceinline__ void block_sort_lds(double* sk, int* si) {
    constexpr int PER = TILE / ENM_SORT_THREADS;
    __syncthreads();
    const int e0 = threadIdx.x * PER;
    double k[PER];
    int x[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) { k[e] = sk[e0 + e]; x[e] = si[e0 + e]; }
    sort_regs<PER>(k, x);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) { sk[e0 + e] = k[e]; si[e0 + e] = x[e]; }
    __syncthreads();
    for (int run = PER; run < TILE; run <<= 1) {
        const int pb = (e0 / (2 * run)) * (2 * run);
        const double* ka = sk + pb;
        const int* ia = si + pb;
        const double* kb = ka + run;
        const int* ib = ia + run;
        const int d = e0 - pb;
        int a = co_rank(d, ka, ia, run, kb, ib, run);
        int b = d - a;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            // Clamped reads stay inside the pair; the flags decide.
            const int ac = min(a, run - 1), bc = min(b, run - 1);
            const bool take_a = b >= run || (a < run && kless(ka[ac], ia[ac], kb[bc], ib[bc]));
            const int src = take_a ? ac : run + bc;
            k[e] = ka[src];
            x[e] = ia[src];
            if (take_a) ++a; else ++b;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PER; ++e) { sk[e0 + e] = k[e]; si[e0 + e] = x[e]; }
        __syncthreads();
    }
}

// Sort one TILE chunk of a segment: keys arena[src + off + t] with their
// point indices p0 + off + t, into the sort buffers.
template <int TILE>
__global__ __launch_bounds__(ENM_SORT_THREADS)
void block_sort_kernel(const SortSeg* __restrict__ segs, int nsegs,
                       const double* __restrict__ arena,
                       double* __restrict__ kout, int* __restrict__ iout) {
    __shared__ double sk[TILE];
    __shared__ int si[TILE];
    const int tile = blockIdx.x;
    const int s = seg_of_tile(segs, nsegs, tile);
    const SortSeg sg = segs[s];
    const int off = (tile - sg.first_tile) * TILE;
    const int len = min(TILE, sg.len - off);
    for (int t = threadIdx.x; t < TILE; t += ENM_SORT_THREADS) {
        if (t < len) {
            sk[t] = arena[sg.src + off + t];
            si[t] = sg.p0 + off + t;
        } else {
            sk[t] = INFINITY;
            si[t] = 0x7fffffff;
        }
    }
    block_sort_lds<TILE>(sk, si);
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        kout[sg.sbase + off + t] = sk[t];
        iout[sg.sbase + off + t] = si[t];
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
    __shared__ double sk[TILE];
    __shared__ int si[TILE];
    const GatherItem it = items[blockIdx.x];
    const SortSeg sg = segs[it.sseg];
    const OrderDev& od = orders[it.order];
    const int len = sg.len;
    for (int t = threadIdx.x; t < TILE; t += ENM_SORT_THREADS) {
        if (t < len) {
            sk[t] = col64[sg.src + t];
            si[t] = sg.p0 + t;
        } else {
            sk[t] = INFINITY;
            si[t] = 0x7fffffff;
        }
    }
    block_sort_lds<TILE>(sk, si);
    const int64_t n = od.n;
    const int ncol = od.ncol;
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const int idx = si[t];
        const int64_t p = it.pos0 + t;
        perm[od.perm + p] = perm[od.base_perm + idx];
        for (int c = 0; c < ncol; ++c) {
            col64[od.c64 + c * n + p] = col64[od.base_c64 + c * n + idx];
            if (with32) col32[od.c32 + c * n + p] = col32[od.base_c32 + c * n + idx];
        }
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
                    const double* __restrict__ kin, const int* __restrict__ iin, int* __restrict__ split) {
    const int tile = blockIdx.x * ENM_SORT_THREADS + threadIdx.x;
    if (tile >= ntiles) return;
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
                       double* __restrict__ kout, int* __restrict__ iout) {
    __shared__ double sk[ENM_SORT_TILE];
    __shared__ int si[ENM_SORT_TILE];
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
        if (t < la) { sk[t] = ka[i0 + t]; si[t] = ia[i0 + t]; }
        else        { sk[t] = kb[j0 + t - la]; si[t] = ib[j0 + t - la]; }
    }
    __syncthreads();
    // Each thread merges 8 consecutive outputs into registers; they go back
    // to LDS and out with coalesced stores.
    constexpr int PER = ENM_SORT_TILE / ENM_SORT_THREADS;
    const int e0 = threadIdx.x * PER;
    const int ne = min(PER, la + lb - e0);
    double ok[PER];
    int oi[PER];
    if (ne > 0) {
        int a = co_rank(e0, sk, si, la, sk + la, si + la, lb);
        int b = e0 - a;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            if (e < ne) {
                // Clamped reads stay inside the staged slices; the flags decide.
                const int ac = max(0, min(a, la - 1)), bc = la + max(0, min(b, lb - 1));
                const bool take_a = lb == 0 || b >= lb || (a < la && kless(sk[ac], si[ac], sk[bc], si[bc]));
                const int src = take_a ? ac : bc;
                ok[e] = sk[src];
                oi[e] = si[src];
                if (take_a) ++a; else ++b;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        if (e < ne) {
            sk[e0 + e] = ok[e];
            si[e0 + e] = oi[e];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < la + lb; t += ENM_SORT_THREADS) {
        kout[base + d0 + t] = sk[t];
        iout[base + d0 + t] = si[t];
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
    __shared__ double red[ENM_SORT_THREADS / ENM_WAVE];
    const GatherItem it = items[blockIdx.x];
    const SortSeg sg = segs[it.sseg];
    // Fields read one by one: a local copy of the whole descriptor, indexed
    // by a loop counter, would live in scratch memory.
    const OrderDev& od = orders[it.order];
    const int len = min(ENM_SORT_TILE, sg.len - it.off);
    const int64_t n = od.n;
    const bool indirect = od.base_perm >= 0;
    double m = 0.0;
    for (int t = threadIdx.x; t < len; t += ENM_SORT_THREADS) {
        const int64_t sp = sg.sbase + it.off + t;
        const int64_t p = it.pos0 + it.off + t;
        const int idx = sidx[sp];
        if (indirect) {
            // idx is a base-order position in the same strip: read locally.
            perm[od.perm + p] = perm[od.base_perm + idx];
            for (int c = 0; c < od.ncol; ++c) {
                col64[od.c64 + c * n + p] = col64[od.base_c64 + c * n + idx];
                if (with32) col32[od.c32 + c * n + p] = col32[od.base_c32 + c * n + idx];
            }
            continue;
        }
        perm[od.perm + p] = idx;
        const double key = skey[sp];
        col64[od.c64 + p] = key;
        if (with32) col32[od.c32 + p] = (float)key;
        m = fmax(m, fabs(key));
        for (int c = 1; c < od.ncol; ++c) {
            const double v = arena[od.src[c] + idx];
            col64[od.c64 + c * n + p] = v;
            if (with32) col32[od.c32 + c * n + p] = (float)v;
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

hipError_t enm_launch_sort(const SortSeg* d_segs, int nsegs, int ntiles, int64_t maxlen, int tile,
                           const double* arena, double* k0, int* i0, double* k1, int* i1, int* split,
                           int* final_buf, int* n_launch, hipStream_t st) {
    *final_buf = 0;
    if (ntiles <= 0) return hipSuccess;
    // Segments no longer than the tile are sorted by the block sort alone;
    // longer ones (always with 2048-key tiles) continue with merge passes.
    if (tile == 512)
        hipLaunchKernelGGL(block_sort_kernel<512>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, k0, i0);
    else if (tile == 1024)
        hipLaunchKernelGGL(block_sort_kernel<1024>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, k0, i0);
    else if (tile == ENM_SORT_TILE)
        hipLaunchKernelGGL(block_sort_kernel<ENM_SORT_TILE>, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, arena, k0, i0);
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
                           dim3(ENM_SORT_THREADS), 0, st, d_segs, nsegs, ntiles, w, kin, iin, split);
        hipLaunchKernelGGL(merge_pass_kernel, dim3(ntiles), dim3(ENM_SORT_THREADS), 0, st,
                           d_segs, nsegs, w, kin, iin, split, kout, iout);
        *n_launch += 2;
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur ^= 1;
    }
    *final_buf = cur;
    return hipSuccess;
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
