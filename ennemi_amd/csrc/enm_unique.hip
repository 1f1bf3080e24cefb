// Distinct rows and their counts, on the device: np.unique(rows, axis=0,
// return_inverse=True, return_counts=True) for the discrete plug-in
// estimators of the reference (ennemi/_entropy_estimators.py:44-56 entropy,
// :257-289 MI, :291-320 conditional MI), which spend their time there.
//
// A row set is one or more float64 columns of n values.  Each distinct
// column is sorted once (the library's segmented merge sort, keys compared
// as doubles and ties by point index, so -0.0 and 0.0 are one value, as for
// np.unique); a run head is a position whose key differs from the previous
// one, and a value's code is the number of run heads up to it, minus one:
// codes number the distinct values in increasing order.  A row set of
// several columns is keyed by the mixed-radix number of its columns' codes
// (column 0 most significant: the order of np.unique's lexicographic row
// sort), exact in float64 while the product of the columns' distinct counts
// is at most 2^53, and factorised the same way.  Counts are run lengths.
#include "enm_device.h"
#include "enm_launch.h"

#include <algorithm>

namespace {

constexpr int UQ_THREADS = 256;
constexpr int UQ_PER = 4;                       // elements per thread and tile
constexpr int UQ_TILE = UQ_THREADS * UQ_PER;

// One block per segment: codes (by point), run starts, counts and the
// number of distinct values.
__global__ __launch_bounds__(UQ_THREADS)
void uq_codes_kernel(const UqSeg* __restrict__ segs, const double* __restrict__ skeys, const int* __restrict__ sidx,
                     int* __restrict__ codes, int64_t* __restrict__ starts, int64_t* __restrict__ counts,
                     int64_t* __restrict__ distinct) {
    __shared__ int s_wave[UQ_THREADS / ENM_WAVE];
    __shared__ int64_t s_carry;
    const UqSeg sg = segs[blockIdx.x];
    const int lane = threadIdx.x & (ENM_WAVE - 1), wv = threadIdx.x / ENM_WAVE;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t t0 = 0; t0 < sg.n; t0 += UQ_TILE) {
        // Thread-local heads of UQ_PER consecutive positions.
        int heads = 0;
        bool h[UQ_PER];
#pragma unroll
        for (int j = 0; j < UQ_PER; ++j) {
            const int64_t i = t0 + threadIdx.x * UQ_PER + j;
            h[j] = i < sg.n && (i == 0 || skeys[sg.sbase + i] != skeys[sg.sbase + i - 1]);
            heads += h[j];
        }
        // Block-wide inclusive scan of the per-thread head counts.
        int inc = heads;
        for (int off = 1; off < ENM_WAVE; off <<= 1) {
            const int o = __shfl_up(inc, off, ENM_WAVE);
            if (lane >= off) inc += o;
        }
        if (lane == ENM_WAVE - 1) s_wave[wv] = inc;
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < UQ_THREADS / ENM_WAVE; ++w) {
            before += w < wv ? s_wave[w] : 0;
            total += s_wave[w];
        }
        const int64_t carry = s_carry;
        int64_t rank = carry + before + inc - heads - 1;  // code of the position before this thread's first
#pragma unroll
        for (int j = 0; j < UQ_PER; ++j) {
            const int64_t i = t0 + threadIdx.x * UQ_PER + j;
            if (i < sg.n) {
                rank += h[j];
                codes[sg.code_off + sidx[sg.sbase + i]] = (int)rank;
                if (h[j]) starts[sg.start_off + rank] = i;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
    const int64_t m = s_carry;
    if (threadIdx.x == 0) {
        starts[sg.start_off + m] = sg.n;
        distinct[sg.slot] = m;
    }
    __syncthreads();
    for (int64_t r = threadIdx.x; r < m; r += UQ_THREADS)
        counts[sg.count_off + r] = starts[sg.start_off + r + 1] - starts[sg.start_off + r];
}

// Mixed-radix keys of the rows of multi-column row sets (exact doubles).
__global__ void uq_key_kernel(const UqRowKey* __restrict__ rk, int n_rk, const int* __restrict__ ccodes,
                              double* __restrict__ keys) {
    const int s = blockIdx.y;
    if (s >= n_rk) return;
    const UqRowKey r = rk[s];
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < r.n; p += (int64_t)gridDim.x * blockDim.x) {
        int64_t key = 0;
        for (int c = 0; c < r.ncol; ++c) key = key * r.radix[c] + ccodes[r.code_off[c] + p];
        keys[r.key_off + p] = (double)key;
    }
}

}  // namespace

hipError_t enm_launch_uq_codes(const UqSeg* segs, int nsegs, const double* skeys, const int* sidx, int* codes,
                               int64_t* starts, int64_t* counts, int64_t* distinct, hipStream_t st) {
    if (nsegs <= 0) return hipSuccess;
    hipLaunchKernelGGL(uq_codes_kernel, dim3(nsegs), dim3(UQ_THREADS), 0, st, segs, skeys, sidx, codes, starts,
                       counts, distinct);
    return hipGetLastError();
}

hipError_t enm_launch_uq_keys(const UqRowKey* rk, int n_rk, int64_t max_n, const int* ccodes, double* keys,
                              hipStream_t st) {
    if (n_rk <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((max_n + 255) / 256, 4096);
    hipLaunchKernelGGL(uq_key_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1), (unsigned)n_rk), dim3(256), 0, st,
                       rk, n_rk, ccodes, keys);
    return hipGetLastError();
}
