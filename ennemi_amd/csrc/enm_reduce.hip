// Final combination of each instance's per-tile partial sums into the
// estimate, on the device, and the optional export of per-point radii and
// counts in the instance's own point numbering.
//
// One wave per instance sums the partials of each space with a fixed lane
// assignment and butterfly, so the result does not depend on the order in
// which the work queue ran the tiles, nor on the kNN and count tilings: 1-D
// count spaces sum per tile of a key-sorted order (which no tuning changes),
// multi-dimensional spaces and entropy log radii are summed in point order by
// reduce_pt_kernel.  The formulas and the psi(0) = +inf
// broadcasting follow ennemi/_entropy_estimators.py:
//   KSG   :113  psi(N) + psi(k) - mean(psi(nx) + psi(ny))
//   CMI   :156  psi(k) - mean(psi(nxz) + psi(nyz) - psi(nz))
//   ROSS  :199-200 psi(N) + psi(k) - mean(psi(n_full)) - sum_c psi(n_c) n_c / N
//   CROSS :247  as CMI
//   KL    :42   psi(N) - psi(k) + d (mean(log eps) + log 2)
// where a 0 anywhere in a space's counts makes that space's psi a scalar
// +inf (:338-339), which then turns the mean into +-inf or NaN.
#include "enm_device.h"
#include "enm_launch.h"

namespace {

__device__ double psi_int(int64_t v) { return enm_psi((double)v); }

__global__ __launch_bounds__(ENM_WAVE)
void combine_kernel(const InstDev* __restrict__ insts, const double* __restrict__ cls, int first,
                    const double* __restrict__ res64,
                    const int* __restrict__ res32, double* __restrict__ out) {
    const int ii = first + blockIdx.x;
    const InstDev in = insts[ii];
    const int lane = threadIdx.x;
    double sum[ENM_MAX_SPACES] = {0.0, 0.0, 0.0};
    int zero[ENM_MAX_SPACES] = {0, 0, 0};
    for (int s = 0; s < in.nspace; ++s) {
        double acc = 0.0;
        int z = 0;
        for (int t = lane; t < in.ntile[s]; t += ENM_WAVE) {
            acc += res64[in.part[s] + t];
            if (in.zero[s] >= 0) z |= res32[in.zero[s] + t];
        }
        sum[s] = wave_sum(acc);
        zero[s] = wave_or(z);
    }
    // A space summed in point order flags a zero count as a +inf sum.
#pragma unroll
    for (int s = 0; s < ENM_MAX_SPACES; ++s)
        if (s < in.nspace && in.kind != ENM_KIND_KL && sum[s] == INFINITY) zero[s] = 1;
    if (lane != 0) return;
    const double N = (double)in.n;
    const double psik = enm_psi((double)in.k);
    double value;
    switch (in.kind) {
        case ENM_KIND_KSG: {
            const double mean = (zero[0] || zero[1]) ? INFINITY : (sum[0] + sum[1]) / N;
            value = psi_int(in.n) + psik - mean;
            break;
        }
        case ENM_KIND_CMI:
        case ENM_KIND_CROSS: {
            const bool za = zero[0] || zero[1];
            const bool zb = zero[2];
            double mean;
            if (za && zb) mean = NAN;
            else if (za) mean = INFINITY;
            else if (zb) mean = -INFINITY;
            else mean = (sum[0] + sum[1] - sum[2]) / N;
            value = psik - mean;
            break;
        }
        case ENM_KIND_ROSS: {
            const double mean = zero[0] ? INFINITY : sum[0] / N;
            double term = 0.0;
            for (int c = 0; c < in.n_segs; ++c) {
                const double nc = cls[in.seg_len + c];
                term += enm_psi(nc) * (nc / N);
            }
            value = psi_int(in.n) + psik - mean - term;
            break;
        }
        case ENM_KIND_KL:
        default: {
            value = psi_int(in.n) - psik + (double)in.dx * (sum[0] / N + log(2.0));
            break;
        }
    }
    out[ii] = value;
}

// Chunked sums of point-order arrays: one wave per chunk of ENM_RED_CHUNK
// points, lane-strided then butterfly, so each partial (and the instance's
// estimate) depends only on the values in point order.
__global__ __launch_bounds__(ENM_WAVE)
void reduce_pt_kernel(const RedItem* __restrict__ items, const RedDev* __restrict__ descs,
                      double* __restrict__ res64) {
    const RedItem it = items[blockIdx.x];
    const RedDev rd = descs[it.desc];
    const int64_t a = (int64_t)it.chunk * ENM_RED_CHUNK;
    const int64_t b = min(rd.n, a + ENM_RED_CHUNK);
    double acc = 0.0;
    for (int64_t i = a + threadIdx.x; i < b; i += ENM_WAVE) {
        const double v = res64[rd.src + i];
        acc += rd.log ? log(v) : v;
    }
    acc = wave_sum(acc);
    if (threadIdx.x == 0) res64[rd.part + it.chunk] = acc;
}

// Export one order-positioned array into instance numbering.
__global__ void export_eps_kernel(const double* __restrict__ src, const int* __restrict__ perm,
                                  int64_t n, double* __restrict__ dst) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[perm[p]] = src[p];
}

__global__ void export_cnt_kernel(const int* __restrict__ src, const int* __restrict__ perm,
                                  int64_t n, int* __restrict__ dst) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[perm[p]] = src[p];
}

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

hipError_t enm_launch_combine(const InstDev* insts, const double* cls, int first, int count,
                              const double* res64, const int* res32, double* out, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(combine_kernel, dim3(count), dim3(ENM_WAVE), 0, st, insts, cls, first, res64,
                       res32, out);
    return hipGetLastError();
}

hipError_t enm_launch_reduce_pt(const RedItem* items, int n_items, const RedDev* descs, double* res64,
                                hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_pt_kernel, dim3(n_items), dim3(ENM_WAVE), 0, st, items, descs, res64);
    return hipGetLastError();
}

hipError_t enm_launch_export_eps(const double* src, const int* perm, int64_t n, double* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(export_eps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, perm, n, dst);
    return hipGetLastError();
}

hipError_t enm_launch_export_cnt(const int* src, const int* perm, int64_t n, int* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(export_cnt_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, perm, n, dst);
    return hipGetLastError();
}

hipError_t enm_launch_fill_i32(int* p, int64_t n, int v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
