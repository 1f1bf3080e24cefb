// Device helpers for the gfx950 kernels of libennemi_hip.so.
// Wave64 throughout: one work item is one wave, one query per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "enm_internal.h"

#define ENM_WAVE 64

// The reference's truncated digamma series for y >= 1
// (ennemi/_entropy_estimators.py:341-348): psi(1) is SciPy's constant and
// for y >= 2  psi(y) = ln y - y^-6 (y^2 (y^2 (y/2 + 1/12) - 1/120) + 1/252).
// psi(0) = +inf is handled by the callers as a per-space flag, because the
// reference turns the whole array into a scalar inf (:338-339).
__device__ __forceinline__ double enm_psi(double y) {
    if (y == 1.0) return -0.5772156649015331;
    const double y2 = y * y;
    const double inv6 = 1.0 / (y2 * y2 * y2);
    return log(y) - inv6 * (y2 * (y2 * (y / 2.0 + 1.0 / 12.0) - 1.0 / 120.0) + 1.0 / 252.0);
}

// Smallest float >= x and largest float <= x (x may be +-inf).
__device__ __forceinline__ float f32_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
__device__ __forceinline__ float f32_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

// FP32 guard band of an instance from the largest |coordinate| M:
// |fl32(a)-fl32(b) - (a-b)| <= 2^-24 (2|a| + 2|b|) (1 + 2^-24) and
// |fl64(a-b) - (a-b)| <= 2^-53 |a-b|, so a guard of 2^-21 M (twice the bound)
// plus 2^-126 for FP32 subnormal spacing covers every per-axis distance.
// The filter is used only when M is comfortably inside the FP32 range.
__device__ __forceinline__ double enm_guard(uint64_t maxabs_bits, bool* f32ok) {
    const double m = __longlong_as_double((long long)maxabs_bits);
    *f32ok = (m < 1e30);
    return m * 0x1p-21 + 0x1p-126;
}

__device__ __forceinline__ bool wave_any(bool p) {
    return __builtin_amdgcn_ballot_w64(p) != 0;
}
__device__ __forceinline__ bool wave_all(bool p) {
    return __builtin_amdgcn_ballot_w64(!p) == 0;
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Pop the next work-item ticket of a persistent kernel, wave-uniform.  Every
// lane takes part in the atomic (lane 0 adds 1, the others add 0), so no
// lane-0-only branch precedes the broadcast: a value set inside such a branch
// and broadcast after it can be read with part of the wave already split off
// by the control-flow structurizer.  Lane 0's old value is the ticket
// whatever order the lanes' additions take, as the others add 0.
__device__ __forceinline__ unsigned pop_ticket(unsigned* counter) {
    const unsigned t = atomicAdd(counter, threadIdx.x == 0 ? 1u : 0u);
    return (unsigned)__builtin_amdgcn_readfirstlane((int)t);
}

// Butterfly sums and extrema over the 64 lanes.  The pairing is fixed, so
// the result is bit-reproducible for given lane values.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, ENM_WAVE);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, ENM_WAVE);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, ENM_WAVE));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, ENM_WAVE));
    return v;
}
__device__ __forceinline__ int wave_or(int v) {
    return __builtin_amdgcn_ballot_w64(v != 0) != 0;
}

// Insert v into the ascending list L[0..P) keeping the P smallest.  Slots
// below P-K1 hold -inf, so L[P-1] is the K1-th smallest value seen.
template <int P>
__device__ __forceinline__ void list_insert(double (&L)[P], double v) {
#pragma unroll
    for (int t = P - 1; t > 0; --t) L[t] = fmin(L[t], fmax(v, L[t - 1]));
    L[0] = fmin(L[0], v);
}

__device__ __forceinline__ double uni_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void cex(double& a, double& b) {
    const double lo = fmin(a, b), hi = fmax(a, b);
    a = lo;
    b = hi;
}

// Sort 8 values ascending (19 compare-exchanges, Batcher's odd-even merge
// network for n = 8).
__device__ __forceinline__ void sort8(double (&e)[8]) {
    cex(e[0], e[1]); cex(e[2], e[3]); cex(e[4], e[5]); cex(e[6], e[7]);
    cex(e[0], e[2]); cex(e[1], e[3]); cex(e[4], e[6]); cex(e[5], e[7]);
    cex(e[1], e[2]); cex(e[5], e[6]);
    cex(e[0], e[4]); cex(e[1], e[5]); cex(e[2], e[6]); cex(e[3], e[7]);
    cex(e[2], e[4]); cex(e[3], e[5]);
    cex(e[1], e[2]); cex(e[3], e[4]); cex(e[5], e[6]);
}

// Keep the P smallest of the ascending list L[0..P) and the ascending
// values e[0..8) (P = 4 or 8), L ascending afterwards: min(L[t], e[P-1-t])
// selects them as a bitonic sequence, which a bitonic merge then sorts.
template <int P>
__device__ __forceinline__ void list_merge8(double (&L)[P], const double (&e)[8]) {
    static_assert(P == 4 || P == 8, "merge network for 4 or 8 slots");
#pragma unroll
    for (int t = 0; t < P; ++t) L[t] = fmin(L[t], e[P - 1 - t]);
#pragma unroll
    for (int half = P / 2; half > 0; half >>= 1) {
#pragma unroll
        for (int t = 0; t < P; ++t)
            if ((t & half) == 0) cex(L[t], L[t + half]);
    }
}

// First position in [lo, hi) whose key is >= x (keys ascending), with a
// wave-uniform key array and bounds: every lane computes the same answer.
__device__ __forceinline__ int lower_bound_u(const double* __restrict__ v, int lo, int hi, double x) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// First position in [lo, hi) whose key is > x.
__device__ __forceinline__ int upper_bound_u(const double* __restrict__ v, int lo, int hi, double x) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
