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

// FP64 min / max of values that are never NaN (distances, +-inf).  fmin /
// fmax make the compiler quiet possible signalling NaNs first (one extra
// v_max_f64 per operand carried across loop iterations, under the IEEE mode
// compute kernels run in); these are the bare instructions.
__device__ __forceinline__ double vmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double vmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// max(a, |b|) of values that are never NaN: one instruction, the absolute
// value taken by a source modifier.
__device__ __forceinline__ double vmax_abs(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, |%2|" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Holds v in registers here: the load that produces it stays where it is
// written, with its neighbours, instead of being moved into the branch that
// first uses it, where it would wait out its own round trip.
__device__ __forceinline__ void pin(double& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(int& v) { asm volatile("" : "+v"(v)); }

// Insert v into the ascending list L[0..P) keeping the P smallest.  Slots
// below P-K1 hold -inf, so L[P-1] is the K1-th smallest value seen.
template <int P>
__device__ __forceinline__ void list_insert(double (&L)[P], double v) {
#pragma unroll
    for (int t = P - 1; t > 0; --t) L[t] = vmin(L[t], vmax(v, L[t - 1]));
    L[0] = vmin(L[0], v);
}

__device__ __forceinline__ double uni_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void cex(double& a, double& b) {
    const double lo = vmin(a, b), hi = vmax(a, b);
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
    for (int t = 0; t < P; ++t) L[t] = vmin(L[t], e[P - 1 - t]);
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

// Both bounds of [lb, ub] in the ascending keys v[lo, hi) at once, by the
// whole wave: a = first position with v >= lb, b = first with v > ub (the
// predicates of lower_bound_u / upper_bound_u).  Lanes 0-31 search a, lanes
// 32-63 b; each round probes 32 evenly spaced positions of the interval
// still open, and the ballot of "before the answer" (a prefix of the
// probes, the keys being sorted) cuts it 32-fold: 2 rounds of independent
// loads for a 512-key strip instead of 2 x 9 dependent ones.
__device__ __forceinline__ void wave_bounds(const double* __restrict__ v, int lo, int hi, double lb, double ub,
                                            int& a, int& b) {
    const int lane = (int)(threadIdx.x & (ENM_WAVE - 1));
    const bool upper = lane >= 32;
    const int l = lane & 31;
    int L = lo, H = hi;  // the answer lies in [L, H] (H: no key qualifies)
    for (;;) {
        const int len = H - L;
        if (wave_all(len == 0)) break;
        const int step = (len + 31) >> 5;
        const int p = L + l * step;
        const bool in = len > 0 && p < H;
        const double val = in ? v[p] : 0.0;
        const bool before = in && (upper ? (val <= ub) : (val < lb));
        const unsigned long long m = __builtin_amdgcn_ballot_w64(before);
        const int c = __popc(upper ? (unsigned)(m >> 32) : (unsigned)m);
        if (len > 0) {
            const int nl = c == 0 ? L : L + (c - 1) * step + 1;
            H = min(H, L + c * step);
            L = nl;
        }
    }
    a = __builtin_amdgcn_readlane(L, 0);
    b = __builtin_amdgcn_readlane(L, 32);
}

// Positions [a, b) of the strip [s0, s1) of a y-sorted column that can
// hold a candidate j with fl(|y_j - y_i|) <= T_i for some live lane i: the
// union of the lanes' intervals [y_i - T_i, y_i + T_i], each widened by a
// relative 2^-50 so that the rounding of the bounds (far below the
// widening) cannot exclude such a candidate.  Wave-uniform result.
__device__ __forceinline__ void strip_window(const double* __restrict__ yc, int s0, int s1, double yq, double T,
                                         bool live, int& a, int& b) {
    const double slack = 0x1p-50 * (fabs(yq) + T);
    const double lq = live ? (yq - T * (1.0 + 0x1p-50)) - slack : INFINITY;
    const double uq = live ? (yq + T * (1.0 + 0x1p-50)) + slack : -INFINITY;
    const double lb = uni_f64(wave_min(lq));
    const double ub = uni_f64(wave_max(uq));
    if (!(lb <= ub)) { a = b = s0; return; }
    wave_bounds(yc, s0, s1, lb, ub, a, b);
}

// The block number that block b of an nb-block launch works on when
// consecutive numbers should share an XCD's L2: the hardware deals
// workgroups to the ENM_NQ XCDs round-robin, so this gives XCD x the x-th of
// ENM_NQ nearly equal contiguous ranges.  A bijection of [0, nb).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t q = nb / ENM_NQ, r = nb % ENM_NQ;
    const int64_t x = b % ENM_NQ, k = b / ENM_NQ;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Persistent work loop of a wave64 workgroup over ENM_NQ queues.  The host
// places each instance's items in one queue (large instances split into a
// few contiguous blocks), balancing the queues by work; a workgroup starts on
// queue blockIdx.x % ENM_NQ, which the round-robin dispatch of workgroups
// over the XCDs maps to one XCD, so an instance's columns are read into one
// XCD's L2 only.  A drained queue sends the workgroup on to the next one, so
// every queue empties whatever the placement.  A ticket is `chunk`
// consecutive items: the atomics on a queue's counter (each counter on its
// own 128-byte line, ENM_CSTRIDE slots apart) serialize where the line
// lives, so one per item would bound the whole launch at about ten
// nanoseconds per item.  Each counter only grows, so a workgroup that found
// a queue empty never finds work there again: every wave leaves after at
// most ENM_NQ empty pops.
template <class F>
__device__ __forceinline__ void for_each_wave_item(const int* __restrict__ qoff, unsigned* counter, int chunk,
                                                   F&& f) {
    const int home = (int)(blockIdx.x % ENM_NQ);
    for (int r = 0; r < ENM_NQ; ++r) {
        const int q = uni((home + r) % ENM_NQ);
        const int qb = uni(qoff[q]);
        const int qn = uni(qoff[q + 1]) - qb;
        for (;;) {
            const unsigned t = pop_ticket(counter + q * ENM_CSTRIDE);
            const long long first = (long long)t * chunk;
            if (first >= qn) break;
            const int last = (int)min(first + chunk, (long long)qn);
            for (int k = (int)first; k < last; ++k) f(qb + k);
        }
    }
}

// The same for a multi-wave workgroup, one item per ticket: thread 0 pops,
// the ticket is passed through LDS between two barriers and made
// wave-uniform.
template <class F>
__device__ __forceinline__ void for_each_block_item(const int* __restrict__ qoff, unsigned* counter,
                                                    unsigned* s_ticket, F&& f) {
    const int home = (int)(blockIdx.x % ENM_NQ);
    for (int r = 0; r < ENM_NQ; ++r) {
        const int q = uni((home + r) % ENM_NQ);
        const int qb = uni(qoff[q]);
        const int qn = uni(qoff[q + 1]) - qb;
        for (;;) {
            if (threadIdx.x == 0) *s_ticket = atomicAdd(counter + q * ENM_CSTRIDE, 1u);
            __syncthreads();
            const unsigned t = (unsigned)__builtin_amdgcn_readfirstlane((int)*s_ticket);
            __syncthreads();
            if (t >= (unsigned)qn) break;
            f(qb + (int)t);
        }
    }
}
