// Sort hints for batches prepared on the device (enm_prep_hints): the
// ascending order of every task column, derived from one sort per series
// instead of one sort per column.  They stand in for the per-task kd-tree
// builds of the reference (ennemi/_entropy_estimators.py:36,100-102,135-139),
// which enm_run_batch otherwise replaces by a sort of every column.
//
// A task column is a window of a raw series with some rows dropped (mask,
// NaN), standardised as a = (v - mean) / std -- monotone in v, rounding
// included -- plus the reference's noise N (|N| <= M, the largest noise
// magnitude drawn): s = a + N.  So:
//   1. each series' rows in the windows' union are sorted once by (value,
//      row), NaN last (a NaN row is never kept);
//   2. a column's sorted series rows, filtered to its task's kept rows (the
//      kept-row map enm_prep_write made) and renumbered to positions in the
//      task, are its values' order before the noise; a is recomputed there
//      from the sorted raw value, bit for bit as the preparation did.  The
//      filter cuts each column into segments of ENM_HINT_SEG sorted rows
//      (a workgroup each): their kept rows are counted, scanned per column,
//      then compacted in order from each segment's offset;
//   3. the noise can only swap values whose a differ by about 2M or less:
//      the column cuts into runs wherever neighbours differ by
//      a - a' > 2M + 2^-48 (|a| + |a'| + 2M + 1), more than the noise and
//      the rounding of a + N on both sides, so no value crosses a cut.  The
//      filter marks every position not cut from the one before it (a
//      continuation bit), and each run of two to ENM_HINT_RUN values is
//      re-sorted by (s, position) in place.  A longer run (ties or
//      near-ties in the raw data) leaves the column without a hint
//      (hint[out] = -1) and the batch sorts it.
// The batch checks every hint in full before it trusts it, so a hint is an
// accelerator, never a source of results.
#include "enm_device.h"
#include "enm_launch.h"

#include <algorithm>

namespace {

constexpr int HT = 256;                           // threads per block
constexpr int HPER = ENM_HINT_BLOCK / HT;         // sorted rows per thread

// Keys of every series range: the raw values, NaN as +inf.
__global__ __launch_bounds__(HT)
void hint_keys_kernel(const HintSeries* __restrict__ series, const double* __restrict__ raw,
                      double* __restrict__ keys) {
    const HintSeries S = series[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * HT + threadIdx.x;
    if (i >= S.len) return;
    const double v = raw[S.col + S.lo + i];
    keys[S.kbase + i] = isnan(v) ? INFINITY : v;
}

// The task position of series row R (relative to the series range) in
// column C, or -1 when the row is outside the window or dropped.  The task's
// kept-row map is either in LDS, as its words' bits and the prefix of every
// other word (6 bytes per 32 rows, so eight blocks fit a CU), or read in
// place.
__device__ __forceinline__ int kept_pos(const HintCol& C, int64_t R, bool in_lds, const uint32_t* lbits,
                                        const int* lpre, const KeptWord* __restrict__ gmap) {
    const int64_t r = R - C.ws;
    if (r < 0 || r >= C.w) return -1;
    const int64_t w = r >> 5;
    const uint32_t bit = 1u << (r & 31);
    if (in_lds) {
        const uint32_t bits = lbits[w];
        if (!(bits & bit)) return -1;
        return lpre[w >> 1] + ((w & 1) ? __popc(lbits[w - 1]) : 0) + __popc(bits & (bit - 1u));
    }
    const KeptWord kw = gmap[w];
    return (kw.bits & bit) ? kw.prefix + __popc(kw.bits & (bit - 1u)) : -1;
}

// Whether neighbours with values a' <= a before the noise are far enough
// apart that the noise cannot swap anything across them (m2 = 2M).
__device__ __forceinline__ bool is_cut(double ap, double a, double m2) {
    return a - ap > m2 + 0x1p-48 * (fabs(a) + fabs(ap) + m2 + 1.0);
}


// The task's kept-row map of column C into LDS when it fits (lds_words
// words: their bits, then the prefix of every other word); false when it is
// to be read in place.  Ends with a barrier.
__device__ __forceinline__ bool stage_map(const HintCol& C, const KeptWord* __restrict__ kept, int lds_words,
                                          uint32_t* lbits, int* lpre) {
    const int64_t words = (C.w + 31) / 32;
    const bool in_lds = words <= lds_words;
    if (in_lds)
        for (int64_t t = threadIdx.x; t < words; t += HT) {
            const KeptWord kw = kept[C.pm + t];
            lbits[t] = kw.bits;
            if (!(t & 1)) lpre[t >> 1] = kw.prefix;
        }
    __syncthreads();
    return in_lds;
}

// The column's standardisation, as the preparation applied it (a constant
// variable is left as it is): a = scaled ? (v - mean) / sd : v.
struct Scaling {
    bool scaled;
    double mean, sd;
    __device__ __forceinline__ double operator()(double v) const { return scaled ? (v - mean) / sd : v; }
};
__device__ __forceinline__ Scaling col_scaling(const HintCol& C, const PrepTaskDev* __restrict__ tasks,
                                               const double* __restrict__ stats, int rescale) {
    const PrepTaskDev& T = tasks[C.task];
    const unsigned cbit = C.slot == 0 ? 1u : C.slot == 1 ? 2u : 4u;
    Scaling sc;
    sc.scaled = rescale && !(T.flags & cbit);
    sc.mean = sc.scaled ? stats[T.stat + 2 * C.slot] : 0.0;
    sc.sd = sc.scaled ? stats[T.stat + 2 * C.slot + 1] : 1.0;
    return sc;
}

// Filter pass 1, one block per segment of SEG sorted rows of a column
// (segments of a column, and columns of a series, on one XCD): the
// segment's kept rows, and the value before the noise of its last one.
__global__ __launch_bounds__(HT, 8)
void hint_seg_count_kernel(const HintCol* __restrict__ cols, int segs_per_col, int64_t seg, const int* __restrict__ sidx,
                           const double* __restrict__ skey, const KeptWord* __restrict__ kept,
                           const PrepTaskDev* __restrict__ tasks, const double* __restrict__ stats, int rescale,
                           int lds_words, int* __restrict__ seg_cnt, double* __restrict__ seg_last) {
    extern __shared__ uint32_t smap[];
    __shared__ int s_cnt[HT / ENM_WAVE];
    __shared__ long long s_top[HT / ENM_WAVE];
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int ci = (int)(b / segs_per_col);
    const HintCol C = cols[ci];
    const int64_t e_begin = (b - (int64_t)ci * segs_per_col) * seg;
    if (e_begin >= C.len) return;   // uniform over the block
    const int64_t e_end = min((int64_t)C.len, e_begin + seg);
    uint32_t* lbits = smap;
    int* lpre = reinterpret_cast<int*>(smap + lds_words);
    const bool in_lds = stage_map(C, kept, lds_words, lbits, lpre);
    const KeptWord* gmap = kept + C.pm;
    int cnt = 0;
    long long top = -1;   // the last kept sorted row of this thread
    for (int64_t e0 = e_begin; e0 < e_end; e0 += ENM_HINT_BLOCK) {
        int R[HPER];   // a chunk's rows loaded together
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            const int64_t e = e0 + j * HT + threadIdx.x;
            R[j] = e < e_end ? sidx[C.sb + e] : -1;
        }
#pragma unroll
        for (int j = 0; j < HPER; ++j)
            if (R[j] >= 0 && kept_pos(C, R[j], in_lds, lbits, lpre, gmap) >= 0) {
                ++cnt;
                top = e0 + j * HT + threadIdx.x;
            }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_xor(cnt, off, ENM_WAVE);
        top = max(top, (long long)__shfl_xor(top, off, ENM_WAVE));
    }
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) {
        s_cnt[threadIdx.x / ENM_WAVE] = cnt;
        s_top[threadIdx.x / ENM_WAVE] = top;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        long long t = -1;
        for (int w = 0; w < HT / ENM_WAVE; ++w) { tot += s_cnt[w]; t = max(t, s_top[w]); }
        const int64_t idx = (int64_t)ci * segs_per_col + (e_begin / seg);
        seg_cnt[idx] = tot;
        seg_last[idx] = t >= 0 ? col_scaling(C, tasks, stats, rescale)(skey[C.sb + t]) : 0.0;
    }
}

// Filter pass 2, one wave per column: each segment's first output position
// and the value before the noise of the last kept row before it (with
// whether there is one); a column whose kept rows do not add up to n gets
// no hint.
__global__ __launch_bounds__(ENM_WAVE)
void hint_seg_scan_kernel(const HintCol* __restrict__ cols, int n_cols, int segs_per_col, int64_t seg,
                          const int* __restrict__ seg_cnt, const double* __restrict__ seg_last,
                          int64_t* __restrict__ seg_off, double* __restrict__ seg_cin, int* __restrict__ seg_chas,
                          int* __restrict__ bad) {
    const int ci = blockIdx.x;
    if (ci >= n_cols || threadIdx.x != 0) return;
    const HintCol C = cols[ci];
    const int64_t nseg = (C.len + seg - 1) / seg, b0 = (int64_t)ci * segs_per_col;
    int64_t at = 0;
    double carry = 0.0;
    int has = 0;
    for (int64_t sg = 0; sg < nseg; ++sg) {
        seg_off[b0 + sg] = at;
        seg_cin[b0 + sg] = carry;
        seg_chas[b0 + sg] = has;
        const int c = seg_cnt[b0 + sg];
        if (c > 0) { carry = seg_last[b0 + sg]; has = 1; }
        at += c;
    }
    bad[ci] = at != C.n ? 1 : 0;
}

// Filter pass 3, one block per segment: the segment's sorted rows in chunks
// of ENM_HINT_BLOCK, each compacted in order (ballots, wave offsets) with
// the running offset carried from chunk to chunk, from the segment's first
// output position.  With one segment per column (many columns: they fill
// the GPU by themselves) passes 1 and 2 are skipped: the column starts at
// position 0 with nothing before it, and the block checks its kept rows
// against n itself.  Writes the kept rows' positions, in sorted order, to
// the hint; a kept row whose value before the noise is not cut from the one
// before it (in the column's order) sets its continuation bit (cont, the
// column's words from C.cbits; zeroed before).
__global__ __launch_bounds__(HT, 8)   // eight waves per SIMD
void hint_filter_kernel(const HintCol* __restrict__ cols, int segs_per_col, int64_t seg, const int* __restrict__ sidx,
                        const double* __restrict__ skey, const KeptWord* __restrict__ kept,
                        const PrepTaskDev* __restrict__ tasks, const double* __restrict__ stats, int rescale,
                        const unsigned long long* __restrict__ noise_max, int lds_words,
                        const int64_t* __restrict__ seg_off, const double* __restrict__ seg_cin,
                        const int* __restrict__ seg_chas, int* __restrict__ bad, int* __restrict__ hint,
                        unsigned* __restrict__ cont) {
    constexpr int NW = HT / ENM_WAVE;
    extern __shared__ uint32_t smap[];   // lds_words bits, then (lds_words + 1) / 2 prefixes
    __shared__ int s_wcnt[HPER][NW];
    __shared__ double s_last[HPER][NW];  // value of each (round, wave)'s last kept row
    __shared__ double s_prev[HPER][NW];  // value of the last kept row before each (round, wave)
    __shared__ int s_has[HPER][NW];      // whether there is one
    __shared__ double s_carry;
    __shared__ int s_hascarry;
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int ci = (int)(b / segs_per_col);
    const HintCol C = cols[ci];
    const bool single = segs_per_col == 1;
    const int64_t e_begin = (b - (int64_t)ci * segs_per_col) * seg;
    if (e_begin >= C.len || (!single && bad[ci])) return;   // uniform over the block
    const int64_t e_end = min((int64_t)C.len, e_begin + seg);
    const int64_t sidx_seg = (int64_t)ci * segs_per_col + e_begin / seg;
    uint32_t* lbits = smap;
    int* lpre = reinterpret_cast<int*>(smap + lds_words);
    if (threadIdx.x == 0) {
        s_carry = single ? 0.0 : seg_cin[sidx_seg];
        s_hascarry = single ? 0 : seg_chas[sidx_seg];
    }
    const bool in_lds = stage_map(C, kept, lds_words, lbits, lpre);
    const KeptWord* gmap = kept + C.pm;
    const Scaling sc = col_scaling(C, tasks, stats, rescale);
    const double m2 = 2.0 * __longlong_as_double((long long)*noise_max);
    const int lane = threadIdx.x & (ENM_WAVE - 1), wv = threadIdx.x / ENM_WAVE;
    int64_t at = single ? 0 : seg_off[sidx_seg];
    // Rows and keys of a chunk are loaded together, one chunk ahead.
    int nR[HPER];
    double nkey[HPER];
    auto load = [&](int64_t e0) {
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            const int64_t e = e0 + j * HT + threadIdx.x;
            const bool in = e < e_end;
            nR[j] = in ? sidx[C.sb + e] : -1;
            nkey[j] = in ? skey[C.sb + e] : 0.0;
        }
    };
    load(e_begin);
    for (int64_t e0 = e_begin; e0 < e_end; e0 += ENM_HINT_BLOCK) {
        int pos[HPER];
        double av[HPER];
        int R[HPER];
        double key[HPER];
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            R[j] = nR[j];
            key[j] = nkey[j];
        }
        load(e0 + ENM_HINT_BLOCK);
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            pos[j] = R[j] >= 0 ? kept_pos(C, R[j], in_lds, lbits, lpre, gmap) : -1;
            av[j] = sc(key[j]);
        }
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            const unsigned long long m = __builtin_amdgcn_ballot_w64(pos[j] >= 0);
            const int top = m ? 63 - __clzll(m) : -1;   // the wave's last kept lane
            const double last = __shfl(av[j], top < 0 ? 0 : top, ENM_WAVE);
            if (lane == 0) {
                s_wcnt[j][wv] = __popcll(m);
                s_last[j][wv] = last;
            }
        }
        __syncthreads();
        if (threadIdx.x < ENM_WAVE) {
            // The value before each (round, wave) in the column's order: lane
            // l = round * NW + wave takes the last (round, wave) below it with
            // kept rows, else the carry from the previous chunk.
            static_assert(HPER * NW <= ENM_WAVE, "one lane per (round, wave)");
            const int l = threadIdx.x;
            const bool mine = l < HPER * NW && s_wcnt[l / NW][l % NW] > 0;
            const unsigned long long any = __builtin_amdgcn_ballot_w64(mine);
            const unsigned long long lower = any & ((1ull << l) - 1ull);
            const double carry = s_carry;
            const int hascarry = s_hascarry;
            if (l < HPER * NW) {
                const int src = lower ? 63 - __clzll(lower) : -1;
                s_prev[l / NW][l % NW] = src >= 0 ? s_last[src / NW][src % NW] : carry;
                s_has[l / NW][l % NW] = src >= 0 ? 1 : hascarry;
            }
            // Every lane read the old carry before lane 0 replaces it (the
            // wave runs in lockstep and the reads came first).
            __builtin_amdgcn_wave_barrier();
            if (l == 0 && any) {
                const int top = 63 - __clzll(any);
                s_carry = s_last[top / NW][top % NW];
                s_hascarry = 1;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < HPER; ++j) {
            int before = 0, total = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                before += w < wv ? s_wcnt[j][w] : 0;
                total += s_wcnt[j][w];
            }
            // The kept row before this one: a lower kept lane of the wave,
            // else the last one before the wave.
            const unsigned long long lower =
                __builtin_amdgcn_ballot_w64(pos[j] >= 0) & ((1ull << lane) - 1ull);
            const int64_t o = at + before + __popcll(lower);
            const int pl = lower ? 63 - __clzll(lower) : 0;
            const double pa = __shfl(av[j], pl, ENM_WAVE);
            if (pos[j] >= 0 && o < C.n) {
                hint[C.out + o] = pos[j];
                const bool has = lower ? true : s_has[j][wv] != 0;
                const double ap = lower ? pa : s_prev[j][wv];
                if (has && !is_cut(ap, av[j], m2)) atomicOr(cont + C.cbits + (o >> 5), 1u << (o & 31));
            }
            at += total;
        }
        __syncthreads();   // s_wcnt, s_last, s_prev are rewritten by the next chunk
    }
    if (single && threadIdx.x == 0) bad[ci] = at != C.n ? 1 : 0;
}

// M: the largest noise magnitude (0 without noise), folded into *out as
// IEEE bits (|v| >= 0, so the bit patterns order like the values); *out is
// zeroed before.
__global__ __launch_bounds__(HT)
void hint_noise_max_kernel(const double* __restrict__ noise, int64_t len, unsigned long long* __restrict__ out) {
    __shared__ double red[HT / ENM_WAVE];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * HT + threadIdx.x; i < len; i += (int64_t)gridDim.x * HT)
        m = fmax(m, fabs(noise[i]));
    m = wave_max(m);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) red[threadIdx.x / ENM_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = red[0];
        for (int w = 1; w < HT / ENM_WAVE; ++w) mm = fmax(mm, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(mm));
    }
}

// Re-sort each run of two to ENM_HINT_RUN positions (a position without
// its continuation bit followed by positions with it) by (value, position)
// in place, by the thread of the word holding its start; a longer run marks
// the column.  Only run members' values are gathered.
__global__ __launch_bounds__(HT)
void hint_run_kernel(const HintCol* __restrict__ cols, int64_t wstride, const double* __restrict__ arena,
                     const unsigned* __restrict__ cont, int* __restrict__ hint, int* __restrict__ bad) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int ci = (int)(b / wstride);
    const HintCol C = cols[ci];
    const int64_t t = (b - ci * wstride) * HT + threadIdx.x;   // the word
    const int64_t words = (C.n + 31) / 32;
    if (t >= words || bad[ci]) return;
    const unsigned* cw = cont + C.cbits;
    const uint64_t both = ((uint64_t)(t + 1 < words ? cw[t + 1] : 0u) << 32) | cw[t];
    uint64_t starts = ~both & (both >> 1) & 0xffffffffull;   // no bit at i, bit at i + 1
    if (!starts) return;
    const double* v = arena + C.out;
    int* h = hint + C.out;
    auto is_cont = [&](int64_t p) { return (cw[p >> 5] >> (p & 31)) & 1u; };
    while (starts) {
        const int64_t s = 32 * t + __builtin_ctzll(starts);
        starts &= starts - 1;
        int64_t j = s + 1;
        while (j < C.n && j - s <= ENM_HINT_RUN && is_cont(j)) ++j;
        if (j - s > ENM_HINT_RUN) {
            bad[ci] = 1;
            return;
        }
        for (int64_t u = s + 1; u < j; ++u) {
            const int hu = h[u];
            const double su = v[hu];
            int64_t q = u - 1;
            for (; q >= s; --q) {
                const int hq = h[q];
                const double sq = v[hq];
                if (sq < su || (sq == su && hq < hu)) break;
                h[q + 1] = hq;
            }
            h[q + 1] = hu;
        }
    }
}

// Columns without a hint: hint[out] = -1.
__global__ void hint_mark_kernel(const HintCol* __restrict__ cols, int n_cols, const int* __restrict__ bad,
                                 int* __restrict__ hint) {
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci < n_cols && bad[ci] && cols[ci].n > 0) hint[cols[ci].out] = -1;
}

}  // namespace

hipError_t enm_launch_hint_keys(const HintSeries* series, int n_series, int64_t maxlen, const double* raw,
                                double* keys, hipStream_t st) {
    if (n_series <= 0 || maxlen <= 0) return hipSuccess;
    hipLaunchKernelGGL(hint_keys_kernel, dim3((unsigned)((maxlen + HT - 1) / HT), (unsigned)n_series), dim3(HT), 0,
                       st, series, raw, keys);
    return hipGetLastError();
}

int hint_segments(int n_cols, int64_t max_len, int target_blocks) {
    if (n_cols <= 0 || max_len <= 0 || 2 * (int64_t)n_cols >= target_blocks) return 1;
    const int64_t most = (max_len + ENM_HINT_BLOCK - 1) / ENM_HINT_BLOCK;   // one chunk each
    return (int)std::max<int64_t>(1, std::min<int64_t>(most, (target_blocks + n_cols - 1) / n_cols));
}

int64_t hint_segment_rows(int64_t max_len, int segs_per_col) {
    const int64_t chunks = (max_len + ENM_HINT_BLOCK - 1) / ENM_HINT_BLOCK;
    return std::max<int64_t>(1, (chunks + segs_per_col - 1) / segs_per_col) * ENM_HINT_BLOCK;
}

hipError_t enm_launch_hint_order(const HintCol* cols, int n_cols, int64_t max_w, int64_t max_n, int64_t max_len,
                                 int target_blocks,
                                 const int* sidx, const double* skey, const KeptWord* kept,
                                 const PrepTaskDev* tasks, const double* stats, int rescale, const double* noise,
                                 int64_t noise_len, double* noise_max, unsigned* cont, int64_t cont_words,
                                 const HintSegs& sg, const double* arena, int* hint, int* bad, hipStream_t st) {
    if (n_cols <= 0) return hipSuccess;
    unsigned long long* nm = reinterpret_cast<unsigned long long*>(noise_max);
    hipError_t e = enm_launch_fill_u64(nm, 1, 0ull, st);
    if (e == hipSuccess) e = enm_launch_fill_i32(reinterpret_cast<int*>(cont), cont_words, 0, st);
    if (e != hipSuccess) return e;
    const int64_t len = noise ? noise_len : 0;
    if (len > 0) {
        const int64_t blocks = std::min<int64_t>(1024, (len + 16 * HT - 1) / (16 * HT));
        hipLaunchKernelGGL(hint_noise_max_kernel, dim3((unsigned)blocks), dim3(HT), 0, st, noise, len, nm);
    }
    // Kept-row maps of up to 64 KB go to LDS; larger ones are read in place.
    const int64_t words = (max_w + 31) / 32;
    const size_t lds = (size_t)(words + (words + 1) / 2) * 4;
    const int lds_words = lds <= 65536 ? (int)words : 0;
    // Segments: one per column when the columns fill the GPU's resident
    // workgroups; otherwise enough (of whole chunks) for about that many.
    const int segs_per_col = hint_segments(n_cols, max_len, target_blocks);
    const int64_t seg = hint_segment_rows(max_len, segs_per_col);
    const int64_t nseg = (int64_t)segs_per_col * n_cols;
    if (segs_per_col > 1) {
        hipLaunchKernelGGL(hint_seg_count_kernel, dim3((unsigned)nseg), dim3(HT), lds_words ? lds : 0, st, cols,
                           segs_per_col, seg, sidx, skey, kept, tasks, stats, rescale, lds_words, sg.cnt, sg.last);
        hipLaunchKernelGGL(hint_seg_scan_kernel, dim3((unsigned)n_cols), dim3(ENM_WAVE), 0, st, cols, n_cols,
                           segs_per_col, seg, sg.cnt, sg.last, sg.off, sg.cin, sg.chas, bad);
    }
    hipLaunchKernelGGL(hint_filter_kernel, dim3((unsigned)nseg), dim3(HT), lds_words ? lds : 0, st, cols,
                       segs_per_col, seg, sidx, skey, kept, tasks, stats, rescale, nm, lds_words, sg.off, sg.cin,
                       sg.chas, bad, hint, cont);
    const int64_t wstride = ((max_n + 31) / 32 + HT - 1) / HT;
    if (wstride > 0)
        hipLaunchKernelGGL(hint_run_kernel, dim3((unsigned)(wstride * n_cols)), dim3(HT), 0, st, cols, wstride, arena,
                           cont, hint, bad);
    hipLaunchKernelGGL(hint_mark_kernel, dim3((unsigned)((n_cols + 63) / 64)), dim3(64), 0, st, cols, n_cols, bad,
                       hint);
    return hipGetLastError();
}
