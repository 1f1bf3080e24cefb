// Device-side preparation of estimate_mi tasks: the lag windows, mask and
// NaN compaction, and the rescaling of ennemi's _lagged_mi
// (ennemi/_driver.py:795-806 windows, :834-854 _apply_masks, :871-902
// _rescale_data), producing the float64 instance columns of a batch directly
// in HBM from the raw series.
//
// Every value is bit-identical to NumPy's:
//   * compaction keeps the rows in order (a stable prefix scan per block);
//   * x.mean() and x.std() of a column (and of a one-column condition) sum
//     in NumPy's order: 8192-element chunks, each by its pairwise recursion
//     (halves cut at a multiple of 8, leaves of at most 128 elements summed
//     by 8 interleaved accumulators), the chunk sums added in order to -0.0;
//   * zs.mean(axis=0) and zs.std(axis=0) of a condition of two or more
//     columns accumulate each column row by row;
//   * std = sqrt(sum((v - mean)^2) / n); the scaled value is
//     ((v - mean) / std) + noise, each operation rounded once (the library
//     is built without FP contraction).
// The noise is the reference's: every task draws from a fresh
// default_rng(2_718281828) in the order x, y, z (z row-major), so its draws
// are prefixes of one stream, which the host generates once per call.  A
// variable with std < 1e-20 is left as it is and draws nothing (for the
// condition: if any of its columns is constant).
#include "enm_device.h"
#include "enm_launch.h"

namespace {

constexpr int PB_THREADS = 256;
constexpr int PB_ROWS = ENM_PREP_BLOCK;          // rows per block
constexpr int PB_ROUNDS = PB_ROWS / PB_THREADS;  // rows per thread
constexpr int PW_CHUNK = 8192;                   // NumPy's reduction buffer
constexpr int PW_LEAF = 128;                     // NumPy's pairwise block

// Window row r of task t in two steps, so that a thread issues the loads of
// all its rows before it uses any: row_load reads x, y and the mask byte;
// row_keep says whether the row is kept (the mask first, then, with
// drop_nan, no NaN in x, y or the condition) and whether a row that passes
// the mask holds a NaN.  Rows at or past the window's end are neither kept
// nor NaN (row 0 is read instead, so every load stays inside the series).
__device__ __forceinline__ void row_load(const PrepTaskDev& t, const double* __restrict__ raw,
                                         const unsigned char* __restrict__ mask, int64_t r0,
                                         double (&xv)[PB_ROUNDS], double (&yv)[PB_ROUNDS], int (&mk)[PB_ROUNDS]) {
#pragma unroll
    for (int j = 0; j < PB_ROUNDS; ++j) {
        const int64_t r = r0 + j * PB_THREADS + threadIdx.x;
        const int64_t rr = r < t.w ? r : 0;
        xv[j] = raw[t.x_col + t.x0 + rr];
        yv[j] = raw[t.y_col + t.y0 + rr];
    }
    if (mask != nullptr) {   // uniform: the bytes' loads go out together too
#pragma unroll
        for (int j = 0; j < PB_ROUNDS; ++j) {
            const int64_t r = r0 + j * PB_THREADS + threadIdx.x;
            mk[j] = mask[t.y0 + (r < t.w ? r : 0)];
        }
    } else {
#pragma unroll
        for (int j = 0; j < PB_ROUNDS; ++j) mk[j] = 1;
    }
}

__device__ __forceinline__ bool row_keep(const PrepTaskDev& t, const double* __restrict__ raw, int drop_nan,
                                         int64_t r, double xv, double yv, int mk, bool& has_nan) {
    const bool in = r < t.w;
    const int64_t rr = in ? r : 0;
    const bool pass = in && mk != 0;
    bool nan = isnan(xv) || isnan(yv);
    for (int c = 0; c < t.dz; ++c) nan = nan || isnan(raw[t.z_col[c] + t.z0[c] + rr]);
    has_nan = pass && nan;
    return pass && !(drop_nan && nan);
}

// Blocks are laid out `stride` per task (the most any task needs; the
// rest of a shorter task's blocks find no rows), so a block's task is a
// division, not a search.

// Phase 1: kept rows per block, and a NaN among the masked rows per task.
// Every row's loads are issued before any of them is used.
__global__ __launch_bounds__(PB_THREADS)
void prep_count_kernel(const PrepTaskDev* __restrict__ tasks, int64_t stride, const double* __restrict__ raw,
                       const unsigned char* __restrict__ mask, int drop_nan, int* __restrict__ blk_count,
                       unsigned* __restrict__ nan_flag) {
    __shared__ int s_cnt[PB_THREADS / ENM_WAVE];
    __shared__ int s_nan;
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);   // a task's blocks on one XCD
    const int ti = (int)(b / stride);
    const PrepTaskDev& t = tasks[ti];   // read in place: a copy indexed by c would sit in scratch
    const int64_t r0 = (b - ti * stride) * PB_ROWS;
    if (threadIdx.x == 0) s_nan = 0;
    __syncthreads();
    int cnt = 0;
    bool any_nan = false;
    double xv[PB_ROUNDS], yv[PB_ROUNDS];
    int mk[PB_ROUNDS];
    row_load(t, raw, mask, r0, xv, yv, mk);
#pragma unroll
    for (int j = 0; j < PB_ROUNDS; ++j) {
        bool nan;
        cnt += row_keep(t, raw, drop_nan, r0 + j * PB_THREADS + threadIdx.x, xv[j], yv[j], mk[j], nan) ? 1 : 0;
        any_nan = any_nan || nan;
    }
    const int lane = threadIdx.x & (ENM_WAVE - 1), wv = threadIdx.x / ENM_WAVE;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, ENM_WAVE);
    if (lane == 0) s_cnt[wv] = cnt;
    if (wave_any(any_nan) && lane == 0) s_nan = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < PB_THREADS / ENM_WAVE; ++w) tot += s_cnt[w];
        blk_count[b] = tot;
        if (s_nan && !drop_nan) atomicOr(nan_flag + ti, 1u);
    }
}

// Per task: block offsets (exclusive scan of the block counts, in place in
// blk_off) and the kept rows n.  One wave per task.
__global__ __launch_bounds__(ENM_WAVE)
void prep_scan_kernel(const PrepTaskDev* __restrict__ tasks, int64_t stride, int n_tasks,
                      const int* __restrict__ blk_count, int64_t* __restrict__ blk_off, int64_t* __restrict__ n_out) {
    const int ti = blockIdx.x;
    if (ti >= n_tasks) return;
    const int64_t b0 = ti * stride, nb = tasks[ti].nblk;
    const int lane = threadIdx.x;
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += ENM_WAVE) {
        const int64_t v = base + lane < nb ? blk_count[b0 + base + lane] : 0;
        int64_t inc = v;
        for (int off = 1; off < ENM_WAVE; off <<= 1) {
            const int64_t o = __shfl_up(inc, off, ENM_WAVE);
            if (lane >= off) inc += o;
        }
        if (base + lane < nb) blk_off[b0 + base + lane] = carry + inc - v;
        carry += __shfl(inc, ENM_WAVE - 1, ENM_WAVE);
    }
    if (lane == 0) n_out[ti] = carry;
}

// Phase 2a: the kept rows of every block, in row order, to the task's
// arena columns (x, y, then the condition columns, n values each); with a
// kept-row map, each 32 window rows' kept bits and the kept rows before
// them, for the sort hints.  Every
// row's state and x, y values are read first; one barrier then gives each
// wave the kept rows of the waves before it, round by round.
__global__ __launch_bounds__(PB_THREADS)
void prep_compact_kernel(const PrepTaskDev* __restrict__ tasks, int64_t stride,
                         const double* __restrict__ raw, const unsigned char* __restrict__ mask, int drop_nan,
                         const int64_t* __restrict__ blk_off, double* __restrict__ arena,
                         KeptWord* __restrict__ kept) {
    constexpr int NW = PB_THREADS / ENM_WAVE;
    __shared__ int s_wcnt[PB_ROUNDS][NW];
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);   // a task's blocks on one XCD
    const int ti = (int)(b / stride);
    const PrepTaskDev& t = tasks[ti];   // read in place: a copy indexed by c would sit in scratch
    const int64_t r0 = (b - ti * stride) * PB_ROWS;
    const int lane = threadIdx.x & (ENM_WAVE - 1), wv = threadIdx.x / ENM_WAVE;
    bool keep[PB_ROUNDS];
    double xv[PB_ROUNDS], yv[PB_ROUNDS];
    int below[PB_ROUNDS];
    int mk[PB_ROUNDS];
    row_load(t, raw, mask, r0, xv, yv, mk);
#pragma unroll
    for (int j = 0; j < PB_ROUNDS; ++j) {
        bool nan;
        keep[j] = row_keep(t, raw, drop_nan, r0 + j * PB_THREADS + threadIdx.x, xv[j], yv[j], mk[j], nan);
    }
    unsigned long long wm[PB_ROUNDS];
#pragma unroll
    for (int j = 0; j < PB_ROUNDS; ++j) {
        const unsigned long long m = __builtin_amdgcn_ballot_w64(keep[j]);
        wm[j] = m;
        below[j] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[j][wv] = __popcll(m);
    }
    __syncthreads();
    int64_t pos = blk_off[b];
    double* out = arena + t.arena_off;
#pragma unroll
    for (int j = 0; j < PB_ROUNDS; ++j) {
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            before += w < wv ? s_wcnt[j][w] : 0;
            total += s_wcnt[j][w];
        }
        const int64_t r = r0 + j * PB_THREADS + threadIdx.x;
        const int64_t p = pos + before + below[j];
        if (keep[j]) {
            out[p] = xv[j];
            out[t.n + p] = yv[j];
            for (int c = 0; c < t.dz; ++c) out[(2 + c) * t.n + p] = raw[t.z_col[c] + t.z0[c] + r];
        }
        if (kept != nullptr && lane < 2) {
            // The wave's 64 rows are two words of the map.
            const int64_t row = r0 + j * PB_THREADS + wv * ENM_WAVE + 32 * lane;
            if (row < t.w) {
                const uint32_t lo = (uint32_t)wm[j];
                const uint32_t bits = lane ? (uint32_t)(wm[j] >> 32) : lo;
                const int64_t pre = pos + before + (lane ? __popc(lo) : 0);
                kept[t.pmap + row / 32] = KeptWord{bits, (int32_t)pre};
            }
        }
        pos += total;
    }
}

// One term of a sum: the value, or its squared deviation from `mean`.
struct Term {
    const double* v;
    double mean;
    bool dev;
    __device__ __forceinline__ double operator()(int64_t i) const {
        if (!dev) return v[i];
        const double d = v[i] - mean;
        return d * d;
    }
};

// NumPy's pairwise recursion over n <= PW_CHUNK values as a heap-ordered
// binary tree (node 1 the root, nodes 2i and 2i+1 the halves of node i): a
// node of more than PW_LEAF values splits at half its length rounded down
// to a multiple of 8; smaller nodes are leaves.  Right halves are at most
// half their parent plus 8 values, so every leaf of an 8191-value chunk is
// within depth 7: heap slots 1 .. PW_HEAP - 1.
constexpr int PW_HEAP = 256;
constexpr int PW_DEPTH = 7;

// The sum of a full PW_CHUNK-value chunk in NumPy's order, by one wave: the
// recursion's 64 leaves of 128 values form a balanced tree.  Lane l takes
// accumulator l % 8 of leaf 8 * pass + l / 8 (sixteen values, added in
// order), so one load instruction reads eight 64-byte runs.  Butterfly
// shuffles then combine the accumulators as the leaf does,
// ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)), and the pass's eight
// leaves pairwise; IEEE addition commutes, so both lanes of a pair hold the
// same bits.  The eight passes' subtree sums end the tree.
__device__ __forceinline__ double pw_full_chunk(const Term& f) {
    const int lane = threadIdx.x & (ENM_WAVE - 1);
    const int j = lane & 7, g = lane >> 3;
    double part[8];
#pragma unroll
    for (int pass = 0; pass < 8; ++pass) {
        const int64_t base = (int64_t)(pass * 8 + g) * PW_LEAF + j;
        double v[PW_LEAF / 8];
#pragma unroll
        for (int i = 0; i < PW_LEAF / 8; ++i) v[i] = f(base + 8 * i);
        double acc = v[0];
#pragma unroll
        for (int i = 1; i < PW_LEAF / 8; ++i) acc += v[i];
        acc += __shfl_xor(acc, 1, ENM_WAVE);
        acc += __shfl_xor(acc, 2, ENM_WAVE);
        acc += __shfl_xor(acc, 4, ENM_WAVE);
        acc += __shfl_xor(acc, 8, ENM_WAVE);
        acc += __shfl_xor(acc, 16, ENM_WAVE);
        acc += __shfl_xor(acc, 32, ENM_WAVE);
        part[pass] = acc;
    }
    return ((part[0] + part[1]) + (part[2] + part[3])) + ((part[4] + part[5]) + (part[6] + part[7]));
}

// Phase 2b: one wave per (array, chunk): the pairwise sum of a chunk of
// values (pass 0) or of squared deviations from the array's mean (pass 1).
// A full chunk takes pw_full_chunk; otherwise the wave builds the
// recursion's tree in LDS level by level (slot lengths; -1 for no node),
// sums its leaves, eight lanes to a leaf, and adds every node's two halves
// bottom-up, left + right, as NumPy does.
__global__ __launch_bounds__(ENM_WAVE)
void prep_chunk_kernel(const PrepSumDev* __restrict__ sums, const int* __restrict__ chunk_sum_of,
                       const int* __restrict__ chunk_index, const int* __restrict__ list, int n_list,
                       const double* __restrict__ arena, const double* __restrict__ stats, int pass,
                       double* __restrict__ chunk_val) {
    __shared__ int s_lo[PW_HEAP], s_len[PW_HEAP];
    __shared__ double s_val[PW_HEAP];
    __shared__ short s_leaf[PW_HEAP];
    __shared__ int s_nl;
    if ((int)blockIdx.x >= n_list) return;
    const int c = list[blockIdx.x];
    const PrepSumDev& sd = sums[chunk_sum_of[c]];
    const int64_t first = (int64_t)chunk_index[c] * PW_CHUNK;
    const int m = (int)min((int64_t)PW_CHUNK, sd.n - first);
    const Term f{arena + sd.src + first, pass ? stats[sd.stat] : 0.0, pass != 0};
    if (m == PW_CHUNK) {
        const double sum = pw_full_chunk(f);
        if (threadIdx.x == 0) chunk_val[c] = sum;
        return;
    }
    const int lane = (int)threadIdx.x;
    // The tree, top-down: slot 1 is the whole chunk; a slot's halves exist
    // when it has more than PW_LEAF values.
    for (int i = lane; i < PW_HEAP; i += ENM_WAVE) s_len[i] = -1;
    if (lane == 0) {
        s_lo[1] = 0;
        s_len[1] = m;
        s_nl = 0;
    }
    __syncthreads();
    for (int d = 0; d < PW_DEPTH; ++d) {
        for (int i = (1 << d) + lane; i < (2 << d); i += ENM_WAVE) {
            const int len = s_len[i];
            if (len > PW_LEAF) {
                int n2 = len / 2;
                n2 -= n2 % 8;
                s_lo[2 * i] = s_lo[i];
                s_len[2 * i] = n2;
                s_lo[2 * i + 1] = s_lo[i] + n2;
                s_len[2 * i + 1] = len - n2;
            }
        }
        __syncthreads();
    }
    // The leaves, listed in any order (each is summed on its own).
    for (int i = 1 + lane; i < PW_HEAP; i += ENM_WAVE) {
        const int len = s_len[i];
        if (len >= 0 && len <= PW_LEAF) s_leaf[atomicAdd(&s_nl, 1)] = (short)i;
    }
    __syncthreads();
    // A leaf is NumPy's pairwise_sum leaf: fewer than 8 values summed from
    // 0.0; up to 128 by 8 interleaved accumulators combined as
    // ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), the remainder added after.
    // Eight lanes per leaf, eight leaves per round: lane j of a group takes
    // accumulator j (its up to sixteen values loaded at once, added in
    // order), butterfly shuffles within the group combine the accumulators,
    // and the group's first lane adds the remainder.
    const int j = lane & 7;
    const int nl = s_nl;
    for (int l0 = 0; l0 < nl; l0 += ENM_WAVE / 8) {
        const int l = l0 + (lane >> 3);
        const int slot = l < nl ? s_leaf[l] : 0;
        const int lo = l < nl ? s_lo[slot] : 0, len = l < nl ? s_len[slot] : 0;
        const int full = len >= 8 ? len - len % 8 : 0;
        double v[PW_LEAF / 8];
#pragma unroll
        for (int i = 0; i < PW_LEAF / 8; ++i) v[i] = 8 * i < full ? f(lo + 8 * i + j) : 0.0;
        double acc = v[0];
#pragma unroll
        for (int i = 1; i < PW_LEAF / 8; ++i)
            if (8 * i < full) acc += v[i];
        acc += __shfl_xor(acc, 1, ENM_WAVE);
        acc += __shfl_xor(acc, 2, ENM_WAVE);
        acc += __shfl_xor(acc, 4, ENM_WAVE);
        if (j == 0 && l < nl) {
            double res;
            if (len < 8) {
                res = 0.0;
                for (int t = 0; t < len; ++t) res += f(lo + t);
            } else {
                res = acc;
                for (int t = full; t < len; ++t) res += f(lo + t);
            }
            s_val[slot] = res;
        }
    }
    __syncthreads();
    // Bottom-up: every split node is its left half plus its right half.
    for (int d = PW_DEPTH - 1; d >= 0; --d) {
        for (int i = (1 << d) + lane; i < (2 << d); i += ENM_WAVE)
            if (s_len[i] > PW_LEAF) s_val[i] = s_val[2 * i] + s_val[2 * i + 1];
        __syncthreads();
    }
    if (lane == 0) chunk_val[c] = s_val[1];
}

// Full chunks only, one per wave, four waves per block (no barriers: each
// wave sums its own chunk).
__global__ __launch_bounds__(4 * ENM_WAVE)
void prep_full_chunk_kernel(const PrepSumDev* __restrict__ sums, const int* __restrict__ chunk_sum_of,
                            const int* __restrict__ chunk_index, const int* __restrict__ list, int n_list,
                            const double* __restrict__ arena, const double* __restrict__ stats, int pass,
                            double* __restrict__ chunk_val) {
    const int idx = (int)blockIdx.x * 4 + (int)(threadIdx.x / ENM_WAVE);
    if (idx >= n_list) return;
    const int c = list[idx];
    const PrepSumDev& sd = sums[chunk_sum_of[c]];
    const int64_t first = (int64_t)chunk_index[c] * PW_CHUNK;
    const Term f{arena + sd.src + first, pass ? stats[sd.stat] : 0.0, pass != 0};
    const double sum = pw_full_chunk(f);
    if ((threadIdx.x & (ENM_WAVE - 1)) == 0) chunk_val[c] = sum;
}

// Rows per LDS tile of a row-by-row sum, and rows each lane loads per tile.
constexpr int SEQ_TILE = 512;
constexpr int SEQ_PER = SEQ_TILE / ENM_WAVE;

// Phase 2c: per array, the chunk sums added in order to -0.0 (pairwise
// arrays), or the whole column accumulated row by row (columns of a
// multi-column condition); then mean = sum / n (pass 0) or
// std = sqrt(sum / n) (pass 1).  One wave per array.  The row-by-row sum is
// one chain of dependent additions.  The wave loads 512 rows at a time
// (coalesced, the next 512 in flight meanwhile), forms the terms, stages
// them in LDS and every lane adds them in row order from broadcast LDS
// reads issued a batch ahead: the same additions, in the same order, with
// nothing but the additions on the chain.  The sum starts at -0.0, which
// leaves the first row's term unchanged (x + -0.0 == x for every x), as a
// sum that starts from the first row does.
__global__ __launch_bounds__(ENM_WAVE)
void prep_finish_kernel(const PrepSumDev* __restrict__ sums, int n_sums, const double* __restrict__ arena,
                        const double* __restrict__ chunk_val, int pass, double* __restrict__ stats) {
    __shared__ double2 s_rows[SEQ_TILE / 2];
    const int s = blockIdx.x;
    if (s >= n_sums) return;
    const PrepSumDev& sd = sums[s];
    const int lane = threadIdx.x;
    double acc = -0.0;
    if (sd.sequential) {
        const double* v = arena + sd.src;
        const int64_t n = sd.n;
        const double mean = pass ? stats[sd.stat] : 0.0;
        auto term = [&](int64_t i) {
            double t = i < n ? v[i] : 0.0;
            if (pass) {
                t = t - mean;
                t = t * t;
            }
            return t;
        };
        double cur[SEQ_PER];
#pragma unroll
        for (int j = 0; j < SEQ_PER; ++j) cur[j] = term(j * ENM_WAVE + lane);
        double* rows = reinterpret_cast<double*>(s_rows);
        for (int64_t base = 0; base < n; base += SEQ_TILE) {
            double nxt[SEQ_PER];
#pragma unroll
            for (int j = 0; j < SEQ_PER; ++j) nxt[j] = term(base + SEQ_TILE + j * ENM_WAVE + lane);
#pragma unroll
            for (int j = 0; j < SEQ_PER; ++j) rows[j * ENM_WAVE + lane] = cur[j];
            __syncthreads();
            if (n - base >= SEQ_TILE) {
                // Sixteen rows per batch; the next batch's reads are issued
                // before this batch's additions.
                double2 w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = s_rows[q];
#pragma unroll
                for (int r = 0; r < SEQ_TILE / 2; r += 8) {
                    double2 u[8];
                    if (r + 8 < SEQ_TILE / 2) {
#pragma unroll
                        for (int q = 0; q < 8; ++q) u[q] = s_rows[r + 8 + q];
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        acc = acc + w[q].x;
                        acc = acc + w[q].y;
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) w[q] = u[q];
                }
            } else {
                const int m = (int)(n - base);
                for (int r = 0; r < m; ++r) acc = acc + rows[r];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SEQ_PER; ++j) cur[j] = nxt[j];
        }
    } else {
        for (int c = 0; c < sd.n_chunks; ++c) acc += chunk_val[sd.chunk0 + c];
    }
    if (lane != 0) return;
    const double q = acc / (double)sd.n;
    if (pass == 0) stats[sd.stat] = q;
    else stats[sd.stat + 1] = sqrt(q);
}

// Phase 2d: which variables are rescaled, where their noise starts, and
// their flag bits (bit 0 x, bit 1 y, bit 2 z constant).  One thread per task.
__global__ void prep_noise_kernel(PrepTaskDev* __restrict__ tasks, int n_tasks, const double* __restrict__ stats,
                                  unsigned* __restrict__ const_flag) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n_tasks) return;
    PrepTaskDev& t = tasks[ti];
    unsigned fl = 0;
    if (fabs(stats[t.stat + 1]) < 1e-20) fl |= 1u;
    if (fabs(stats[t.stat + 3]) < 1e-20) fl |= 2u;
    for (int c = 0; c < t.dz; ++c)
        if (fabs(stats[t.stat + 4 + 2 * c + 1]) < 1e-20) fl |= 4u;
    int64_t drawn = 0;
    t.noise_x = drawn;
    if (!(fl & 1u)) drawn += t.n;
    t.noise_y = drawn;
    if (!(fl & 2u)) drawn += t.n;
    t.noise_z = drawn;
    t.flags = fl;
    const_flag[ti] = fl;
}

// Phase 2e: (v - mean) / std + noise for every rescaled value, in place.
__global__ __launch_bounds__(PB_THREADS)
void prep_scale_kernel(const PrepTaskDev* __restrict__ tasks, int64_t stride, const double* __restrict__ stats,
                       const double* __restrict__ noise, double* __restrict__ arena) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);   // a task's blocks on one XCD
    const int ti = (int)(b / stride);
    const PrepTaskDev& t = tasks[ti];   // read in place: a copy indexed by c would sit in scratch
    const int64_t i = (b - ti * stride) * PB_THREADS + threadIdx.x;
    if (i >= t.n) return;
    double* col = arena + t.arena_off;
    if (!(t.flags & 1u)) {
        const double v = (col[i] - stats[t.stat]) / stats[t.stat + 1];
        col[i] = v + noise[t.noise_x + i];
    }
    if (!(t.flags & 2u)) {
        const double v = (col[t.n + i] - stats[t.stat + 2]) / stats[t.stat + 3];
        col[t.n + i] = v + noise[t.noise_y + i];
    }
    if (!(t.flags & 4u)) {
        for (int c = 0; c < t.dz; ++c) {
            const double mean = stats[t.stat + 4 + 2 * c], sd = stats[t.stat + 5 + 2 * c];
            const double v = (col[(2 + c) * t.n + i] - mean) / sd;
            col[(2 + c) * t.n + i] = v + noise[t.noise_z + i * t.dz + c];
        }
    }
}

}  // namespace

hipError_t enm_launch_prep_count(const PrepTaskDev* tasks, int64_t stride, int n_tasks,
                                 const double* raw, const unsigned char* mask, int drop_nan, int* blk_count,
                                 int64_t* blk_off, unsigned* nan_flag, int64_t* n_out, hipStream_t st) {
    const int64_t n_blocks = stride * n_tasks;
    if (n_tasks <= 0) return hipSuccess;
    hipError_t e = enm_launch_fill_i32(reinterpret_cast<int*>(nan_flag), n_tasks, 0, st);
    if (e != hipSuccess) return e;
    if (n_blocks > 0)
        hipLaunchKernelGGL(prep_count_kernel, dim3((unsigned)n_blocks), dim3(PB_THREADS), 0, st, tasks, stride,
                           raw, mask, drop_nan, blk_count, nan_flag);
    hipLaunchKernelGGL(prep_scan_kernel, dim3((unsigned)n_tasks), dim3(ENM_WAVE), 0, st, tasks, stride, n_tasks,
                       blk_count, blk_off, n_out);
    return hipGetLastError();
}

hipError_t enm_launch_prep_write(PrepTaskDev* tasks, int64_t stride, int n_tasks,
                                 const double* raw, const unsigned char* mask, int drop_nan, int rescale,
                                 const int64_t* blk_off, const PrepSumDev* sums, int n_sums,
                                 const int* chunk_sum_of, const int* chunk_index, const int* chunk_list,
                                 int n_chunks, int n_full, double* chunk_val, double* stats, int64_t sstride,
                                 const double* noise, unsigned* const_flag, double* arena, KeptWord* kept,
                                 hipStream_t st) {
    if (n_tasks <= 0) return hipSuccess;
    const int64_t n_blocks = stride * n_tasks, n_sblocks = sstride * n_tasks;
    if (n_blocks > 0)
        hipLaunchKernelGGL(prep_compact_kernel, dim3((unsigned)n_blocks), dim3(PB_THREADS), 0, st, tasks, stride,
                           raw, mask, drop_nan, blk_off, arena, kept);
    if (!rescale) return hipGetLastError();
    for (int pass = 0; pass < 2; ++pass) {
        if (n_chunks > 0) {
            // Full chunks four to a block; the other (last) chunks one each.
            if (n_full > 0)
                hipLaunchKernelGGL(prep_full_chunk_kernel, dim3((unsigned)((n_full + 3) / 4)), dim3(4 * ENM_WAVE), 0,
                                   st, sums, chunk_sum_of, chunk_index, chunk_list, n_full, arena, stats, pass,
                                   chunk_val);
            if (n_chunks > n_full)
                hipLaunchKernelGGL(prep_chunk_kernel, dim3((unsigned)(n_chunks - n_full)), dim3(ENM_WAVE), 0, st,
                                   sums, chunk_sum_of, chunk_index, chunk_list + n_full, n_chunks - n_full, arena,
                                   stats, pass, chunk_val);
        }
        hipLaunchKernelGGL(prep_finish_kernel, dim3((unsigned)n_sums), dim3(ENM_WAVE), 0, st, sums, n_sums, arena,
                           chunk_val, pass, stats);
    }
    hipLaunchKernelGGL(prep_noise_kernel, dim3((unsigned)((n_tasks + 63) / 64)), dim3(64), 0, st, tasks, n_tasks,
                       stats, const_flag);
    if (n_sblocks > 0)
        hipLaunchKernelGGL(prep_scale_kernel, dim3((unsigned)n_sblocks), dim3(PB_THREADS), 0, st, tasks, sstride,
                           stats, noise, arena);
    return hipGetLastError();
}
