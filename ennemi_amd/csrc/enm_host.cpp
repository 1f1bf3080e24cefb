// C ABI and host planner of libennemi_hip.so.
//
// enm_run_batch replaces the task fan-out of the reference
// (_map_maybe_parallel over _lagged_mi, ennemi/_driver.py:492,715,736-785):
// every instance of the call is planned into one set of work lists, and one
// launch sequence on one stream runs them all:
//
//   fill     zero the counters, maxima and statistics (kernels, so the
//            sequence can be captured in a hipGraph)
//   sort     segmented sort of each distinct key column (block sort + merges)
//   gather   each order's columns in sorted order, FP64 + FP32, magnitudes
//   knn      persistent work-queue kernels, one per (dimension, list depth)
//   count    persistent work-queue kernels, 1-D and per dimension
//   combine  one wave per instance: psi sums -> estimate
//   export   optional per-point radii and counts
//
// Plans are cached by the bytes of the call (instances, class sizes, device
// pointers, flags, tuning).  With ENM_F_GRAPH the launch sequence of a plan is
// captured once into a hipGraph and replayed on later identical calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "enm_internal.h"
#include "enm_launch.h"
#include "../../include/ennemi_hip.h"

namespace {

std::string g_init_error;

// A device buffer: owned (hipMalloc) or borrowed from the static workspace.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    hipError_t ensure(size_t b) {
        if (b <= bytes && p && !borrowed) return hipSuccess;
        release();
        if (b == 0) b = 256;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        else p = nullptr;
        return e;
    }
    void borrow(void* q, size_t b) {
        release();
        p = q;
        bytes = b;
        borrowed = true;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

struct Group {
    int is_count;
    int dclass;
    int pclass;
    int first;      // first item in the kernel's item list
    int n_items;
    int counter;    // first of its ENM_NQ + ENM_NDEF counters in the counter array,
                    // ENM_CSTRIDE slots apart (queue tickets, then the
                    // deferred 1-D query counts)
    int qoff;       // first of its ENM_NQ + 1 queue offsets in the offset table
    int inst;       // instance (per-instance mode) or -1
    int walk;       // kNN items on grid orders (per-lane walk kernel)
    int k1max;      // largest k+1 of its kNN items (sizes heap lists)
    int waves;      // persistent grid: waves per CU
    int ticket;     // items per queue ticket (0: automatic)
    int smax;       // strip items: the largest strip (LDS slots)
};

struct ExportSpace {
    int64_t counts;   // res32 offset
    int64_t perm;     // perm offset of the order
};

struct ExportInst {
    int64_t point_off;
    int64_t n;
    int64_t eps;      // res64 offset (order A positions)
    int64_t permA;    // perm offset of order A
    int nspace;
    ExportSpace sp[ENM_MAX_SPACES];
};

struct Plan {
    std::string key;
    // host tables
    std::vector<SortSeg> segs, segs2;        // stage 1 (arena keys), stage 2 (grid strips)
    std::vector<GatherItem> gitems, gitems2;
    std::vector<OrderDev> orders;
    std::vector<KnnDev> kdescs;
    std::vector<CountDev> cdescs;
    std::vector<InstDev> insts;
    std::vector<RedDev> rdescs;
    std::vector<RedItem> ritems;
    std::vector<std::pair<int, int>> inst_ritems;   // per instance: first item, count
    std::vector<double> cls;
    std::vector<Item> kitems, citems;
    std::vector<Item> sitems;                        // psi tiles of 1-D spaces
    std::vector<std::pair<int, int>> inst_sitems;    // per instance: first item, count
    std::vector<int> qoffs;                          // queue offsets of every group
    std::vector<Group> groups;
    std::vector<ExportInst> exports;
    int ntiles = 0, ntiles2 = 0;
    int tile2 = ENM_SORT_TILE;   // block-sort tile of the second stage
    // Sample sort of the first stage (segments longer than one block-sort
    // tile, none longer than ENM_SS_MAXLEN); otherwise merge passes.
    bool sample_sort = false;
    std::vector<SampleSeg> ssegs;    // per first-stage segment
    std::vector<SortSeg> samp_segs;  // sample ranges to sort
    std::vector<SsCons> cons;        // orders laid out from each segment
    std::vector<int> qbase;          // run queue offsets (ENM_NQ + 1)
    int samp_tiles = 0;
    int64_t samp_maxlen = 0, n_samp = 0, n_counts = 0, n_buckets = 0;
    int64_t maxlen = 0, maxlen2 = 0;
    int64_t sort_elems = 0, sort_elems2 = 0;
    int64_t n_col64 = 0, n_col32 = 0, n_perm = 0, n_res64 = 0, n_res32 = 0;
    int64_t psi_tab = 0, psi_n = 0;   // res64 offset of psi(0..psi_n)
    int64_t max_c1_queries = 0;       // 1-D count queries of the largest group
    unsigned deferred_seg = 0;        // capacity of one deferred 1-D query list segment
    int n_inst = 0;
    bool f32 = true, brute = false, per_instance = false;
    bool want_eps = false;   // the call exports radii (in point order)
    bool hinted = false;     // the call passes sort hints (first-stage segments of whole columns)
    // Kernels that read the FP32 column copies: the brute-force kNN and
    // multi-D counts off the grid (band filter).  Only then are they made.
    bool need32 = false;
    // device
    int64_t heap_bytes = 0;   // global heap lists of kNN groups with k+1 > ENM_LIST_LDS_MAX
    int64_t far_bytes = 0;    // far list of the strip kNN (queries reaching past the staged strips)
    DevBuf far,
        tables, sk0, si0, sk1, si1, split, col64, col32, perm, res64, res32, maxabs, counters, stats,
        deferred, heap, samp_k0, samp_k1, samp_i0, samp_i1, samp_split, bid, ss_counts, ss_tiles, ss_long, ss_ctl,
        hint_bad;
    size_t off_segs = 0, off_gitems = 0, off_orders = 0, off_kdescs = 0, off_cdescs = 0, off_insts = 0,
           off_cls = 0, off_kitems = 0, off_citems = 0, off_segs2 = 0, off_gitems2 = 0, off_rdescs = 0,
           off_ritems = 0, off_sitems = 0, off_qoffs = 0, off_ssegs = 0, off_samp_segs = 0, off_cons = 0,
           off_qbase = 0, table_bytes = 0;
    void* pinned = nullptr;
    hipEvent_t upload_done = nullptr;
    hipGraphExec_t graph = nullptr;
    const void* graph_ptrs[5] = {};   // arena, out, eps, counts, hints the graph was captured with
    int n_launch = 0;
    bool timing = false;       // events recorded by the last enqueue
    hipEvent_t ev[ENM_N_STAGES + 1] = {};
    ~Plan() {
        for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (upload_done) (void)hipEventDestroy(upload_done);
        if (pinned) (void)hipHostFree(pinned);
    }
};

}  // namespace

// Workspace of enm_run_batch_static.  Every GPU that loads this code object
// holds its own copy until the process ends.  In host code the name is only
// a stand-in: the copy of the current device is at the address
// hipGetSymbolAddress gives (kept per device in g_scratch), and that address
// is what kernels and copies receive.
static __device__ __align__(256) unsigned char scratch[ENM_STATIC_WORKSPACE_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch;   // device -> address of its scratch copy
}  // namespace

struct enm_ctx {
    int device = 0;
    int num_cus = 256;
    std::string err;
    // Tile shapes: the default, and the shapes enm_autotune picked per
    // (joint dimension class, list class, floor(log2 n)).  Plans are cached
    // by tuning_gen, which every change bumps.
    enm_tuning tuning{1024, 32, 0, 0};
    // Until a tuning is set, instances of three or more joint columns take
    // wide strips: their grid sweeps scan the union of a wave's windows,
    // which narrows as strips grow.
    bool wide_multid = true;
    std::map<std::tuple<int, int, int>, enm_tuning> tuned;
    uint64_t tuning_gen = 0;
    std::list<std::unique_ptr<Plan>> cache;   // most recent first
    Plan* last = nullptr;
    hipStream_t last_stream = nullptr;
    // Device-side task preparation: the state enm_prep_count leaves for
    // enm_prep_write, and their workspace.
    struct Prep {
        std::vector<PrepTaskDev> tasks;
        std::vector<int64_t> n;
        int64_t stride = 0;       // compaction blocks per task (the most any task needs)
        const void* raw = nullptr;
        const void* mask = nullptr;
        uint32_t drop_nan = 0;
        bool counted = false;
        std::vector<char> blob1, blob2;   // host tables of the last uploads
        DevBuf tables1, tables2, blk_count, blk_off, nan, nout, chunk_val, stats;
        // Sort hints: the kept-row map written by the last enm_prep_write
        // with ENM_P_HINTS (pm_total words), and what that call was given.
        int64_t pm_total = 0;
        size_t o_tasks = 0;       // the task table in tables2 (constant flags set there)
        bool written = false;
        std::vector<int64_t> arena_off;
        const void* arena = nullptr;
        const void* noise = nullptr;
        int64_t noise_len = 0;
        uint32_t flags = 0;
        std::vector<char> blob3;
        DevBuf kept, tables3, hkeys, hk0, hk1, hi0, hi1, hsplit, hcont, hbad, hmax, sg_cnt, sg_last, sg_off,
            sg_cin, sg_chas;
    } prep;
    // Workspace of enm_unique_rows.
    struct Unique {
        DevBuf tables, k0, k1, i0, i1, split, ccodes, cstarts, ccounts, cdist, rkeys, rstarts, rdist;
        std::vector<char> blob1, blob2;
    } uq;
    void* scratch = nullptr;          // this device's static workspace
    std::string scratch_err;          // why it is unavailable, if it is
    std::unique_ptr<Plan> static_plan;   // plan of the last enm_run_batch_static
};

namespace {

const size_t kMaxCachedPlans = 4;

int fail(enm_ctx* ctx, int code, const std::string& msg) {
    ctx->err = msg;
    return code;
}

int hip_fail(enm_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, ENM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Planning
// ---------------------------------------------------------------------------

struct Planner {
    Plan& P;
    const enm_instance* inst;
    const int64_t* seg_len;
    std::map<std::tuple<int64_t, int64_t, int>, int> sort_groups;  // (key src, n, segmented) -> group
    std::vector<std::vector<int>> group_segs;                      // group -> sort segment ids
    std::vector<std::vector<int64_t>> group_pos0;                  // group -> order position of each segment
    std::map<std::string, int> order_ids;
    std::vector<int> order_group;

    explicit Planner(Plan& p, const enm_instance* in, const int64_t* sl) : P(p), inst(in), seg_len(sl) {}

    int sort_group(int ii, int64_t key_src, bool segmented) {
        const enm_instance& I = inst[ii];
        const int nseg = segmented ? I.n_segs : 1;
        // Class segments make the key for distinct class layouts differ.
        auto key = std::make_tuple(key_src, (int64_t)I.n, segmented ? (int)(I.seg_index * 64 + nseg) : -1);
        auto it = sort_groups.find(key);
        if (it != sort_groups.end()) return it->second;
        const int g = (int)group_segs.size();
        sort_groups[key] = g;
        group_segs.emplace_back();
        group_pos0.emplace_back();
        int64_t p0 = 0;
        for (int s = 0; s < nseg; ++s) {
            const int64_t len = segmented ? seg_len[I.seg_index + s] : I.n;
            if (len > 0) {
                SortSeg sg;
                sg.src = key_src + p0;
                sg.sbase = P.sort_elems;
                sg.len = (int32_t)len;
                sg.first_tile = P.ntiles;
                sg.p0 = (int32_t)p0;
                // A whole column's order can come from the caller's hints.
                sg.hint = P.hinted && !segmented ? 1 : 0;
                P.ntiles += (int)((len + ENM_SORT_TILE - 1) / ENM_SORT_TILE);
                P.sort_elems += len;
                P.maxlen = std::max(P.maxlen, len);
                group_segs[g].push_back((int)P.segs.size());
                group_pos0[g].push_back(p0);
                P.segs.push_back(sg);
            }
            p0 += len;
        }
        return g;
    }

    // An order: the instance's points sorted by cols_src[0] within each scope
    // range, with the columns cols_src gathered.
    int order(int ii, const std::vector<int64_t>& cols_src, bool segmented) {
        const enm_instance& I = inst[ii];
        const int g = sort_group(ii, cols_src[0], segmented);
        std::string key((const char*)cols_src.data(), cols_src.size() * sizeof(int64_t));
        key += std::string((const char*)&g, sizeof(g));
        auto it = order_ids.find(key);
        if (it != order_ids.end()) return it->second;
        OrderDev od;
        std::memset(&od, 0, sizeof(od));
        od.n = I.n;
        od.ncol = (int32_t)cols_src.size();
        od.inst = ii;
        od.c64 = P.n_col64;
        od.c32 = P.n_col32;
        od.perm = P.n_perm;
        od.base_perm = -1;
        od.base_c64 = -1;
        od.base_c32 = -1;
        od.xy = -1;
        od.xy_only = 0;
        for (size_t c = 0; c < cols_src.size(); ++c) od.src[c] = cols_src[c];
        P.n_col64 += od.ncol * I.n;
        P.n_col32 += od.ncol * I.n;
        P.n_perm += I.n;
        const int id = (int)P.orders.size();
        P.orders.push_back(od);
        order_ids[key] = id;
        order_group.push_back(g);
        for (size_t t = 0; t < group_segs[g].size(); ++t) {
            const int s = group_segs[g][t];
            for (int off = 0; off < P.segs[s].len; off += ENM_SORT_TILE) {
                GatherItem gi;
                gi.order = id;
                gi.sseg = s;
                gi.off = off;
                gi.pos0 = (int32_t)group_pos0[g][t];
                P.gitems.push_back(gi);
            }
        }
        return id;
    }

    // A grid order over base order `base`: each scope range cut into strips of
    // S consecutive base positions, each strip sorted by the base order's
    // column 1 (second-stage sort), all columns gathered in that order.
    int grid_order(int base, int S) {
        std::string key = "grid";
        key += std::string((const char*)&base, sizeof base) + std::string((const char*)&S, sizeof S);
        auto it = order_ids.find(key);
        if (it != order_ids.end()) return it->second;
        const OrderDev bo = P.orders[base];
        OrderDev od = bo;
        od.xy = -1;
        od.xy_only = 0;
        od.c64 = P.n_col64;
        od.c32 = P.n_col32;
        od.perm = P.n_perm;
        od.base_perm = bo.perm;
        od.base_c64 = bo.c64;
        od.base_c32 = bo.c32;
        P.n_col64 += od.ncol * od.n;
        P.n_col32 += od.ncol * od.n;
        P.n_perm += od.n;
        const int id = (int)P.orders.size();
        P.orders.push_back(od);
        order_ids[key] = id;
        order_group.push_back(order_group[base]);
        for (auto rg : ranges(base)) {
            for (int st = rg.first; st < rg.second; st += S) {
                const int len = std::min(S, rg.second - st);
                SortSeg sg;
                sg.src = bo.c64 + 1 * bo.n + st;  // column 1 of the base order, in col64
                sg.sbase = P.sort_elems2;
                sg.len = len;
                sg.first_tile = P.ntiles2;
                sg.p0 = st;
                sg.hint = 0;
                P.ntiles2 += (len + P.tile2 - 1) / P.tile2;
                P.sort_elems2 += len;
                P.maxlen2 = std::max<int64_t>(P.maxlen2, len);
                const int sid = (int)P.segs2.size();
                P.segs2.push_back(sg);
                for (int off = 0; off < len; off += ENM_SORT_TILE)
                    P.gitems2.push_back(GatherItem{id, sid, off, st});
            }
        }
        return id;
    }

    // Scope ranges [lo, hi) of an order (class segments or all points).
    std::vector<std::pair<int, int>> ranges(int order_id) {
        std::vector<std::pair<int, int>> r;
        const int g = order_group[order_id];
        for (size_t t = 0; t < group_segs[g].size(); ++t) {
            const int lo = (int)group_pos0[g][t];
            r.emplace_back(lo, lo + P.segs[group_segs[g][t]].len);
        }
        return r;
    }
};

// Chunked point-order sum of n values at res64 offset src; returns the
// partials' offset and sets the chunk count.
int64_t add_reduction(Plan& P, int64_t src, int64_t n, int log_values, int* nchunks) {
    RedDev rd;
    std::memset(&rd, 0, sizeof(rd));
    rd.src = src;
    rd.n = n;
    rd.log = log_values;
    rd.part = P.n_res64;
    const int chunks = (int)((n + ENM_RED_CHUNK - 1) / ENM_RED_CHUNK);
    P.n_res64 += chunks;
    const int d = (int)P.rdescs.size();
    P.rdescs.push_back(rd);
    for (int c = 0; c < chunks; ++c) P.ritems.push_back(RedItem{d, c});
    *nchunks = chunks;
    return rd.part;
}

struct PendingItem {
    Item it;
    int64_t cost;
    int inst;
    int dclass, pclass;
    int walk = 0;   // kNN on a grid order: the per-lane walk kernel
    int k1 = 0;     // k+1 of its instance (kNN items)
    int waves = 32, ticket = 0;   // its instance's tile shape
    int strip = 0;  // strip items: the strip size
};

int grid_for(const enm_ctx* ctx, int waves_per_cu, int n_items) {
    const int g = ctx->num_cus * waves_per_cu;
    return std::max(1, std::min(n_items, g));
}

std::tuple<int, int, int> shape_key(int dims, int k, int64_t n) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << (b + 1)) <= n) ++b;
    return std::make_tuple(enm_knn_dim_class(dims), enm_knn_list_class(k + 1), b);
}

// The tile shape the planner uses for an instance of `dims` joint columns.
enm_tuning tuning_for(const enm_ctx* ctx, int dims, int k, int64_t n) {
    auto it = ctx->tuned.find(shape_key(dims, k, n));
    if (it != ctx->tuned.end()) return it->second;
    enm_tuning t = ctx->tuning;
    if (ctx->wide_multid && dims >= 3 && t.strip_points > 0) t.strip_points = ENM_STRIP_WIDE;
    return t;
}

int plan_batch(enm_ctx* ctx, Plan& P, const enm_instance* inst, int n_inst, const int64_t* seg_len) {
    Planner pl(P, inst, seg_len);
    std::vector<PendingItem> kpend, cpend;
    P.n_inst = n_inst;
    // Every instance's tile shape; grid strips up to 2048 points are sorted
    // by one block each, with the smallest block tile that holds the
    // largest strip.
    std::vector<enm_tuning> tunes(n_inst);
    std::set<int> cols_read;   // orders whose column-major columns a kernel reads
    int max_strip = 0;
    for (int ii = 0; ii < n_inst; ++ii) {
        const int cols = std::max(0, inst[ii].dx) + std::max(0, inst[ii].dy) + std::max(0, inst[ii].dz);
        tunes[ii] = tuning_for(ctx, cols, inst[ii].k, inst[ii].n);
        if (cols >= 2) max_strip = std::max(max_strip, (int)tunes[ii].strip_points);
    }
    P.tile2 = max_strip <= 512 ? 512 : max_strip <= 1024 ? 1024 : ENM_SORT_TILE;
    int64_t point_off = 0;
    for (int ii = 0; ii < n_inst; ++ii) {
        const enm_instance& I = inst[ii];
        const int dx = I.dx, dy = I.dy, dz = I.dz;
        const int ncols = dx + dy + dz;
        char msg[256];
        auto bad = [&](const char* what) {
            std::snprintf(msg, sizeof msg, "instance %d: %s", ii, what);
            return fail(ctx, ENM_E_ARG, msg);
        };
        if (I.kind < ENM_KIND_KSG || I.kind > ENM_KIND_KL) return bad("unknown kind");
        if (I.k < 1) return bad("k must be >= 1");
        if (I.n < 1 || I.n > 0x7fffffffLL) return bad("n out of range");
        if (dx < 0 || dy < 0 || dz < 0 || ncols < 1 || ncols > ENM_MAX_COLS) return bad("bad column counts");
        const bool classes = I.kind == ENM_KIND_ROSS || I.kind == ENM_KIND_CROSS;
        if (I.n_segs < 1 || (!classes && I.n_segs != 1)) return bad("bad n_segs");
        if (classes) {
            int64_t tot = 0;
            for (int s = 0; s < I.n_segs; ++s) tot += seg_len[I.seg_index + s];
            if (tot != I.n) return bad("class sizes do not sum to n");
        }
        switch (I.kind) {
            case ENM_KIND_KSG: if (dx < 1 || dy < 1 || dz != 0) return bad("KSG needs dx, dy >= 1, dz = 0"); break;
            case ENM_KIND_CMI: if (dx < 1 || dy < 1 || dz < 1) return bad("CMI needs dx, dy, dz >= 1"); break;
            case ENM_KIND_ROSS: if (dx < 1 || dy != 0 || dz != 0) return bad("ROSS needs dx >= 1 only"); break;
            case ENM_KIND_CROSS: if (dx < 1 || dy != 0 || dz < 1) return bad("CROSS needs dx, dz >= 1, dy = 0"); break;
            case ENM_KIND_KL: if (dx < 1 || dy != 0 || dz != 0) return bad("KL needs dx >= 1 only"); break;
        }
        std::vector<int64_t> X(I.col, I.col + dx), Y(I.col + dx, I.col + dx + dy), Z(I.col + dx + dy, I.col + ncols);

        InstDev id;
        std::memset(&id, 0, sizeof(id));
        id.kind = I.kind;
        id.k = I.k;
        id.n = I.n;
        id.dx = dx;
        id.n_segs = I.n_segs;
        id.seg_len = -1;
        for (int s = 0; s < ENM_MAX_SPACES; ++s) { id.part[s] = 0; id.zero[s] = -1; id.ntile[s] = 0; }

        ExportInst ex;
        std::memset(&ex, 0, sizeof(ex));
        ex.point_off = point_off;
        ex.n = I.n;
        point_off += I.n;

        // Orders and spaces of this kind.  The kNN order's columns put the
        // condition's first two columns first (CMI, conditional Ross), so the
        // key and the grid's second axis are shared by every count space;
        // each space lists its columns in ascending order (the max-norm does
        // not depend on column order), so it starts with columns 0 and 1
        // whenever it contains both.
        int A = -1;
        std::vector<int64_t> acols;
        std::vector<int> posX, posY, posZ;
        const bool segA = classes;
        auto put = [&](std::vector<int>& pos, int64_t col) { pos.push_back((int)acols.size()); acols.push_back(col); };
        if (I.kind == ENM_KIND_CMI || I.kind == ENM_KIND_CROSS) {
            put(posZ, Z[0]);
            if (dz >= 2) put(posZ, Z[1]);
            for (int64_t c : X) put(posX, c);
            for (int64_t c : Y) put(posY, c);
            for (int t = 2; t < dz; ++t) put(posZ, Z[t]);
        } else {
            for (int64_t c : X) put(posX, c);
            for (int64_t c : Y) put(posY, c);
        }
        A = pl.order(ii, acols, segA);
        // The kNN runs on a grid order (strips of the key-sorted order, each
        // sorted by the second column) when there are two or more columns.
        const enm_tuning& T = tunes[ii];
        const int S = T.strip_points;
        const bool grid = !P.brute && S > 0 && acols.size() >= 2;
        const int K = grid ? pl.grid_order(A, S) : A;
        if (grid && acols.size() == 2 && P.orders[K].xy < 0) {
            // The per-lane walk reads (column 0, column 1) pairs.
            P.n_col64 += P.n_col64 & 1;
            P.orders[K].xy = P.n_col64;
            P.n_col64 += 2 * I.n;
        }

        struct SpaceSpec { int order; std::vector<int> cols; bool segmented; };
        std::vector<SpaceSpec> spaces;
        auto iota = [](int a, int b) { std::vector<int> v; for (int t = a; t < b; ++t) v.push_back(t); return v; };
        auto join = [](std::vector<int> a, const std::vector<int>& b) {
            a.insert(a.end(), b.begin(), b.end());
            std::sort(a.begin(), a.end());
            return a;
        };
        bool need_orig = false;
        if (I.kind == ENM_KIND_KSG) {
            spaces.push_back({A, posX, false});
            const int Yo = pl.order(ii, Y, false);
            spaces.push_back({Yo, iota(0, dy), false});
        } else if (I.kind == ENM_KIND_CMI) {
            spaces.push_back({A, join(posX, posZ), false});
            spaces.push_back({A, join(posY, posZ), false});
            spaces.push_back({A, posZ, false});
        } else if (I.kind == ENM_KIND_ROSS) {
            const int B = I.n_segs > 1 ? pl.order(ii, X, false) : A;
            spaces.push_back({B, iota(0, dx), false});
        } else if (I.kind == ENM_KIND_CROSS) {
            const int B = I.n_segs > 1 ? pl.order(ii, acols, false) : A;
            spaces.push_back({B, iota(0, (int)acols.size()), false});
            spaces.push_back({A, posZ, true});
            spaces.push_back({B, posZ, false});
        }
        // A multi-dimensional space whose first two columns are the order's
        // columns 0 and 1 counts on the grid order over that order.
        std::vector<int> space_strip(spaces.size(), 0);
        std::vector<int64_t> space_base_key(spaces.size(), -1);
        for (size_t t = 0; t < spaces.size(); ++t) {
            SpaceSpec& sp = spaces[t];
            if (grid && sp.cols.size() >= 2 && sp.cols[0] == 0 && sp.cols[1] == 1) {
                space_base_key[t] = P.orders[sp.order].c64;
                sp.order = pl.grid_order(sp.order, S);
                space_strip[t] = S;
            }
        }
        // The radius in kNN order is kept for spaces on that order and for
        // the export; other spaces read it through the point numbering.
        bool need_k = P.want_eps;
        for (auto& s : spaces) {
            if (s.order != K) need_orig = true;
            else need_k = true;
        }
        // Entropy sums log radii in point order, whatever the kNN order.
        if (I.kind == ENM_KIND_KL) need_orig = true;

        // kNN descriptor and items.
        KnnDev kd;
        std::memset(&kd, 0, sizeof(kd));
        kd.order = K;
        kd.inst = ii;
        kd.strip = grid ? S : 0;
        kd.base_key = P.orders[A].c64;
        kd.k = I.k;
        kd.kl = I.kind == ENM_KIND_KL;
        kd.eps = -1;
        if (need_k) { kd.eps = P.n_res64; P.n_res64 += I.n; }
        kd.eps_orig = -1;
        if (need_orig) { kd.eps_orig = P.n_res64; P.n_res64 += I.n; }
        kd.part = -1;
        const int kdi = (int)P.kdescs.size();
        const int D = (int)acols.size();
        const int dcl = enm_knn_dim_class(D);
        // A register list may be given more slots than k+1 needs (the
        // extra ones stay pinned at -inf).
        const int pcl = I.k + 1 <= 32 && T.list_slots > 0 ? enm_knn_list_class(std::max(I.k + 1, (int)T.list_slots))
                                                          : enm_knn_list_class(I.k + 1);
        if (dcl < 0) return bad("joint dimension above 16");
        if (I.k > 0x3fffffff) return bad("k out of range");
        int ktiles = 0;
        std::vector<PendingItem> mine;
        // Two-column grid orders with register lists and strips that three
        // LDS slots hold: a strip per workgroup.
        const bool strip_items = grid && D == 2 && pcl > 0 && S <= ENM_STRIP_LDS_MAX;
        for (auto rg : pl.ranges(K)) {
            if (strip_items) {
                for (int t0 = rg.first; t0 < rg.second; t0 += S) {
                    PendingItem pi;
                    pi.it = Item{kdi, ktiles++, rg.first, rg.second, t0, std::min(t0 + S, rg.second)};
                    pi.cost = (std::min(t0 + S, rg.second) - t0 + ENM_TILE - 1) / ENM_TILE;
                    pi.inst = ii; pi.dclass = dcl; pi.pclass = pcl; pi.walk = 2;
                    pi.k1 = I.k + 1;
                    pi.waves = T.waves_per_cu; pi.ticket = T.items_per_ticket;
                    pi.strip = S;
                    mine.push_back(pi);
                }
                continue;
            }
            // Tiles never straddle a strip: strips start at lo + s*strip and
            // the strip size is a multiple of the tile size.
            for (int t0 = rg.first; t0 < rg.second; t0 += ENM_TILE) {
                PendingItem pi;
                pi.it = Item{kdi, ktiles++, rg.first, rg.second, t0, std::min(t0 + ENM_TILE, rg.second)};
                // Work per tile: the scope range in brute force, about
                // constant with window pruning.
                pi.cost = P.brute ? (int64_t)(rg.second - rg.first) : 1;
                // Per-lane walks on two-column grids; with more columns the
                // gridded two prune less and tiles share candidates.
                pi.inst = ii; pi.dclass = dcl; pi.pclass = pcl; pi.walk = grid && D == 2 ? 1 : 0;
                pi.k1 = I.k + 1;
                pi.waves = T.waves_per_cu; pi.ticket = T.items_per_ticket;
                mine.push_back(pi);
            }
        }
        const int red_first = (int)P.ritems.size();
        const int sit_first = (int)P.sitems.size();
        if (kd.kl) {
            // Entropy: log radii summed in point order.
            int chunks = 0;
            id.nspace = 1;
            id.part[0] = add_reduction(P, kd.eps_orig, I.n, 1, &chunks);
            id.zero[0] = -1;
            id.ntile[0] = chunks;
        }
        P.kdescs.push_back(kd);
        if (!strip_items) cols_read.insert(K);
        kpend.insert(kpend.end(), mine.begin(), mine.end());
        ex.eps = kd.eps;
        ex.permA = P.orders[K].perm;

        // Count descriptors and items.
        if (!kd.kl) id.nspace = (int)spaces.size();
        for (size_t s = 0; s < spaces.size(); ++s) {
            const SpaceSpec& sp = spaces[s];
            CountDev cd;
            std::memset(&cd, 0, sizeof(cd));
            cd.strip = space_strip[s];
            cd.base_key = space_base_key[s];
            cd.order = sp.order;
            cd.inst = ii;
            cd.ncol = (int)sp.cols.size();
            for (int c = 0; c < cd.ncol; ++c) cd.cols[c] = sp.cols[c];
            cd.same_order = sp.order == K;
            cd.eps = cd.same_order ? kd.eps : kd.eps_orig;
            cd.eps_perm = P.orders[sp.order].perm;
            cd.counts = P.n_res32;
            P.n_res32 += I.n;
            if (cd.ncol >= 2 && cd.strip == 0) P.need32 = true;
            const int cdi = (int)P.cdescs.size();
            const int ccl = enm_count_dim_class(cd.ncol);
            cd.psi_pt = -1;
            int ctiles = 0;
            std::vector<std::pair<int, int>> rgs;
            if (sp.segmented) rgs = pl.ranges(sp.order);
            else rgs.emplace_back(0, (int)I.n);
            // 1-D counts take ENM_C1_ITEM positions per item (and psi tiles
            // of ENM_PSI_TILE), multi-D ENM_TILE.
            const int step = ccl == 1 ? ENM_C1_ITEM : ENM_TILE;
            for (auto rg : rgs) {
                for (int t0 = rg.first; t0 < rg.second; t0 += step) {
                    PendingItem pi;
                    pi.it = Item{cdi, ccl == 1 ? 0 : ctiles++, rg.first, rg.second, t0,
                                 std::min(t0 + step, rg.second)};
                    pi.cost = P.brute && ccl != 1 ? (int64_t)(rg.second - rg.first) : (int64_t)step;
                    pi.inst = ii; pi.dclass = ccl; pi.pclass = 0;
                    pi.waves = T.waves_per_cu; pi.ticket = T.items_per_ticket;
                    cpend.push_back(pi);
                }
                if (ccl == 1)
                    for (int t0 = rg.first; t0 < rg.second; t0 += ENM_PSI_TILE)
                        P.sitems.push_back(Item{cdi, ctiles++, rg.first, rg.second, t0,
                                                std::min(t0 + ENM_PSI_TILE, rg.second)});
            }
            if (ccl == 1) {
                // 1-D spaces sum per tile of their key-sorted order.
                cd.part = P.n_res64;
                P.n_res64 += ctiles;
                cd.zero = P.n_res32;
                P.n_res32 += ctiles;
                id.part[s] = cd.part;
                id.zero[s] = cd.zero;
                id.ntile[s] = ctiles;
            } else {
                // Multi-dimensional spaces write psi in point order.
                int chunks = 0;
                cd.psi_pt = P.n_res64;
                P.n_res64 += I.n;
                cd.part = -1;
                cd.zero = -1;
                id.part[s] = add_reduction(P, cd.psi_pt, I.n, 0, &chunks);
                id.zero[s] = -1;
                id.ntile[s] = chunks;
            }
            ex.sp[s].counts = cd.counts;
            ex.sp[s].perm = cd.eps_perm;
            P.cdescs.push_back(cd);
        }
        ex.nspace = (int)spaces.size();
        P.inst_ritems.emplace_back(red_first, (int)P.ritems.size() - red_first);
        P.inst_sitems.emplace_back(sit_first, (int)P.sitems.size() - sit_first);
        P.psi_n = std::max<int64_t>(P.psi_n, I.n);
        if (I.kind == ENM_KIND_ROSS) {
            id.seg_len = (int64_t)P.cls.size();
            for (int s = 0; s < I.n_segs; ++s) P.cls.push_back((double)seg_len[I.seg_index + s]);
        }
        P.insts.push_back(id);
        P.exports.push_back(ex);
    }

    // Grid orders read through their interleaved copy only (by the strip
    // kNN) skip the column-major layout.
    for (const CountDev& cd : P.cdescs) cols_read.insert(cd.order);
    for (size_t o = 0; o < P.orders.size(); ++o)
        if (P.orders[o].xy >= 0 && !cols_read.count((int)o)) P.orders[o].xy_only = 1;

    // Work lists: grouped by kernel class (or per instance in per-instance
    // mode).  Inside a group the items go to ENM_NQ queues, one per XCD: each
    // instance's items in blocks of consecutive tiles (one block unless the
    // instance is a large share of the group), blocks placed largest first
    // on the least-loaded queue.
    auto build = [&](std::vector<PendingItem>& pend, std::vector<Item>& out, int is_count) {
        std::stable_sort(pend.begin(), pend.end(), [&](const PendingItem& a, const PendingItem& b) {
            if (P.per_instance && a.inst != b.inst) return a.inst < b.inst;
            if (a.dclass != b.dclass) return a.dclass < b.dclass;
            if (a.pclass != b.pclass) return a.pclass < b.pclass;
            if (a.walk != b.walk) return a.walk < b.walk;
            return a.cost > b.cost;
        });
        // A deferred 1-D query code holds the item index in its upper bits.
        const size_t max_group = (size_t)1 << (32 - ENM_C1_OFF_BITS);
        size_t i = 0;
        while (i < pend.size()) {
            size_t j = i;
            while (j < pend.size() && pend[j].dclass == pend[i].dclass && pend[j].pclass == pend[i].pclass &&
                   pend[j].walk == pend[i].walk && (!P.per_instance || pend[j].inst == pend[i].inst) &&
                   j - i < max_group)
                ++j;
            Group g;
            g.is_count = is_count;
            g.dclass = pend[i].dclass;
            g.pclass = pend[i].pclass;
            g.first = (int)out.size();
            g.n_items = (int)(j - i);
            g.counter = (int)P.groups.size() * (ENM_NQ + ENM_NDEF) * ENM_CSTRIDE;
            g.inst = P.per_instance ? pend[i].inst : -1;
            g.walk = pend[i].walk;
            g.k1max = 1;
            g.smax = 0;
            for (size_t t = i; t < j; ++t) {
                g.k1max = std::max(g.k1max, pend[t].k1);
                g.smax = std::max(g.smax, pend[t].strip);
            }
            // The group runs with the shape of its costliest item's instance.
            g.waves = pend[i].waves;
            g.ticket = pend[i].ticket;
            if (!is_count && g.pclass == ENM_LIST_HEAP_GLOBAL) {
                const int hg = enm_knn_heap_grid(grid_for(ctx, g.waves, g.n_items), g.k1max);
                P.heap_bytes = std::max<int64_t>(P.heap_bytes, (int64_t)hg * ENM_TILE * g.k1max * 8);
            }
            if (!is_count && g.walk == 2) {
                // Room for one query in 16 of the group (about 1 % reach past
                // their neighbours on the lag scans), at most 64 MiB; the
                // groups run one after the other and share it.
                const int64_t rec = 16 + 8 * (int64_t)g.pclass;
                const int64_t want = (int64_t)g.n_items * g.smax / 16 * rec;
                P.far_bytes = std::max<int64_t>(P.far_bytes, std::min<int64_t>(want, (int64_t)64 << 20));
            }
            const size_t blk = std::max<size_t>(64, (j - i + 4 * ENM_NQ - 1) / (4 * ENM_NQ));
            struct Block { int64_t cost; size_t a, b; };
            std::vector<Block> blocks;
            for (size_t t = i; t < j;) {
                size_t u = t;
                int64_t c = 0;
                while (u < j && pend[u].inst == pend[t].inst && u - t < blk) c += pend[u++].cost;
                blocks.push_back(Block{c, t, u});
                t = u;
            }
            std::stable_sort(blocks.begin(), blocks.end(),
                             [](const Block& a, const Block& b) { return a.cost > b.cost; });
            std::vector<std::vector<size_t>> queue(ENM_NQ);
            std::vector<int64_t> load(ENM_NQ, 0);
            for (const Block& b : blocks) {
                const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[q] += b.cost;
                for (size_t t = b.a; t < b.b; ++t) queue[q].push_back(t);
            }
            g.qoff = (int)P.qoffs.size();
            int64_t queries = 0;
            for (int q = 0; q < ENM_NQ; ++q) {
                P.qoffs.push_back((int)(out.size() - g.first));
                for (size_t t : queue[q]) {
                    out.push_back(pend[t].it);
                    queries += pend[t].it.t1 - pend[t].it.t0;
                }
            }
            P.qoffs.push_back((int)(out.size() - g.first));
            if (is_count && g.dclass == 1) P.max_c1_queries = std::max(P.max_c1_queries, queries);
            P.groups.push_back(g);
            i = j;
        }
    };
    build(kpend, P.kitems, 0);
    build(cpend, P.citems, 1);
    // Stage-1 gathers read their columns at random positions of one
    // instance: interleave the items so that block b (dispatched to XCD
    // b % ENM_NQ) only serves the instances of queue b % ENM_NQ, and each
    // instance's columns are cached in one XCD's L2.
    {
        std::vector<std::vector<GatherItem>> lists(ENM_NQ);
        for (const GatherItem& gi : P.gitems) lists[P.orders[gi.order].inst % ENM_NQ].push_back(gi);
        std::vector<GatherItem> mixed;
        mixed.reserve(P.gitems.size());
        for (size_t k = 0; mixed.size() < P.gitems.size(); ++k)
            for (int q = 0; q < ENM_NQ; ++q)
                if (k < lists[q].size()) mixed.push_back(lists[q][k]);
        P.gitems.swap(mixed);
    }
    // First stage: sample sort when segments need more than one block-sort
    // tile and none is longer than ENM_SS_MAXLEN.
    P.sample_sort = P.maxlen > ENM_SORT_TILE && P.maxlen <= ENM_SS_MAXLEN;
    if (P.sample_sort) {
        // The orders each segment is laid out into (the first-stage gather
        // items, one per order and segment), and the run queue of the
        // segment: by instance, as the gather items are interleaved.
        std::vector<std::vector<SsCons>> seg_cons(P.segs.size());
        std::set<std::pair<int, int>> seen;
        for (const GatherItem& gi : P.gitems) {
            if (!seen.insert({gi.sseg, gi.order}).second) continue;
            SsCons c;
            c.pos0 = gi.pos0;
            c.order = gi.order;
            c.pad = 0;
            seg_cons[gi.sseg].push_back(c);
        }
        std::vector<int64_t> qload(ENM_NQ, 0);
        for (size_t si = 0; si < P.segs.size(); ++si) {
            const SortSeg& sg = P.segs[si];
            SampleSeg q;
            std::memset(&q, 0, sizeof(q));
            std::sort(seg_cons[si].begin(), seg_cons[si].end(),
                      [](const SsCons& a, const SsCons& b) { return a.order < b.order; });
            q.cons0 = (int32_t)P.cons.size();
            q.ncons = (int32_t)seg_cons[si].size();
            P.cons.insert(P.cons.end(), seg_cons[si].begin(), seg_cons[si].end());
            q.q = seg_cons[si].empty() ? 0 : P.orders[seg_cons[si][0].order].inst % ENM_NQ;
            q.nb = (int32_t)std::min<int64_t>(ENM_SS_MAXB, (sg.len + ENM_SS_BUCKET - 1) / ENM_SS_BUCKET);
            q.ns = q.nb > 1 ? ENM_SS_OVER * q.nb : 0;   // < len: nb > 1 means len > ENM_SS_BUCKET
            q.samp = (int32_t)P.n_samp;
            q.cbase = P.n_counts;
            if (q.ns > 0) {
                SortSeg s2;
                s2.src = q.samp;
                s2.sbase = q.samp;
                s2.len = q.ns;
                s2.first_tile = P.samp_tiles;
                s2.p0 = 0;
                s2.hint = 0;
                P.samp_tiles += (q.ns + ENM_SORT_TILE - 1) / ENM_SORT_TILE;
                P.samp_maxlen = std::max<int64_t>(P.samp_maxlen, q.ns);
                P.samp_segs.push_back(s2);
            }
            P.n_samp += q.ns;
            P.n_counts += (int64_t)((sg.len + ENM_SORT_TILE - 1) / ENM_SORT_TILE) * q.nb;
            P.n_buckets += q.nb;
            qload[q.q] += q.nb;   // a queue holds at most its segments' buckets
            P.ssegs.push_back(q);
        }
        P.qbase.assign(ENM_NQ + 1, 0);
        for (int q = 0; q < ENM_NQ; ++q) P.qbase[q + 1] = P.qbase[q] + (int)qload[q];
    }
    // psi(0..max n) for the count kernels.
    P.psi_tab = P.n_res64;
    P.n_res64 += P.psi_n + 1;
    // Deferred 1-D queries: about 1 % on Gaussian-like data; a full list
    // makes the rest search in place, so the capacity only bounds memory.
    P.deferred_seg = (unsigned)std::min<int64_t>(
        std::max<int64_t>(P.max_c1_queries / 8 / ENM_NDEF, 1024), (1 << 24) / ENM_NDEF);
    return ENM_OK;
}

template <class T>
size_t append(std::vector<char>& blob, const std::vector<T>& v) {
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// The plan's device tables, packed into one host blob (256-byte aligned).
std::vector<char> pack_tables(Plan& P) {
    std::vector<char> blob;
    P.off_segs = append(blob, P.segs);
    P.off_gitems = append(blob, P.gitems);
    P.off_orders = append(blob, P.orders);
    P.off_kdescs = append(blob, P.kdescs);
    P.off_cdescs = append(blob, P.cdescs);
    P.off_insts = append(blob, P.insts);
    P.off_cls = append(blob, P.cls);
    P.off_kitems = append(blob, P.kitems);
    P.off_citems = append(blob, P.citems);
    P.off_segs2 = append(blob, P.segs2);
    P.off_gitems2 = append(blob, P.gitems2);
    P.off_rdescs = append(blob, P.rdescs);
    P.off_ritems = append(blob, P.ritems);
    P.off_sitems = append(blob, P.sitems);
    P.off_qoffs = append(blob, P.qoffs);
    P.off_ssegs = append(blob, P.ssegs);
    P.off_samp_segs = append(blob, P.samp_segs);
    P.off_cons = append(blob, P.cons);
    P.off_qbase = append(blob, P.qbase);
    P.table_bytes = blob.size();
    return blob;
}

struct BufSpec {
    DevBuf* buf;
    size_t bytes;
    const char* what;
};

// Every device buffer of a plan and its size in bytes.
std::vector<BufSpec> workspace_layout(Plan& P) {
    std::vector<BufSpec> v;
    const int64_t se = std::max(P.sort_elems, P.sort_elems2);
    // FP32 column copies exist only when a kernel reads them.
    const bool with32 = P.f32 && P.need32;
    v.push_back({&P.tables, P.table_bytes, "tables"});
    v.push_back({&P.sk0, (size_t)se * 8, "sort keys"});
    v.push_back({&P.sk1, (size_t)se * 8, "sort keys"});
    v.push_back({&P.si0, (size_t)se * 4, "sort index"});
    v.push_back({&P.si1, (size_t)se * 4, "sort index"});
    v.push_back({&P.split, (size_t)std::max(P.ntiles, P.ntiles2) * 4 + 4, "merge splits"});
    v.push_back({&P.col64, (size_t)P.n_col64 * 8, "columns"});
    v.push_back({&P.col32, with32 ? (size_t)P.n_col32 * 4 : 256, "FP32 columns"});
    v.push_back({&P.perm, (size_t)P.n_perm * 4, "permutations"});
    v.push_back({&P.res64, (size_t)P.n_res64 * 8, "radii"});
    v.push_back({&P.res32, (size_t)P.n_res32 * 4, "counts"});
    v.push_back({&P.maxabs, (size_t)P.n_inst * 8, "maxima"});
    v.push_back({&P.counters, P.groups.size() * (ENM_NQ + ENM_NDEF) * ENM_CSTRIDE * 4 + 4, "counters"});
    v.push_back({&P.deferred, (size_t)P.deferred_seg * ENM_NDEF * 4, "deferred queries"});
    v.push_back({&P.stats, ST_N * 8, "stats"});
    if (P.heap_bytes > 0) v.push_back({&P.heap, (size_t)P.heap_bytes, "heap lists"});
    if (P.far_bytes > 0) v.push_back({&P.far, (size_t)P.far_bytes, "strip kNN far list"});
    if (P.sample_sort) {
        v.push_back({&P.samp_k0, (size_t)P.n_samp * 8 + 8, "samples"});
        v.push_back({&P.samp_k1, (size_t)P.n_samp * 8 + 8, "samples"});
        v.push_back({&P.samp_i0, (size_t)P.n_samp * 4 + 4, "samples"});
        v.push_back({&P.samp_i1, (size_t)P.n_samp * 4 + 4, "samples"});
        v.push_back({&P.samp_split, (size_t)P.samp_tiles * 4 + 4, "sample splits"});
        v.push_back({&P.bid, (size_t)P.sort_elems * 2 + 2, "buckets"});
        v.push_back({&P.ss_counts, (size_t)P.n_counts * 4 + 4, "bucket counts"});
        v.push_back({&P.ss_tiles, (size_t)P.n_buckets * sizeof(SsTile), "bucket runs"});
        v.push_back({&P.ss_long, (size_t)P.n_buckets * sizeof(SsTile), "long bucket runs"});
        v.push_back({&P.ss_ctl, (2 * ENM_NQ + 1) * ENM_CSTRIDE * 4, "run queue counters"});
        if (P.hinted) v.push_back({&P.hint_bad, P.segs.size() * 4 + 4, "hint checks"});
    }
    return v;
}

int allocate_and_upload(enm_ctx* ctx, Plan& P, hipStream_t st) {
    std::vector<char> blob = pack_tables(P);
    hipError_t e;
    if (P.upload_done) {
        e = hipEventSynchronize(P.upload_done);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipEventSynchronize");
    } else {
        e = hipEventCreateWithFlags(&P.upload_done, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipEventCreate");
    }
    if (P.pinned) { (void)hipHostFree(P.pinned); P.pinned = nullptr; }
    e = hipHostMalloc(&P.pinned, std::max<size_t>(P.table_bytes, 256), hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipHostMalloc");
    std::memcpy(P.pinned, blob.data(), P.table_bytes);

    for (const BufSpec& b : workspace_layout(P)) {
        if (b.buf->ensure(b.bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(ctx, ENM_E_NOMEM, std::string("hipMalloc failed for ") + b.what);
        }
    }
    e = hipMemcpyAsync(P.tables.p, P.pinned, P.table_bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(tables)");
    e = hipEventRecord(P.upload_done, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    return ENM_OK;
}

template <class T>
const T* tab(const Plan& P, size_t off) {
    return reinterpret_cast<const T*>(P.tables.as<char>() + off);
}

// Enqueue the whole launch sequence of a plan on `st`.  With `timing`, an
// event is recorded on `st` before the first and after every stage.
int enqueue(enm_ctx* ctx, Plan& P, const double* arena, const int* hint, double* out, double* d_eps, int* d_counts,
            hipStream_t st, bool timing) {
    hipError_t e;
    int nl = 0;
    auto chk = [&](hipError_t er, const char* what) -> int {
        if (er != hipSuccess) return hip_fail(ctx, er, what);
        return ENM_OK;
    };
    int rc;
    P.timing = timing;
    if (timing) {
        for (hipEvent_t& ev : P.ev)
            if (!ev && (e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(ctx, e, "hipEventCreate");
    }
    auto mark = [&](int stage) -> int {
        if (!timing) return ENM_OK;
        return chk(hipEventRecord(P.ev[stage], st), "hipEventRecord");
    };
    if ((rc = mark(0))) return rc;
    unsigned long long* stats = P.stats.as<unsigned long long>();
    if ((rc = chk(enm_launch_fill_u64(stats, ST_N, 0ull, st), "fill"))) return rc;
    if ((rc = chk(enm_launch_fill_u64(P.maxabs.as<unsigned long long>(), P.n_inst, 0ull, st), "fill"))) return rc;
    if ((rc = chk(enm_launch_fill_i32(P.counters.as<int>(),
                                      (int64_t)P.groups.size() * (ENM_NQ + ENM_NDEF) * ENM_CSTRIDE, 0, st),
                  "fill")))
        return rc;
    if ((rc = chk(enm_launch_psi_table(P.res64.as<double>() + P.psi_tab, P.psi_n, st), "fill"))) return rc;
    nl += 4;
    const bool hints = P.hinted && P.sample_sort && hint != nullptr;
    if (hints) {
        if ((rc = chk(enm_launch_fill_i32(P.hint_bad.as<int>(), (int64_t)P.segs.size() + 1, 0, st), "fill")))
            return rc;
        ++nl;
    }
    if ((rc = mark(1))) return rc;
    int fb = 0;
    if (P.sample_sort) {
        SampleSortArgs a;
        a.segs = tab<SortSeg>(P, P.off_segs);
        a.nsegs = (int)P.segs.size();
        a.ntiles = P.ntiles;
        a.ss = tab<SampleSeg>(P, P.off_ssegs);
        a.samp_segs = tab<SortSeg>(P, P.off_samp_segs);
        a.n_samp_segs = (int)P.samp_segs.size();
        a.samp_tiles = P.samp_tiles;
        a.samp_maxlen = P.samp_maxlen;
        a.samp_k0 = P.samp_k0.as<double>();
        a.samp_k1 = P.samp_k1.as<double>();
        a.samp_i0 = P.samp_i0.as<int>();
        a.samp_i1 = P.samp_i1.as<int>();
        a.samp_split = P.samp_split.as<int>();
        a.bid = P.bid.as<unsigned short>();
        a.counts = P.ss_counts.as<int>();
        a.tiles = P.ss_tiles.as<SsTile>();
        a.qbase = tab<int>(P, P.off_qbase);
        a.qctl = P.ss_ctl.as<unsigned>();
        a.long_runs = P.ss_long.as<SsTile>();
        a.long_grid = std::max(1, ctx->num_cus);
        // ss_tile_kernel holds 50 KB of LDS: three blocks per CU.
        a.tile_grid = std::max(1, ctx->num_cus * 3);
        a.cons = tab<SsCons>(P, P.off_cons);
        a.orders = tab<OrderDev>(P, P.off_orders);
        a.col64 = P.col64.as<double>();
        a.col32 = P.col32.as<float>();
        a.perm = P.perm.as<int>();
        a.maxabs = P.maxabs.as<unsigned long long>();
        a.with32 = P.f32 && P.need32 ? 1 : 0;
        a.hint = hints ? hint : nullptr;
        a.hint_bad = hints ? P.hint_bad.as<int>() : nullptr;
        a.all_hinted = hints && std::all_of(P.segs.begin(), P.segs.end(), [](const SortSeg& g) { return g.hint != 0; });
        if ((rc = chk(enm_launch_sample_sort(a, arena, P.sk0.as<double>(), P.si0.as<int>(), P.sk1.as<double>(),
                                             P.si1.as<int>(), &nl, st), "sort")))
            return rc;
    } else if ((rc = chk(enm_launch_sort(tab<SortSeg>(P, P.off_segs), (int)P.segs.size(), P.ntiles, P.maxlen,
                                         ENM_SORT_TILE, arena, nullptr,
                                         P.sk0.as<double>(), P.si0.as<int>(), P.sk1.as<double>(), P.si1.as<int>(),
                                         P.split.as<int>(), &fb, &nl, st), "sort"))) {
        return rc;
    }
    if ((rc = mark(2))) return rc;
    const double* skey = fb ? P.sk1.as<double>() : P.sk0.as<double>();
    const int* sidx = fb ? P.si1.as<int>() : P.si0.as<int>();
    // (The sample sort lays its segments out into the orders itself.)
    if (!P.sample_sort && (rc = chk(enm_launch_gather(tab<GatherItem>(P, P.off_gitems), (int)P.gitems.size(),
                                    tab<SortSeg>(P, P.off_segs), tab<OrderDev>(P, P.off_orders), arena, skey,
                                    sidx, P.col64.as<double>(), P.col32.as<float>(), P.perm.as<int>(),
                                    P.maxabs.as<unsigned long long>(), P.f32 && P.need32 ? 1 : 0, &nl, st),
                      "gather")))
        return rc;
    if (!P.segs2.empty()) {
        // Second stage: sort each grid strip by the base order's column 1
        // (keys read from col64) and lay out the grid orders; strips that
        // fit one block do both in one kernel.
        if (P.maxlen2 <= ENM_SORT_TILE) {
            if ((rc = chk(enm_launch_strip_sort(tab<GatherItem>(P, P.off_gitems2), (int)P.gitems2.size(),
                                                tab<SortSeg>(P, P.off_segs2), tab<OrderDev>(P, P.off_orders),
                                                P.tile2, P.col64.as<double>(), P.col32.as<float>(),
                                                P.perm.as<int>(), P.f32 && P.need32 ? 1 : 0, &nl, st),
                          "sort")))
                return rc;
        } else if ((rc = chk(enm_launch_sort(tab<SortSeg>(P, P.off_segs2), (int)P.segs2.size(), P.ntiles2, P.maxlen2,
                                      P.tile2, P.col64.as<double>(), nullptr, P.sk0.as<double>(), P.si0.as<int>(),
                                      P.sk1.as<double>(), P.si1.as<int>(), P.split.as<int>(), &fb, &nl, st),
                      "sort"))) {
            return rc;
        } else {
            const double* skey2 = fb ? P.sk1.as<double>() : P.sk0.as<double>();
            const int* sidx2 = fb ? P.si1.as<int>() : P.si0.as<int>();
            if ((rc = chk(enm_launch_gather(tab<GatherItem>(P, P.off_gitems2), (int)P.gitems2.size(),
                                            tab<SortSeg>(P, P.off_segs2), tab<OrderDev>(P, P.off_orders), arena,
                                            skey2, sidx2, P.col64.as<double>(), P.col32.as<float>(),
                                            P.perm.as<int>(), P.maxabs.as<unsigned long long>(),
                                            P.f32 && P.need32 ? 1 : 0, &nl, st),
                          "gather")))
                return rc;
        }
    }
    if ((rc = mark(3))) return rc;
    const OrderDev* orders = tab<OrderDev>(P, P.off_orders);
    const unsigned long long* maxabs = P.maxabs.as<unsigned long long>();
    unsigned* counters = P.counters.as<unsigned>();
    // Per-instance mode interleaves each instance's kNN, counts and combine.
    const int* qoffs = tab<int>(P, P.off_qoffs);
    const double* psi_tab = P.res64.as<double>() + P.psi_tab;
    auto run_group = [&](const Group& g) -> int {
        if (!g.is_count) {
            int grid = g.pclass == ENM_LIST_HEAP_GLOBAL ? enm_knn_heap_grid(grid_for(ctx, g.waves, g.n_items), g.k1max)
                                                        : grid_for(ctx, g.waves, g.n_items);
            if (g.walk == 2) {
                // Strip workgroups: as many per CU as the waves per CU and
                // the LDS (three slots of the largest strip) allow.
                const int nt = std::min(512, ((g.smax + 63) / 64) * 64);
                const int by_waves = std::max(1, g.waves * 64 / nt);
                const int by_lds = std::max(1, (160 * 1024) / (3 * g.smax * 16));
                grid = std::max(1, std::min(g.n_items, ctx->num_cus * std::min(by_waves, by_lds)));
            }
            e = enm_launch_knn(tab<Item>(P, P.off_kitems) + g.first, qoffs + g.qoff, g.n_items,
                               counters + g.counter,
                               tab<KnnDev>(P, P.off_kdescs), orders, maxabs, P.perm.as<int>(),
                               P.col64.as<double>(), P.col32.as<float>(), P.res64.as<double>(), stats,
                               g.dclass, g.pclass, P.f32 && P.brute, P.brute, g.walk, grid,
                               g.k1max, P.heap.as<double>(), g.ticket, g.smax,
                               g.walk == 2 && P.far_bytes > 0 ? P.far.as<char>() : nullptr,
                               counters + g.counter + ENM_NQ * ENM_CSTRIDE, P.far_bytes, st);
            nl += g.walk == 2 && P.far_bytes > 0 ? 2 : 1;
            return chk(e, "knn");
        }
        e = enm_launch_count(tab<Item>(P, P.off_citems) + g.first, qoffs + g.qoff, g.n_items,
                             counters + g.counter, tab<CountDev>(P, P.off_cdescs), orders, maxabs,
                             P.perm.as<int>(), psi_tab, P.col64.as<double>(), P.col32.as<float>(),
                             P.res64.as<double>(), P.res32.as<int>(), stats, g.dclass, P.f32, P.brute,
                             grid_for(ctx, g.waves, g.n_items * (g.dclass == 1 ? ENM_C1_THREADS / 64 : 1)),
                             P.deferred.as<unsigned>(), counters + g.counter + ENM_NQ * ENM_CSTRIDE,
                             P.deferred_seg, g.ticket, st);
        nl += g.dclass == 1 ? 2 : 1;
        return chk(e, "count");
    };
    auto psi_tiles = [&](int first, int count) -> int {
        if (count <= 0) return ENM_OK;
        ++nl;
        return chk(enm_launch_psi_tiles(tab<Item>(P, P.off_sitems) + first, count, tab<CountDev>(P, P.off_cdescs),
                                        psi_tab, P.res32.as<int>(), P.res64.as<double>(), P.res32.as<int>(), st),
                   "psi");
    };
    if (P.per_instance) {
        for (int ii = 0; ii < P.n_inst; ++ii) {
            for (const Group& g : P.groups)
                if (g.inst == ii && !g.is_count && (rc = run_group(g))) return rc;
            for (const Group& g : P.groups)
                if (g.inst == ii && g.is_count && (rc = run_group(g))) return rc;
            if ((rc = psi_tiles(P.inst_sitems[ii].first, P.inst_sitems[ii].second))) return rc;
            const auto ri = P.inst_ritems[ii];
            if ((rc = chk(enm_launch_reduce_pt(tab<RedItem>(P, P.off_ritems) + ri.first, ri.second,
                                               tab<RedDev>(P, P.off_rdescs), P.res64.as<double>(), st),
                          "reduce")))
                return rc;
            if (ri.second > 0) ++nl;
            if ((rc = chk(enm_launch_combine(tab<InstDev>(P, P.off_insts), tab<double>(P, P.off_cls), ii, 1,
                                             P.res64.as<double>(), P.res32.as<int>(), out, st), "combine")))
                return rc;
            ++nl;
        }
        if ((rc = mark(4))) return rc;
        if ((rc = mark(5))) return rc;
    } else {
        for (const Group& g : P.groups)
            if (!g.is_count && (rc = run_group(g))) return rc;
        if ((rc = mark(4))) return rc;
        for (const Group& g : P.groups)
            if (g.is_count && (rc = run_group(g))) return rc;
        if ((rc = psi_tiles(0, (int)P.sitems.size()))) return rc;
        if ((rc = chk(enm_launch_reduce_pt(tab<RedItem>(P, P.off_ritems), (int)P.ritems.size(),
                                           tab<RedDev>(P, P.off_rdescs), P.res64.as<double>(), st),
                      "reduce")))
            return rc;
        if (!P.ritems.empty()) ++nl;
        if ((rc = mark(5))) return rc;
        if ((rc = chk(enm_launch_combine(tab<InstDev>(P, P.off_insts), tab<double>(P, P.off_cls), 0, P.n_inst,
                                         P.res64.as<double>(), P.res32.as<int>(), out, st), "combine")))
            return rc;
        ++nl;
    }
    if (d_eps || d_counts) {
        for (const ExportInst& ex : P.exports) {
            if (d_eps) {
                if ((rc = chk(enm_launch_export_eps(P.res64.as<double>() + ex.eps, P.perm.as<int>() + ex.permA,
                                                    ex.n, d_eps + ex.point_off, st), "export")))
                    return rc;
                ++nl;
            }
            if (d_counts) {
                int* base = d_counts + 3 * ex.point_off;
                if ((rc = chk(enm_launch_fill_i32(base, 3 * ex.n, 0, st), "export"))) return rc;
                ++nl;
                for (int s = 0; s < ex.nspace; ++s) {
                    if ((rc = chk(enm_launch_export_cnt(P.res32.as<int>() + ex.sp[s].counts,
                                                        P.perm.as<int>() + ex.sp[s].perm, ex.n,
                                                        base + s * ex.n, st), "export")))
                        return rc;
                    ++nl;
                }
            }
        }
    }
    if ((rc = mark(6))) return rc;
    P.n_launch = nl;
    return ENM_OK;
}

// Plans are cached by the shape of the call: the instances (column offsets,
// not the arena's address), class sizes, flags, which exports are asked for
// and the tuning.  Device pointers are kernel arguments of every enqueue, so
// a plan serves any arena and outputs; a captured graph holds the pointers
// it was captured with and is captured again when they change.
std::string plan_key(const enm_instance* inst, int n_inst, const int64_t* seg_len, bool eps, bool cnt,
                     bool hinted, uint32_t flags, uint64_t gen) {
    std::string k;
    k.append((const char*)inst, sizeof(enm_instance) * (size_t)n_inst);
    for (int i = 0; i < n_inst; ++i)
        if (inst[i].n_segs > 1)
            k.append((const char*)(seg_len + inst[i].seg_index), sizeof(int64_t) * (size_t)inst[i].n_segs);
    const uint64_t misc[5] = {flags, gen, eps ? 1u : 0u, cnt ? 1u : 0u, hinted ? 1u : 0u};
    k.append((const char*)misc, sizeof misc);
    return k;
}

}  // namespace

extern "C" {

int enm_init(int device, enm_ctx** out_ctx) {
    if (!out_ctx) { g_init_error = "out_ctx is NULL"; return ENM_E_ARG; }
    *out_ctx = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        g_init_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return ENM_E_HIP;
    }
    if (device < 0 || device >= ndev) { g_init_error = "device index out of range"; return ENM_E_ARG; }
    e = hipSetDevice(device);
    if (e != hipSuccess) { g_init_error = std::string("hipSetDevice: ") + hipGetErrorString(e); return ENM_E_HIP; }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        g_init_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return ENM_E_HIP;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_init_error = std::string("libennemi_hip is built for gfx950, device is ") + prop.gcnArchName;
        return ENM_E_UNSUPPORTED;
    }
    enm_ctx* c = new enm_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    {
        // The static workspace's address on this device, looked up once per
        // device with the device current.
        std::lock_guard<std::mutex> guard(g_scratch_mu);
        auto it = g_scratch.find(device);
        if (it != g_scratch.end()) {
            c->scratch = it->second;
        } else {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e == hipSuccess && p) {
                g_scratch[device] = p;
                c->scratch = p;
            } else {
                c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
                (void)hipGetLastError();
            }
        }
    }
    *out_ctx = c;
    return ENM_OK;
}

void enm_shutdown(enm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->last_stream) (void)hipStreamSynchronize(ctx->last_stream);
    (void)hipDeviceSynchronize();
    ctx->cache.clear();
    delete ctx;
}

const char* enm_last_error(const enm_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_init_error.c_str();
}

int enm_run_batch(enm_ctx* ctx, const enm_instance* inst, int32_t n_inst, const int64_t* seg_len,
                  const double* d_arena, double* d_out, double* d_eps, int32_t* d_counts, uint32_t flags,
                  void* stream) {
    return enm_run_batch_hinted(ctx, inst, n_inst, seg_len, d_arena, nullptr, d_out, d_eps, d_counts, flags,
                                stream);
}

int enm_run_batch_hinted(enm_ctx* ctx, const enm_instance* inst, int32_t n_inst, const int64_t* seg_len,
                         const double* d_arena, const int32_t* d_hint, double* d_out, double* d_eps,
                         int32_t* d_counts, uint32_t flags, void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    if (!inst || n_inst <= 0) return fail(ctx, ENM_E_ARG, "no instances");
    if (!d_arena || !d_out) return fail(ctx, ENM_E_ARG, "d_arena and d_out must be device pointers");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    if (ctx->last_stream != st && ctx->last_stream != nullptr) {
        // Plans reuse their workspace; finish work queued on another stream first.
        e = hipStreamSynchronize(ctx->last_stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    }
    ctx->last_stream = st;
    const bool use_graph = (flags & ENM_F_GRAPH) != 0;
    const bool timing = (flags & ENM_F_TIMING) != 0;
    const std::string key = plan_key(inst, n_inst, seg_len, d_eps != nullptr, d_counts != nullptr,
                                     d_hint != nullptr, flags & ~(uint32_t)ENM_F_TIMING, ctx->tuning_gen);
    const void* ptrs[5] = {d_arena, d_out, d_eps, d_counts, d_hint};
    // Capture the plan's launch sequence with these pointers and launch it.
    auto capture_and_launch = [&](Plan& P) -> int {
        if (P.graph) { (void)hipGraphExecDestroy(P.graph); P.graph = nullptr; }
        hipGraph_t graph = nullptr;
        hipError_t er = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (er != hipSuccess) return hip_fail(ctx, er, "hipStreamBeginCapture");
        int r = enqueue(ctx, P, d_arena, d_hint, d_out, d_eps, d_counts, st, false);
        hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e2 != hipSuccess) return hip_fail(ctx, e2, "hipStreamEndCapture");
        er = hipGraphInstantiate(&P.graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (er != hipSuccess) { P.graph = nullptr; return hip_fail(ctx, er, "hipGraphInstantiate"); }
        std::memcpy(P.graph_ptrs, ptrs, sizeof ptrs);
        er = hipGraphLaunch(P.graph, st);
        if (er != hipSuccess) return hip_fail(ctx, er, "hipGraphLaunch");
        return ENM_OK;
    };
    // Reuse a cached plan of the same shape.
    for (auto it = ctx->cache.begin(); it != ctx->cache.end(); ++it) {
        if ((*it)->key == key) {
            Plan* P = it->get();
            if (it != ctx->cache.begin()) ctx->cache.splice(ctx->cache.begin(), ctx->cache, it);
            ctx->last = P;
            P->timing = false;
            if (use_graph) {
                if (P->graph && std::memcmp(P->graph_ptrs, ptrs, sizeof ptrs) == 0) {
                    e = hipGraphLaunch(P->graph, st);
                    if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
                    P->n_launch = 0;
                    return ENM_OK;
                }
                return capture_and_launch(*P);
            }
            return enqueue(ctx, *P, d_arena, d_hint, d_out, d_eps, d_counts, st, timing);
        }
    }
    std::unique_ptr<Plan> up(new Plan());
    Plan& P = *up;
    P.key = key;
    P.f32 = (flags & ENM_F_FP64) == 0;
    P.brute = (flags & ENM_F_BRUTE) != 0;
    P.need32 = P.brute;
    P.per_instance = (flags & ENM_F_PER_INSTANCE) != 0;
    P.want_eps = d_eps != nullptr;
    P.hinted = d_hint != nullptr;
    int rc = plan_batch(ctx, P, inst, n_inst, seg_len);
    if (rc) return rc;
    // Evict before allocating so a cache full of large plans frees memory first.
    auto evict = [&](size_t keep) {
        while (ctx->cache.size() > keep) {
            if (ctx->cache.back().get() == ctx->last) ctx->last = nullptr;
            (void)hipStreamSynchronize(st);
            ctx->cache.pop_back();
        }
    };
    evict(kMaxCachedPlans - 1);
    rc = allocate_and_upload(ctx, P, st);
    if (rc == ENM_E_NOMEM && !ctx->cache.empty()) {
        // Out of device memory with other plans resident: drop them all and
        // try once more.
        evict(0);
        rc = allocate_and_upload(ctx, P, st);
    }
    if (rc) return rc;
    if (use_graph) {
        rc = capture_and_launch(P);
        if (rc) return rc;
    } else {
        rc = enqueue(ctx, P, d_arena, d_hint, d_out, d_eps, d_counts, st, timing);
        if (rc) return rc;
    }
    ctx->last = up.get();
    ctx->cache.push_front(std::move(up));
    return ENM_OK;
}

int enm_run_batch_static(enm_ctx* ctx, const enm_instance* inst, int32_t n_inst, const int64_t* seg_len,
                         const double* d_arena, double* d_out, double* d_eps, int32_t* d_counts, uint32_t flags,
                         void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    if (!inst || n_inst <= 0) return fail(ctx, ENM_E_ARG, "no instances");
    if (!d_arena || !d_out) return fail(ctx, ENM_E_ARG, "d_arena and d_out must be device pointers");
    if (flags & (ENM_F_GRAPH | ENM_F_TIMING))
        return fail(ctx, ENM_E_ARG, "enm_run_batch_static takes neither ENM_F_GRAPH nor ENM_F_TIMING");
    if (!ctx->scratch) return fail(ctx, ENM_E_HIP, "static workspace unavailable: " + ctx->scratch_err);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    if (ctx->last_stream != st && ctx->last_stream != nullptr) {
        e = hipStreamSynchronize(ctx->last_stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    }
    ctx->last_stream = st;
    std::unique_ptr<Plan> up(new Plan());
    Plan& P = *up;
    P.f32 = (flags & ENM_F_FP64) == 0;
    P.brute = (flags & ENM_F_BRUTE) != 0;
    P.need32 = P.brute;
    P.per_instance = (flags & ENM_F_PER_INSTANCE) != 0;
    P.want_eps = d_eps != nullptr;
    int rc = plan_batch(ctx, P, inst, n_inst, seg_len);
    if (rc) return rc;
    const std::vector<char> blob = pack_tables(P);
    const std::vector<BufSpec> layout = workspace_layout(P);
    auto rounded = [](size_t b) { return (std::max<size_t>(b, 256) + 255) & ~(size_t)255; };
    size_t total = 0;
    for (const BufSpec& b : layout) total += rounded(b.bytes);
    if (total > (size_t)ENM_STATIC_WORKSPACE_BYTES) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "the batch needs %zu bytes of workspace; enm_run_batch_static holds at most %llu",
                      total, (unsigned long long)ENM_STATIC_WORKSPACE_BYTES);
        return fail(ctx, ENM_E_NOMEM, msg);
    }
    char* base = static_cast<char*>(ctx->scratch);
    size_t off = 0;
    for (const BufSpec& b : layout) {
        b.buf->borrow(base + off, b.bytes);
        off += rounded(b.bytes);
    }
    // The tables go up from pageable host memory; the copy is complete
    // before the blob is released.
    e = hipMemcpyAsync(P.tables.p, blob.data(), P.table_bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(tables)");
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    rc = enqueue(ctx, P, d_arena, nullptr, d_out, d_eps, d_counts, st, false);
    if (rc) return rc;
    ctx->static_plan = std::move(up);
    ctx->last = ctx->static_plan.get();
    return ENM_OK;
}

int enm_prep_count(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks, const double* d_raw,
                   const uint8_t* d_mask, uint32_t flags, int64_t* h_n, uint32_t* h_nan, void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    auto& S = ctx->prep;
    S.counted = false;
    if (!tasks || n_tasks <= 0) return fail(ctx, ENM_E_ARG, "no tasks");
    if (!d_raw || !h_n || !h_nan) return fail(ctx, ENM_E_ARG, "d_raw, h_n and h_nan are required");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    S.tasks.assign(n_tasks, PrepTaskDev{});
    S.written = false;
    int64_t stride = 0, pm = 0;
    for (int t = 0; t < n_tasks; ++t) {
        const enm_prep_task& T = tasks[t];
        if (T.w < 1 || T.w > 0x7fffffffLL || T.dz < 0 || T.dz > ENM_PREP_MAXZ || T.x0 < 0 || T.y0 < 0) {
            char msg[96];
            std::snprintf(msg, sizeof msg, "task %d: bad window or condition columns", t);
            return fail(ctx, ENM_E_ARG, msg);
        }
        PrepTaskDev& D = S.tasks[t];
        std::memset(&D, 0, sizeof D);
        D.x_col = T.x_col; D.y_col = T.y_col; D.x0 = T.x0; D.y0 = T.y0; D.w = T.w; D.dz = T.dz;
        for (int c = 0; c < T.dz; ++c) { D.z_col[c] = T.z_col[c]; D.z0[c] = T.z0[c]; }
        D.nblk = (int32_t)((T.w + ENM_PREP_BLOCK - 1) / ENM_PREP_BLOCK);
        D.pmap = pm;
        pm += (T.w + 31) / 32;
        stride = std::max<int64_t>(stride, D.nblk);
    }
    const int64_t nb = stride * n_tasks;
    S.stride = stride;
    S.pm_total = pm;
    S.blob1.clear();
    const size_t off_t = append(S.blob1, S.tasks);
    if (S.tables1.ensure(S.blob1.size()) != hipSuccess || S.blk_count.ensure(nb * 4 + 4) != hipSuccess ||
        S.blk_off.ensure(nb * 8 + 8) != hipSuccess || S.nan.ensure((size_t)n_tasks * 4) != hipSuccess ||
        S.nout.ensure((size_t)n_tasks * 8) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for the preparation workspace");
    }
    e = hipMemcpyAsync(S.tables1.p, S.blob1.data(), S.blob1.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(prep tables)");
    const PrepTaskDev* dt = reinterpret_cast<const PrepTaskDev*>(S.tables1.as<char>() + off_t);
    e = enm_launch_prep_count(dt, stride, n_tasks, d_raw, d_mask, (flags & ENM_P_DROP_NAN) ? 1 : 0,
                              S.blk_count.as<int>(), S.blk_off.as<int64_t>(), S.nan.as<unsigned>(),
                              S.nout.as<int64_t>(), st);
    if (e != hipSuccess) return hip_fail(ctx, e, "prep count");
    e = hipMemcpyAsync(h_n, S.nout.p, (size_t)n_tasks * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(h_nan, S.nan.p, (size_t)n_tasks * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "prep count read-back");
    S.n.assign(h_n, h_n + n_tasks);
    S.raw = d_raw;
    S.mask = d_mask;
    S.drop_nan = flags & ENM_P_DROP_NAN;
    S.counted = true;
    return ENM_OK;
}

int enm_prep_write(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks, const double* d_raw,
                   const uint8_t* d_mask, const double* d_noise, int64_t noise_len, uint32_t flags,
                   const int64_t* arena_off, double* d_arena, uint32_t* d_const, void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    auto& S = ctx->prep;
    if (!S.counted || (int)S.tasks.size() != n_tasks || S.raw != d_raw || S.mask != d_mask ||
        S.drop_nan != (flags & ENM_P_DROP_NAN))
        return fail(ctx, ENM_E_ARG, "enm_prep_write must follow enm_prep_count of the same tasks and inputs");
    if (!tasks || !arena_off || !d_arena || !d_const) return fail(ctx, ENM_E_ARG, "missing argument");
    const bool rescale = (flags & ENM_P_RESCALE) != 0;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    std::vector<PrepSumDev> sums;
    std::vector<int> chunk_sum_of, chunk_index;
    int64_t sstride = 0, need_noise = 0;
    int stat = 0;
    for (int t = 0; t < n_tasks; ++t) {
        const enm_prep_task& T = tasks[t];
        PrepTaskDev& D = S.tasks[t];
        if (T.x_col != D.x_col || T.y_col != D.y_col || T.x0 != D.x0 || T.y0 != D.y0 || T.w != D.w || T.dz != D.dz)
            return fail(ctx, ENM_E_ARG, "enm_prep_write must follow enm_prep_count of the same tasks");
        D.arena_off = arena_off[t];
        D.n = S.n[t];
        D.stat = stat;
        D.flags = 0;
        const int64_t n = D.n;
        need_noise = std::max<int64_t>(need_noise, n * (2 + D.dz));
        sstride = std::max<int64_t>(sstride, (n + 255) / 256);
        // x, y and a one-column condition: pairwise sums; the columns of a
        // multi-column condition: row-by-row sums.
        const int narr = 2 + D.dz;
        for (int a = 0; a < narr; ++a) {
            PrepSumDev sd;
            std::memset(&sd, 0, sizeof sd);
            sd.src = D.arena_off + a * n;
            sd.n = n;
            sd.stat = stat + 2 * a;
            sd.sequential = (a >= 2 && D.dz >= 2) ? 1 : 0;
            sd.chunk0 = (int32_t)chunk_sum_of.size();
            sd.n_chunks = 0;
            if (!sd.sequential && n > 0) {
                sd.n_chunks = (int32_t)((n + 8191) / 8192);
                for (int c = 0; c < sd.n_chunks; ++c) {
                    chunk_sum_of.push_back((int)sums.size());
                    chunk_index.push_back(c);
                }
            }
            sums.push_back(sd);
        }
        stat += 2 * narr;
    }
    if (rescale && (!d_noise || noise_len < need_noise)) {
        char msg[128];
        std::snprintf(msg, sizeof msg, "the noise holds %lld values, the tasks draw up to %lld",
                      (long long)noise_len, (long long)need_noise);
        return fail(ctx, ENM_E_ARG, msg);
    }
    // Chunks listed full ones first (8192 values: one wave each, four to a
    // block), then the last chunk of each array.
    std::vector<int> chunk_list;
    chunk_list.reserve(chunk_sum_of.size());
    for (int pass = 0; pass < 2; ++pass)
        for (size_t c = 0; c < chunk_sum_of.size(); ++c) {
            const bool full = ((int64_t)chunk_index[c] + 1) * 8192 <= sums[chunk_sum_of[c]].n;
            if (full == (pass == 0)) chunk_list.push_back((int)c);
        }
    int n_full = 0;
    for (size_t c = 0; c < chunk_sum_of.size(); ++c)
        n_full += ((int64_t)chunk_index[c] + 1) * 8192 <= sums[chunk_sum_of[c]].n ? 1 : 0;
    S.blob2.clear();
    const size_t o_t = append(S.blob2, S.tasks);
    const size_t o_s = append(S.blob2, sums);
    const size_t o_cs = append(S.blob2, chunk_sum_of);
    const size_t o_ci = append(S.blob2, chunk_index);
    const size_t o_cl = append(S.blob2, chunk_list);
    const bool hints = (flags & ENM_P_HINTS) != 0;
    if (S.tables2.ensure(S.blob2.size()) != hipSuccess ||
        S.chunk_val.ensure(chunk_sum_of.size() * 8 + 8) != hipSuccess ||
        S.stats.ensure((size_t)stat * 8 + 8) != hipSuccess ||
        (hints && S.kept.ensure((size_t)S.pm_total * sizeof(KeptWord) + 8) != hipSuccess)) {
        (void)hipGetLastError();
        return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for the preparation workspace");
    }
    e = hipMemcpyAsync(S.tables2.p, S.blob2.data(), S.blob2.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(prep tables)");
    char* base = S.tables2.as<char>();
    e = enm_launch_prep_write(reinterpret_cast<PrepTaskDev*>(base + o_t), S.stride, n_tasks, d_raw, d_mask,
                              S.drop_nan ? 1 : 0, rescale ? 1 : 0, S.blk_off.as<int64_t>(),
                              reinterpret_cast<const PrepSumDev*>(base + o_s), (int)sums.size(),
                              reinterpret_cast<const int*>(base + o_cs), reinterpret_cast<const int*>(base + o_ci),
                              reinterpret_cast<const int*>(base + o_cl), (int)chunk_sum_of.size(), n_full,
                              S.chunk_val.as<double>(), S.stats.as<double>(), sstride,
                              d_noise, d_const, d_arena, hints ? S.kept.as<KeptWord>() : nullptr, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "prep write");
    if (!rescale) {
        e = enm_launch_fill_i32(reinterpret_cast<int*>(d_const), n_tasks, 0, st);
        if (e != hipSuccess) return hip_fail(ctx, e, "prep write");
    }
    S.counted = false;
    S.written = hints;
    S.o_tasks = o_t;
    S.arena_off.assign(arena_off, arena_off + n_tasks);
    S.arena = d_arena;
    S.noise = rescale ? d_noise : nullptr;
    S.noise_len = rescale ? noise_len : 0;
    S.flags = flags;
    return ENM_OK;
}

int enm_prep_hints(enm_ctx* ctx, const enm_prep_task* tasks, int32_t n_tasks, const double* d_raw,
                   const uint8_t* d_mask, const double* d_noise, int64_t noise_len, uint32_t flags,
                   const int64_t* arena_off, const double* d_arena, int32_t* d_hint, void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    auto& S = ctx->prep;
    if (!tasks || !arena_off || !d_arena || !d_hint) return fail(ctx, ENM_E_ARG, "missing argument");
    bool same = S.written && (int)S.tasks.size() == n_tasks && S.raw == d_raw && S.mask == d_mask &&
                S.arena == d_arena && S.flags == flags &&
                S.noise == ((flags & ENM_P_RESCALE) ? (const void*)d_noise : nullptr) &&
                S.noise_len == ((flags & ENM_P_RESCALE) ? noise_len : 0);
    for (int t = 0; same && t < n_tasks; ++t) {
        const enm_prep_task& T = tasks[t];
        const PrepTaskDev& D = S.tasks[t];
        same = T.x_col == D.x_col && T.y_col == D.y_col && T.x0 == D.x0 && T.y0 == D.y0 && T.w == D.w &&
               T.dz == D.dz && arena_off[t] == S.arena_off[t];
        for (int c = 0; same && c < T.dz; ++c) same = T.z_col[c] == D.z_col[c] && T.z0[c] == D.z0[c];
    }
    if (!same)
        return fail(ctx, ENM_E_ARG,
                    "enm_prep_hints must follow an enm_prep_write with ENM_P_HINTS of the same tasks and inputs");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    // The series: every raw column a window reads, over the union of its windows.
    std::map<int64_t, std::pair<int64_t, int64_t>> range;
    auto take = [&](int64_t col, int64_t first, int64_t w) {
        auto it = range.find(col);
        if (it == range.end()) range[col] = {first, first + w};
        else it->second = {std::min(it->second.first, first), std::max(it->second.second, first + w)};
    };
    int max_cols = 0;
    for (int t = 0; t < n_tasks; ++t) {
        const PrepTaskDev& D = S.tasks[t];
        take(D.x_col, D.x0, D.w);
        take(D.y_col, D.y0, D.w);
        for (int c = 0; c < D.dz; ++c) take(D.z_col[c], D.z0[c], D.w);
        max_cols = std::max(max_cols, 2 + D.dz);
    }
    std::vector<HintSeries> series;
    std::vector<SortSeg> segs;
    std::map<int64_t, int> sid;
    int64_t nkeys = 0, maxlen = 0;
    int ntiles = 0;
    for (const auto& kv : range) {
        const int64_t len = kv.second.second - kv.second.first;
        if (len > 0x7fffffffLL) return fail(ctx, ENM_E_UNSUPPORTED, "a series range of 2^31 rows or more");
        sid[kv.first] = (int)series.size();
        series.push_back(HintSeries{kv.first, kv.second.first, len, nkeys});
        SortSeg sg;
        std::memset(&sg, 0, sizeof sg);
        sg.src = nkeys;
        sg.sbase = nkeys;
        sg.len = (int32_t)len;
        sg.first_tile = ntiles;
        sg.p0 = 0;
        sg.hint = 0;
        segs.push_back(sg);
        ntiles += (int)((len + ENM_SORT_TILE - 1) / ENM_SORT_TILE);
        nkeys += len;
        maxlen = std::max(maxlen, len);
    }
    // The columns, by column slot and then by task (consecutive filter
    // blocks, one XCD's share, read one series).
    std::vector<HintCol> cols;
    int64_t max_w = 0, max_n = 0, cont_words = 0, max_len = 0;
    for (int c = 0; c < max_cols; ++c) {
        for (int t = 0; t < n_tasks; ++t) {
            const PrepTaskDev& D = S.tasks[t];
            if (c >= 2 + D.dz) continue;
            const int64_t col = c == 0 ? D.x_col : c == 1 ? D.y_col : D.z_col[c - 2];
            const int64_t first = c == 0 ? D.x0 : c == 1 ? D.y0 : D.z0[c - 2];
            const HintSeries& H = series[sid[col]];
            HintCol C;
            std::memset(&C, 0, sizeof C);
            C.sb = H.kbase;
            C.ws = first - H.lo;
            C.w = D.w;
            C.pm = D.pmap;
            C.n = S.n[t];
            C.out = S.arena_off[t] + c * C.n;
            C.len = (int32_t)H.len;
            C.task = t;
            C.slot = c;
            C.cbits = cont_words;
            cont_words += (C.n + 31) / 32;
            max_w = std::max<int64_t>(max_w, D.w);
            max_n = std::max<int64_t>(max_n, C.n);
            max_len = std::max<int64_t>(max_len, H.len);
            cols.push_back(C);
        }
    }
    S.blob3.clear();
    const size_t o_ser = append(S.blob3, series);
    const size_t o_seg = append(S.blob3, segs);
    const size_t o_col = append(S.blob3, cols);
    const size_t ncols = cols.size();
    const int target_blocks = 8 * ctx->num_cus;   // resident filter workgroups
    const size_t nsg = (size_t)hint_segments((int)ncols, max_len, target_blocks) * ncols + 1;
    if (S.tables3.ensure(S.blob3.size()) != hipSuccess || S.hkeys.ensure((size_t)nkeys * 8 + 8) != hipSuccess ||
        S.hk0.ensure((size_t)nkeys * 8 + 8) != hipSuccess || S.hk1.ensure((size_t)nkeys * 8 + 8) != hipSuccess ||
        S.hi0.ensure((size_t)nkeys * 4 + 4) != hipSuccess || S.hi1.ensure((size_t)nkeys * 4 + 4) != hipSuccess ||
        S.hsplit.ensure((size_t)ntiles * 4 + 4) != hipSuccess || S.hcont.ensure((size_t)cont_words * 4 + 4) != hipSuccess ||
        S.hbad.ensure(ncols * 4 + 4) != hipSuccess || S.hmax.ensure(8) != hipSuccess ||
        S.sg_cnt.ensure(nsg * 4) != hipSuccess || S.sg_last.ensure(nsg * 8) != hipSuccess ||
        S.sg_off.ensure(nsg * 8) != hipSuccess || S.sg_cin.ensure(nsg * 8) != hipSuccess ||
        S.sg_chas.ensure(nsg * 4) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for the sort-hint workspace");
    }
    e = hipMemcpyAsync(S.tables3.p, S.blob3.data(), S.blob3.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(hint tables)");
    const char* base = S.tables3.as<char>();
    const HintSeries* dser = reinterpret_cast<const HintSeries*>(base + o_ser);
    const SortSeg* dseg = reinterpret_cast<const SortSeg*>(base + o_seg);
    const HintCol* dcol = reinterpret_cast<const HintCol*>(base + o_col);
    e = enm_launch_hint_keys(dser, (int)series.size(), maxlen, d_raw, S.hkeys.as<double>(), st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hint keys");
    int fb = 0, nl = 0;
    e = enm_launch_sort(dseg, (int)segs.size(), ntiles, maxlen, ENM_SORT_TILE, S.hkeys.as<double>(), nullptr,
                        S.hk0.as<double>(), S.hi0.as<int>(), S.hk1.as<double>(), S.hi1.as<int>(),
                        S.hsplit.as<int>(), &fb, &nl, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hint series sort");
    const int* sidx = fb ? S.hi1.as<int>() : S.hi0.as<int>();
    const double* skey = fb ? S.hk1.as<double>() : S.hk0.as<double>();
    const PrepTaskDev* dtasks = reinterpret_cast<const PrepTaskDev*>(S.tables2.as<char>() + S.o_tasks);
    const HintSegs hsg{S.sg_cnt.as<int>(), S.sg_last.as<double>(), S.sg_off.as<int64_t>(), S.sg_cin.as<double>(),
                       S.sg_chas.as<int>()};
    e = enm_launch_hint_order(dcol, (int)ncols, max_w, max_n, max_len, target_blocks, sidx, skey,
                              S.kept.as<KeptWord>(), dtasks,
                              S.stats.as<double>(), (flags & ENM_P_RESCALE) ? 1 : 0,
                              (flags & ENM_P_RESCALE) ? d_noise : nullptr, noise_len, S.hmax.as<double>(),
                              S.hcont.as<unsigned>(), cont_words, hsg, d_arena, d_hint, S.hbad.as<int>(), st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hint order");
    return ENM_OK;
}

int enm_unique_rows(enm_ctx* ctx, const enm_rowset* sets, int32_t n_sets, const double* d_values,
                    int32_t* d_codes, int64_t* d_counts, int64_t* h_distinct, void* stream) {
    if (!ctx) { g_init_error = "ctx is NULL"; return ENM_E_ARG; }
    ctx->err.clear();
    if (!sets || n_sets <= 0 || !d_values || !d_codes || !d_counts || !h_distinct)
        return fail(ctx, ENM_E_ARG, "missing argument");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    auto& U = ctx->uq;
    // Distinct columns, each sorted once.
    std::map<std::pair<int64_t, int64_t>, int> colid;
    std::vector<SortSeg> segs1;
    std::vector<UqSeg> useg1;
    std::vector<int64_t> set_off(n_sets);
    int64_t tot1 = 0, tot2 = 0, maxlen1 = 0, maxlen2 = 0, out_tot = 0;
    int tiles1 = 0;
    for (int s = 0; s < n_sets; ++s) {
        const enm_rowset& R = sets[s];
        if (R.n < 1 || R.n > 0x7fffffffLL || R.ncol < 1 || R.ncol > ENM_MAX_COLS)
            return fail(ctx, ENM_E_ARG, "bad row set");
        set_off[s] = out_tot;
        out_tot += R.n;
        if (R.ncol > 1) { tot2 += R.n; maxlen2 = std::max(maxlen2, R.n); }
        for (int c = 0; c < R.ncol; ++c) {
            auto key = std::make_pair((int64_t)R.col[c], (int64_t)R.n);
            if (colid.count(key)) continue;
            const int id = (int)segs1.size();
            colid[key] = id;
            SortSeg sg;
            sg.src = R.col[c];
            sg.sbase = tot1;
            sg.len = (int32_t)R.n;
            sg.first_tile = tiles1;
            sg.p0 = 0;
            sg.hint = 0;
            tiles1 += (int)((R.n + ENM_SORT_TILE - 1) / ENM_SORT_TILE);
            UqSeg u;
            std::memset(&u, 0, sizeof u);
            u.sbase = tot1;
            u.n = R.n;
            u.code_off = tot1;
            u.start_off = tot1 + id;   // n + 1 slots each
            u.count_off = tot1;
            u.slot = id;
            segs1.push_back(sg);
            useg1.push_back(u);
            tot1 += R.n;
            maxlen1 = std::max(maxlen1, R.n);
        }
    }
    const int ncols = (int)segs1.size();
    const int64_t sort_n = std::max(tot1, tot2);
    const int tiles_max = std::max<int>(tiles1, (int)((tot2 + ENM_SORT_TILE - 1) / ENM_SORT_TILE) + n_sets);
    auto ensure_all = [&]() -> bool {
        return U.k0.ensure(sort_n * 8) == hipSuccess && U.k1.ensure(sort_n * 8) == hipSuccess &&
               U.i0.ensure(sort_n * 4) == hipSuccess && U.i1.ensure(sort_n * 4) == hipSuccess &&
               U.split.ensure((size_t)tiles_max * 4 + 4) == hipSuccess &&
               U.ccodes.ensure(tot1 * 4) == hipSuccess && U.cstarts.ensure((tot1 + ncols) * 8) == hipSuccess &&
               U.ccounts.ensure(tot1 * 8) == hipSuccess && U.cdist.ensure((size_t)ncols * 8) == hipSuccess &&
               U.rkeys.ensure(tot2 * 8 + 8) == hipSuccess && U.rstarts.ensure((tot2 + n_sets) * 8 + 8) == hipSuccess &&
               U.rdist.ensure((size_t)n_sets * 8) == hipSuccess;
    };
    if (!ensure_all()) {
        (void)hipGetLastError();
        return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for the distinct-row workspace");
    }
    U.blob1.clear();
    const size_t o_s1 = append(U.blob1, segs1), o_u1 = append(U.blob1, useg1);
    if (U.tables.ensure(U.blob1.size()) != hipSuccess) return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for tables");
    e = hipMemcpyAsync(U.tables.p, U.blob1.data(), U.blob1.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(tables)");
    int fb = 0, nl = 0;
    e = enm_launch_sort(reinterpret_cast<const SortSeg*>(U.tables.as<char>() + o_s1), ncols, tiles1, maxlen1,
                        ENM_SORT_TILE, d_values, nullptr, U.k0.as<double>(), U.i0.as<int>(), U.k1.as<double>(),
                        U.i1.as<int>(), U.split.as<int>(), &fb, &nl, st);
    if (e == hipSuccess)
        e = enm_launch_uq_codes(reinterpret_cast<const UqSeg*>(U.tables.as<char>() + o_u1), ncols,
                                fb ? U.k1.as<double>() : U.k0.as<double>(), fb ? U.i1.as<int>() : U.i0.as<int>(),
                                U.ccodes.as<int>(), U.cstarts.as<int64_t>(), U.ccounts.as<int64_t>(),
                                U.cdist.as<int64_t>(), st);
    std::vector<int64_t> cdist(ncols);
    if (e == hipSuccess) e = hipMemcpyAsync(cdist.data(), U.cdist.p, (size_t)ncols * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "distinct values of the columns");
    // Multi-column row sets: mixed-radix keys, sorted and numbered.
    std::vector<UqRowKey> rks;
    std::vector<SortSeg> segs2;
    std::vector<UqSeg> useg2;
    std::vector<int> multi_set;
    int64_t koff = 0;
    int tiles2 = 0;
    for (int s = 0; s < n_sets; ++s) {
        const enm_rowset& R = sets[s];
        if (R.ncol == 1) continue;
        UqRowKey rk;
        std::memset(&rk, 0, sizeof rk);
        rk.n = R.n;
        rk.key_off = koff;
        rk.ncol = R.ncol;
        double prod = 1.0;
        for (int c = 0; c < R.ncol; ++c) {
            const int id = colid[std::make_pair((int64_t)R.col[c], (int64_t)R.n)];
            rk.code_off[c] = useg1[id].code_off;
            rk.radix[c] = cdist[id];
            prod *= (double)cdist[id];
        }
        if (prod > 9007199254740992.0)
            return fail(ctx, ENM_E_UNSUPPORTED, "too many distinct value combinations for float64 row keys");
        SortSeg sg;
        sg.src = koff;
        sg.sbase = koff;
        sg.len = (int32_t)R.n;
        sg.first_tile = tiles2;
        sg.p0 = 0;
        sg.hint = 0;
        tiles2 += (int)((R.n + ENM_SORT_TILE - 1) / ENM_SORT_TILE);
        UqSeg u;
        std::memset(&u, 0, sizeof u);
        u.sbase = koff;
        u.n = R.n;
        u.code_off = set_off[s];
        u.start_off = koff + (int64_t)useg2.size();
        u.count_off = set_off[s];
        u.slot = (int32_t)useg2.size();
        rks.push_back(rk);
        segs2.push_back(sg);
        useg2.push_back(u);
        multi_set.push_back(s);
        koff += R.n;
    }
    std::vector<int64_t> rdist(useg2.size());
    if (!rks.empty()) {
        U.blob2.clear();
        const size_t o_r = append(U.blob2, rks), o_s2 = append(U.blob2, segs2), o_u2 = append(U.blob2, useg2);
        if (U.tables.ensure(std::max(U.blob1.size(), U.blob2.size())) != hipSuccess)
            return fail(ctx, ENM_E_NOMEM, "hipMalloc failed for tables");
        e = hipMemcpyAsync(U.tables.p, U.blob2.data(), U.blob2.size(), hipMemcpyHostToDevice, st);
        const char* tb = U.tables.as<char>();
        if (e == hipSuccess)
            e = enm_launch_uq_keys(reinterpret_cast<const UqRowKey*>(tb + o_r), (int)rks.size(), maxlen2,
                                   U.ccodes.as<int>(), U.rkeys.as<double>(), st);
        fb = 0;
        if (e == hipSuccess)
            e = enm_launch_sort(reinterpret_cast<const SortSeg*>(tb + o_s2), (int)segs2.size(), tiles2, maxlen2,
                                ENM_SORT_TILE, U.rkeys.as<double>(), nullptr, U.k0.as<double>(), U.i0.as<int>(),
                                U.k1.as<double>(), U.i1.as<int>(), U.split.as<int>(), &fb, &nl, st);
        if (e == hipSuccess)
            e = enm_launch_uq_codes(reinterpret_cast<const UqSeg*>(tb + o_u2), (int)useg2.size(),
                                    fb ? U.k1.as<double>() : U.k0.as<double>(), fb ? U.i1.as<int>() : U.i0.as<int>(),
                                    d_codes, U.rstarts.as<int64_t>(), d_counts, U.rdist.as<int64_t>(), st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rdist.data(), U.rdist.p, rdist.size() * 8, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(ctx, e, "distinct rows");
    }
    // Single-column row sets take their column's codes and counts.
    for (int s = 0; s < n_sets; ++s) {
        const enm_rowset& R = sets[s];
        if (R.ncol != 1) continue;
        const int id = colid[std::make_pair((int64_t)R.col[0], (int64_t)R.n)];
        e = hipMemcpyAsync(d_codes + set_off[s], U.ccodes.as<int>() + useg1[id].code_off, (size_t)R.n * 4,
                           hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_counts + set_off[s], U.ccounts.as<int64_t>() + useg1[id].count_off,
                               (size_t)cdist[id] * 8, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(codes)");
        h_distinct[s] = cdist[id];
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    for (size_t t = 0; t < multi_set.size(); ++t) h_distinct[multi_set[t]] = rdist[t];
    return ENM_OK;
}

int enm_get_stats(enm_ctx* ctx, enm_stats* out) {
    if (!ctx || !out) return ENM_E_ARG;
    std::memset(out, 0, sizeof(*out));
    if (!ctx->last) return fail(ctx, ENM_E_ARG, "no batch has run");
    hipError_t e = hipStreamSynchronize(ctx->last_stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    unsigned long long h[ST_N];
    e = hipMemcpy(h, ctx->last->stats.p, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(stats)");
    out->knn_pairs = h[ST_KNN_PAIRS];
    out->knn_rechecks = h[ST_KNN_RECHECK];
    out->count_pairs = h[ST_CNT_PAIRS];
    out->count_rechecks = h[ST_CNT_RECHECK];
    out->count_queries_1d = h[ST_CNT_1D];
    out->sort_keys = (uint64_t)(ctx->last->sort_elems + ctx->last->sort_elems2);
    out->n_launches = (uint32_t)ctx->last->n_launch;
    out->graph_replayed = ctx->last->graph != nullptr && ctx->last->n_launch == 0;
    if (ctx->last->timing) {
        for (int s = 0; s < ENM_N_STAGES; ++s) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, ctx->last->ev[s], ctx->last->ev[s + 1]);
            if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
            out->stage_ms[s] = ms;
        }
        float tot = 0.f;
        e = hipEventElapsedTime(&tot, ctx->last->ev[0], ctx->last->ev[ENM_N_STAGES]);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
        out->total_ms = tot;
    }
    return ENM_OK;
}

namespace {
int check_tuning(enm_ctx* ctx, const enm_tuning& t) {
    if (t.strip_points < 0 || t.strip_points > 65536 || t.strip_points % ENM_TILE != 0)
        return fail(ctx, ENM_E_ARG, "strip_points must be 0 or a multiple of 64 up to 65536");
    if (t.waves_per_cu < 1 || t.waves_per_cu > 64) return fail(ctx, ENM_E_ARG, "waves_per_cu must be in 1..64");
    if (!(t.list_slots == 0 || t.list_slots == 4 || t.list_slots == 8 || t.list_slots == 16 || t.list_slots == 32))
        return fail(ctx, ENM_E_ARG, "list_slots must be 0, 4, 8, 16 or 32");
    if (t.items_per_ticket < 0 || t.items_per_ticket > 64)
        return fail(ctx, ENM_E_ARG, "items_per_ticket must be in 0..64");
    return ENM_OK;
}
}  // namespace

int enm_set_tuning(enm_ctx* ctx, int32_t strip_points, int32_t waves_per_cu) {
    if (!ctx) return ENM_E_ARG;
    const enm_tuning t{strip_points, waves_per_cu, 0, 0};
    return enm_set_tuning_ex(ctx, &t);
}

int enm_set_tuning_ex(enm_ctx* ctx, const enm_tuning* tuning) {
    if (!ctx || !tuning) return ENM_E_ARG;
    int rc = check_tuning(ctx, *tuning);
    if (rc) return rc;
    ctx->tuning = *tuning;
    ctx->wide_multid = false;
    ctx->tuned.clear();
    ++ctx->tuning_gen;
    return ENM_OK;
}

int enm_get_tuning(const enm_ctx* ctx, int32_t* strip_points, int32_t* waves_per_cu) {
    if (!ctx) return ENM_E_ARG;
    if (strip_points) *strip_points = ctx->tuning.strip_points;
    if (waves_per_cu) *waves_per_cu = ctx->tuning.waves_per_cu;
    return ENM_OK;
}

int enm_get_tuning_for(const enm_ctx* ctx, int32_t dims, int32_t k, int64_t n, enm_tuning* out) {
    if (!ctx || !out || dims < 1 || k < 1 || n < 1) return ENM_E_ARG;
    *out = tuning_for(ctx, dims, k, n);
    return ENM_OK;
}

int enm_autotune(enm_ctx* ctx, const enm_instance* inst, int32_t n_inst, const int64_t* seg_len,
                 const double* d_arena, double* d_out, uint32_t flags, void* stream) {
    if (!ctx) return ENM_E_ARG;
    if (!inst || n_inst <= 0) return fail(ctx, ENM_E_ARG, "no instances");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return fail(ctx, ENM_E_HIP, "hipEventCreate");
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return fail(ctx, ENM_E_HIP, "hipEventCreate"); }
    // The batch's shapes run with the default while it is swept.
    std::set<std::tuple<int, int, int>> shapes;
    for (int i = 0; i < n_inst; ++i)
        shapes.insert(shape_key(inst[i].dx + inst[i].dy + inst[i].dz, inst[i].k, inst[i].n));
    for (const auto& sk : shapes) ctx->tuned.erase(sk);
    const enm_tuning saved = ctx->tuning;
    const bool saved_wide = ctx->wide_multid;
    ctx->wide_multid = false;
    const uint32_t fl = flags & ~(uint32_t)(ENM_F_GRAPH | ENM_F_TIMING);
    float best_ms = 1e30f;
    enm_tuning best = saved;
    int rc = ENM_OK;
    auto trial = [&](const enm_tuning& t) -> int {
        ctx->tuning = t;
        ++ctx->tuning_gen;
        // The first call plans and allocates; the second is timed.
        int r = enm_run_batch(ctx, inst, n_inst, seg_len, d_arena, d_out, nullptr, nullptr, fl, stream);
        if (r) return r;
        (void)hipEventRecord(a, st);
        r = enm_run_batch(ctx, inst, n_inst, seg_len, d_arena, d_out, nullptr, nullptr, fl, stream);
        (void)hipEventRecord(b, st);
        if (r) return r;
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best_ms) { best_ms = ms; best = t; }
        return ENM_OK;
    };
    // Coordinate descent: strip, then waves per CU, then list slots, then
    // items per ticket, each swept with the best of the ones before.
    const int strips[] = {0, 256, 512, 1024, 2048, 4096, 8192};
    const int waves[] = {8, 16, 32};
    const int slots[] = {0, 8, 16};
    const int tickets[] = {0, 1, 2, 4};
    enm_tuning cur = saved;
    cur.list_slots = 0;
    cur.items_per_ticket = 0;
    for (int v : strips) { enm_tuning t = cur; t.strip_points = v; if ((rc = trial(t))) break; }
    cur = best;
    if (!rc) for (int v : waves) { enm_tuning t = cur; t.waves_per_cu = v; if ((rc = trial(t))) break; }
    cur = best;
    if (!rc) for (int v : slots) { enm_tuning t = cur; t.list_slots = v; if ((rc = trial(t))) break; }
    cur = best;
    if (!rc) for (int v : tickets) { enm_tuning t = cur; t.items_per_ticket = v; if ((rc = trial(t))) break; }
    // The winner is kept for each of the batch's shapes; other shapes keep
    // the default.
    ctx->tuning = saved;
    ctx->wide_multid = saved_wide;
    if (!rc)
        for (const auto& sk : shapes) ctx->tuned[sk] = best;
    ++ctx->tuning_gen;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // extern "C"
