// Host side of the C ABI (include/grid2grid_amd.h): layout validation, the
// grid cover, plan objects, launches of the planning and gather kernels,
// batches and hipGraph replay.
#include "grid2grid_amd.h"
#include "g2g_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace g2g;

// ------------------------------------------------------------- errors

namespace {

thread_local std::string t_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_last_error = buf;
    return code;
}

#define G2G_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(G2G_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

#define G2G_TRY(call)            \
    do {                         \
        int rc_ = (call);        \
        if (rc_ != G2G_OK) return rc_; \
    } while (0)

uint32_t dtype_size(int dtype) {
    switch (dtype) {
    case G2G_FLOAT32: return 4;
    case G2G_FLOAT64: return 8;
    case G2G_COMPLEX64: return 8;
    case G2G_COMPLEX128: return 16;
    default: return 0;
    }
}

bool dtype_complex(int dtype) { return dtype == G2G_COMPLEX64 || dtype == G2G_COMPLEX128; }

// Device allocation owned by an object; freed on destruction.
template <typename T>
struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(G2G_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                        hipGetErrorString(e));
        }
        n = count;
        return G2G_OK;
    }
    int upload(const T *h, size_t count, hipStream_t s) {
        G2G_TRY(alloc(count));
        if (count) G2G_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return G2G_OK;
    }
    void swap(dbuf &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

// ------------------------------------------------------ work counters
//
// A persistent launch pulls work indices from kQueueWords device counters,
// reset by a kernel at the head of the launch. Launches on one stream run in
// order, so every stream gets one set of counters of its own (created on the
// stream's first persistent launch): launches of one plan or batch on
// several streams never share a counter. A captured graph owns a set of its
// own. Sets are carved from chunks that live until the process ends.
constexpr int kCounterStride = 32;    // words per set (128 bytes: one set per cache line)
constexpr int kCountersPerChunk = 256;

struct counter_pool {
    std::mutex mu;
    std::vector<std::pair<std::pair<int, hipStream_t>, uint32_t *>> by_stream;
    std::vector<std::pair<int, uint32_t *>> free_sets;  // (device, set)
};
counter_pool &counters() {
    static counter_pool *pool = new counter_pool();  // never destroyed: kernels may still run at exit
    return *pool;
}

// A fresh set on the current device (pool lock held).
int counter_set_locked(counter_pool &pool, int dev, uint32_t **out) {
    for (size_t k = 0; k < pool.free_sets.size(); ++k)
        if (pool.free_sets[k].first == dev) {
            *out = pool.free_sets[k].second;
            pool.free_sets.erase(pool.free_sets.begin() + (std::ptrdiff_t)k);
            return G2G_OK;
        }
    uint32_t *chunk = nullptr;
    const size_t bytes = sizeof(uint32_t) * kCounterStride * kCountersPerChunk;
    hipError_t e = hipMalloc(&chunk, bytes);
    if (e != hipSuccess) return fail(G2G_ERR_NOMEM, "work counters: hipMalloc failed: %s", hipGetErrorString(e));
    e = hipMemset(chunk, 0, bytes);
    if (e != hipSuccess) return fail(G2G_ERR_HIP, "work counters: hipMemset failed: %s", hipGetErrorString(e));
    for (int k = kCountersPerChunk - 1; k >= 1; --k) pool.free_sets.push_back({dev, chunk + k * kCounterStride});
    *out = chunk;
    return G2G_OK;
}

// The counters of launches on stream s.
int stream_counters(hipStream_t s, uint32_t **out) {
    int dev = 0;
    G2G_HIP(hipGetDevice(&dev));
    counter_pool &pool = counters();
    std::lock_guard<std::mutex> g(pool.mu);
    for (auto &e : pool.by_stream)
        if (e.first.first == dev && e.first.second == s) {
            *out = e.second;
            return G2G_OK;
        }
    G2G_TRY(counter_set_locked(pool, dev, out));
    pool.by_stream.push_back({{dev, s}, *out});
    return G2G_OK;
}

// A set owned by one object (a captured graph); give it back with release_counters.
int private_counters(uint32_t **out) {
    int dev = 0;
    G2G_HIP(hipGetDevice(&dev));
    counter_pool &pool = counters();
    std::lock_guard<std::mutex> g(pool.mu);
    return counter_set_locked(pool, dev, out);
}

void release_counters(uint32_t *set) {
    if (!set) return;
    int dev = 0;
    (void)hipGetDevice(&dev);
    counter_pool &pool = counters();
    std::lock_guard<std::mutex> g(pool.mu);
    pool.free_sets.push_back({dev, set});
}

// Tuning knobs (g2g_set_tuning).
std::atomic<int64_t> g_tile_bytes{0};
std::atomic<int> g_launch_mode{-1};

int device_cus() {
    static std::once_flag once;
    static int cus = 256;
    std::call_once(once, [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
    });
    return cus;
}

// Work-item size: large enough to amortise per-item setup, small enough to
// give every CU several items.
uint64_t choose_tile_bytes(uint64_t total) {
    const int64_t forced = g_tile_bytes.load();
    if (forced > 0) return (uint64_t)forced;
    uint64_t tb = total / (uint64_t)(device_cus() * 8);
    tb = std::max<uint64_t>(tb, 16384);
    tb = std::min<uint64_t>(tb, 262144);
    return (tb + 4095) & ~uint64_t(4095);
}

// ---------------------------------------------------------- layouts

// A validated layout in the working orientation (transposed when the
// transform transposes, as grid_layout::transpose_or_conjugate does in the
// reference, grid_layout.hpp:18-25), with one block per cell.
struct wlayout {
    std::vector<int32_t> rs, cs;  // working splits
    int32_t nbr = 0, nbc = 0;     // working block rows / cols
    std::vector<cell> cells;      // [i * nbc + j]
    int32_t n_ranks = 0;
};

int check_splits(const char *what, const int32_t *s, int32_t n) {
    if (n < 2 || !s) return fail(G2G_ERR_INVALID, "%s: need at least 2 split entries", what);
    if (s[0] < 0) return fail(G2G_ERR_INVALID, "%s: splits must be non-negative", what);
    if (n == 2 && s[1] == s[0]) return G2G_OK;  // an empty dimension: one empty interval
    for (int32_t i = 1; i < n; ++i)
        if (s[i] <= s[i - 1])
            return fail(G2G_ERR_INVALID, "%s: splits must be strictly increasing (entry %d)", what, i);
    return G2G_OK;
}

int load_layout(const char *what, const g2g_layout *l, uint32_t es, bool transpose, wlayout &w) {
    if (!l) return fail(G2G_ERR_INVALID, "%s layout is NULL", what);
    G2G_TRY(check_splits(what, l->rows_split, l->n_rows_split));
    G2G_TRY(check_splits(what, l->cols_split, l->n_cols_split));
    if (l->n_ranks < 1 || l->n_ranks > 65535)
        return fail(G2G_ERR_INVALID, "%s: n_ranks must be in [1, 65535]", what);
    const int32_t br = l->n_rows_split - 1, bc = l->n_cols_split - 1;
    if (!l->owners) return fail(G2G_ERR_INVALID, "%s: owners is NULL", what);
    if (l->n_blocks < 0 || (l->n_blocks > 0 && !l->blocks))
        return fail(G2G_ERR_INVALID, "%s: bad blocks array", what);
    std::vector<cell> sc((size_t)br * bc);
    std::vector<uint8_t> seen((size_t)br * bc, 0);
    for (int64_t k = 0; k < (int64_t)br * bc; ++k) {
        const int32_t o = l->owners[k];
        if (o < 0 || o >= l->n_ranks)
            return fail(G2G_ERR_INVALID, "%s: owner %d of cell %lld outside [0, %d)", what, o, (long long)k,
                        l->n_ranks);
        sc[k].owner = o;
        sc[k].pad = 0;
    }
    for (int32_t b = 0; b < l->n_blocks; ++b) {
        const g2g_block &blk = l->blocks[b];
        if (blk.row < 0 || blk.row >= br || blk.col < 0 || blk.col >= bc)
            return fail(G2G_ERR_INVALID,
                        "%s: block %d coordinates (%d, %d) do not belong to the grid", what, b, blk.row,
                        blk.col);
        const int64_t k = (int64_t)blk.row * bc + blk.col;
        if (seen[k]) return fail(G2G_ERR_INVALID, "%s: cell (%d, %d) has two blocks", what, blk.row, blk.col);
        seen[k] = 1;
        const int64_t rows = l->rows_split[blk.row + 1] - l->rows_split[blk.row];
        if (blk.stride < rows || blk.stride > 0xffffffffll)
            return fail(G2G_ERR_INVALID, "%s: block (%d, %d) stride %lld invalid for %lld rows", what, blk.row,
                        blk.col, (long long)blk.stride, (long long)rows);
        if (!blk.data) return fail(G2G_ERR_INVALID, "%s: block (%d, %d) data is NULL", what, blk.row, blk.col);
        if (((uintptr_t)blk.data) % (es >= 8 ? 8 : es))
            return fail(G2G_ERR_INVALID, "%s: block (%d, %d) data not aligned to its element", what, blk.row,
                        blk.col);
        sc[k].ptr = (uint64_t)(uintptr_t)blk.data;
        sc[k].ld = (uint64_t)blk.stride;
    }
    for (int64_t k = 0; k < (int64_t)br * bc; ++k)
        if (!seen[k])
            return fail(G2G_ERR_INVALID, "%s: cell (%lld, %lld) has no block", what, (long long)(k / bc),
                        (long long)(k % bc));
    w.n_ranks = l->n_ranks;
    if (!transpose) {
        w.rs.assign(l->rows_split, l->rows_split + l->n_rows_split);
        w.cs.assign(l->cols_split, l->cols_split + l->n_cols_split);
        w.nbr = br;
        w.nbc = bc;
        w.cells = std::move(sc);
    } else {
        // assigned_grid2D::transpose (reference grid2D.cpp:47-50, 93-110)
        w.rs.assign(l->cols_split, l->cols_split + l->n_cols_split);
        w.cs.assign(l->rows_split, l->rows_split + l->n_rows_split);
        w.nbr = bc;
        w.nbc = br;
        w.cells.resize(sc.size());
        for (int32_t i = 0; i < br; ++i)
            for (int32_t j = 0; j < bc; ++j) w.cells[(size_t)j * br + i] = sc[(size_t)i * bc + j];
    }
    return G2G_OK;
}

// The grid cover along one dimension: the union of the final and initial
// splits as intervals, each tagged with the final and initial block that
// contain it. This is the interval table equivalent of the reference's
// get_decomp_cover (grid_cover.cpp:54-121), which lists, per block of one
// grid, the range of blocks of the other grid it overlaps.
int union_spans(const std::vector<int32_t> &f, const std::vector<int32_t> &i, const char *dim,
                std::vector<span> &out) {
    if (f.front() != i.front() || f.back() != i.back())
        return fail(G2G_ERR_INVALID,
                    "%s of the initial layout (after op) [%d, %d) and the final layout [%d, %d) differ", dim,
                    i.front(), i.back(), f.front(), f.back());
    out.clear();
    size_t a = 0, b = 0;
    int32_t pos = f.front();
    while (pos < f.back()) {
        const int32_t next = std::min(f[a + 1], i[b + 1]);
        out.push_back(span{pos, next, (int32_t)a, (int32_t)b});
        pos = next;
        if (f[a + 1] == pos) ++a;
        if (i[b + 1] == pos) ++b;
    }
    return G2G_OK;
}

}  // namespace

// ------------------------------------------------------------- plans

struct stage_table {
    dbuf<piece> pieces;
    dbuf<uint64_t> pend;
    dbuf<uint32_t> first;
    int64_t n = 0;
    uint64_t total = 0;
    uint64_t tile_bytes = 0;
    uint32_t n_tiles = 0;
    uint32_t split = 0;        // 0, or the divisor of the last round's work items (retile)
    uint32_t split_from = 0;   // item_range
    uint32_t split_bytes = 0;
    launch_cfg cfg;
};

struct g2g_plan_s {
    int dtype = 0;
    char trans = 'N';
    uint32_t es = 0;
    int32_t P = 0;
    int64_t m = 0, n = 0;
    std::vector<span> urows, ucols;  // host copy of the interval table

    stage_table fused;               // all pieces in union-cell order (wire order, plain form)
    stage_table alt;                 // the gather's stream when it is not `fused` itself (use_alt)
    bool use_alt = false;
    uint32_t alt_w = 0;              // alt: columns per chunk (0: whole pieces)
    uint32_t alt_order = 0;          // alt: G2G_ORDER_* of the stream
    uint32_t alt_h = 0;              // alt: rows per chunk (0: whole columns)
    uint64_t stream_gen = 0;         // changes whenever the gather's piece table changes
    uint64_t class_key = 0;          // shape class of the tuning table
    double tuned_ms = 0;             // time of the tuned configuration (0 = not tuned)
    dbuf<uint32_t> ranks;            // (src << 16) | dst per piece
    // guards what is built on first use (wire order, staging, chunks) and
    // the plan's captured graph
    std::mutex mu;
    // g2g_transform_graph: the launch captured for one (alpha, beta), its
    // stream table and configuration; the graph owns its work counters
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t *gcounters = nullptr;
    gather_args gargs;
    launch_cfg gcfg;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~g2g_plan_s() {
        drop_graph();
        release_counters(gcounters);
    }

    // wire order (computed on demand)
    bool wire_ready = false;
    dbuf<uint64_t> within;
    std::vector<uint64_t> counts;    // [s * P + d] elements
    std::vector<uint64_t> sbase, sdspl, rbase, rdspl;
    int64_t remote_elems = 0;

    // plain form (allocated on demand)
    bool staged_ready = false;
    dbuf<unsigned char> send_arena, recv_arena;
    stage_table sa, sb, sc;
};

namespace {

// The table the fused gather runs over: the plan's selected stream (column
// chunks and/or destination order) when it has one, else the pieces
// themselves in union-cell order.
stage_table &gtab(g2g_plan p) { return p->use_alt ? p->alt : p->fused; }

// The table a request runs over: scaled requests take the pieces in their own
// order, whose rows and columns are the whole piece's (the scaled epilogue's
// operand order depends on an element's position in the piece).
stage_table &xtab(g2g_plan p, const kreqs &kr) {
    return ((kr.r[0].flags | kr.r[1].flags) & REQ_SCALE) ? p->fused : gtab(p);
}

bool persistent_mode(uint32_t n_tiles) {
    const int p = g_launch_mode.load();
    if (p >= 0) return p == G2G_LAUNCH_PERSISTENT;
    return n_tiles > (uint32_t)device_cus() * 64u;
}

void set_launch_mode(launch_cfg &c, int mode) {
    c.persistent = mode == G2G_LAUNCH_PERSISTENT || mode == G2G_LAUNCH_PERSISTENT_XCD;
    c.xcd_ranges = mode == G2G_LAUNCH_DIRECT_XCD || mode == G2G_LAUNCH_PERSISTENT_XCD;
}

int launch_mode_of(const launch_cfg &c) {
    if (c.persistent) return c.xcd_ranges ? G2G_LAUNCH_PERSISTENT_XCD : G2G_LAUNCH_PERSISTENT;
    return c.xcd_ranges ? G2G_LAUNCH_DIRECT_XCD : G2G_LAUNCH_DIRECT;
}

int grid_limit() { return device_cus() * 8; }

// Work items of a table for a given size: first piece per item.
// Work items of a stream: tile_bytes each; with split = 2, 4 or 8, the last
// round of a launch -- the stream's last min(total / 2, CUs x 4 x tile_bytes)
// bytes, the items before them whole -- in items of tile_bytes / split, so
// that the workgroups of a launch finish closer together.
int retile(stage_table &t, uint64_t tile_bytes, hipStream_t s, uint32_t split = 0) {
    if (split != 0 && split != 2 && split != 4 && split != 8)
        return fail(G2G_ERR_INVALID, "tail_split %u: 0, 2, 4 or 8", split);
    uint64_t full = (t.total + tile_bytes - 1) / tile_bytes, nt = full, sb = tile_bytes;
    if (split) {
        sb = tile_bytes / split;
        const uint64_t tail = std::min(t.total / 2, (uint64_t)device_cus() * 4 * tile_bytes);
        full = (t.total - tail) / tile_bytes;
        nt = full + (t.total - full * tile_bytes + sb - 1) / sb;
    }
    if (nt > 0xffffffffull) return fail(G2G_ERR_UNSUPPORTED, "too many work items");
    t.tile_bytes = tile_bytes;
    t.split = split;
    t.n_tiles = (uint32_t)nt;
    t.split_from = (uint32_t)full;
    t.split_bytes = (uint32_t)sb;
    G2G_TRY(t.first.alloc(nt));
    G2G_HIP(launch_first_piece(t.pend.p, t.n, t.tile_bytes, t.split_from, t.split_bytes, (int64_t)t.n_tiles, t.first.p,
                               s));
    G2G_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}

// ---- tuning table: best configuration per shape class
struct tuned_cfg {
    uint64_t tile_bytes;
    launch_cfg cfg;
    double ms;
    uint32_t chunk_w = 0;  // > 0: the gather runs over column chunks of this width
    uint32_t order = 0;    // G2G_ORDER_* of the gather's stream
    uint32_t chunk_h = 0;  // > 0: ... and row chunks of this height
    uint32_t split = 0;    // retile's split of the last round's work items
};
std::mutex g_tune_mu;
std::vector<std::pair<uint64_t, tuned_cfg>> g_tune_table;

int ilog2(uint64_t v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        ++r;
    }
    return r;
}

// element size, transposition, size class and piece-size class
uint64_t shape_class(uint32_t es, char trans, uint64_t total, int64_t n_pieces) {
    const uint64_t avg = n_pieces > 0 ? total / (uint64_t)n_pieces : total;
    return ((uint64_t)es << 24) | ((uint64_t)(trans != 'N') << 16) | ((uint64_t)ilog2(total | 1) << 8) |
           (uint64_t)ilog2(avg | 1);
}

bool lookup_tuned(uint64_t key, tuned_cfg &out) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (auto &e : g_tune_table)
        if (e.first == key) {
            out = e.second;
            return true;
        }
    return false;
}

void store_tuned(uint64_t key, const tuned_cfg &c) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (auto &e : g_tune_table)
        if (e.first == key) {
            e.second = c;
            return;
        }
    g_tune_table.push_back({key, c});
}

launch_cfg default_cfg(uint32_t n_tiles) {
    launch_cfg c;
    if (g_launch_mode.load() >= 0)
        set_launch_mode(c, g_launch_mode.load());
    else
        c.persistent = persistent_mode(n_tiles);
    return c;
}

// bytes -> inclusive scan -> first piece per tile for a stage table whose
// pieces are already on the device.
int finalize_table(stage_table &t, const kreqs &kr, uint64_t total_hint, hipStream_t s) {
    G2G_TRY(t.pend.alloc((size_t)t.n));
    G2G_HIP(launch_piece_bytes(t.pieces.p, t.n, kr, t.pend.p, s));
    dbuf<uint64_t> scratch;
    G2G_TRY(scratch.alloc(scan_scratch_elems(t.n)));
    G2G_HIP(inclusive_scan_u64(t.pend.p, t.n, scratch.p, s));
    uint64_t total = 0;
    if (t.n) G2G_HIP(hipMemcpyAsync(&total, t.pend.p + t.n - 1, sizeof total, hipMemcpyDeviceToHost, s));
    G2G_HIP(hipStreamSynchronize(s));
    if (total_hint && total != total_hint)
        return fail(G2G_ERR_INVALID, "internal: piece bytes %llu != matrix bytes %llu",
                    (unsigned long long)total, (unsigned long long)total_hint);
    t.total = total;
    G2G_TRY(retile(t, choose_tile_bytes(total), s));
    t.cfg = default_cfg(t.n_tiles);
    return G2G_OK;
}

bool is_one(int dtype, const void *a) {
    switch (dtype) {
    case G2G_FLOAT32: return *(const float *)a == 1.0f;
    case G2G_FLOAT64: return *(const double *)a == 1.0;
    case G2G_COMPLEX64: return ((const float *)a)[0] == 1.0f && ((const float *)a)[1] == 0.0f;
    default: return ((const double *)a)[0] == 1.0 && ((const double *)a)[1] == 0.0;
    }
}

bool is_zero(int dtype, const void *a) {
    switch (dtype) {
    case G2G_FLOAT32: return *(const float *)a == 0.0f;
    case G2G_FLOAT64: return *(const double *)a == 0.0;
    case G2G_COMPLEX64: return ((const float *)a)[0] == 0.0f && ((const float *)a)[1] == 0.0f;
    default: return ((const double *)a)[0] == 0.0 && ((const double *)a)[1] == 0.0;
    }
}

// Request for a plan: the transform's op, and scaling unless (alpha, beta)
// is exactly (1, 0), the reference's copy shortcut (memory_utils.hpp:32-35).
request make_request(int dtype, char trans, const void *alpha, const void *beta, bool with_op,
                     bool with_scale) {
    request r;
    std::memset(&r, 0, sizeof r);
    r.dtype = (uint32_t)dtype;
    r.esize = dtype_size(dtype);
    if (with_op && (trans == 'T' || trans == 'C')) r.flags |= REQ_TRANS;
    if (with_op && trans == 'C' && dtype_complex(dtype)) r.flags |= REQ_CONJ;
    const void *a = alpha, *b = beta;
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    static const float onef[2] = {1.0f, 0.0f}, zerof[2] = {0.0f, 0.0f};
    const bool single = dtype == G2G_FLOAT32 || dtype == G2G_COMPLEX64;
    if (!a) a = single ? (const void *)onef : (const void *)one;
    if (!b) b = single ? (const void *)zerof : (const void *)zero;
    if (with_scale && !(is_one(dtype, a) && is_zero(dtype, b))) r.flags |= REQ_SCALE;
    std::memcpy(r.alpha, a, r.esize);
    std::memcpy(r.beta, b, r.esize);
    return r;
}

int ensure_wire(g2g_plan p, hipStream_t s) {
    if (p->wire_ready) return G2G_OK;
    const int32_t P = p->P;
    // per-pair tables are P*P entries (host and device): bounded here
    if (P > kMaxWireRanks)
        return fail(G2G_ERR_UNSUPPORTED, "the plain form supports at most %d ranks (P = %d)", kMaxWireRanks, P);
    const int64_t n = p->fused.n;
    dbuf<uint64_t> pc;
    G2G_TRY(pc.alloc((size_t)P * P));
    G2G_TRY(p->within.alloc((size_t)n));
    if ((int64_t)P * P * 8 <= 64 * 1024) {  // pair table in LDS: multi-split wavefront scan
        const int64_t n_chunks = std::max<int64_t>(1, (n + 2047) / 2048);
        dbuf<uint64_t> hist;
        G2G_TRY(hist.alloc((size_t)P * P * n_chunks));
        G2G_HIP(launch_pair_offsets(p->ranks.p, p->fused.pieces.p, n, P, hist.p, n_chunks, p->within.p, pc.p,
                                    s));
    } else {
        G2G_TRY(launch_pair_offsets_sorted(p->ranks.p, p->fused.pieces.p, n, P, p->within.p, pc.p, s));
    }
    p->counts.assign((size_t)P * P, 0);
    G2G_HIP(hipMemcpyAsync(p->counts.data(), pc.p, sizeof(uint64_t) * P * P, hipMemcpyDeviceToHost, s));
    G2G_HIP(hipStreamSynchronize(s));
    // communication_data.cpp:87-89: dspls = exclusive prefix of counts,
    // per rank, over peers; the per-rank buffers are laid out rank after rank.
    p->sbase.assign(P, 0);
    p->rbase.assign(P, 0);
    p->sdspl.assign((size_t)P * P, 0);
    p->rdspl.assign((size_t)P * P, 0);
    uint64_t acc = 0;
    for (int32_t sr = 0; sr < P; ++sr) {
        p->sbase[sr] = acc;
        uint64_t d_acc = 0;
        for (int32_t d = 0; d < P; ++d) {
            p->sdspl[(size_t)sr * P + d] = d_acc;
            d_acc += p->counts[(size_t)sr * P + d];
        }
        acc += d_acc;
    }
    p->remote_elems = (int64_t)acc;
    acc = 0;
    for (int32_t d = 0; d < P; ++d) {
        p->rbase[d] = acc;
        uint64_t s_acc = 0;
        for (int32_t sr = 0; sr < P; ++sr) {
            p->rdspl[(size_t)d * P + sr] = s_acc;
            s_acc += p->counts[(size_t)sr * P + d];
        }
        acc += s_acc;
    }
    p->wire_ready = true;
    return G2G_OK;
}

int ensure_staged(g2g_plan p, hipStream_t s) {
    if (p->staged_ready) return G2G_OK;
    G2G_TRY(ensure_wire(p, s));
    const int32_t P = p->P;
    const uint64_t es = p->es;
    G2G_TRY(p->send_arena.alloc((size_t)p->remote_elems * es));
    G2G_TRY(p->recv_arena.alloc((size_t)p->remote_elems * es));
    dbuf<uint64_t> d_sbase, d_sdspl, d_rbase, d_rdspl;
    G2G_TRY(d_sbase.upload(p->sbase.data(), P, s));
    G2G_TRY(d_sdspl.upload(p->sdspl.data(), (size_t)P * P, s));
    G2G_TRY(d_rbase.upload(p->rbase.data(), P, s));
    G2G_TRY(d_rdspl.upload(p->rdspl.data(), (size_t)P * P, s));
    const int64_t n = p->fused.n;
    G2G_TRY(p->sa.pieces.alloc((size_t)n));
    G2G_TRY(p->sc.pieces.alloc((size_t)n));
    p->sa.n = p->sc.n = n;
    G2G_HIP(launch_stage_pieces(p->fused.pieces.p, p->ranks.p, p->within.p, n, P, d_sbase.p, d_sdspl.p,
                                d_rbase.p, d_rdspl.p, (uint64_t)(uintptr_t)p->send_arena.p,
                                (uint64_t)(uintptr_t)p->recv_arena.p, p->es, p->sa.pieces.p,
                                p->sc.pieces.p, s));
    // Stage B: one contiguous segment per (source, destination) pair with a
    // nonzero count, as MPI_Isend/MPI_Irecv of transform.cpp:53-87 move them.
    std::vector<piece> seg;
    const uint64_t send0 = (uint64_t)(uintptr_t)p->send_arena.p, recv0 = (uint64_t)(uintptr_t)p->recv_arena.p;
    for (int32_t sr = 0; sr < P; ++sr)
        for (int32_t d = 0; d < P; ++d) {
            uint64_t cnt = p->counts[(size_t)sr * P + d];
            uint64_t so = p->sbase[sr] + p->sdspl[(size_t)sr * P + d];
            uint64_t ro = p->rbase[d] + p->rdspl[(size_t)d * P + sr];
            // A contiguous segment viewed as R-element columns (any factoring
            // of a contiguous range is the same copy), so the gather spreads it
            // over whole workgroups; the remainder is one short column.
            const uint64_t R = 1024;
            while (cnt) {
                piece pc;
                std::memset(&pc, 0, sizeof pc);
                pc.src = send0 + so * es;
                pc.dst = recv0 + ro * es;
                uint64_t len;
                if (cnt >= R) {
                    const uint64_t cols = std::min<uint64_t>(cnt / R, 1u << 30);
                    pc.lds = pc.ldd = pc.rows = (uint32_t)R;
                    pc.cols = (uint32_t)cols;
                    len = cols * R;
                } else {
                    pc.lds = pc.ldd = pc.rows = (uint32_t)cnt;
                    pc.cols = 1;
                    len = cnt;
                }
                pc.vmul = div_magic((pc.rows * (uint32_t)es + 15) / 16);
                seg.push_back(pc);
                cnt -= len;
                so += len;
                ro += len;
            }
        }
    p->sb.n = (int64_t)seg.size();
    G2G_TRY(p->sb.pieces.upload(seg.data(), seg.size(), s));
    kreqs kr;
    std::memset(&kr, 0, sizeof kr);
    kr.r[0] = make_request(p->dtype, p->trans, nullptr, nullptr, false, false);
    kr.r[1] = kr.r[0];
    G2G_TRY(finalize_table(p->sa, kr, (uint64_t)p->m * p->n * es, s));
    G2G_TRY(finalize_table(p->sb, kr, (uint64_t)p->remote_elems * es, s));
    G2G_TRY(finalize_table(p->sc, kr, (uint64_t)p->remote_elems * es, s));
    p->staged_ready = true;
    return G2G_OK;
}

work_desc make_desc(const stage_table &t, const kreqs &kr, uint32_t tile_base) {
    work_desc d;
    std::memset(&d, 0, sizeof d);
    d.pieces = t.pieces.p;
    d.pend = t.pend.p;
    d.first_piece = t.first.p;
    d.tile_bytes = t.tile_bytes;
    d.total_bytes = t.total;
    d.n_tiles = t.n_tiles;
    d.split_from = t.split_from;
    d.split_bytes = t.split_bytes;
    d.tile_base = tile_base;
    d.n_pieces = (uint32_t)t.n;
    d.kr = kr;
    return d;
}

// The launch of one table: arguments and configuration. `queue`: the work
// counters of a persistent launch (the stream's when null).
int table_launch(const stage_table &t, const kreqs &kr, uint32_t *queue, hipStream_t s, gather_args &a,
                 launch_cfg &c) {
    std::memset(&a, 0, sizeof a);
    a.one = make_desc(t, kr, 0);
    a.total_tiles = t.n_tiles;
    c = t.cfg;
    if (g_launch_mode.load() >= 0) set_launch_mode(c, g_launch_mode.load());
    c.features = std::max(request_features(kr.r[0]), request_features(kr.r[1]));
    if (c.persistent && t.n_tiles) {
        if (!queue) G2G_TRY(stream_counters(s, &queue));
        a.queue = queue;
    }
    return G2G_OK;
}

int run_table(const stage_table &t, const kreqs &kr, hipStream_t s, uint32_t *queue = nullptr) {
    gather_args a;
    launch_cfg c;
    G2G_TRY(table_launch(t, kr, queue, s, a, c));
    G2G_HIP(launch_gather(a, c, device_cus(), s));
    return G2G_OK;
}

}  // namespace

extern "C" {

const char *g2g_last_error(void) { return t_last_error.c_str(); }

int g2g_version(void) { return G2G_ABI_VERSION; }

int g2g_synchronize(void *stream) {
    G2G_HIP(hipStreamSynchronize((hipStream_t)stream));
    return G2G_OK;
}

int g2g_set_tuning(int64_t tile_bytes, int launch_mode) {
    if (tile_bytes != 0 && (tile_bytes < 4096 || tile_bytes > (1ll << 26)))
        return fail(G2G_ERR_INVALID, "tile_bytes %lld outside [4096, 2^26]", (long long)tile_bytes);
    if (launch_mode > G2G_LAUNCH_PERSISTENT_XCD) return fail(G2G_ERR_INVALID, "unknown launch mode %d", launch_mode);
    g_tile_bytes.store(tile_bytes);
    g_launch_mode.store(launch_mode < 0 ? -1 : launch_mode);
    return G2G_OK;
}

namespace {

// Sub-pieces of the column-chunked stream of width w (n_ur per chunk).
uint64_t chunk_count(const g2g_plan_s *p, uint32_t w) {
    uint64_t g = 0;
    for (const span &c : p->ucols) g += ((uint64_t)(c.end - c.start) + w - 1) / w;
    return g * p->urows.size();
}

// The gather's stream of a plan: the pieces (w = 0) or the pieces cut into
// chunks of at most w final columns, ordered chunk after chunk down each
// column interval (w > 0), untransposed streams optionally cut again into
// chunks of at most h rows (h > 0, each piece's row chunks in row order), in
// that order (G2G_ORDER_UNION) or ordered by destination address
// (G2G_ORDER_DESTINATION). Built on first use, rebuilt for another (w, h,
// order). Column chunks: vertically adjacent pieces then move the same
// columns at about the same time: their destination segments, contiguous
// within a final block column, are written together (longer DRAM runs, and
// the cache line two pieces share when a segment does not end on a 128-byte
// line is merged in L2 instead of reaching memory partly written, twice).
// Row chunks: a tall piece's long column runs become runs of h rows, so that
// a work item moves a few-KiB-high band of many columns at a large stride
// rather than a few long contiguous runs. Destination order: work items that
// run at the same time write neighbouring destination addresses, final block
// after final block.
int select_stream(g2g_plan p, uint32_t w, uint32_t order, uint32_t h, hipStream_t s) {
    if ((w == 0 && h == 0 && order == G2G_ORDER_UNION) || p->fused.n == 0) {
        if (p->use_alt) ++p->stream_gen;
        p->use_alt = false;
        return G2G_OK;
    }
    if (h && p->trans != 'N') return fail(G2G_ERR_INVALID, "row chunks apply to untransposed plans only");
    if (p->alt.n && p->alt_w == w && p->alt_order == order && p->alt_h == h) {
        if (!p->use_alt) ++p->stream_gen;
        p->use_alt = true;
        return G2G_OK;
    }
    ++p->stream_gen;  // batches built on the old table must not run it
    p->use_alt = false;
    p->alt.pieces.release();
    p->alt.pend.release();
    p->alt.first.release();
    p->alt.n = 0;
    p->alt.tile_bytes = 0;
    p->alt_w = w;
    p->alt_order = order;
    p->alt_h = h;
    // the stream's pieces before ordering: the pieces, column chunks of
    // them, row chunks of either; `bytes` of a built stage
    const piece *cur = p->fused.pieces.p;
    int64_t n = p->fused.n;
    dbuf<piece> cbuf, rbuf;
    dbuf<uint64_t> cbytes, rbytes;
    dbuf<piece> *last = nullptr;
    dbuf<uint64_t> *last_bytes = nullptr;
    if (w) {
        const int64_t n_ur = (int64_t)p->urows.size();
        std::vector<uint32_t> cuj, cc0;
        for (size_t uj = 0; uj < p->ucols.size(); ++uj) {
            const int32_t cols = p->ucols[uj].end - p->ucols[uj].start;
            for (int32_t c0 = 0; c0 < cols; c0 += (int32_t)w) {
                cuj.push_back((uint32_t)uj);
                cc0.push_back((uint32_t)c0);
            }
        }
        const int64_t n_chunks = (int64_t)cuj.size();
        n = n_chunks * n_ur;
        // sub-piece indices, plus the 64 examined at once by a wave, stay within 32 bits
        if (n > 0xffffff00ll)
            return fail(G2G_ERR_UNSUPPORTED, "%lld column chunks of %u columns: too many", (long long)n, w);
        dbuf<uint32_t> d_cuj, d_cc0;
        G2G_TRY(d_cuj.upload(cuj.data(), cuj.size(), s));
        G2G_TRY(d_cc0.upload(cc0.data(), cc0.size(), s));
        G2G_TRY(cbuf.alloc((size_t)n));
        G2G_TRY(cbytes.alloc((size_t)n));
        G2G_HIP(launch_chunk_pieces(p->fused.pieces.p, (int32_t)n_ur, d_cuj.p, d_cc0.p, n_chunks, w, p->es,
                                    p->trans != 'N', cbuf.p, cbytes.p, s));
        G2G_HIP(hipStreamSynchronize(s));  // the chunk tables go out of scope
        cur = cbuf.p;
        last = &cbuf;
        last_bytes = &cbytes;
    }
    if (h) {
        dbuf<uint64_t> ends, scratch;
        G2G_TRY(ends.alloc((size_t)n));
        G2G_HIP(launch_row_counts(cur, n, h, ends.p, s));
        G2G_TRY(scratch.alloc(scan_scratch_elems(n)));
        G2G_HIP(inclusive_scan_u64(ends.p, n, scratch.p, s));
        uint64_t m = 0;
        G2G_HIP(hipMemcpyAsync(&m, ends.p + n - 1, sizeof m, hipMemcpyDeviceToHost, s));
        G2G_HIP(hipStreamSynchronize(s));
        if (m > 0xffffff00ull)
            return fail(G2G_ERR_UNSUPPORTED, "%llu row chunks of %u rows: too many", (unsigned long long)m, h);
        G2G_TRY(rbuf.alloc((size_t)m));
        G2G_TRY(rbytes.alloc((size_t)m));
        G2G_HIP(launch_row_chunks(cur, n, h, p->es, ends.p, rbuf.p, rbytes.p, s));
        G2G_HIP(hipStreamSynchronize(s));  // ends goes out of scope
        cur = rbuf.p;
        n = (int64_t)m;
        last = &rbuf;
        last_bytes = &rbytes;
    }
    if (order == G2G_ORDER_DESTINATION) {
        G2G_TRY(p->alt.pieces.alloc((size_t)n));
        G2G_TRY(p->alt.pend.alloc((size_t)n));
        G2G_TRY(sort_pieces_by_dst(cur, n, p->es, p->alt.pieces.p, p->alt.pend.p, s));
    } else {  // chunks in their own order: the last stage's tables become the stream's
        p->alt.pieces.swap(*last);
        p->alt.pend.swap(*last_bytes);
    }
    dbuf<uint64_t> scratch;
    G2G_TRY(scratch.alloc(scan_scratch_elems(n)));
    G2G_HIP(inclusive_scan_u64(p->alt.pend.p, n, scratch.p, s));
    uint64_t check = 0;
    if (n) G2G_HIP(hipMemcpyAsync(&check, p->alt.pend.p + n - 1, sizeof check, hipMemcpyDeviceToHost, s));
    G2G_HIP(hipStreamSynchronize(s));
    if (check != p->fused.total)
        return fail(G2G_ERR_INVALID, "internal: the stream covers %llu bytes, matrix has %llu",
                    (unsigned long long)check, (unsigned long long)p->fused.total);
    p->alt.n = n;  // built
    p->alt.total = p->fused.total;
    p->alt.tile_bytes = 0;  // not tiled yet
    p->alt.cfg = p->fused.cfg;
    p->use_alt = true;
    return G2G_OK;
}

}  // namespace

int g2g_plan_create(const g2g_layout *initial, const g2g_layout *final_, int dtype, char trans,
                    void *stream, g2g_plan *out) {
    if (!out) return fail(G2G_ERR_INVALID, "out is NULL");
    *out = nullptr;
    const uint32_t es = dtype_size(dtype);
    if (!es) return fail(G2G_ERR_INVALID, "unknown dtype %d", dtype);
    trans = (char)std::toupper((unsigned char)trans);
    if (trans != 'N' && trans != 'T' && trans != 'C')
        return fail(G2G_ERR_INVALID, "trans must be 'N', 'T' or 'C'");
    hipStream_t s = (hipStream_t)stream;
    wlayout wi, wf;
    G2G_TRY(load_layout("initial", initial, es, trans != 'N', wi));
    G2G_TRY(load_layout("final", final_, es, false, wf));
    g2g_plan p = new (std::nothrow) g2g_plan_s();
    if (!p) return fail(G2G_ERR_NOMEM, "out of host memory");
    struct guard {
        g2g_plan &p;
        ~guard() { delete p; }
    } g{p};
    p->dtype = dtype;
    p->trans = trans;
    p->es = es;
    p->P = std::max(wi.n_ranks, wf.n_ranks);
    p->m = wf.rs.back() - wf.rs.front();
    p->n = wf.cs.back() - wf.cs.front();
    G2G_TRY(union_spans(wf.rs, wi.rs, "rows", p->urows));
    G2G_TRY(union_spans(wf.cs, wi.cs, "columns", p->ucols));
    const int64_t n_ur = (int64_t)p->urows.size(), n_uc = (int64_t)p->ucols.size();
    const int64_t np = n_ur * n_uc;
    // piece indices, plus the 64 examined at once by a wave, stay within 32 bits
    if (np > 0xffffff00ll) return fail(G2G_ERR_UNSUPPORTED, "too many pieces (%lld)", (long long)np);
    // the gather rounds a column's bytes up to 16 in 32-bit arithmetic
    for (const span &r : p->urows)
        if ((uint64_t)(r.end - r.start) * es > 0xffffffffull - 32)
            return fail(G2G_ERR_UNSUPPORTED, "piece column exceeds 4 GiB - 32 bytes");
    dbuf<span> d_ur, d_uc;
    dbuf<int32_t> d_irs, d_ics, d_frs, d_fcs;
    dbuf<cell> d_ic, d_fc;
    G2G_TRY(d_ur.upload(p->urows.data(), p->urows.size(), s));
    G2G_TRY(d_uc.upload(p->ucols.data(), p->ucols.size(), s));
    G2G_TRY(d_irs.upload(wi.rs.data(), wi.rs.size(), s));
    G2G_TRY(d_ics.upload(wi.cs.data(), wi.cs.size(), s));
    G2G_TRY(d_frs.upload(wf.rs.data(), wf.rs.size(), s));
    G2G_TRY(d_fcs.upload(wf.cs.data(), wf.cs.size(), s));
    G2G_TRY(d_ic.upload(wi.cells.data(), wi.cells.size(), s));
    G2G_TRY(d_fc.upload(wf.cells.data(), wf.cells.size(), s));
    p->fused.n = np;
    G2G_TRY(p->fused.pieces.alloc((size_t)np));
    G2G_TRY(p->fused.pend.alloc((size_t)np));
    G2G_TRY(p->ranks.alloc((size_t)np));
    G2G_HIP(launch_plan_pieces(d_ur.p, (int32_t)n_ur, d_uc.p, (int32_t)n_uc, d_irs.p, d_ics.p, d_frs.p, d_fcs.p,
                               d_ic.p, wi.nbc, d_fc.p, wf.nbc, es, trans != 'N', p->fused.pieces.p,
                               p->fused.pend.p, p->ranks.p, s));
    dbuf<uint64_t> scratch;
    G2G_TRY(scratch.alloc(scan_scratch_elems(np)));
    G2G_HIP(inclusive_scan_u64(p->fused.pend.p, np, scratch.p, s));
    const uint64_t total = (uint64_t)p->m * p->n * es;
    p->fused.total = total;
    uint64_t check = 0;
    if (np) G2G_HIP(hipMemcpyAsync(&check, p->fused.pend.p + np - 1, sizeof check, hipMemcpyDeviceToHost, s));
    G2G_HIP(hipStreamSynchronize(s));
    if (check != total)
        return fail(G2G_ERR_INVALID, "internal: pieces cover %llu bytes, matrix has %llu",
                    (unsigned long long)check, (unsigned long long)total);
    // configuration: the tuned one of this shape class (stream included),
    // else the defaults
    p->class_key = shape_class(es, trans, total, np);
    tuned_cfg tc;
    if (g_tile_bytes.load() == 0 && lookup_tuned(p->class_key, tc)) {
        if (tc.chunk_w || tc.order || tc.chunk_h) G2G_TRY(select_stream(p, tc.chunk_w, tc.order, tc.chunk_h, s));
        stage_table &gt = gtab(p);
        G2G_TRY(retile(gt, tc.tile_bytes, s, tc.split));
        gt.cfg = tc.cfg;
    } else {
        stage_table &gt = gtab(p);
        G2G_TRY(retile(gt, choose_tile_bytes(total), s));
        gt.cfg = default_cfg(gt.n_tiles);
    }
    // scaled requests run over the pieces themselves: tiled from the start,
    // so that no launch has to build anything
    if (p->use_alt) {
        G2G_TRY(retile(p->fused, choose_tile_bytes(total), s));
        p->fused.cfg = default_cfg(p->fused.n_tiles);
    }
    *out = p;
    p = nullptr;  // released from the guard
    return G2G_OK;
}

int g2g_plan_destroy(g2g_plan plan) {
    delete plan;
    return G2G_OK;
}

int g2g_plan_get_tuning(g2g_plan p, g2g_tuning *t) {
    if (!p || !t) return fail(G2G_ERR_INVALID, "NULL argument");
    const stage_table &gt = gtab(p);
    t->tile_bytes = (int64_t)gt.tile_bytes;
    t->threads = gt.cfg.threads;
    t->unroll_bytes = gt.cfg.unroll_bytes;
    t->nontemporal = gt.cfg.nontemporal;
    t->launch_mode = launch_mode_of(gt.cfg);
    t->column_chunks = p->use_alt ? (int32_t)p->alt_w : 0;
    t->stream_order = p->use_alt ? (int32_t)p->alt_order : G2G_ORDER_UNION;
    t->row_chunks = p->use_alt ? (int32_t)p->alt_h : 0;
    t->tail_split = (int32_t)gt.split;
    t->ms = p->tuned_ms;
    return G2G_OK;
}

int g2g_plan_set_tuning(g2g_plan p, const g2g_tuning *t) {
    if (!p || !t) return fail(G2G_ERR_INVALID, "NULL argument");
    if (t->tile_bytes < 4096 || t->tile_bytes > (1ll << 26) || (t->threads != 256 && t->threads != 512) ||
        (t->unroll_bytes != 64 && t->unroll_bytes != 128) || t->launch_mode < 0 ||
        t->launch_mode > G2G_LAUNCH_PERSISTENT_XCD)
        return fail(G2G_ERR_INVALID, "tuning outside the compiled variants");
    if (t->column_chunks < 0) return fail(G2G_ERR_INVALID, "column_chunks %d < 0", (int)t->column_chunks);
    if (t->stream_order != G2G_ORDER_UNION && t->stream_order != G2G_ORDER_DESTINATION)
        return fail(G2G_ERR_INVALID, "unknown stream_order %d", (int)t->stream_order);
    if (t->row_chunks < 0) return fail(G2G_ERR_INVALID, "row_chunks %d < 0", (int)t->row_chunks);
    if (t->row_chunks && p->trans != 'N') return fail(G2G_ERR_INVALID, "row chunks apply to untransposed plans only");
    if (t->tail_split != 0 && t->tail_split != 2 && t->tail_split != 4 && t->tail_split != 8)
        return fail(G2G_ERR_INVALID, "tail_split %d: 0, 2, 4 or 8", (int)t->tail_split);
    std::lock_guard<std::mutex> lock(p->mu);
    G2G_TRY(select_stream(p, (uint32_t)t->column_chunks, (uint32_t)t->stream_order, (uint32_t)t->row_chunks, nullptr));
    stage_table &gt = gtab(p);
    if ((uint64_t)t->tile_bytes != gt.tile_bytes || (uint32_t)t->tail_split != gt.split)
        G2G_TRY(retile(gt, (uint64_t)t->tile_bytes, nullptr, (uint32_t)t->tail_split));
    gt.cfg.threads = t->threads;
    gt.cfg.unroll_bytes = t->unroll_bytes;
    gt.cfg.nontemporal = t->nontemporal != 0;
    set_launch_mode(gt.cfg, t->launch_mode);
    return G2G_OK;
}

}  // extern "C"

namespace {

// Launch timing for the tuners. Launches that move less than kColdBytes,
// run back to back on the same buffers, find their data in the 256 MB
// Infinity Cache, and the fastest configuration there is not the fastest on
// data in HBM, where a caller's data normally is between calls (other work
// has used the cache since). Such launches are timed one at a time, each
// after a kFlushBytes write of a scratch buffer with ordinary (allocating)
// stores has evicted the cache; larger ones back to back.
constexpr uint64_t kColdBytes = 1ull << 30;
constexpr uint64_t kFlushBytes = 512ull << 20;

struct launch_timer {
    hipStream_t s = nullptr;
    bool cold = false;
    dbuf<unsigned char> flush;
    std::vector<hipEvent_t> ev;
    launch_timer() = default;
    launch_timer(const launch_timer &) = delete;
    launch_timer &operator=(const launch_timer &) = delete;
    ~launch_timer() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    // moved_bytes: bytes one launch reads and writes; max_reps: most timed launches per call of time()
    int init(hipStream_t stream, uint64_t moved_bytes, int max_reps) {
        s = stream;
        cold = moved_bytes < kColdBytes;
        // without room for the eviction buffer the launches are timed back
        // to back (on cache-resident data) rather than not at all
        if (cold && flush.alloc(kFlushBytes) != G2G_OK) {
            cold = false;
            (void)hipGetLastError();
        }
        ev.assign(cold ? 2 * (size_t)std::max(1, max_reps) : 2, nullptr);
        for (hipEvent_t &e : ev) G2G_HIP(hipEventCreate(&e));
        return G2G_OK;
    }
    // milliseconds per launch of `launch` (returns a G2G status) over n timed
    // launches after one untimed one
    template <class F>
    int time(F &&launch, int n, double &ms_out) {
        G2G_TRY(launch());
        float total = 0;
        if (!cold) {
            G2G_HIP(hipEventRecord(ev[0], s));
            for (int r = 0; r < n; ++r) G2G_TRY(launch());
            G2G_HIP(hipEventRecord(ev[1], s));
            G2G_HIP(hipEventSynchronize(ev[1]));
            G2G_HIP(hipEventElapsedTime(&total, ev[0], ev[1]));
        } else {
            n = std::max(1, std::min<int>(n, (int)(ev.size() / 2)));
            for (int r = 0; r < n; ++r) {
                G2G_HIP(launch_stream(3, nullptr, flush.p, kFlushBytes, device_cus() * 16, s));
                G2G_HIP(hipEventRecord(ev[2 * r], s));
                G2G_TRY(launch());
                G2G_HIP(hipEventRecord(ev[2 * r + 1], s));
            }
            G2G_HIP(hipEventSynchronize(ev[2 * n - 1]));
            for (int r = 0; r < n; ++r) {
                float ms = 0;
                G2G_HIP(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
                total += ms;
            }
        }
        ms_out = total / std::max(1, n);
        return G2G_OK;
    }
};

}  // namespace

extern "C" {

int g2g_plan_autotune(g2g_plan p, void *stream, int reps, g2g_tuning *best) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    reps = std::max(1, reps);
    launch_timer timer;
    G2G_TRY(timer.init(s, 2 * p->fused.total, 3 * reps));
    kreqs kr;
    std::memset(&kr, 0, sizeof kr);
    kr.r[0] = make_request(p->dtype, p->trans, nullptr, nullptr, true, true);
    kr.r[1] = kr.r[0];
    // The sweep runs over the pieces, in their own order and in destination
    // order; the eight fastest configurations of each order, at their own and
    // the neighbouring work-item sizes, are then timed
    // again on it and on column-chunked streams of several widths in the same
    // order (transposed plans: whole groups of 256 bytes of final columns,
    // the width of a transpose tile of 64 final rows), and on tall pieces
    // also cut into row chunks.
    std::lock_guard<std::mutex> lock(p->mu);
    if (p->use_alt) ++p->stream_gen;  // batches built on another stream must not run it
    p->use_alt = false;
    const uint64_t total = p->fused.total;
    std::vector<uint64_t> sizes;
    for (uint64_t tb : {32768u, 40960u, 49152u, 65536u, 98304u, 131072u, 262144u, 524288u, 1048576u})
        if (sizes.empty() || tb < total * 2) sizes.push_back(tb);
    // time one configuration on the current stream (retiled to tb when it differs)
    auto time_cfg = [&](uint64_t tb, const launch_cfg &c, int n, double &ms_out, uint32_t split = 0) -> int {
        stage_table &gt = gtab(p);
        if (gt.tile_bytes != tb || gt.split != split) G2G_TRY(retile(gt, tb, s, split));
        gt.cfg = c;
        gather_args a;
        std::memset(&a, 0, sizeof a);
        a.one = make_desc(gt, kr, 0);
        a.total_tiles = gt.n_tiles;
        if (c.persistent) G2G_TRY(stream_counters(s, &a.queue));
        return timer.time(
            [&]() -> int {
                G2G_HIP(launch_gather(a, c, device_cus(), s));
                return G2G_OK;
            },
            n, ms_out);
    };
    std::vector<uint32_t> orders{G2G_ORDER_UNION};
    if (p->fused.n > 1) orders.push_back(G2G_ORDER_DESTINATION);
    std::vector<uint32_t> widths{0};
    const uint32_t q = 256 / p->es;
    for (uint32_t w : (p->trans == 'N' ? std::vector<uint32_t>{1, 2, 4, 8, 16, 32, 64}
                                       : std::vector<uint32_t>{q, 2 * q, 4 * q}))
        if (chunk_count(p, w) <= (16ull << 20)) widths.push_back(w);
    // (columns, rows) of each chunked stream timed: the widths above, and for
    // untransposed plans whose pieces' columns average 2 KiB or more, row
    // chunks of 1 and 2 KiB on chunks of 8 and 32 columns
    std::vector<std::pair<uint32_t, uint32_t>> streams;
    for (uint32_t w : widths) streams.push_back({w, 0u});
    const uint64_t mean_colb = p->urows.empty() ? 0 : (uint64_t)p->m / p->urows.size() * p->es;
    if (p->trans == 'N' && mean_colb >= 2048)
        for (uint32_t hb : {1024u, 2048u})
            for (uint32_t w : {8u, 32u})
                if (chunk_count(p, w) * (mean_colb / hb + 1) <= (16ull << 20)) streams.push_back({w, hb / p->es});
    tuned_cfg win{p->fused.tile_bytes, p->fused.cfg, 0};
    double best_ms = 1e30;
    int rc = G2G_OK;
    for (uint32_t order : orders) {
        // sweep every variant, then time the few fastest again with more
        // runs, so that one noisy measurement does not pick the winner
        if ((rc = select_stream(p, 0, order, 0, s)) != G2G_OK) break;
        std::vector<tuned_cfg> cand;
        for (uint64_t tb : sizes) {
            for (int th : {256, 512})
                for (int ub : {64, 128})
                    for (int nt = 0; nt < 2; ++nt)
                        for (int mode : {G2G_LAUNCH_DIRECT, G2G_LAUNCH_PERSISTENT, G2G_LAUNCH_DIRECT_XCD,
                                         G2G_LAUNCH_PERSISTENT_XCD}) {
                            launch_cfg c;
                            c.threads = th;
                            c.unroll_bytes = ub;
                            c.nontemporal = nt != 0;
                            set_launch_mode(c, mode);
                            c.features = request_features(kr.r[0]);
                            double ms = 0;
                            if ((rc = time_cfg(tb, c, reps, ms)) != G2G_OK) break;
                            cand.push_back(tuned_cfg{tb, c, ms});
                        }
            if (rc != G2G_OK) break;
        }
        if (rc != G2G_OK || cand.empty()) break;
        std::sort(cand.begin(), cand.end(), [](const tuned_cfg &x, const tuned_cfg &y) { return x.ms < y.ms; });
        // the eight fastest, each also at the neighbouring work-item sizes
        // (a chunked stream's best size can be one step from the pieces')
        std::vector<tuned_cfg> retime;
        for (size_t k = 0; k < std::min<size_t>(8, cand.size()); ++k) {
            const size_t i = (size_t)(std::find(sizes.begin(), sizes.end(), cand[k].tile_bytes) - sizes.begin());
            for (size_t j = (i ? i - 1 : i); j <= std::min(i + 1, sizes.size() - 1); ++j) {
                tuned_cfg c = cand[k];
                c.tile_bytes = sizes[j];
                bool dup = false;
                for (const tuned_cfg &r : retime)
                    dup = dup || (r.tile_bytes == c.tile_bytes && r.cfg.threads == c.cfg.threads &&
                                  r.cfg.unroll_bytes == c.cfg.unroll_bytes && r.cfg.nontemporal == c.cfg.nontemporal &&
                                  r.cfg.persistent == c.cfg.persistent && r.cfg.xcd_ranges == c.cfg.xcd_ranges);
                if (!dup) retime.push_back(c);
            }
        }
        for (const auto &wh : streams) {
            if ((rc = select_stream(p, wh.first, order, wh.second, s)) != G2G_OK) break;
            for (size_t k = 0; k < retime.size() && rc == G2G_OK; ++k) {
                double ms = 0;
                if ((rc = time_cfg(retime[k].tile_bytes, retime[k].cfg, 3 * reps, ms)) != G2G_OK) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    win = tuned_cfg{retime[k].tile_bytes, retime[k].cfg, ms, wh.first, order, wh.second};
                }
            }
            if (rc != G2G_OK) break;
        }
        if (rc != G2G_OK) break;
    }
    if (rc != G2G_OK) return rc;
    // the winner timed again as it is and with its last round of work items
    // split (retile), in turn, twice each; a split has to gain 0.5 % to be
    // taken (its gains are small, a few tenths of a percent is noise)
    G2G_TRY(select_stream(p, win.chunk_w, win.order, win.chunk_h, s));
    const uint32_t splits[3] = {0u, 2u, 4u};
    double split_ms[3] = {1e30, 1e30, 1e30};
    for (int round = 0; round < 2; ++round)
        for (int k = 0; k < 3; ++k) {
            double ms = 0;
            G2G_TRY(time_cfg(win.tile_bytes, win.cfg, 3 * reps, ms, splits[k]));
            split_ms[k] = std::min(split_ms[k], ms);
        }
    win.ms = split_ms[0];
    win.split = 0;
    for (int k = 1; k < 3; ++k)
        if (split_ms[k] < 0.995 * split_ms[0] && split_ms[k] < win.ms) {
            win.ms = split_ms[k];
            win.split = splits[k];
        }
    stage_table &gt = gtab(p);
    G2G_TRY(retile(gt, win.tile_bytes, s, win.split));
    gt.cfg = win.cfg;
    p->tuned_ms = win.ms;
    store_tuned(p->class_key, win);
    if (best) G2G_TRY(g2g_plan_get_tuning(p, best));
    return G2G_OK;
}

int g2g_plan_get_info(g2g_plan p, g2g_plan_info *info) {
    if (!p || !info) return fail(G2G_ERR_INVALID, "NULL argument");
    std::memset(info, 0, sizeof *info);
    info->m = p->m;
    info->n = p->n;
    info->n_pieces = p->fused.n;
    info->bytes = (int64_t)p->fused.total;
    info->remote_bytes = p->wire_ready ? p->remote_elems * (int64_t)p->es : -1;
    info->n_tiles = gtab(p).n_tiles;
    info->tile_bytes = (int64_t)gtab(p).tile_bytes;
    info->n_ranks = p->P;
    info->dtype = p->dtype;
    info->trans = p->trans;
    return G2G_OK;
}

int g2g_transform(g2g_plan p, const void *alpha, const void *beta, void *stream) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    if ((alpha == nullptr) != (beta == nullptr)) return fail(G2G_ERR_INVALID, "give both alpha and beta or neither");
    kreqs kr;
    std::memset(&kr, 0, sizeof kr);
    kr.r[0] = make_request(p->dtype, p->trans, alpha, beta, true, true);
    kr.r[1] = kr.r[0];
    return run_table(xtab(p, kr), kr, (hipStream_t)stream);
}

int g2g_transform_graph(g2g_plan p, const void *alpha, const void *beta, void *stream) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    if ((alpha == nullptr) != (beta == nullptr)) return fail(G2G_ERR_INVALID, "give both alpha and beta or neither");
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(G2G_ERR_INVALID, "graph replay needs a non-null stream");
    kreqs kr;
    std::memset(&kr, 0, sizeof kr);
    kr.r[0] = make_request(p->dtype, p->trans, alpha, beta, true, true);
    kr.r[1] = kr.r[0];
    std::lock_guard<std::mutex> lock(p->mu);
    if (!p->gcounters) G2G_TRY(private_counters(&p->gcounters));
    gather_args a;
    launch_cfg c;
    G2G_TRY(table_launch(xtab(p, kr), kr, p->gcounters, s, a, c));
    // the graph holds the launch's arguments: capture again when the request,
    // the table or its configuration changed since
    const bool same = p->gexec && std::memcmp(&a, &p->gargs, sizeof a) == 0 && c.threads == p->gcfg.threads &&
                      c.unroll_bytes == p->gcfg.unroll_bytes && c.nontemporal == p->gcfg.nontemporal &&
                      c.persistent == p->gcfg.persistent && c.xcd_ranges == p->gcfg.xcd_ranges &&
                      c.features == p->gcfg.features;
    if (!same) {
        p->drop_graph();
        G2G_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const hipError_t le = launch_gather(a, c, device_cus(), s);
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &graph);
        if (le != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(G2G_ERR_HIP, "graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ee));
        }
        p->graph = graph;
        G2G_HIP(hipGraphInstantiate(&p->gexec, graph, nullptr, nullptr, 0));
        p->gargs = a;
        p->gcfg = c;
    }
    G2G_HIP(hipGraphLaunch(p->gexec, s));
    return G2G_OK;
}

int g2g_transform_staged(g2g_plan p, const void *alpha, const void *beta, void *stream) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    if ((alpha == nullptr) != (beta == nullptr)) return fail(G2G_ERR_INVALID, "give both alpha and beta or neither");
    hipStream_t s = (hipStream_t)stream;
    {
        std::lock_guard<std::mutex> lock(p->mu);
        G2G_TRY(ensure_staged(p, s));
    }
    kreqs ka, kb, kc;
    std::memset(&ka, 0, sizeof ka);
    ka.r[0] = make_request(p->dtype, p->trans, alpha, beta, true, false);  // pack
    ka.r[1] = make_request(p->dtype, p->trans, alpha, beta, true, true);   // local copies
    kb = ka;
    kb.r[0] = make_request(p->dtype, 'N', nullptr, nullptr, false, false);  // exchange
    kb.r[1] = kb.r[0];
    kc = ka;
    kc.r[0] = make_request(p->dtype, 'N', alpha, beta, false, true);        // unpack
    kc.r[1] = kc.r[0];
    G2G_TRY(run_table(p->sa, ka, s));
    G2G_TRY(run_table(p->sb, kb, s));
    G2G_TRY(run_table(p->sc, kc, s));
    return G2G_OK;
}

int g2g_plan_counts(g2g_plan p, int64_t *counts, int64_t capacity) {
    if (!p || !counts) return fail(G2G_ERR_INVALID, "NULL argument");
    if (capacity < (int64_t)p->P * p->P) return fail(G2G_ERR_INVALID, "capacity below P*P");
    std::lock_guard<std::mutex> lock(p->mu);
    G2G_TRY(ensure_wire(p, nullptr));
    for (size_t k = 0; k < p->counts.size(); ++k) counts[k] = (int64_t)p->counts[k];
    return G2G_OK;
}

int g2g_plan_wire(g2g_plan p, int32_t *row0, int32_t *col0, int32_t *rows, int32_t *cols,
                  int32_t *src_rank, int32_t *dst_rank, int64_t *send_off, int64_t *recv_off,
                  int64_t capacity) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    const int64_t n = p->fused.n;
    if (capacity < n) return fail(G2G_ERR_INVALID, "capacity below n_pieces");
    std::lock_guard<std::mutex> lock(p->mu);
    G2G_TRY(ensure_wire(p, nullptr));
    std::vector<uint32_t> rk((size_t)n);
    std::vector<uint64_t> w((size_t)n);
    if (n) {
        G2G_HIP(hipMemcpy(rk.data(), p->ranks.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        G2G_HIP(hipMemcpy(w.data(), p->within.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    }
    const int64_t n_ur = (int64_t)p->urows.size();
    const int32_t P = p->P;
    for (int64_t q = 0; q < n; ++q) {
        const span &R = p->urows[q % n_ur];
        const span &C = p->ucols[q / n_ur];
        const int32_t sr = (int32_t)(rk[q] >> 16), d = (int32_t)(rk[q] & 0xffffu);
        if (row0) row0[q] = R.start;
        if (col0) col0[q] = C.start;
        if (rows) rows[q] = R.end - R.start;
        if (cols) cols[q] = C.end - C.start;
        if (src_rank) src_rank[q] = sr;
        if (dst_rank) dst_rank[q] = d;
        const bool remote = sr != d;
        if (send_off) send_off[q] = remote ? (int64_t)(p->sdspl[(size_t)sr * P + d] + w[q]) : -1;
        if (recv_off) recv_off[q] = remote ? (int64_t)(p->rdspl[(size_t)d * P + sr] + w[q]) : -1;
    }
    return G2G_OK;
}

int g2g_plan_send_buffer(g2g_plan p, int rank, const void **data, int64_t *elems) {
    if (!p || !data || !elems) return fail(G2G_ERR_INVALID, "NULL argument");
    if (!p->staged_ready) return fail(G2G_ERR_STATE, "the plain form has not run on this plan");
    if (rank < 0 || rank >= p->P) return fail(G2G_ERR_INVALID, "rank %d outside [0, %d)", rank, p->P);
    uint64_t tot = 0;
    for (int32_t d = 0; d < p->P; ++d) tot += p->counts[(size_t)rank * p->P + d];
    *data = p->send_arena.p + p->sbase[rank] * p->es;
    *elems = (int64_t)tot;
    return G2G_OK;
}

// ---------------------------------------------------------------- batches

}  // extern "C"

struct g2g_batch_s {
    std::vector<g2g_plan> plans;
    std::vector<work_desc> host_descs;  // requests as the batch runs them
    dbuf<work_desc> descs;
    dbuf<uint32_t> first;               // per plan: first piece of each batch work item
    dbuf<work_rec> work;  // per global work index: transform and first piece
    std::vector<const stage_table *> tabs;  // per plan: the table its request runs over
    uint32_t *gcounters = nullptr;      // work counters of the captured graph
    uint32_t total_tiles = 0;
    uint64_t total_bytes = 0;
    uint64_t tile_bytes = 0;
    launch_cfg cfg;
    double tuned_ms = 0;
    std::vector<uint64_t> plan_gen;     // each plan's stream generation when its desc was made
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~g2g_batch_s() {
        drop_graph();
        release_counters(gcounters);
    }
};

namespace {

// A batch holds its plans' piece tables by address: a plan retuned to
// another stream since (g2g_plan_set_tuning, g2g_plan_autotune) has freed or
// left the table the batch would read.
int batch_check(g2g_batch b) {
    for (size_t k = 0; k < b->plans.size(); ++k)
        if (b->plans[k]->stream_gen != b->plan_gen[k])
            return fail(G2G_ERR_INVALID, "plan %zu was retuned after the batch was built: build the batch again", k);
    return G2G_OK;
}

// queue: the work counters (the stream's when null)
int batch_launch(g2g_batch b, hipStream_t s, uint32_t *queue = nullptr) {
    G2G_TRY(batch_check(b));
    gather_args a;
    std::memset(&a, 0, sizeof a);
    a.descs = b->descs.p;
    a.work = b->work.p;
    a.total_tiles = b->total_tiles;
    if (!queue && b->total_tiles) G2G_TRY(stream_counters(s, &queue));
    a.queue = queue;
    launch_cfg c = b->cfg;
    c.features = FEAT_COPY;
    for (const work_desc &d : b->host_descs)
        c.features = std::max(c.features, std::max(request_features(d.kr.r[0]), request_features(d.kr.r[1])));
    G2G_HIP(launch_gather(a, c, device_cus(), s));
    return G2G_OK;
}

// Work items of a batch: every plan's piece stream cut at the batch's own
// work-item size (a batch of small transforms wants larger items than each
// transform run alone), numbered plan after plan.
int batch_retile(g2g_batch b, uint64_t tile_bytes, hipStream_t s) {
    b->drop_graph();
    b->tile_bytes = tile_bytes;
    std::vector<uint64_t> off(b->plans.size());
    uint64_t tiles = 0;
    for (size_t k = 0; k < b->plans.size(); ++k) {
        off[k] = tiles;
        tiles += (b->tabs[k]->total + tile_bytes - 1) / tile_bytes;
    }
    if (tiles > 0xffffffffull) return fail(G2G_ERR_UNSUPPORTED, "too many work items in one launch");
    b->total_tiles = (uint32_t)tiles;
    G2G_TRY(b->first.alloc(tiles));
    G2G_TRY(b->work.alloc(tiles));
    for (size_t k = 0; k < b->plans.size(); ++k) {
        const stage_table &t = *b->tabs[k];
        work_desc &d = b->host_descs[k];
        d.first_piece = b->first.p + off[k];
        d.tile_bytes = tile_bytes;
        d.n_tiles = (uint32_t)((t.total + tile_bytes - 1) / tile_bytes);
        d.split_from = d.n_tiles;  // a batch's own items are not split
        d.split_bytes = (uint32_t)tile_bytes;
        d.tile_base = (uint32_t)off[k];
        G2G_HIP(launch_first_piece(t.pend.p, t.n, tile_bytes, d.split_from, d.split_bytes, (int64_t)d.n_tiles,
                                   b->first.p + off[k], s));
    }
    G2G_TRY(b->descs.upload(b->host_descs.data(), b->host_descs.size(), s));
    G2G_HIP(launch_work_records(b->descs.p, (int32_t)b->host_descs.size(), b->work.p, s));
    G2G_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}


}  // namespace

namespace g2g {

int internal_desc(g2g_plan_s *p, const void *alpha, const void *beta, uint32_t tile_base, work_desc &out,
                  launch_cfg &cfg) {
    if (!p) return fail(G2G_ERR_INVALID, "plan is NULL");
    if ((alpha == nullptr) != (beta == nullptr)) return fail(G2G_ERR_INVALID, "alpha/beta mismatch");
    kreqs kr;
    std::memset(&kr, 0, sizeof kr);
    kr.r[0] = make_request(p->dtype, p->trans, alpha, beta, true, true);
    kr.r[1] = kr.r[0];
    const stage_table &t = xtab(p, kr);
    out = make_desc(t, kr, tile_base);
    cfg = t.cfg;
    cfg.features = request_features(kr.r[0]);
    return G2G_OK;
}

int internal_descs(int n, g2g_plan_s *const *plans, const void *const *alphas, const void *const *betas,
                   std::vector<work_desc> &out, uint32_t &total_tiles, uint64_t &total_bytes) {
    out.clear();
    uint64_t tiles = 0, bytes = 0;
    for (int k = 0; k < n; ++k) {
        if (!plans[k]) return fail(G2G_ERR_INVALID, "plan %d is NULL", k);
        const void *a = alphas ? alphas[k] : nullptr, *be = betas ? betas[k] : nullptr;
        if ((a == nullptr) != (be == nullptr)) return fail(G2G_ERR_INVALID, "plan %d: alpha/beta mismatch", k);
        kreqs kr;
        std::memset(&kr, 0, sizeof kr);
        kr.r[0] = make_request(plans[k]->dtype, plans[k]->trans, a, be, true, true);
        kr.r[1] = kr.r[0];
        const stage_table &t = xtab(plans[k], kr);
        out.push_back(make_desc(t, kr, (uint32_t)tiles));
        tiles += t.n_tiles;
        bytes += t.total;
        if (tiles > 0xffffffffull) return fail(G2G_ERR_UNSUPPORTED, "too many work items in one launch");
    }
    total_tiles = (uint32_t)tiles;
    total_bytes = bytes;
    return G2G_OK;
}

int internal_grid_limit() { return device_cus(); }

int internal_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_last_error = buf;
    return code;
}

int internal_check_grid(const char *what, const g2g_layout *l) {
    if (!l) return fail(G2G_ERR_INVALID, "%s layout is NULL", what);
    G2G_TRY(check_splits(what, l->rows_split, l->n_rows_split));
    G2G_TRY(check_splits(what, l->cols_split, l->n_cols_split));
    if (l->n_ranks < 1 || l->n_ranks > 65535)
        return fail(G2G_ERR_INVALID, "%s: n_ranks must be in [1, 65535]", what);
    if (!l->owners) return fail(G2G_ERR_INVALID, "%s: owners is NULL", what);
    const int64_t cells = (int64_t)(l->n_rows_split - 1) * (l->n_cols_split - 1);
    for (int64_t k = 0; k < cells; ++k)
        if (l->owners[k] < 0 || l->owners[k] >= l->n_ranks)
            return fail(G2G_ERR_INVALID, "%s: owner %d of cell %lld outside [0, %d)", what, l->owners[k],
                        (long long)k, l->n_ranks);
    return G2G_OK;
}

int internal_union_spans(const std::vector<int32_t> &f, const std::vector<int32_t> &i, const char *dim,
                         std::vector<span> &out) {
    return union_spans(f, i, dim, out);
}

}  // namespace g2g


extern "C" {

int g2g_batch_create(int n, const g2g_plan *plans, const void *const *alphas,
                     const void *const *betas, void *stream, g2g_batch *out) {
    if (!out || (n > 0 && !plans) || n < 0) return fail(G2G_ERR_INVALID, "bad arguments");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    g2g_batch b = new (std::nothrow) g2g_batch_s();
    if (!b) return fail(G2G_ERR_NOMEM, "out of host memory");
    struct guard {
        g2g_batch &b;
        ~guard() { delete b; }
    } g{b};
    uint32_t tiles = 0;
    G2G_TRY(internal_descs(n, plans, alphas, betas, b->host_descs, tiles, b->total_bytes));
    b->plans.assign(plans, plans + n);
    for (int k = 0; k < n; ++k) {
        b->plan_gen.push_back(plans[k]->stream_gen);
        b->tabs.push_back(&xtab(plans[k], b->host_descs[k].kr));
    }
    // batches always run persistent; g2g_set_tuning may ask for per-XCD ranges
    set_launch_mode(b->cfg, g_launch_mode.load() == G2G_LAUNCH_PERSISTENT_XCD ? G2G_LAUNCH_PERSISTENT_XCD
                                                                           : G2G_LAUNCH_PERSISTENT);
    const int64_t forced = g_tile_bytes.load();
    G2G_TRY(batch_retile(b, forced > 0 ? (uint64_t)forced : (uint64_t)131072, s));
    *out = b;
    b = nullptr;
    return G2G_OK;
}

int g2g_batch_execute(g2g_batch b, void *stream) {
    if (!b) return fail(G2G_ERR_INVALID, "batch is NULL");
    return batch_launch(b, (hipStream_t)stream);
}

int g2g_batch_execute_graph(g2g_batch b, void *stream) {
    if (!b) return fail(G2G_ERR_INVALID, "batch is NULL");
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(G2G_ERR_INVALID, "graph replay needs a non-null stream");
    G2G_TRY(batch_check(b));
    if (!b->exec) {
        if (!b->gcounters) G2G_TRY(private_counters(&b->gcounters));
        G2G_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = batch_launch(b, s, b->gcounters);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc != G2G_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(G2G_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        b->graph = graph;
        G2G_HIP(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
    }
    G2G_HIP(hipGraphLaunch(b->exec, s));
    return G2G_OK;
}

int g2g_batch_autotune(g2g_batch b, void *stream, int reps, g2g_tuning *best) {
    if (!b) return fail(G2G_ERR_INVALID, "batch is NULL");
    hipStream_t s = (hipStream_t)stream;
    reps = std::max(1, reps);
    launch_timer timer;
    G2G_TRY(timer.init(s, 2 * b->total_bytes, 3 * reps));
    auto time_cfg = [&](uint64_t tb, const launch_cfg &c, int n, double &ms_out) -> int {
        if (b->tile_bytes != tb) G2G_TRY(batch_retile(b, tb, s));
        b->cfg = c;
        return timer.time([&]() -> int { return batch_launch(b, s); }, n, ms_out);
    };
    // sweep, then time the few fastest again with more runs
    std::vector<std::pair<uint64_t, std::pair<launch_cfg, double>>> cand;
    int rc = G2G_OK;
    if (b->tile_bytes != 32768) rc = batch_retile(b, 32768, s);
    for (uint64_t tb : {32768u, 40960u, 49152u, 65536u, 98304u, 131072u, 262144u, 524288u, 1048576u}) {
        if (rc != G2G_OK) break;
        for (int th : {256, 512})
            for (int ub : {64, 128})
                for (int nt = 0; nt < 2 && rc == G2G_OK; ++nt)
                    for (int mode : {G2G_LAUNCH_PERSISTENT, G2G_LAUNCH_PERSISTENT_XCD}) {
                        launch_cfg c = b->cfg;
                        c.threads = th;
                        c.unroll_bytes = ub;
                        c.nontemporal = nt != 0;
                        set_launch_mode(c, mode);
                        double ms = 0;
                        if ((rc = time_cfg(tb, c, reps, ms)) != G2G_OK) break;
                        cand.push_back({tb, {c, ms}});
                    }
    }
    double best_ms = 1e30;
    uint64_t best_tb = b->tile_bytes;
    launch_cfg best_cfg = b->cfg;
    if (rc == G2G_OK && !cand.empty()) {
        std::sort(cand.begin(), cand.end(),
                  [](const auto &x, const auto &y) { return x.second.second < y.second.second; });
        for (size_t k = 0; k < std::min<size_t>(4, cand.size()) && rc == G2G_OK; ++k) {
            double ms = 0;
            if ((rc = time_cfg(cand[k].first, cand[k].second.first, 3 * reps, ms)) != G2G_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best_tb = cand[k].first;
                best_cfg = cand[k].second.first;
            }
        }
    }
    if (rc != G2G_OK) return rc;
    G2G_TRY(batch_retile(b, best_tb, s));
    b->cfg = best_cfg;
    b->tuned_ms = best_ms;
    if (best) {
        best->tile_bytes = (int64_t)b->tile_bytes;
        best->threads = b->cfg.threads;
        best->unroll_bytes = b->cfg.unroll_bytes;
        best->nontemporal = b->cfg.nontemporal;
        best->launch_mode = launch_mode_of(b->cfg);
        best->column_chunks = 0;
        best->stream_order = G2G_ORDER_UNION;
        best->row_chunks = 0;
        best->tail_split = 0;
        best->ms = b->tuned_ms;
    }
    return G2G_OK;
}

int g2g_batch_get_tuning(g2g_batch b, g2g_tuning *t) {
    if (!b || !t) return fail(G2G_ERR_INVALID, "NULL argument");
    t->tile_bytes = (int64_t)b->tile_bytes;
    t->threads = b->cfg.threads;
    t->unroll_bytes = b->cfg.unroll_bytes;
    t->nontemporal = b->cfg.nontemporal;
    t->launch_mode = launch_mode_of(b->cfg);
    t->column_chunks = 0;
    t->stream_order = G2G_ORDER_UNION;
    t->row_chunks = 0;
    t->tail_split = 0;
    t->ms = b->tuned_ms;
    return G2G_OK;
}

int g2g_batch_set_tuning(g2g_batch b, const g2g_tuning *t, void *stream) {
    if (!b || !t) return fail(G2G_ERR_INVALID, "NULL argument");
    if (t->tile_bytes < 4096 || t->tile_bytes > (1ll << 26) || (t->threads != 256 && t->threads != 512) ||
        (t->unroll_bytes != 64 && t->unroll_bytes != 128) ||
        (t->launch_mode != G2G_LAUNCH_PERSISTENT && t->launch_mode != G2G_LAUNCH_PERSISTENT_XCD))
        return fail(G2G_ERR_INVALID, "batch tuning outside the compiled variants (batches launch persistent)");
    if (t->column_chunks != 0 || t->stream_order != G2G_ORDER_UNION || t->row_chunks != 0 || t->tail_split != 0)
        return fail(G2G_ERR_INVALID,
                    "a batch runs each plan's own stream in its own work items: column_chunks, stream_order, "
                    "row_chunks and tail_split must be 0");
    if ((uint64_t)t->tile_bytes != b->tile_bytes) G2G_TRY(batch_retile(b, (uint64_t)t->tile_bytes, (hipStream_t)stream));
    b->drop_graph();
    b->cfg.threads = t->threads;
    b->cfg.unroll_bytes = t->unroll_bytes;
    b->cfg.nontemporal = t->nontemporal != 0;
    set_launch_mode(b->cfg, t->launch_mode);
    return G2G_OK;
}

int g2g_batch_bytes(g2g_batch b, int64_t *bytes) {
    if (!b || !bytes) return fail(G2G_ERR_INVALID, "NULL argument");
    *bytes = (int64_t)b->total_bytes;
    return G2G_OK;
}

int g2g_batch_destroy(g2g_batch b) {
    delete b;
    return G2G_OK;
}

// ------------------------------------------------------ test support

static int fill_table(const g2g_layout *l, std::vector<fill_block> &fb, int64_t &n_cols) {
    if (!l || !l->rows_split || !l->cols_split || l->n_rows_split < 2 || l->n_cols_split < 2)
        return fail(G2G_ERR_INVALID, "bad layout");
    const int32_t br = l->n_rows_split - 1, bc = l->n_cols_split - 1;
    fb.clear();
    for (int32_t b = 0; b < l->n_blocks; ++b) {
        const g2g_block &x = l->blocks[b];
        if (x.row < 0 || x.row >= br || x.col < 0 || x.col >= bc || !x.data)
            return fail(G2G_ERR_INVALID, "bad block %d", b);
        fill_block f;
        f.ptr = (uint64_t)(uintptr_t)x.data;
        f.ld = (uint64_t)x.stride;
        f.r0 = l->rows_split[x.row];
        f.r1 = l->rows_split[x.row + 1];
        f.c0 = l->cols_split[x.col];
        f.c1 = l->cols_split[x.col + 1];
        fb.push_back(f);
    }
    n_cols = l->cols_split[bc];
    return G2G_OK;
}

int g2g_fill_layout(const g2g_layout *layout, int dtype, uint64_t seed, void *stream) {
    if (!dtype_size(dtype)) return fail(G2G_ERR_INVALID, "unknown dtype");
    std::vector<fill_block> fb;
    int64_t n_cols = 0;
    G2G_TRY(fill_table(layout, fb, n_cols));
    hipStream_t s = (hipStream_t)stream;
    dbuf<fill_block> d;
    G2G_TRY(d.upload(fb.data(), fb.size(), s));
    G2G_HIP(launch_fill(d.p, (int32_t)fb.size(), dtype, seed, n_cols, s));
    G2G_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}

int g2g_scale_layout(const g2g_layout *layout, int dtype, const void *beta, void *stream) {
    if (!dtype_size(dtype) || !beta) return fail(G2G_ERR_INVALID, "bad arguments");
    std::vector<fill_block> fb;
    int64_t n_cols = 0;
    G2G_TRY(fill_table(layout, fb, n_cols));
    const request rq = make_request(dtype, 'N', beta, beta, false, false);
    if (is_one(dtype, beta)) return G2G_OK;  // grid_layout.hpp:50, block.cpp:293: beta == 1 is a no-op
    hipStream_t s = (hipStream_t)stream;
    dbuf<fill_block> d;
    G2G_TRY(d.upload(fb.data(), fb.size(), s));
    G2G_HIP(launch_scale(d.p, (int32_t)fb.size(), rq, s));
    G2G_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}

int g2g_copy_bytes(void *dst, const void *src, int64_t bytes) {
    if (bytes < 0 || (bytes && (!dst || !src))) return fail(G2G_ERR_INVALID, "bad arguments");
    if (bytes) G2G_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
    return G2G_OK;
}

int g2g_verify_layout(const g2g_layout *final_, int dtype, uint64_t seed, char trans,
                      int64_t a_n_cols, void *stream, int64_t *mismatches) {
    if (!dtype_size(dtype) || !mismatches) return fail(G2G_ERR_INVALID, "bad arguments");
    std::vector<fill_block> fb;
    int64_t n_cols = 0;
    G2G_TRY(fill_table(final_, fb, n_cols));
    hipStream_t s = (hipStream_t)stream;
    dbuf<fill_block> d;
    dbuf<unsigned long long> bad;
    G2G_TRY(d.upload(fb.data(), fb.size(), s));
    G2G_TRY(bad.alloc(1));
    G2G_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s));
    const int t = (trans == 'T' || trans == 'C' || trans == 't' || trans == 'c');
    const int c = (trans == 'C' || trans == 'c') && dtype_complex(dtype);
    G2G_HIP(launch_verify(d.p, (int32_t)fb.size(), dtype, seed, t, c, a_n_cols, bad.p, s));
    unsigned long long h = 0;
    G2G_HIP(hipMemcpyAsync(&h, bad.p, sizeof h, hipMemcpyDeviceToHost, s));
    G2G_HIP(hipStreamSynchronize(s));
    *mismatches = (int64_t)h;
    return G2G_OK;
}

int g2g_stream_probe(int kind, const void *src, void *dst, int64_t bytes, int reps, void *stream,
                     double *ms) {
    if (kind < 0 || kind > 2 || !ms || reps < 1 || bytes <= 0 || (bytes & 15) || !dst || (kind != 2 && !src) ||
        ((uintptr_t)src & 15) || ((uintptr_t)dst & 15))
        return fail(G2G_ERR_INVALID, "stream probe: kind 0-2, 16-byte aligned buffers, bytes > 0 and a multiple of 16");
    hipStream_t s = (hipStream_t)stream;
    const int grid = device_cus() * 16;
    hipEvent_t e0, e1;
    G2G_HIP(hipEventCreate(&e0));
    G2G_HIP(hipEventCreate(&e1));
    hipError_t he = launch_stream(kind, src, dst, (uint64_t)bytes, grid, s);  // warm-up
    if (he == hipSuccess) he = hipEventRecord(e0, s);
    for (int r = 0; r < reps && he == hipSuccess; ++r) he = launch_stream(kind, src, dst, (uint64_t)bytes, grid, s);
    if (he == hipSuccess) he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float t = 0;
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (he != hipSuccess) return fail(G2G_ERR_HIP, "stream probe: %s", hipGetErrorString(he));
    *ms = (double)t / reps;
    return G2G_OK;
}

}  // extern "C"
