// Internal declarations shared by the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

struct g2g_plan_s;
struct g2g_layout;  // include/grid2grid_amd.h

namespace g2g {

constexpr int kThreads = 256;    // threads per workgroup of the planning kernels
constexpr int kQueueWords = 8;   // work counters of a persistent launch (one per XCD)
constexpr int kMaxWireRanks = 4096;  // ranks of the plain form (P*P pair tables)

enum : uint32_t {
    REQ_TRANS = 1u,  // destination(i, j) = source(j, i)
    REQ_CONJ = 2u,   // flip the sign bit of the imaginary part
    REQ_SCALE = 4u,  // dst = beta*dst + alpha*op(src)
};

// piece.req: bit 0 selects the launch's second request; bits 1-2 name the
// reference call site that scales the piece (only read when the request
// scales): the per-column loop of copy2D_and_scale, its one loop over a
// contiguous piece, or copy_transpose_and_scale of a local transposed piece
// (reference memory_utils.hpp:79-122, 186-256).
enum : uint32_t {
    PIECE_REQ1 = 1u,
    PIECE_SITE_SHIFT = 1u,
    PIECE_SITE_MASK = 3u << 1,
    PIECE_SITE_COL = 0u,
    PIECE_SITE_FLAT = 1u,
    PIECE_SITE_TRANS = 2u
};

// Per-request parameters: element type, operation and scalars (raw bits).
struct request {
    uint32_t dtype;
    uint32_t flags;
    uint32_t esize;
    uint32_t pad;
    uint64_t alpha[2];
    uint64_t beta[2];
};
static_assert(sizeof(request) == 48, "request layout");

// One rectangle moved by the gather: the intersection of an initial and a
// final cell ("piece"), or a staging segment. Final orientation: rows x cols.
// Source element (i, j) lives at src + (j*lds + i)*esize, or at
// src + (i*lds + j)*esize when the request transposes; destination element
// (i, j) at dst + (j*ldd + i)*esize.
struct piece {
    uint64_t src;
    uint64_t dst;
    uint32_t lds;
    uint32_t ldd;
    uint32_t rows;
    uint32_t cols;
    uint32_t req;
    uint32_t vmul;   // division magic for the 16-byte vectors per column, ceil(rows*esize/16)
};

// Magic multiplier of unsigned 32-bit division by d (Granlund & Montgomery,
// "Division by invariant integers using multiplication", 1994, fig. 4.1):
// q = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0), t = mulhi(magic, n),
// l = ceil(log2(d)).
__host__ __device__ inline uint32_t div_ceil_log2(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    return l;
}
__host__ __device__ inline uint32_t div_magic(uint32_t d) {
    const uint32_t l = div_ceil_log2(d);
    return (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
}
static_assert(sizeof(piece) == 40, "piece layout");

// Device block table entry (one per grid cell).
struct cell {
    uint64_t ptr;   // byte address of the block's element (0, 0)
    uint64_t ld;    // leading dimension in elements
    int32_t owner;  // virtual rank
    int32_t pad;
};
static_assert(sizeof(cell) == 24, "cell layout");

// Interval of the union of two split vectors: [start, end) lies inside
// final block `fblk` and initial block `iblk` (indices along this dimension).
struct span {
    int32_t start;
    int32_t end;
    int32_t fblk;
    int32_t iblk;
};

// Request arguments passed by value; used when no device request table is given.
struct kreqs {
    request r[2];
};

// ------------------------------------------------------------- launchers

// pieces[p] for every union cell p = uj * n_ur + ui, plus bytes[p], ranks.
hipError_t launch_plan_pieces(const span *urows, int32_t n_ur, const span *ucols, int32_t n_uc,
                              const int32_t *irs, const int32_t *ics, const int32_t *frs,
                              const int32_t *fcs, const cell *iblk, int32_t n_ic,
                              const cell *fblk, int32_t n_fc, uint32_t esize, int trans,
                              piece *pieces, uint64_t *bytes, uint32_t *ranks,
                              hipStream_t stream);

// The pieces cut into chunks of at most w final columns and ordered chunk
// after chunk down each column interval: sub-piece k = g * n_ur + ui holds
// final columns [chunk_c0[g], + w) of piece (ui, chunk_uj[g]) (pieces indexed
// uj * n_ur + ui; trans: the pieces' sources are transposed); bytes[k] its size.
hipError_t launch_chunk_pieces(const piece *pieces, int32_t n_ur, const uint32_t *chunk_uj,
                               const uint32_t *chunk_c0, int64_t n_chunks, uint32_t w, uint32_t esize,
                               int trans, piece *out, uint64_t *bytes, hipStream_t stream);

// Untransposed pieces cut into chunks of at most h rows: cnt[p] = chunks of
// piece p; with ends = the inclusive scan of cnt, piece p's chunks go to
// out[ends[p-1] ..] in row order, bytes[k] = out[k]'s bytes.
hipError_t launch_row_counts(const piece *in, int64_t n, uint32_t h, uint64_t *cnt, hipStream_t stream);
hipError_t launch_row_chunks(const piece *in, int64_t n, uint32_t h, uint32_t es, const uint64_t *ends, piece *out,
                             uint64_t *bytes, hipStream_t stream);

// out = the n pieces of `in` ordered by destination address, bytes[k] =
// out[k].rows * cols * es. Returns a g2g_status; synchronous.
int sort_pieces_by_dst(const piece *in, int64_t n, uint32_t es, piece *out, uint64_t *bytes, hipStream_t stream);

// bytes[p] = rows*cols*esize of the piece's request.
hipError_t launch_piece_bytes(const piece *pieces, int64_t n, const kreqs &kr, uint64_t *bytes,
                              hipStream_t stream);

// In-place inclusive scan of n uint64 values (scratch: see scan_scratch_elems).
size_t scan_scratch_elems(int64_t n);
hipError_t inclusive_scan_u64(uint64_t *data, int64_t n, uint64_t *scratch, hipStream_t stream);

// first_piece[t] = the piece containing the first byte of work item t of the
// piece stream (item_range: tile_bytes-long items, from split_from on
// split_bytes-long ones).
hipError_t launch_first_piece(const uint64_t *pend, int64_t n, uint64_t tile_bytes, uint32_t split_from,
                              uint32_t split_bytes, int64_t n_tiles, uint32_t *first_piece, hipStream_t stream);

// Wire order of the plain form: per piece, the element offset inside its
// (source, destination) pair's segment, pieces ordered by position p
// (= the reference's (cols interval, rows interval) order); totals per pair.
hipError_t launch_pair_offsets(const uint32_t *ranks, const piece *pieces, int64_t n,
                               int32_t n_ranks, uint64_t *hist, int64_t n_chunks,
                               uint64_t *within, uint64_t *pair_counts, hipStream_t stream);

// The same by a stable radix sort of the pieces by pair, for rank counts whose
// P*P pair table does not fit in LDS. Returns a g2g_status; synchronous.
int launch_pair_offsets_sorted(const uint32_t *ranks, const piece *pieces, int64_t n, int32_t P,
                               uint64_t *within, uint64_t *pair_counts, hipStream_t stream);

// Stage tables of the plain form from pieces, ranks, within-pair offsets and
// per-rank bases/displacements (elements).
hipError_t launch_stage_pieces(const piece *pieces, const uint32_t *ranks, const uint64_t *within,
                               int64_t n, int32_t n_ranks, const uint64_t *sbase,
                               const uint64_t *sdspl, const uint64_t *rbase,
                               const uint64_t *rdspl, uint64_t send_arena, uint64_t recv_arena,
                               uint32_t esize, piece *stage_a, piece *stage_c,
                               hipStream_t stream);

// One transform (or one stage of the plain form) as the gather sees it: a
// finalized piece table, its work items and the requests its pieces use.
struct work_desc {
    const piece *pieces;
    const uint64_t *pend;         // inclusive byte scan over pieces
    const uint32_t *first_piece;  // per work item
    uint64_t tile_bytes;
    uint64_t total_bytes;
    uint32_t n_tiles;
    uint32_t tile_base;           // first global work index of this desc in a launch
    uint32_t n_pieces;
    uint32_t split_from;          // work items from this one on are split_bytes long (item_range)
    uint32_t split_bytes;
    uint32_t pad;
    kreqs kr;                     // piece.req selects kr.r[0] or kr.r[1] (same element type)
};
static_assert(sizeof(work_desc) == 160, "work_desc layout");

// One work item, everything the gather needs to start on it: the owning
// transform's piece table and byte scan, the item's byte range [lo, hi) of
// the piece stream, its first piece and the last piece that can hold one of
// its columns (the next work item's first piece, or the table's last), and
// (multi-transform launches) the index of the owning desc, whose requests
// apply.
struct work_rec {
    const piece *pieces;
    const uint64_t *pend;
    uint64_t lo, hi;
    uint32_t first_piece;
    uint32_t n_pieces;
    uint32_t desc;
    uint32_t last_piece;
};
static_assert(sizeof(work_rec) == 48, "work_rec layout");

// A launch: one desc passed by value, or a device table of descs with a
// record for every global work index (multi-transform launches).
struct gather_args {
    const work_desc *descs;       // device table, or null: `one` is used
    const work_rec *work;         // per global work index
    work_desc one;
    uint32_t total_tiles;
    uint32_t xcd_ranges;          // each XCD takes one contiguous range of work items
    uint32_t *queue;              // persistent-mode work counters, kQueueWords of them (reset by the launcher)
};
// Compiled variants of the gather: workgroup size, bytes each thread keeps in
// flight, non-temporal hint, launch mode.
// Feature levels: the smallest kernel that covers every request of a launch
// (fewer paths, fewer registers, more waves per CU).
enum : int {
    FEAT_COPY = 0,   // untransposed, unscaled pieces only
    FEAT_TRANS = 1,  // plus transposed / conjugated pieces
    FEAT_ALL = 2     // plus scaling
};
struct launch_cfg {
    int threads = 256;       // 256 or 512
    int unroll_bytes = 64;   // 64 or 128
    bool nontemporal = false;
    bool persistent = false;
    bool xcd_ranges = false;  // contiguous work ranges per XCD (persistent: one counter per XCD)
    int features = FEAT_ALL;
};

// Feature level a request needs.
inline int request_features(const request &r) {
    if (r.flags & REQ_SCALE) return FEAT_ALL;
    if (r.flags & (REQ_TRANS | REQ_CONJ)) return FEAT_TRANS;
    return FEAT_COPY;
}
// `cus` = compute units of the device (persistent grids fill them).
hipError_t launch_gather(const gather_args &a, const launch_cfg &c, int cus, hipStream_t stream);

// work[w] = the record of work index w for w in [descs[k].tile_base, +
// n_tiles), k < n.
hipError_t launch_work_records(const work_desc *descs, int32_t n, work_rec *work, hipStream_t stream);

// ------------------------------------------------ host helpers (g2g_host.cpp)

// Work descriptors of n plans for one multi-transform launch; tile_base is
// the running prefix of work items. Returns a g2g_status.
int internal_descs(int n, g2g_plan_s *const *plans, const void *const *alphas,
                   const void *const *betas, std::vector<work_desc> &out, uint32_t &total_tiles,
                   uint64_t &total_bytes);
// The work descriptor of one plan's request (work items numbered from
// tile_base) and the plan's launch configuration with the request's feature
// level. Allocates nothing. Returns a g2g_status.
int internal_desc(g2g_plan_s *p, const void *alpha, const void *beta, uint32_t tile_base, work_desc &out,
                  launch_cfg &cfg);
// Compute units of the device (persistent grids are sized from it).
int internal_grid_limit();
// Records the thread's last error message and returns code.
int internal_fail(int code, const char *fmt, ...);
// Validates the splits, owners and n_ranks of a layout (its blocks are not read).
int internal_check_grid(const char *what, const struct g2g_layout *l);
// The union of two split vectors covering the same range, as intervals
// tagged with the block of each vector they lie in (f: `fblk`, i: `iblk`).
int internal_union_spans(const std::vector<int32_t> &f, const std::vector<int32_t> &i, const char *dim,
                         std::vector<span> &out);

// Test-support kernels.
struct fill_block {
    uint64_t ptr;
    uint64_t ld;
    int32_t r0, r1, c0, c1;
};
hipError_t launch_fill(const fill_block *blocks, int32_t n_blocks, int dtype, uint64_t seed,
                       int64_t n_cols, hipStream_t stream);
// every element of every block times beta (rq.beta, rq.dtype, rq.esize)
hipError_t launch_scale(const fill_block *blocks, int32_t n_blocks, const request &rq, hipStream_t stream);
hipError_t launch_verify(const fill_block *blocks, int32_t n_blocks, int dtype, uint64_t seed,
                         int trans, int conj, int64_t a_n_cols, unsigned long long *bad,
                         hipStream_t stream);
// Streaming copy (kind 0), read (1) or write (2) of `bytes` (a multiple of
// 16; src and dst 16-byte aligned) with `grid` workgroups.
hipError_t launch_stream(int kind, const void *src, void *dst, uint64_t bytes, int grid, hipStream_t stream);

}  // namespace g2g
