// Index arithmetic of the fused gather, shared by the kernels (g2g_gather.hpp)
// and the host-side replay that checks it (tests/cpp/gather_replay.cpp):
// which columns of a piece a work item moves, how a piece is queued, which
// bytes each queue slot moves, and how the transpose tiles address their LDS
// and global memory. Everything here is __host__ __device__ and allocation
// free, so the replay runs exactly the arithmetic the GPU runs.
#pragma once

#include "g2g_internal.hpp"

namespace g2g {

__host__ __device__ __forceinline__ uint32_t imin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Work item t of a piece stream of `total` bytes cut into items of tb bytes,
// the items from split_from on into shorter ones of split_bytes (the last
// round of a launch, so that the workgroups finish closer together): its
// byte range [lo, hi). Unsplit streams have split_from = the item count.
__host__ __device__ __forceinline__ void item_range(uint64_t tb, uint32_t split_from, uint32_t split_bytes,
                                                    uint64_t total, uint32_t t, uint64_t &lo, uint64_t &hi) {
    if (t < split_from) {
        lo = (uint64_t)t * tb;
        hi = lo + tb;
    } else {
        lo = (uint64_t)split_from * tb + (uint64_t)(t - split_from) * split_bytes;
        hi = lo + split_bytes;
    }
    if (hi > total) hi = total;
}

__host__ __device__ __forceinline__ uint32_t ceil_log2(uint32_t v) {
    return v <= 1 ? 0u : 32u - (uint32_t)__builtin_clz(v - 1);
}

// n / d for a divisor with magic multiplier mul = div_magic(d) (Granlund &
// Montgomery, "Division by invariant integers using multiplication", 1994,
// fig. 4.1).
__host__ __device__ __forceinline__ uint32_t magic_div(uint32_t n, uint32_t d, uint32_t mul) {
    const uint32_t l = ceil_log2(d);
    const uint32_t t = (uint32_t)(((uint64_t)mul * n) >> 32);
    return l ? (t + ((n - t) >> 1)) >> (l - 1) : n;
}

// The piece where initial cell I meets final cell F along the union intervals
// R (rows) and C (columns); fr0/fc0 and ir0/ic0 are the first row and column
// of the final and initial cells (in the working orientation: for a
// transposed initial layout its storage rows run along final columns). The
// source pointer follows block::subblock (reference block.cpp:165-196).
// piece.req records the reference call that would scale it: a local
// transposed piece goes through copy_transpose_and_scale
// (communication_data.cpp:197-210), every other piece through
// copy2D_and_scale, as one loop when the piece is contiguous on both sides
// (the packed receive buffer has ld = rows; a local piece reads the initial
// block) and column by column otherwise (memory_utils.hpp:79-122).
__host__ __device__ __forceinline__ piece make_piece(const span &R, const span &C, const cell &F, const cell &I,
                                                     int32_t fr0, int32_t fc0, int32_t ir0, int32_t ic0, uint32_t es,
                                                     bool trans) {
    const uint32_t rows = (uint32_t)(R.end - R.start), cols = (uint32_t)(C.end - C.start);
    piece pc;
    pc.dst = F.ptr + ((uint64_t)(C.start - fc0) * F.ld + (uint64_t)(R.start - fr0)) * es;
    if (!trans)
        pc.src = I.ptr + ((uint64_t)(C.start - ic0) * I.ld + (uint64_t)(R.start - ir0)) * es;
    else
        pc.src = I.ptr + ((uint64_t)(R.start - ir0) * I.ld + (uint64_t)(C.start - ic0)) * es;
    pc.lds = (uint32_t)I.ld;
    pc.ldd = (uint32_t)F.ld;
    pc.rows = rows;
    pc.cols = cols;
    const bool local = I.owner == F.owner;
    uint32_t site;
    if (local && trans)
        site = PIECE_SITE_TRANS;
    else
        site = (F.ld == rows && (!local || I.ld == rows)) ? PIECE_SITE_FLAT : PIECE_SITE_COL;
    pc.req = site << PIECE_SITE_SHIFT;
    pc.vmul = div_magic((rows * es + 15) / 16);
    return pc;
}

// Final columns [c0, c0 + w) of a piece (fewer at its end): the destination
// advances by whole columns; so does the source of an untransposed piece,
// while a transposed piece's source storage rows run along final columns.
__host__ __device__ __forceinline__ piece chunk_piece(piece pc, uint32_t c0, uint32_t w, uint32_t es, bool trans) {
    const uint32_t nc = imin(w, pc.cols - c0);
    pc.src += trans ? (uint64_t)c0 * es : (uint64_t)c0 * pc.lds * es;
    pc.dst += (uint64_t)c0 * pc.ldd * es;
    pc.cols = nc;
    return pc;
}

// Final rows [r0, r0 + h) of an untransposed piece (fewer at its end): source
// and destination advance by r0 elements, the columns stay.
__host__ __device__ __forceinline__ piece row_chunk_piece(piece pc, uint32_t r0, uint32_t h, uint32_t es) {
    const uint32_t nr = imin(h, pc.rows - r0);
    pc.src += (uint64_t)r0 * es;
    pc.dst += (uint64_t)r0 * es;
    pc.rows = nr;
    pc.vmul = div_magic((nr * es + 15) / 16);
    return pc;
}

// Columns [c0, c1) of a piece that belong to the work item covering stream
// bytes [lo, hi), for a piece occupying stream bytes [ps, pe): the columns
// (groups of 256/es columns for transposed pieces) whose first byte lies in
// [lo, hi). Work items therefore partition every piece's columns.
__host__ __device__ __forceinline__ void piece_cols(uint32_t rows, uint32_t cols, uint32_t flags, uint32_t es,
                                                    uint64_t ps, uint64_t pe, uint64_t lo, uint64_t hi,
                                                    uint32_t &c0, uint32_t &c1) {
    c0 = 0;
    c1 = cols;
    if (ps < lo || pe > hi) {  // piece shared with neighbouring work items
        const uint64_t q = (flags & REQ_TRANS) ? 256u / es : 1u;
        const uint64_t qb = (uint64_t)rows * es * q;
        const uint64_t a0 = lo > ps ? ((lo - ps + qb - 1) / qb) * q : 0;
        const uint64_t a1 = ((hi - ps + qb - 1) / qb) * q;
        c0 = (uint32_t)(a0 < cols ? a0 : cols);
        c1 = (uint32_t)(a1 < cols ? a1 : cols);
    }
}

// How a piece's columns are moved.
enum : uint32_t {
    PK_NONE = 0,   // not moved by this launch (or no columns in the work item)
    PK_VEC = 1,    // copy queue: 16-byte slots down each column
    PK_ELEM = 2,   // element queue: small transposed pieces
    PK_DIRECT = 3  // by the whole workgroup: LDS transpose tiles, or the scaled path
};

// Transposed pieces whose final columns hold fewer bytes than this (counting
// at most 8 bytes an element: f32 below 128 rows, the other types below 64),
// or whose final column ranges are narrower (bytes), go to the element queue
// instead of LDS tiles.
#ifndef G2G_TRANS_ELEM_ROWB
#define G2G_TRANS_ELEM_ROWB 512
#endif
#ifndef G2G_TRANS_ELEM_COLB
#define G2G_TRANS_ELEM_COLB 256
#endif
constexpr uint32_t kTransElemRowBytes = G2G_TRANS_ELEM_ROWB;
constexpr uint32_t kTransElemColBytes = G2G_TRANS_ELEM_COLB;

// The move of a piece's columns [c0, c1) for a launch of feature level FEAT
// (FEAT_COPY kernels leave transposed and scaled pieces alone: the host
// never sends them one).
__host__ __device__ __forceinline__ uint32_t piece_kind(int feat, uint32_t flags, uint32_t rows, uint32_t c0,
                                                        uint32_t c1, uint32_t es) {
    if (c0 >= c1) return PK_NONE;
    if (feat == FEAT_COPY) return (flags & (REQ_TRANS | REQ_SCALE)) ? PK_NONE : PK_VEC;
    if (flags & REQ_SCALE) return PK_DIRECT;
    if (flags & REQ_TRANS)
        return ((uint64_t)rows * imin(es, 8u) < kTransElemRowBytes || (uint64_t)(c1 - c0) * es < kTransElemColBytes)
                   ? PK_ELEM : PK_DIRECT;
    return PK_VEC;
}

// Slots per column (copy queue: 16-byte vectors; element queue: groups of
// 16/es final rows, or rows for 16-byte elements) and groups of columns per
// queued piece (element queue: groups of 16/es final columns).
__host__ __device__ __forceinline__ uint32_t piece_vpc(uint32_t kind, uint32_t rows, uint32_t es) {
    if (kind == PK_VEC) return (rows * es + 15) >> 4;
    return es < 16 ? (rows * es + 15) >> 4 : rows;
}
__host__ __device__ __forceinline__ uint32_t piece_groups(uint32_t kind, uint32_t ncols, uint32_t es) {
    return kind == PK_ELEM ? (ncols + (16u / es) - 1) / (16u / es) : ncols;
}

// Runs queued for one flat copy.
// SEG_VEC: columns of colb bytes starting at s (source, leading dimension
//   lds_b bytes) and d (destination, ldd_b), each cut into vpc = ceil(colb/16)
//   16-byte slots; when colb is not a multiple of 16 the last slot is the
//   column's last 16 bytes (overlapping the slot before it), or, in a column
//   shorter than 16 bytes, its 1-3 dwords.
// SEG_TRANS / SEG_CONJ (a queue of their own): a small transposed piece of
//   ncols final columns: destination element (i, c) at d + c*ldd_b + i*es
//   comes from s + i*lds_b + c*es, its imaginary sign bit flipped for
//   SEG_CONJ. One slot per group of 16/es final columns and, for 4- and 8-byte
//   elements, group of 16/es final rows (vpc = ceil(rows*es/16)), or final
//   row (16-byte elements, vpc = final rows).
// slot0 numbers the segment's first slot in its queue; the entry after the
// last one of a queue holds only slot0 = the queue's slot count (sentinel).
enum : uint32_t { SEG_VEC = 0, SEG_TRANS = 1, SEG_CONJ = 2 };
struct seg_entry {
    uint64_t s, d;
    uint64_t lds_b, ldd_b;
    uint32_t colb;
    uint32_t vpc;
    uint32_t slot0;
    uint32_t vmul;  // division magic for vpc
    uint32_t kind;
    uint32_t ncols;  // element segments: final columns of the piece
};
constexpr int kMaxSeg = 64;

// The queue entry of a piece's columns [c0, c0 + ncols) (kind PK_VEC or PK_ELEM).
__host__ __device__ __forceinline__ seg_entry make_seg(const piece &pc, uint32_t kind, uint32_t flags, uint32_t es,
                                                       uint32_t c0, uint32_t ncols, uint32_t vpc, uint32_t slot0) {
    const bool v = kind == PK_VEC;
    const uint64_t ldd_b = (uint64_t)pc.ldd * es, lds_b = (uint64_t)pc.lds * es;
    seg_entry e;
    e.s = pc.src + (uint64_t)c0 * (v ? lds_b : es);
    e.d = pc.dst + (uint64_t)c0 * ldd_b;
    e.lds_b = lds_b;
    e.ldd_b = ldd_b;
    e.colb = pc.rows * es;
    e.vpc = vpc;
    e.slot0 = slot0;
    e.vmul = v ? pc.vmul : div_magic(vpc);
    e.kind = v ? SEG_VEC : (flags & REQ_CONJ) ? SEG_CONJ : SEG_TRANS;
    e.ncols = ncols;
    return e;
}

// Copy-queue slot rel of a segment: `dwords` dwords (4, or 1-3 in a column
// shorter than 16 bytes) from byte so of the source to byte dof of the
// destination, both relative to the segment's s and d.
__host__ __device__ __forceinline__ void vec_slot(const seg_entry &e, uint32_t rel, uint64_t &so, uint64_t &dof,
                                                  uint32_t &dwords) {
    const uint32_t c = magic_div(rel, e.vpc, e.vmul);
    const uint32_t v = rel - c * e.vpc;
    // column byte offset of the slot (v*16 < colb: no wrap)
    uint32_t off = v * 16;
    if (e.colb >= 16 && off > e.colb - 16) off = e.colb - 16;
    so = (uint64_t)c * e.lds_b + off;
    dof = (uint64_t)c * e.ldd_b + off;
    dwords = e.colb < 16 ? e.colb >> 2 : 4;
}

// Element-queue register tile of ES-byte elements (4 or 8), VE = 16/ES: final
// columns [c, c + nc) by final rows [i0, i0 + nr), nc, nr <= VE.
template <uint32_t ES>
__host__ __device__ __forceinline__ void tile_slot(const seg_entry &e, uint32_t rel, uint32_t &c, uint32_t &i0,
                                                   uint32_t &nc, uint32_t &nr) {
    constexpr uint32_t VE = 16 / ES;
    const uint32_t g = magic_div(rel, e.vpc, e.vmul);
    const uint32_t h = rel - g * e.vpc;
    c = g * VE;
    i0 = h * VE;
    nc = imin(VE, e.ncols - c);
    nr = imin(VE, e.colb / ES - i0);
}

// Element-queue slot of 16-byte elements: final row v of final columns
// [c, c + cnt), cnt <= 1.
template <uint32_t ES>
__host__ __device__ __forceinline__ void elem_slot(const seg_entry &e, uint32_t rel, uint32_t &c, uint32_t &v,
                                                   uint32_t &cnt) {
    constexpr uint32_t VE = 16 / ES;
    const uint32_t g = magic_div(rel, e.vpc, e.vmul);
    v = rel - g * e.vpc;
    c = g * VE;
    cnt = imin(VE, e.ncols - c);
}

// ------------------------------------------------------ transpose tiles
//
// A tile is TC final rows by TR = kTileBytes / (TC*ES) final columns; LDS
// holds it in storage order with rows of LDW = TR + 1 elements. Thread q's
// k-th vector: phase 1 (global -> LDS) reads VE consecutive elements of
// source storage column c starting at element r0 (final row c, final columns
// r0..), phase 2 (LDS -> global) writes VE consecutive final rows starting at
// w*VE of final column r.
constexpr uint32_t kTileBytes = 16384;                  // data bytes of one transpose tile
constexpr uint32_t kSmemBytes = kTileBytes + 64 * 16;   // plus one padding element per tile row

template <uint32_t ES, uint32_t TC, uint32_t THREADS>
struct tile_geom {
    static constexpr uint32_t VE = 16 / ES;                  // elements per vector
    static constexpr uint32_t TR = kTileBytes / (TC * ES);   // final columns per tile
    static constexpr uint32_t LDW = TR + 1;                  // LDS row length in elements
    static constexpr uint32_t RV = TR / VE;                  // vectors per source run
    static constexpr uint32_t RVL = __builtin_ctz(RV);
    static constexpr uint32_t W = TC * ES / 16;              // vectors per final column of a tile
    static constexpr uint32_t WL = __builtin_ctz(W);
    static constexpr uint32_t PER = kTileBytes / 16 / THREADS;  // vectors per thread per phase
    static constexpr uint32_t TILE_ELEMS = kSmemBytes / ES;     // LDS area of one tile, in elements
    static_assert(TC * LDW * ES <= kSmemBytes, "transpose tile exceeds its LDS");
    // phase 1: source storage column c (final row) and first element r0 (final column)
    __host__ __device__ static void load_index(uint32_t q, uint32_t &c, uint32_t &r0) {
        c = q >> RVL;
        r0 = (q & (RV - 1)) * VE;
    }
    // phase 2: final column r and vector w of it (final rows w*VE ..)
    __host__ __device__ static void store_index(uint32_t q, uint32_t &r, uint32_t &w) {
        r = q >> WL;
        w = q & (W - 1);
    }
};

// Transpose tiles moved per pass of the full-tile path: enough that every
// thread keeps 128 bytes of loads in flight (a transposing workgroup of 512
// threads shares a CU with only one other, so one 16 KiB tile per pass left a
// CU with 32 KiB of loads in flight). The pass's LDS (4 x 17 KiB for 512
// threads, 2 x 17 KiB for 256) still lets 2 (512) or 3 (256) workgroups share
// a CU.
template <int THREADS>
__host__ __device__ constexpr uint32_t trans_tiles() {
    return THREADS >= 512 ? 4u : 2u;
}

}  // namespace g2g
