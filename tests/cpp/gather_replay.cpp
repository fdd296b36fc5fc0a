// gather_replay.cpp -- TEST INFRASTRUCTURE: the fused gather's indexing run on
// the host, for layouts the GPU tests use, with every access checked.
//
// The replay plans a layout pair as g2g_plan_create does (union of the
// splits, then make_piece per union cell, optionally chunk_piece), cuts the
// piece stream into work items as the library does, and then replays
// process_tile for every work item: the classification of each piece
// (piece_cols, piece_kind, piece_vpc), the queue entries (make_seg, slot
// numbering, sentinels), each thread's walk over the queue, every slot of
// the copy queue (vec_slot), the element queue (tile_slot, elem_slot), the
// LDS transpose tiles (tile_geom: both phases, every thread and vector, the
// full-tile path with its 32-bit offsets and the bounds-checked strips) and
// the scaled path. These are the same __host__ __device__ functions the
// kernels call (grid2grid_amd/csrc/g2g_index.hpp).
//
// Every move is checked against the transform's definition, computed
// independently from the layouts: final element (i, j) must receive initial
// element (i, j) ('N') or (j, i) ('T'/'C'); nothing outside the final blocks
// (ld padding, gaps) may be written; every final element is written, and only
// the copy queue's overlapping last slot of a ragged column may write an
// element twice (with the same source). LDS reads must hit cells written by
// the same tile, inside the workgroup's LDS. Queue walks must stay within the
// queue and its sentinel.
#include "g2g_index.hpp"
#include "grid2grid_amd.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace g2g;

namespace {

// one (element size, tile height, workgroup size) transpose family
template <uint32_t ES, uint32_t TC, uint32_t TH>
struct tg_t {
    using G = tile_geom<ES, TC, TH>;
    static constexpr uint32_t TC_ = TC;
};

struct wl {  // a layout in the working orientation
    std::vector<int32_t> rs, cs;
    int32_t nbr = 0, nbc = 0;
    std::vector<cell> cells;
};

bool load(const g2g_layout *l, uint32_t es, bool transpose, wl &w) {
    const int32_t br = l->n_rows_split - 1, bc = l->n_cols_split - 1;
    std::vector<cell> sc((size_t)br * bc);
    std::vector<char> seen(sc.size(), 0);
    for (int64_t k = 0; k < (int64_t)sc.size(); ++k) sc[k].owner = l->owners[k];
    for (int32_t b = 0; b < l->n_blocks; ++b) {
        const g2g_block &x = l->blocks[b];
        const int64_t k = (int64_t)x.row * bc + x.col;
        sc[k].ptr = (uint64_t)(uintptr_t)x.data;
        sc[k].ld = (uint64_t)x.stride;
        seen[k] = 1;
    }
    for (char c : seen)
        if (!c) return false;
    if (!transpose) {
        w.rs.assign(l->rows_split, l->rows_split + l->n_rows_split);
        w.cs.assign(l->cols_split, l->cols_split + l->n_cols_split);
        w.nbr = br;
        w.nbc = bc;
        w.cells = sc;
    } else {
        w.rs.assign(l->cols_split, l->cols_split + l->n_cols_split);
        w.cs.assign(l->rows_split, l->rows_split + l->n_rows_split);
        w.nbr = bc;
        w.nbc = br;
        w.cells.resize(sc.size());
        for (int32_t i = 0; i < br; ++i)
            for (int32_t j = 0; j < bc; ++j) w.cells[(size_t)j * br + i] = sc[(size_t)i * bc + j];
    }
    return true;
}

// the union of two split vectors as intervals tagged with their blocks
std::vector<span> unite(const std::vector<int32_t> &f, const std::vector<int32_t> &i) {
    std::vector<span> out;
    size_t a = 0, b = 0;
    int32_t pos = f.front();
    while (pos < f.back()) {
        const int32_t next = std::min(f[a + 1], i[b + 1]);
        out.push_back(span{pos, next, (int32_t)a, (int32_t)b});
        pos = next;
        if (f[a + 1] == pos) ++a;
        if (i[b + 1] == pos) ++b;
    }
    return out;
}

// Element maps of a layout: for every element of the (storage-orientation)
// layout's blocks, its address -> global key row * n_cols + col.
struct elem_map {
    uint64_t base = 0, span_bytes = 0;
    uint32_t es = 0;
    int64_t n_cols = 0;
    std::vector<int64_t> key;  // per element slot of [base, base + span_bytes), -1 outside blocks
    void build(const g2g_layout *l, uint32_t e) {
        es = e;
        n_cols = l->cols_split[l->n_cols_split - 1];
        uint64_t lo = ~0ull, hi = 0;
        for (int32_t b = 0; b < l->n_blocks; ++b) {
            const g2g_block &x = l->blocks[b];
            const int64_t rows = l->rows_split[x.row + 1] - l->rows_split[x.row];
            const int64_t cols = l->cols_split[x.col + 1] - l->cols_split[x.col];
            if (!rows || !cols) continue;
            const uint64_t p = (uint64_t)(uintptr_t)x.data;
            lo = std::min(lo, p);
            hi = std::max(hi, p + ((uint64_t)(cols - 1) * x.stride + rows) * es);
        }
        if (lo > hi) lo = hi = 0;
        base = lo;
        span_bytes = hi - lo;
        key.assign(span_bytes / es + 1, -1);
        for (int32_t b = 0; b < l->n_blocks; ++b) {
            const g2g_block &x = l->blocks[b];
            const int32_t r0 = l->rows_split[x.row], r1 = l->rows_split[x.row + 1];
            const int32_t c0 = l->cols_split[x.col], c1 = l->cols_split[x.col + 1];
            for (int32_t j = c0; j < c1; ++j)
                for (int32_t i = r0; i < r1; ++i) {
                    const uint64_t a = (uint64_t)(uintptr_t)x.data + ((uint64_t)(j - c0) * x.stride + (i - r0)) * es;
                    key[(a - base) / es] = (int64_t)i * n_cols + j;
                }
        }
    }
    // the element's key, or -1 for an address outside every block (or misaligned)
    int64_t at(uint64_t a) const {
        if (a < base || a >= base + span_bytes || (a - base) % es) return -1;
        return key[(a - base) / es];
    }
};

struct replay {
    uint32_t es = 0;
    bool trans = false;
    int64_t m = 0, n = 0;  // final matrix
    elem_map ini, fin;
    std::vector<uint64_t> ini_addr;  // per initial key: the element's address
    std::vector<uint8_t> writes;     // per final element slot: times written
    std::string msg;
    int64_t errors = 0;
    int64_t overlap_writes = 0;

    void error(const char *fmt, ...) {
        ++errors;
        if (errors > 8) return;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        msg += buf;
        msg += "\n";
    }

    // the final element at dst receives the initial element at src
    void move(uint64_t src, uint64_t dst, bool may_repeat, const char *path) {
        const int64_t fk = fin.at(dst);
        if (fk < 0) {
            error("%s: write at %#llx outside the final blocks (or misaligned)", path, (unsigned long long)dst);
            return;
        }
        const int64_t i = fk / fin.n_cols, j = fk % fin.n_cols;
        const int64_t ik = trans ? j * ini.n_cols + i : i * ini.n_cols + j;
        if (ik < 0 || ik >= (int64_t)ini_addr.size() || ini_addr[ik] != src) {
            error("%s: final (%lld, %lld) at %#llx receives %#llx, not its source %#llx", path, (long long)i,
                  (long long)j, (unsigned long long)dst, (unsigned long long)src,
                  (unsigned long long)(ik >= 0 && ik < (int64_t)ini_addr.size() ? ini_addr[ik] : 0));
            return;
        }
        uint8_t &w = writes[(dst - fin.base) / es];
        if (w && !may_repeat) error("%s: final (%lld, %lld) written twice", path, (long long)i, (long long)j);
        if (w) ++overlap_writes;
        if (w < 255) ++w;
    }
};

}  // namespace

extern "C" {

// Replays the fused gather of (initial -> final, dtype, trans) with work items
// of tile_bytes, workgroups of `threads`, kernel feature level `feat`
// (FEAT_COPY 0, FEAT_TRANS 1, FEAT_ALL 2), column chunks of chunk_w columns
// cut again into chunks of chunk_h rows (untransposed; 0: whole columns), in
// stream order `order` (G2G_ORDER_UNION or G2G_ORDER_DESTINATION)
// (0: the pieces), scaled = 1 for a scaled request (the scaled path moves
// every piece). Returns the number of errors; the first ones are described in
// msg. stats (may be NULL): work items, copy slots, element slots, tile
// vectors, scaled elements, repeated writes.
//
// plant != 0 plants one typical indexing mistake into the replayed arithmetic
// (to show that the checks catch it): 1 the copy queue's last slot not pulled
// back into a ragged column, 2 source and destination leading dimensions
// swapped in the copy queue, 3 an element-queue tile read along the wrong
// axis, 4 a transpose tile's phase 2 one final column off, 5 the LDS row
// padding dropped in phase 2, 6 work items reaching one column into the next.
int replay_gather(const g2g_layout *initial, const g2g_layout *final_, int dtype, char trans, int scaled,
                  int64_t tile_bytes, int threads, int feat, uint32_t chunk_w, int order, uint32_t chunk_h,
                  uint32_t split, int plant, char *msg, int msg_len,
                  int64_t *stats) {
    replay R;
    const uint32_t es = dtype == G2G_FLOAT32 ? 4 : dtype == G2G_COMPLEX128 ? 16 : 8;
    R.es = es;
    R.trans = trans != 'N';
    wl wi, wf;
    if (!load(initial, es, R.trans, wi) || !load(final_, es, false, wf)) {
        std::snprintf(msg, msg_len, "a cell has no block");
        return 1;
    }
    R.ini.build(initial, es);
    R.fin.build(final_, es);
    R.ini_addr.assign((size_t)(initial->rows_split[initial->n_rows_split - 1]) * R.ini.n_cols + 1, ~0ull);
    for (size_t k = 0; k < R.ini.key.size(); ++k)
        if (R.ini.key[k] >= 0) R.ini_addr[R.ini.key[k]] = R.ini.base + k * es;
    R.writes.assign(R.fin.key.size(), 0);

    // planning (g2g_plan_create, k_plan_pieces, k_chunk_pieces)
    const std::vector<span> ur = unite(wf.rs, wi.rs), uc = unite(wf.cs, wi.cs);
    const int64_t n_ur = (int64_t)ur.size();
    std::vector<piece> pieces;
    for (size_t uj = 0; uj < uc.size(); ++uj)
        for (size_t ui = 0; ui < ur.size(); ++ui) {
            const span &r = ur[ui], &c = uc[uj];
            pieces.push_back(make_piece(r, c, wf.cells[(size_t)r.fblk * wf.nbc + c.fblk],
                                        wi.cells[(size_t)r.iblk * wi.nbc + c.iblk], wf.rs[r.fblk], wf.cs[c.fblk],
                                        wi.rs[r.iblk], wi.cs[c.iblk], es, R.trans));
        }
    if (chunk_w) {
        std::vector<piece> ch;
        for (size_t uj = 0; uj < uc.size(); ++uj)
            for (int32_t c0 = 0; c0 < uc[uj].end - uc[uj].start; c0 += (int32_t)chunk_w)
                for (int64_t ui = 0; ui < n_ur; ++ui)
                    ch.push_back(chunk_piece(pieces[uj * n_ur + ui], (uint32_t)c0, chunk_w, es, R.trans));
        pieces.swap(ch);
    }
    // row chunks of h rows (untransposed streams; k_row_chunks)
    if (chunk_h && !R.trans) {
        std::vector<piece> ch;
        for (const piece &pc : pieces)
            for (uint32_t r0 = 0; r0 < pc.rows; r0 += chunk_h) ch.push_back(row_chunk_piece(pc, r0, chunk_h, es));
        pieces.swap(ch);
    }
    // destination order (select_stream, sort_pieces_by_dst): a stable sort by
    // destination address
    if (order == G2G_ORDER_DESTINATION)
        std::stable_sort(pieces.begin(), pieces.end(), [](const piece &a, const piece &b) { return a.dst < b.dst; });
    const int64_t np = (int64_t)pieces.size();
    std::vector<uint64_t> pend(np);
    uint64_t total = 0;
    for (int64_t p = 0; p < np; ++p) pend[p] = total += (uint64_t)pieces[p].rows * pieces[p].cols * es;
    // work items (retile): tb bytes each; with a tail split, the last
    // min(total / 2, 256 CUs x 4 x tb) bytes after whole items in items of tb / split
    const uint64_t tb = (uint64_t)tile_bytes;
    uint64_t full = (total + tb - 1) / tb, n_tiles = full, sb = tb;
    if (split) {
        sb = tb / split;
        const uint64_t tail = std::min(total / 2, (uint64_t)256 * 4 * tb);
        full = (total - tail) / tb;
        n_tiles = full + (total - full * tb + sb - 1) / sb;
    }
    // first piece of every work item (k_first_piece): the first p with
    // pend[p] > lo, found by walking the pieces alongside the items
    std::vector<uint32_t> first(n_tiles, 0);
    {
        int64_t p = 0;
        for (uint64_t t = 0; t < n_tiles; ++t) {
            uint64_t lo, hi;
            item_range(tb, (uint32_t)full, (uint32_t)sb, total, (uint32_t)t, lo, hi);
            while (p < np - 1 && pend[p] <= lo) ++p;
            first[t] = (uint32_t)p;
        }
    }
    request rq;
    std::memset(&rq, 0, sizeof rq);
    rq.flags = (R.trans ? REQ_TRANS : 0u) | (trans == 'C' && (dtype == G2G_COMPLEX64 || dtype == G2G_COMPLEX128) ? REQ_CONJ : 0u) |
               (scaled ? REQ_SCALE : 0u);
    rq.esize = es;

    int64_t st_vec = 0, st_elem = 0, st_tile = 0, st_scaled = 0;
    std::vector<seg_entry> segs(kMaxSeg + 1), tsegs(kMaxSeg + 1);
    const uint32_t tile_areas = threads >= 512 ? trans_tiles<512>() : trans_tiles<256>();

    uint64_t covered = 0;  // the items' ranges must tile the stream in order
    for (uint64_t t = 0; t < n_tiles; ++t) {
        uint64_t lo, hi;
        item_range(tb, (uint32_t)full, (uint32_t)sb, total, (uint32_t)t, lo, hi);
        if (lo != covered || hi <= lo) {
            std::snprintf(msg, msg_len, "work item %llu covers [%llu, %llu) after %llu", (unsigned long long)t,
                          (unsigned long long)lo, (unsigned long long)hi, (unsigned long long)covered);
            return 1;
        }
        covered = hi;
        const uint32_t n = (t + 1 < n_tiles ? first[t + 1] : (uint32_t)np - 1) + 1;  // last_piece + 1
        uint32_t p = first[t];
        for (;;) {
            if (p >= n) break;
            uint32_t kind[64], c0[64], c1[64], vpc[64];
            bool in[64];
            for (uint32_t lane = 0; lane < 64; ++lane) {
                const uint32_t q = p + lane;
                kind[lane] = PK_NONE;
                c0[lane] = c1[lane] = vpc[lane] = 0;
                in[lane] = q < n;
                if (in[lane]) {
                    const piece &pc = pieces[q];
                    const uint64_t ps = q ? pend[q - 1] : 0, pe = pend[q];
                    in[lane] = ps < hi;
                    piece_cols(pc.rows, pc.cols, rq.flags, es, ps, pe, lo, hi, c0[lane], c1[lane]);
                    if (plant == 6 && c1[lane] < pc.cols) ++c1[lane];
                    if (in[lane]) {
                        kind[lane] = piece_kind(feat, rq.flags, pc.rows, c0[lane], c1[lane], es);
                        vpc[lane] = piece_vpc(kind[lane], pc.rows, es);
                    }
                }
            }
            uint32_t end = 64;
            for (uint32_t lane = 0; lane < 64; ++lane)
                if (!in[lane]) {
                    end = lane;
                    break;
                }
            // in-range pieces come first: none may follow an out-of-range one
            for (uint32_t lane = end; lane < 64; ++lane)
                if (in[lane]) R.error("work item %llu: piece %u inside after one outside", (unsigned long long)t, p + lane);
            // queues, in lane order
            uint32_t nseg = 0, ntseg = 0, total_v = 0, total_e = 0;
            for (uint32_t lane = 0; lane < 64; ++lane) {
                if (kind[lane] != PK_VEC && kind[lane] != PK_ELEM) continue;
                const uint32_t slots = piece_groups(kind[lane], c1[lane] - c0[lane], es) * vpc[lane];
                if (kind[lane] == PK_VEC) {
                    segs[nseg++] = make_seg(pieces[p + lane], kind[lane], rq.flags, es, c0[lane], c1[lane] - c0[lane],
                                            vpc[lane], total_v);
                    total_v += slots;
                } else {
                    tsegs[ntseg++] = make_seg(pieces[p + lane], kind[lane], rq.flags, es, c0[lane],
                                              c1[lane] - c0[lane], vpc[lane], total_e);
                    total_e += slots;
                }
            }
            if (nseg > (uint32_t)kMaxSeg || ntseg > (uint32_t)kMaxSeg) R.error("queue overflow");
            segs[nseg].slot0 = total_v;
            tsegs[ntseg].slot0 = total_e;
            // each thread's walk (flat_copy and the element queue): it starts
            // at entry 0, reads entry j + 1's slot0 and moves forward
            auto walk = [&](const std::vector<seg_entry> &q, uint32_t nq, uint32_t total_q, uint32_t stride,
                            auto &&slot) {
                if (!nq) return;
                for (uint32_t tid = 0; tid < (uint32_t)threads; ++tid) {
                    uint32_t j = 0, next0 = q[1].slot0;
                    for (uint32_t k = tid; k < total_q; k += stride) {
                        while (next0 <= k) {
                            ++j;
                            if (j + 1 > nq) {
                                R.error("queue walk past the sentinel (slot %u of %u)", k, total_q);
                                return;
                            }
                            next0 = q[j + 1].slot0;
                        }
                        if (k < q[j].slot0) R.error("slot %u before its entry's first slot %u", k, q[j].slot0);
                        slot(q[j], k - q[j].slot0);
                    }
                }
            };
            const uint32_t ub_threads = (uint32_t)threads;  // slots are dealt round-robin, any unroll
            walk(segs, nseg, total_v, ub_threads, [&](const seg_entry &e, uint32_t rel) {
                uint64_t so, dof;
                uint32_t dwords;
                vec_slot(e, rel, so, dof, dwords);
                if (plant == 1) {
                    const uint32_t c = magic_div(rel, e.vpc, e.vmul), v = rel - c * e.vpc;
                    so = (uint64_t)c * e.lds_b + v * 16;
                    dof = (uint64_t)c * e.ldd_b + v * 16;
                }
                if (plant == 2) {
                    const uint32_t c = magic_div(rel, e.vpc, e.vmul);
                    so += (uint64_t)c * (e.ldd_b - e.lds_b);
                    dof += (uint64_t)c * (e.lds_b - e.ldd_b);
                }
                ++st_vec;
                if (dwords < 1 || dwords > 4 || (uint64_t)dwords * 4 > std::max<uint64_t>(e.colb, 16)) {
                    R.error("copy slot of %u dwords in a %u-byte column", dwords, e.colb);
                    return;
                }
                const bool ragged = e.colb % 16 != 0;
                for (uint32_t b = 0; b < dwords * 4; b += es) R.move(e.s + so + b, e.d + dof + b, ragged, "copy queue");
            });
            if (ntseg && feat != FEAT_COPY) {
                if (es == 4 || es == 8) {
                    walk(tsegs, ntseg, total_e, ub_threads, [&](const seg_entry &e, uint32_t rel) {
                        uint32_t c, i0, nc, nr;
                        if (es == 4) tile_slot<4>(e, rel, c, i0, nc, nr);
                        else tile_slot<8>(e, rel, c, i0, nc, nr);
                        ++st_elem;
                        // VE reads of final row i0 + y (source storage column), VE writes of final column c + x
                        for (uint32_t y = 0; y < nr; ++y)
                            for (uint32_t x = 0; x < nc; ++x)
                                R.move(plant == 3 ? e.s + (uint64_t)(c + x) * e.lds_b + (uint64_t)(i0 + y) * es
                                                  : e.s + (uint64_t)(i0 + y) * e.lds_b + (uint64_t)(c + x) * es,
                                       e.d + (uint64_t)(c + x) * e.ldd_b + (uint64_t)(i0 + y) * es, false,
                                       "element queue tile");
                    });
                } else {
                    walk(tsegs, ntseg, total_e, ub_threads, [&](const seg_entry &e, uint32_t rel) {
                        uint32_t c, v, cnt;
                        elem_slot<16>(e, rel, c, v, cnt);
                        ++st_elem;
                        for (uint32_t x = 0; x < cnt; ++x)
                            R.move(e.s + (uint64_t)v * e.lds_b + (uint64_t)(c + x) * es,
                                   e.d + (uint64_t)(c + x) * e.ldd_b + (uint64_t)v * es, false, "element queue row");
                    });
                }
            }
            // pieces moved by the whole workgroup
            for (uint32_t lane = 0; lane < 64 && feat != FEAT_COPY; ++lane) {
                if (kind[lane] != PK_DIRECT) continue;
                const piece &pc = pieces[p + lane];
                const uint32_t d0 = c0[lane], d1 = c1[lane];
                const uint64_t lds_b = (uint64_t)pc.lds * es, ldd_b = (uint64_t)pc.ldd * es;
                if (rq.flags & REQ_SCALE) {  // copy_scaled: element (r, c) of columns [d0, d1)
                    if (feat != FEAT_ALL) {
                        R.error("scaled piece in a kernel without scaling");
                        continue;
                    }
                    for (uint32_t c = d0; c < d1; ++c)
                        for (uint32_t r = 0; r < pc.rows; ++r) {
                            const uint64_t so = R.trans ? (uint64_t)r * lds_b + (uint64_t)c * es
                                                        : (uint64_t)c * lds_b + (uint64_t)r * es;
                            R.move(pc.src + so, pc.dst + (uint64_t)c * ldd_b + (uint64_t)r * es, false, "scaled");
                            ++st_scaled;
                        }
                    continue;
                }
                // transpose_call -> transpose_tc -> full tiles and strips
                const uint64_t s = pc.src + (uint64_t)d0 * es, d = pc.dst + (uint64_t)d0 * ldd_b;
                const uint32_t rows = pc.rows, ncols = d1 - d0;
                auto run_tc = [&](auto tg) {
                    using TG = decltype(tg);
                    constexpr uint32_t TC = TG::TC_, TR = TG::G::TR, LDW = TG::G::LDW, VE = TG::G::VE;
                    const uint32_t per = kTileBytes / 16 / (uint32_t)threads;
                    const uint32_t lds_elems = tile_areas * kSmemBytes / es;
                    std::vector<int64_t> lds(lds_elems);
                    // one tile through LDS area `area`: rows [ic, ic + TC) x cols [jc, jc + TR)
                    // of the sub-piece at (sp, dp) of rr rows and cc columns; full = no bounds checks
                    auto tile = [&](uint64_t sp, uint64_t dp, uint32_t ic, uint32_t jc, uint32_t rr, uint32_t cc,
                                    uint32_t area, bool full) {
                        std::fill(lds.begin(), lds.end(), -1);
                        const uint32_t abase = area * (kSmemBytes / es);
                        for (uint32_t tid = 0; tid < (uint32_t)threads; ++tid)
                            for (uint32_t k = 0; k < per; ++k) {
                                uint32_t c, r0;
                                TG::G::load_index(tid + k * threads, c, r0);
                                const uint32_t i = ic + c, j0 = jc + r0;
                                uint32_t cnt;
                                uint64_t src;
                                if (full) {  // 32-bit offsets inside the tile, as the kernel keeps them
                                    const uint32_t so = c * (uint32_t)lds_b + r0 * es;
                                    src = sp + ic * lds_b + (uint64_t)jc * es + so;
                                    cnt = VE;
                                } else {
                                    cnt = (i < rr && j0 < cc) ? imin(VE, cc - j0) : 0u;
                                    src = sp + (uint64_t)i * lds_b + (uint64_t)j0 * es;
                                }
                                for (uint32_t e = 0; e < cnt; ++e) {
                                    const uint64_t li = abase + c * LDW + r0 + e;
                                    if (li >= lds_elems || c * LDW + r0 + e >= kSmemBytes / es)
                                        R.error("LDS write %llu outside the tile area", (unsigned long long)li);
                                    else
                                        lds[li] = (int64_t)(src + e * es);
                                }
                            }
                        for (uint32_t tid = 0; tid < (uint32_t)threads; ++tid)
                            for (uint32_t k = 0; k < per; ++k) {
                                uint32_t r, w;
                                TG::G::store_index(tid + k * threads, r, w);
                                const uint32_t rw = plant == 4 ? (r + 1) % TR : r;  // the written column
                                const uint32_t j = jc + rw, i0 = ic + w * VE;
                                uint32_t cnt;
                                uint64_t dst;
                                if (full) {
                                    const uint32_t dof = rw * (uint32_t)ldd_b + w * VE * es;
                                    dst = dp + jc * ldd_b + (uint64_t)ic * es + dof;
                                    cnt = VE;
                                } else {
                                    cnt = (j < cc && i0 < rr) ? imin(VE, rr - i0) : 0u;
                                    dst = dp + (uint64_t)j * ldd_b + (uint64_t)i0 * es;
                                }
                                for (uint32_t e = 0; e < cnt; ++e) {
                                    const uint64_t li = abase + (w * VE + e) * (plant == 5 ? LDW - 1 : LDW) + r;
                                    if (li >= lds_elems || lds[li] < 0) {
                                        R.error("LDS read %llu of a cell this tile did not write", (unsigned long long)li);
                                        continue;
                                    }
                                    R.move((uint64_t)lds[li], dst + e * es, false, "transpose tile");
                                    ++st_tile;
                                }
                            }
                    };
                    const bool small_ld = (uint64_t)TC * lds_b < (1ull << 32) && (uint64_t)TR * ldd_b < (1ull << 32);
                    const uint32_t rf = small_ld ? rows / TC * TC : 0u, cf = ncols / TR * TR;
                    if (rf && cf) {  // copy_transposed_full: G tiles a pass, row tiles first
                        const uint32_t n_ic = rf / TC, n_t = n_ic * (cf / TR);
                        for (uint32_t tt = 0; tt < n_t; ++tt)
                            tile(s, d, (tt % n_ic) * TC, (tt / n_ic) * TR, rf, cf, tt % tile_areas, true);
                    }
                    auto strip = [&](uint64_t sp, uint64_t dp, uint32_t rr, uint32_t cc) {  // copy_transposed
                        const uint32_t n_ic = (rr + TC - 1) / TC, n_t = n_ic * ((cc + TR - 1) / TR);
                        for (uint32_t tt = 0; tt < n_t; ++tt)
                            tile(sp, dp, (tt % n_ic) * TC, (tt / n_ic) * TR, rr, cc, 0, false);
                    };
                    if (rf < rows) strip(s + rf * lds_b, d + (uint64_t)rf * es, rows - rf, ncols);
                    if (rf && cf < ncols) strip(s + (uint64_t)cf * es, d + cf * ldd_b, rf, ncols - cf);
                };
                auto by_tc = [&](auto es_tag) {
                    constexpr uint32_t ES = decltype(es_tag)::value;
                    if (rows > 32) {
                        if (threads >= 512) run_tc(tg_t<ES, 64, 512>{}); else run_tc(tg_t<ES, 64, 256>{});
                    } else if (rows > 16) {
                        if (threads >= 512) run_tc(tg_t<ES, 32, 512>{}); else run_tc(tg_t<ES, 32, 256>{});
                    } else {
                        if (threads >= 512) run_tc(tg_t<ES, 16, 512>{}); else run_tc(tg_t<ES, 16, 256>{});
                    }
                };
                if (es == 4) by_tc(std::integral_constant<uint32_t, 4>{});
                else if (es == 8) by_tc(std::integral_constant<uint32_t, 8>{});
                else by_tc(std::integral_constant<uint32_t, 16>{});
            }
            if (end < 64) break;
            p += 64;
        }
    }
    if (covered != total)
        R.error("work items cover %llu of %llu bytes", (unsigned long long)covered, (unsigned long long)total);
    // every final element written
    int64_t unwritten = 0;
    for (size_t k = 0; k < R.fin.key.size(); ++k)
        if (R.fin.key[k] >= 0 && !R.writes[k] && R.fin.key[k] >= 0) ++unwritten;
    if (unwritten) R.error("%lld final elements never written", (long long)unwritten);
    if (stats) {
        stats[0] = (int64_t)n_tiles;
        stats[1] = st_vec;
        stats[2] = st_elem;
        stats[3] = st_tile;
        stats[4] = st_scaled;
        stats[5] = R.overlap_writes;
    }
    std::snprintf(msg, msg_len, "%s", R.msg.c_str());
    return (int)std::min<int64_t>(R.errors, 1 << 30);
}

}  // extern "C"
