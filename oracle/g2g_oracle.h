/*
 * g2g_oracle.h -- CPU restatement of grid2grid::transform, for testing only.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or as the
 * timed CPU baseline. The product (grid2grid_amd, libgrid2grid_amd.so) never
 * links, loads or calls it.
 *
 * Every function restates the reference algorithm step by step and cites the
 * reference file:line it follows (paths relative to the reference root).
 * Pinning (what checks each function against something other than itself)
 * is listed per function below and exercised by tests/test_oracle_*.py.
 *
 * The layout structs are laid out exactly like the product's g2g_layout /
 * g2g_block (include/grid2grid_amd.h) so one ctypes description serves both,
 * but they are declared here separately so the oracle does not depend on the
 * product headers.
 */
#ifndef G2G_ORACLE_H
#define G2G_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types, same codes as the product ABI */
enum { ORC_F32 = 0, ORC_F64 = 1, ORC_C64 = 2, ORC_C128 = 3 };

typedef struct orc_block {
    int32_t row;    /* block-row index in the grid        */
    int32_t col;    /* block-column index in the grid     */
    int64_t stride; /* leading dimension in elements      */
    void *data;     /* host pointer to element (0,0), column-major */
} orc_block;

typedef struct orc_layout {
    int32_t n_rows_split; /* entries of rows_split (block rows + 1) */
    int32_t n_cols_split;
    const int32_t *rows_split;
    const int32_t *cols_split;
    const int32_t *owners; /* row-major (block rows) x (block cols) */
    int32_t n_ranks;
    int32_t n_blocks;
    const orc_block *blocks; /* blocks of every virtual rank; owners decide the rank */
} orc_layout;

/* One decomposed piece ("message" in the reference), final orientation. */
typedef struct orc_msg {
    int32_t rank;         /* peer rank: destination on the send side, source on the recv side */
    int32_t tag;          /* layout-pair index (batched form) */
    int32_t r0, r1;       /* global row interval [r0, r1)  */
    int32_t c0, c1;       /* global column interval [c0, c1) */
    int64_t offset;       /* element offset in the packed buffer (-1 for local pieces) */
    int64_t stride;       /* leading dimension of the storage the piece points into */
    int64_t data_off;     /* byte offset of the piece origin from its block's data pointer */
    int32_t block_index;  /* index into layout.blocks */
    int32_t transposed;   /* transpose_on_copy */
    int32_t conjugated;   /* conjugate_on_copy */
    int32_t pad;
} orc_msg;

size_t orc_dtype_size(int dtype);

/* grid_cover.cpp:54-121 (get_decomp_cover). out[2k], out[2k+1] = [start,end)
 * of the red intervals overlapping blue interval k. Returns 0 or -1.
 * Pinned: closed form (upper/lower bound) and brute force in tests, and the
 * reference's own get_decomp_cover via oracle/_ref. */
int orc_decomp_cover(const int32_t *blue, int n_blue, const int32_t *red, int n_red,
                     int32_t *out);

/* Number of virtual ranks of a transform: max(P_init, P_final) (utils.hpp:160). */
int orc_n_ranks(const orc_layout *init, const orc_layout *fin);

/* utils.hpp:146-161 + communication_data.cpp:51-99 for one rank.
 * side 0 = send side (prepare_to_send: init decomposed against final),
 * side 1 = recv side (prepare_to_recv: final decomposed against init).
 * Writes up to max_msgs sorted messages (remote and local, in sort order),
 * counts[P] and dspls[P] (elements). Returns the number of messages, or -1.
 * Pinned: the reference's own communication_data via oracle/_ref, and the
 * tiling invariant (pieces cover every final element exactly once). */
int64_t orc_plan_rank(const orc_layout *init, const orc_layout *fin, int dtype,
                      char trans, int rank, int side, orc_msg *msgs, int64_t max_msgs,
                      int64_t *counts, int64_t *dspls);

/* communication_data.cpp:132-142 (copy_to_buffer): the packed send buffer of
 * one rank, total_size elements (sum of counts). Returns elements written or -1.
 * Pinned: byte-compared with the reference's copy_to_buffer via oracle/_ref. */
int64_t orc_pack_rank(const orc_layout *init, const orc_layout *fin, int dtype, char trans,
                      int rank, void *buffer, int64_t buffer_elems);

/* transform.cpp:126-204 with exchange_async (transform.cpp:41-124) run for all
 * P virtual ranks in one process: pack, exchange (memcpy of each (src,dst)
 * segment), local copies, unpack. alpha/beta point to one element of dtype
 * (NULL = 1 and 0). n_threads > 1 runs the ranks' copy loops with OpenMP
 * when compiled with it. Returns 0 or -1.
 * Pinned: closed form F = op(A) on seeded values, and byte equality with the
 * reference pipeline via oracle/_ref. */
int orc_transform(const orc_layout *init, const orc_layout *fin, int dtype, char trans,
                  const void *alpha, const void *beta);

/* Batched form, transform.cpp:166-204 / utils.hpp:163-212: one exchange for
 * all pairs, pieces tagged with the pair index. All pairs share dtype. */
int orc_transform_batched(int n_pairs, const orc_layout *inits, const orc_layout *fins,
                          int dtype, const char *trans, const void *alphas,
                          const void *betas);

/* transform.cpp:9-39 (communication_volume) with utils.hpp:91-139
 * (rank_to_comm_vol_for_block): for every cell (i, j) of the initial grid,
 * owned by rank r, and every final cell it overlaps (found with the cover of
 * grid_cover.cpp:54-121), owned by rank s, the overlap's element count is
 * added to the unordered pair {min(r,s), max(r,s)}. vol is a dense
 * n x n row-major matrix, n = max(init->n_ranks, fin->n_ranks), written in
 * its upper triangle (vol[a*n+b], a <= b; the rest is zeroed). Only the
 * splits, owners and n_ranks of the layouts are read. Returns 0 or -1.
 * Pinned: the reference's own cover and owners via oracle/_ref; the total
 * equals m*n and vol[a*n+a] equals a brute-force count over elements in tests. */
int orc_comm_volume(const orc_layout *init, const orc_layout *fin, int64_t *vol);

/* ranks_reordering.cpp:4-61 (optimal_reordering): greedy matching on the
 * volume vol (dense n x n, upper triangle as written by orc_comm_volume).
 * Candidate edges are the pairs {a, b} with vol > 0. Adjusted weight:
 * vol(a,b) - vol(a,a) - vol(b,b) for a != b, and 2*vol(a,a)+1 - 2*vol(a,a)
 * = 1 for a == b; edges of adjusted weight 0 are dropped. Edges are visited by
 * decreasing adjusted weight; an edge whose two ends are both unvisited sets
 * perm[a] = b and perm[b] = a, and both ends become visited either way.
 * Two points the reference leaves unspecified are fixed here: equal adjusted
 * weights are visited in ascending (a, b) order (the reference sorts on the
 * weight alone, over an unordered_map's iteration order), and a rank that
 * keeps no elements still gets its {a, a} edge of weight 1 when it occurs in
 * any edge (the reference inserts that entry while iterating the map, so
 * whether it is visited depends on the map's buckets).
 * *reordered = 1 when some a != b was matched. Returns 0 or -1.
 * Pinned: the reference's optimal_reordering via oracle/_ref on volumes where
 * the two unspecified points cannot change the result (every rank keeps
 * elements, adjusted weights above 1 distinct), and brute-force properties. */
int orc_optimal_reordering(const int64_t *vol, int n, int32_t *perm, int *reordered);

/* scalapack_layout.cpp:152-177 (line_split). Returns number of splits written. */
int orc_line_split(int begin, int end, int blk_len, int32_t *out, int max_out);

/* scalapack_layout.cpp:178-278 (full get_scalapack_grid overload) for one rank.
 * ordering: 0 = row_major, 1 = column_major (scalapack_layout.hpp:14).
 * ia/ja are 1-based as in the reference. Fills rows_split/cols_split
 * (capacity given), owners (capacity), and the rank's blocks with data
 * offsets (elements, relative to the rank's base) in offsets[].
 * Returns number of local blocks, or -1.
 * Pinned: the KAT fixtures of examples/scalapack_submatrix.cpp and
 * examples/parallel_scalapack_submatrix.cpp. */
int orc_scalapack_grid(int lld, int m, int n, int ia, int ja, int subm, int subn,
                       int mb, int nb, int prow, int pcol, int ordering, int src_row,
                       int src_col, int rank, int32_t *rows_split, int *n_rows_split,
                       int32_t *cols_split, int *n_cols_split, int32_t *owners,
                       int32_t *blk_row, int32_t *blk_col, int64_t *offsets, int max_blocks);

/* Seeded value function v(seed, i, j) (counter-based, exactly in [-1,1)).
 * key = i * n_cols + j. f64: (x>>11)*2^-52 - 1 with x = splitmix64(seed ^ key);
 * f32: (x>>40)*2^-23 - 1; c64: re from x>>40, im from (x>>16)&0xffffff;
 * c128: re from splitmix64(seed ^ 2key), im from splitmix64(seed ^ (2key+1)).
 * Pinned: fixed known-answer values in tests (and numpy restatement). */
uint64_t orc_splitmix64(uint64_t x);
void orc_value(int dtype, uint64_t seed, int64_t i, int64_t j, int64_t n_cols, void *out);

/* Fill every block of a layout with v(seed, gi, gj) over the layout's own
 * global coordinates (storage orientation), n_cols = global columns. */
int orc_fill_layout(const orc_layout *l, int dtype, uint64_t seed);

/* Count final-layout elements that differ bitwise from op(A): for trans 'N'
 * F(i,j) must equal v(seed,i,j) with A's n_cols; 'T' v(seed,j,i); 'C' conj. */
int64_t orc_check_layout(const orc_layout *fin, int dtype, uint64_t seed, char trans,
                         int64_t a_n_cols);

/* grid_layout::scale_by / block::scale_by (grid_layout.hpp:48-56,
 * block.cpp:291-306): every element of every block of l times beta, unless
 * beta == 1. */
int orc_scale_layout(const orc_layout *l, int dtype, const void *beta);

/* Scaled epilogue internals, for tests. orc_cmul: the restated std::complex
 * product l*r of complex<float> (ORC_C64) or complex<double> (ORC_C128), with
 * __mul?c3 taking r's parts first when swap is set; orc_libgcc_mulxc3: libgcc's
 * own __mulsc3 / __muldc3 on l's and r's parts (what the restatement of C11
 * Annex G.5.1 is checked against). orc_debug_order_xor: XORed into every
 * operand order of the scaled epilogue, its ORD_M1S bit also swaps the
 * operands of scale_by (0 = as the reference; a test plants a wrong order
 * with it). */
void orc_cmul(int dtype, const void *l, const void *r, int swap, void *out);
void orc_libgcc_mulxc3(int dtype, const void *l, const void *r, void *out);
void orc_debug_order_xor(int x);

#ifdef __cplusplus
}
#endif
#endif
