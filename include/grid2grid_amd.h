/*
 * grid2grid_amd.h -- C ABI of the MI355X-native grid2grid::transform.
 *
 * The reference redistributes a matrix between two grid-like layouts over P
 * MPI ranks (src/grid2grid/transform.hpp:13-43). Here the P ranks are P
 * virtual ranks on one GPU: every block of every rank is a device buffer
 * (arbitrary leading dimension, arbitrary placement), and the MPI exchange is
 * an on-device data movement. The entry points are plain C (pointers, sizes,
 * status codes) so any host language can bind them; the headers under include/grid2grid are
 * the header-only C++ facade with the reference's own type and function names.
 *
 * Conventions
 *  - Layouts are column-major inside each block (reference block.hpp:59).
 *  - A layout describes the whole matrix: its blocks array holds the blocks of
 *    every virtual rank; owners[i][j] decides which rank owns cell (i, j).
 *    Every cell must have exactly one block.
 *  - Splits are int32 (the reference uses int); element offsets are 64-bit.
 *  - Streams are hipStream_t passed as void* (NULL = the null stream).
 *  - Every function returns G2G_OK (0) or an error code; g2g_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - Results are bit-exact with the reference: elements move as bit patterns
 *    (no float conversion), 'C' flips the sign bit of the imaginary part, and
 *    the scaled form computes beta*dst + alpha*op(src) without FMA contraction.
 */
#ifndef GRID2GRID_AMD_H
#define GRID2GRID_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1: plans, batches, staged form, volume, reordering, serving loop.
 * 2 adds (nothing removed or changed): g2g_transform_layouts and the plan
 * cache, g2g_transform_graph, g2g_scale_layout, g2g_copy_bytes,
 * g2g_batch_get/set_tuning, g2g_transform_plans_noalloc.
 * 3 changes g2g_tuning: stream_order takes the word version 2 reserved,
 * and row_chunks plus a reserved word are added before ms (the struct grows
 * from 40 to 48 bytes); adds G2G_ORDER_*.
 * 4: g2g_tuning.tail_split takes the word version 3 reserved (same size;
 * 0 keeps version 3's work items). */
#define G2G_ABI_VERSION 4

/* Element types; the reference instantiates transform for these four
 * (src/grid2grid/transform.cpp:206-298). */
typedef enum g2g_dtype {
    G2G_FLOAT32 = 0,    /* float                */
    G2G_FLOAT64 = 1,    /* double               */
    G2G_COMPLEX64 = 2,  /* std::complex<float>  */
    G2G_COMPLEX128 = 3  /* std::complex<double> */
} g2g_dtype;

typedef enum g2g_status {
    G2G_OK = 0,
    G2G_ERR_INVALID = 1,     /* bad argument or inconsistent layouts (reference: std::runtime_error) */
    G2G_ERR_HIP = 2,         /* a HIP runtime call failed */
    G2G_ERR_NOMEM = 3,       /* device or host allocation failed */
    G2G_ERR_BUDGET = 4,      /* request rejected by the HBM admission budget */
    G2G_ERR_UNSUPPORTED = 5, /* valid request outside what this build supports */
    G2G_ERR_STATE = 6        /* object used in the wrong state */
} g2g_status;

/* One block of a layout: cell (row, col) of the grid, stored column-major at
 * `data` (device pointer) with leading dimension `stride` (elements).
 * Mirrors grid2grid::block<T> (reference block.hpp:60-146): coordinates,
 * data pointer and stride; the intervals follow from the grid. */
typedef struct g2g_block {
    int32_t row;
    int32_t col;
    int64_t stride;
    void *data;
} g2g_block;

/* A grid-like layout: grid2D splits (reference grid2D.hpp:17-39), owners and
 * rank count (assigned_grid2D, grid2D.hpp:46-125) and the local blocks of all
 * virtual ranks (local_blocks<T>, block.hpp:167-190). */
typedef struct g2g_layout {
    int32_t n_rows_split;      /* entries of rows_split = block rows + 1 */
    int32_t n_cols_split;      /* entries of cols_split = block cols + 1 */
    const int32_t *rows_split; /* strictly increasing, or {x, x}: an empty dimension */
    const int32_t *cols_split; /* strictly increasing, or {x, x}: an empty dimension */
    const int32_t *owners;     /* row-major, (n_rows_split-1) x (n_cols_split-1) */
    int32_t n_ranks;           /* virtual ranks of the layout */
    int32_t n_blocks;          /* entries of blocks */
    const g2g_block *blocks;
} g2g_layout;

typedef struct g2g_plan_s *g2g_plan;

/* Summary of a plan. */
typedef struct g2g_plan_info {
    int64_t m, n;            /* final matrix rows and columns */
    int64_t n_pieces;        /* intersections of initial and final cells */
    int64_t bytes;           /* matrix bytes moved (m*n*sizeof(T)) */
    int64_t remote_bytes;    /* bytes whose source and destination ranks differ */
    int64_t n_tiles;         /* work items of the fused gather */
    int64_t tile_bytes;      /* bytes per work item (the last round shorter with a tail split) */
    int32_t n_ranks;         /* max(P_initial, P_final) */
    int32_t dtype;
    int32_t trans;           /* 'N', 'T' or 'C' */
    int32_t reserved;
} g2g_plan_info;

/* Message of the last failing call on this thread ("" if none). */
const char *g2g_last_error(void);

/* Library version / build identification. */
int g2g_version(void);

/* Wait until all work queued on `stream` finished (hipStreamSynchronize).
 * The reference's transform returns with the data in place
 * (transform.cpp:117-123, MPI_Waitall); callers that need that guarantee
 * after an asynchronous call use this. */
int g2g_synchronize(void *stream);

/* Build a plan for transforming `initial` into `final`, as
 * grid2grid::transform(initial, final, trans, alpha, beta, comm) does
 * (reference transform.cpp:126-164). The planning of the reference --
 * the grid cover (grid_cover.cpp:54-121), the decomposition into pieces
 * (utils.hpp:10-89), the sort by rank and the per-rank counts, displacements
 * and offsets (utils.hpp:141-212, communication_data.cpp:32-99) -- happens
 * here, once per layout pair: the cover is computed on the host and uploaded
 * as an interval table, pieces and offsets are produced on the device.
 * trans: 'N', 'T' or 'C' (final holds A, A^T or A^H; `initial` describes A in
 * its storage orientation, `final` describes op(A)). Unlike the reference the
 * caller's `initial` descriptor is not modified.
 * Planning work is ordered on `stream`; the call returns after it finished.
 * The plan keeps the block pointers; the buffers must outlive the plan, and
 * the initial and final blocks must not overlap in memory. */
int g2g_plan_create(const g2g_layout *initial, const g2g_layout *final, int dtype, char trans,
                    void *stream, g2g_plan *out);

int g2g_plan_destroy(g2g_plan plan);

int g2g_plan_get_info(g2g_plan plan, g2g_plan_info *info);

/* Run the transform: fused gather (pack, exchange and unpack in one kernel).
 * Replaces exchange_async (reference transform.cpp:41-124) and the copy
 * kernels it calls (communication_data.cpp:101-252, memory_utils.hpp:17-309).
 * alpha/beta point to one host element of the plan's dtype (NULL, NULL =
 * 1 and 0 = plain copy). Every element of every final block is written;
 * bytes outside the pieces (ld padding, gaps) are never touched.
 * Asynchronous on `stream`. The work counters of a launch belong to its
 * stream (each stream gets its own on its first launch, which allocates
 * them once), so one plan may run on several streams at once and be used
 * from several host threads; only the plain form (g2g_transform_staged) and
 * the retuning calls (g2g_plan_set_tuning, g2g_plan_autotune) need launches
 * of the same plan to be ordered. The result is the same whatever the
 * launch configuration, including scaled non-finite values: the scaled
 * epilogue reproduces the reference build's x86 arithmetic (NaN operand
 * choice, default NaN, infinities recovered by the complex product). */
int g2g_transform(g2g_plan plan, const void *alpha, const void *beta, void *stream);

/* The same launch replayed from a hipGraph held by the plan: captured on
 * first use and again whenever alpha/beta or the plan's configuration
 * changed. The graph owns its work counters. `stream` must not be the null
 * stream. Replays of one plan's graph must not overlap; the replay runs
 * alongside g2g_transform launches of the same plan on other streams. */
int g2g_transform_graph(g2g_plan plan, const void *alpha, const void *beta, void *stream);

/* The drop-in call: grid2grid::transform(initial, final, trans, alpha, beta,
 * comm) in one call (reference transform.hpp:13-43, transform.cpp:126-164),
 * planned once per layout pair. The plan of (initial, final, dtype, trans) --
 * splits, owners, rank counts, block coordinates, strides and addresses as
 * given -- is kept in a process-wide cache (least recently used first out,
 * 64 entries by default), so the reference's per-call planning
 * (utils.hpp:146-192) happens on the first call only; the call then runs
 * g2g_transform on `stream`. Safe from several host threads. A plan leaving
 * the cache waits for the device to finish before its tables are freed. */
int g2g_transform_layouts(const g2g_layout *initial, const g2g_layout *final, int dtype, char trans,
                          const void *alpha, const void *beta, void *stream);

/* Capacity of the plan cache (entries; 0 disables keeping plans). */
int g2g_plan_cache_set_capacity(int64_t entries);

/* Drop every cached plan and reset the counters. */
int g2g_plan_cache_clear(void);

/* Calls served from the cache, calls that planned, plans held. Any pointer may be NULL. */
int g2g_plan_cache_stats(int64_t *hits, int64_t *misses, int64_t *entries);

/* The plain form, mirroring the reference step for step: pack every remote
 * piece into the per-rank send buffers in the reference's wire order
 * (communication_data.cpp:132-142) together with the local copies
 * (communication_data.cpp:224-252), exchange each (source, destination)
 * segment between virtual-rank buffers (transform.cpp:53-87), then unpack
 * (communication_data.cpp:157-173). Staging buffers are allocated by the
 * plan on first use. Up to 4096 ranks (the per-pair tables hold P*P
 * entries). Asynchronous on `stream`. */
int g2g_transform_staged(g2g_plan plan, const void *alpha, const void *beta, void *stream);

/* Per-rank send counts of the plan: counts[s * P + d] = elements rank s sends
 * to rank d (0 on the diagonal), as communication_data::counts
 * (communication_data.cpp:72-78). `P` is the plan's n_ranks. */
int g2g_plan_counts(g2g_plan plan, int64_t *counts, int64_t capacity);

/* Copy of the staged form's wire layout, for inspection: per piece (in the
 * plan's piece order) its global row/col origin, shape, source and destination
 * ranks and element offsets in the source rank's send buffer and the
 * destination rank's receive buffer (-1 for local pieces). Arrays hold
 * n_pieces entries; any pointer may be NULL. Synchronous. */
int g2g_plan_wire(g2g_plan plan, int32_t *row0, int32_t *col0, int32_t *rows, int32_t *cols,
                  int32_t *src_rank, int32_t *dst_rank, int64_t *send_off, int64_t *recv_off,
                  int64_t capacity);

/* Staging buffers of the plain form after g2g_transform_staged: device
 * pointer of rank r's packed send buffer (reference communication_data::buffer
 * of the send side) and its size in elements. Synchronous. */
int g2g_plan_send_buffer(g2g_plan plan, int rank, const void **data, int64_t *elems);

/* ---------------------------------------------------------------- batches */

typedef struct g2g_batch_s *g2g_batch;

/* Group several plans (any mix of dtypes and 'N'/'T'/'C') so that they run
 * in ONE persistent launch. Counterpart of the reference's vector overloads
 * and transformer<T> (transform.cpp:166-204, transformer.hpp:8-62), without
 * their single-T restriction. alphas[k]/betas[k] point to one element of plan
 * k's dtype, or the arrays are NULL (plain copies). Plans must outlive the
 * batch. The batch runs each plan's piece stream as it was at creation:
 * once a plan's stream changes (column or row chunks switched on or off or
 * to another width or height, or the stream order changed, by
 * g2g_plan_set_tuning or g2g_plan_autotune) the batch's launches fail with
 * G2G_ERR_INVALID until it is built again; a new work-item size, tail split
 * or kernel variant of the plan leaves the batch as it is (a batch cuts the
 * plans' streams into its own work items). */
int g2g_batch_create(int n, const g2g_plan *plans, const void *const *alphas,
                     const void *const *betas, void *stream, g2g_batch *out);

/* Run every transform of the batch in one persistent kernel that pulls
 * (transform, tile) work items from a device queue (the stream's work
 * counters, as for g2g_transform: launches of one batch on several streams
 * may overlap). */
int g2g_batch_execute(g2g_batch batch, void *stream);

/* Same work replayed from a hipGraph captured on first use (the queue reset
 * is a kernel node, not a memset); the graph owns its work counters, so its
 * replays must not overlap each other. */
int g2g_batch_execute_graph(g2g_batch batch, void *stream);

/* Autotuner of a batch (declared with g2g_tuning below): times the batch's own
 * work-item sizes and kernel variants (always persistent) and keeps the
 * fastest. A batch cuts every plan's piece stream at its own work-item size
 * (default 128 KiB), independently of the plans' standalone configuration. */
struct g2g_tuning;
int g2g_batch_autotune(g2g_batch batch, void *stream, int reps, struct g2g_tuning *best);

/* Launch configuration of a batch (column_chunks, stream_order, row_chunks
 * and tail_split are 0: each plan runs its own stream, cut into the batch's
 * work items) and setting one (those four must be 0): launch_mode G2G_LAUNCH_PERSISTENT or
 * G2G_LAUNCH_PERSISTENT_XCD; a new work-item size re-cuts the batch's work
 * items (synchronous on `stream`). The captured graph, if any, is dropped. */
int g2g_batch_get_tuning(g2g_batch batch, struct g2g_tuning *t);
int g2g_batch_set_tuning(g2g_batch batch, const struct g2g_tuning *t, void *stream);

/* Bytes moved by one execution of the batch. */
int g2g_batch_bytes(g2g_batch batch, int64_t *bytes);

int g2g_batch_destroy(g2g_batch batch);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops that must not call
 * hipMalloc). The work of n plans (alphas/betas as for g2g_batch_create) in
 * one persistent gather launch on `stream`, like g2g_batch_create followed
 * by g2g_batch_execute, but the launch's descriptors, per-work-item records
 * and work counters live in one 4 MiB device array of the library (one per
 * GPU, held until the process ends) instead of allocations: the call
 * allocates nothing and copies nothing from the host (descriptors travel as
 * kernel arguments), so it can be captured. Each plan runs over its own work
 * items (its g2g_plan_get_tuning tile size and tail split; the count is
 * g2g_plan_get_info's n_tiles). At most 4096 plans and 73722
 * work items per call; larger requests fail with G2G_ERR_INVALID before
 * anything is launched. Every call on a GPU shares that GPU's array and
 * leaves its contents behind: calls that overlap on one GPU (other streams,
 * other threads, replays of graphs holding such a call) race; order them. */
int g2g_transform_plans_noalloc(int n, const g2g_plan *plans, const void *const *alphas,
                                const void *const *betas, void *stream);

/* ---------------------------------------------------------------- serving */

/* A request loop on its own host thread: transform requests (plans whose
 * buffers and interval tables stay resident) are submitted as they arrive,
 * admitted while the HBM footprint of the requests in flight (2*m*n*sizeof(T)
 * each: source plus destination) stays within `budget_bytes`, and packed
 * into persistent launches together with whatever else is waiting
 * (continuous batching); up to `n_streams` launches overlap on the device.
 * mode G2G_SERVE_BATCHED: fused gather, up to `max_batch` requests per
 * launch. mode G2G_SERVE_PLAIN: one request per launch with the plain staged
 * form, on one stream (the per-request baseline). No counterpart in the
 * reference, whose transform is a blocking collective call. */
typedef struct g2g_server_s *g2g_server;

enum { G2G_SERVE_BATCHED = 0, G2G_SERVE_PLAIN = 1 };

typedef struct g2g_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t rejected;            /* larger than the whole budget */
    int64_t launches;
    int64_t bytes;               /* matrix bytes of completed requests */
    int64_t in_flight_bytes;     /* footprint admitted and not completed */
    int64_t peak_in_flight_bytes;
} g2g_server_stats;

int g2g_server_create(int64_t budget_bytes, int n_streams, int max_batch, int mode,
                      g2g_server *out);

/* Queue plan for execution (alpha/beta as for g2g_transform, copied).
 * Returns a ticket; G2G_ERR_BUDGET if the request can never fit the budget.
 * A plan must not be submitted again, retuned (g2g_plan_set_tuning,
 * g2g_plan_autotune) or destroyed before its previous ticket completed. */
int g2g_server_submit(g2g_server s, g2g_plan plan, const void *alpha, const void *beta,
                      int64_t *ticket);

/* *done = 1 and *latency_s = completion - submission time once finished,
 * with the request's own status (a failed request returns its error). A
 * finished ticket is reported once, by poll or wait: the server then forgets
 * it, and asking again fails with G2G_ERR_STATE. */
int g2g_server_poll(g2g_server s, int64_t ticket, int32_t *done, double *latency_s);

/* Block until the ticket finished; reports it as g2g_server_poll does. */
int g2g_server_wait(g2g_server s, int64_t ticket, double *latency_s);

/* Block until every submitted request finished. */
int g2g_server_drain(g2g_server s);

int g2g_server_get_stats(g2g_server s, g2g_server_stats *stats);

/* Drains, stops the thread and frees the server. */
int g2g_server_destroy(g2g_server s);

/* ------------------------------------------------------------ tuning */

/* Work-item size (bytes) and launch mode used for plans created afterwards:
 * tile_bytes in [4096, 1<<26] (0 = automatic), launch_mode G2G_LAUNCH_* (-1 = automatic). */
int g2g_set_tuning(int64_t tile_bytes, int launch_mode);

/* Launch modes of the gather: one workgroup per work item (work items dealt
 * round-robin over the XCDs by the hardware); a persistent grid pulling work
 * items from a device counter; and the same two with each XCD taking one
 * contiguous range of work items (neighbouring work items share that XCD's
 * L2; persistent: one counter per XCD, idle XCDs then help the others). */
enum {
    G2G_LAUNCH_DIRECT = 0,
    G2G_LAUNCH_PERSISTENT = 1,
    G2G_LAUNCH_DIRECT_XCD = 2,
    G2G_LAUNCH_PERSISTENT_XCD = 3
};

/* Order of a plan's stream (g2g_tuning.stream_order). */
enum { G2G_ORDER_UNION = 0, G2G_ORDER_DESTINATION = 1 };

/* Launch configuration of a plan's fused gather. */
typedef struct g2g_tuning {
    int64_t tile_bytes;    /* bytes of the piece stream per work item */
    int32_t threads;       /* workgroup size: 256 or 512 */
    int32_t unroll_bytes;  /* bytes each thread keeps in flight: 64 or 128 */
    int32_t nontemporal;   /* 1: non-temporal loads/stores */
    int32_t launch_mode;   /* G2G_LAUNCH_* */
    /* 0: the stream is the pieces in their own order; w > 0: the pieces cut
     * into chunks of at most w final columns, chunk after chunk down each
     * column interval, so that vertically adjacent pieces write their
     * destination column segments (contiguous within a final block, sharing
     * a cache line where a segment does not end on one) at the same time.
     * Batches run each plan's stream as the plan is tuned (their own tuning
     * reports 0). */
    int32_t column_chunks;
    /* G2G_ORDER_UNION: the stream in union-cell order (pieces, or chunks of
     * the pieces as above); G2G_ORDER_DESTINATION: the same stream ordered
     * by destination address, so that work items that run at the same time
     * write neighbouring addresses, final block after final block. Batches
     * report and accept only G2G_ORDER_UNION for their own tuning (each
     * plan's stream runs as the plan is tuned). A caller that zeroes the
     * struct gets the union order. */
    int32_t stream_order;
    /* 0: whole columns; h > 0 (untransposed plans only): every piece or chunk
     * of the stream cut again into chunks of at most h rows, so that tall
     * pieces move as bands of runs of h rows across many columns rather than
     * as long contiguous runs. Batches: 0. */
    int32_t row_chunks;
    /* 0: every work item is tile_bytes long; d = 2, 4 or 8: the stream's
     * last min(total / 2, CUs x 4 x tile_bytes) bytes (counted in whole items
     * from the start) are cut into items of tile_bytes / d, so that the
     * workgroups of a launch finish closer together. Batches: 0. */
    int32_t tail_split;
    double ms;             /* measured time per transform (set by g2g_plan_autotune) */
} g2g_tuning;

int g2g_plan_get_tuning(g2g_plan plan, g2g_tuning *t);
int g2g_plan_set_tuning(g2g_plan plan, const g2g_tuning *t);

/* Autotuner: time every compiled variant (workgroup size, bytes in flight,
 * non-temporal hint, launch mode) at several work-item sizes on this plan's
 * pieces, in union-cell and in destination order (`reps` timed runs each,
 * on `stream`, hipEvents), time the eight fastest of each order again (each
 * also at the neighbouring work-item sizes) on the pieces and on
 * column-chunked (and, for tall pieces, row-chunked) streams in that order,
 * time the fastest again as it is and with tail splits 2 and 4 (a split is
 * taken when it is 0.5 % faster), keep the result for the plan and record it for the plan's shape class
 * (element size, operation, size class, piece-size class): plans created
 * later in this process with the same class start from it. The transform
 * runs while tuning (the final layout is written). *best receives the
 * chosen configuration. */
int g2g_plan_autotune(g2g_plan plan, void *stream, int reps, g2g_tuning *best);

/* ------------------------------------ communication volume, rank relabelling */

/* Elements exchanged between one unordered pair of ranks {src, dest},
 * src <= dest (src == dest: elements that stay on the rank). */
typedef struct g2g_edge {
    int32_t src;
    int32_t dest;
    int64_t weight;
} g2g_edge;

/* Replaces grid2grid::communication_volume(assigned_grid2D&, assigned_grid2D&)
 * (transform.hpp:47, transform.cpp:9-39, utils.hpp:91-139): for every cell of
 * the initial grid and every final cell it overlaps, the overlap's element
 * count is added to the pair {initial owner, final owner}. Only the splits,
 * owners and n_ranks of the two layouts are read (blocks may be absent).
 * Computed on the device (one key per cell of the split union, radix sort,
 * reduction by key); synchronous on `stream`. *n_edges receives the number of
 * pairs with a nonzero volume; the first min(capacity, *n_edges) of them are
 * written to `edges` (may be NULL) in ascending (src, dest) order. */
int g2g_comm_volume(const g2g_layout *initial, const g2g_layout *final, g2g_edge *edges,
                    int64_t capacity, int64_t *n_edges, void *stream);

/* Replaces grid2grid::optimal_reordering(comm_volume&, int, bool&)
 * (ranks_reordering.hpp:7, ranks_reordering.cpp:4-61): greedy matching of
 * ranks on the volume. Edge {a, b}, a != b, weighs vol(a,b) - vol(a,a) -
 * vol(b,b); edge {a, a} weighs 1; edges weighing 0 are dropped; edges are taken
 * by decreasing weight when both ends are still free (permutation[a] = b,
 * permutation[b] = a), and both ends are then taken. Equal weights go in
 * ascending (a, b) order, and every rank that occurs in an edge has its {a, a}
 * edge (the reference leaves both points to its hash map). Edges may come in
 * any order, with src > dest and repeats (summed, as comm_volume::operator+=
 * does). Host-only. `permutation` has n_ranks entries; *reordered = 1 when some
 * rank is mapped to another. */
int g2g_optimal_reordering(const g2g_edge *edges, int64_t n_edges, int32_t n_ranks,
                           int32_t *permutation, int32_t *reordered);

/* ------------------------------------------------------ layout utilities */

/* Replaces grid_layout<T>::scale_by and block<T>::scale_by (reference
 * grid_layout.hpp:48-56, block.hpp:136, block.cpp:291-306): every element of
 * every block of `layout` (its rows and columns by the splits, at the block's
 * stride) is multiplied by *beta, one element of `dtype`; nothing happens
 * when *beta == 1. The products are the reference build's (data[offset] *=
 * beta with its x86 arithmetic, NaN payloads and complex infinities
 * included). Padding and gaps are not touched. Synchronous on `stream`
 * (the block table is uploaded for the call). */
int g2g_scale_layout(const g2g_layout *layout, int dtype, const void *beta, void *stream);

/* Synchronous copy of `bytes` between any host or device addresses
 * (hipMemcpy, direction inferred). The facade's block<T>::local_element
 * (reference block.hpp:130-131) reads and writes one element of device
 * memory through it. */
int g2g_copy_bytes(void *dst, const void *src, int64_t bytes);

/* ------------------------------------------------ test-support kernels */

/* Fill every block of `layout` with the seeded value v(seed, i, j) of its
 * global element (i, j), i.e. splitmix64(seed ^ (i * n_cols + j)) mapped to
 * [-1, 1) exactly (f64: (x>>11)*2^-52-1; f32: (x>>40)*2^-23-1; c64: real from
 * x>>40, imaginary from (x>>16)&0xffffff; c128: real/imaginary from keys 2k
 * and 2k+1), n_cols = the layout's global columns. Asynchronous. */
int g2g_fill_layout(const g2g_layout *layout, int dtype, uint64_t seed, void *stream);

/* Count elements of `final` that differ bitwise from op(A) where A(i, j) =
 * v(seed, i, j) with a_n_cols columns. Synchronous; result in *mismatches. */
int g2g_verify_layout(const g2g_layout *final, int dtype, uint64_t seed, char trans,
                      int64_t a_n_cols, void *stream, int64_t *mismatches);

/* Streaming reference rate of the device, for context beside a transform's
 * bandwidth: `bytes` (> 0, a multiple of 16; buffers 16-byte aligned) copied
 * from src to dst (kind 0), read from src (1; dst receives at most 16 bytes)
 * or written to dst (2), by 16 workgroups per CU moving 16-byte vectors with
 * non-temporal hints; average milliseconds over `reps` launches after one
 * warm-up (hipEvents on `stream`) in *ms. Overwrites dst. */
int g2g_stream_probe(int kind, const void *src, void *dst, int64_t bytes, int reps, void *stream,
                     double *ms);

#ifdef __cplusplus
}
#endif
#endif /* GRID2GRID_AMD_H */
