// Per rank-pair work done by sorting on the device: the communication volume
// between the virtual ranks of two grids and the greedy rank relabelling
// computed from it (g2g_comm_volume, g2g_optimal_reordering in
// include/grid2grid_amd.h), and the wire order of the plain form for rank
// counts whose pair table does not fit in LDS (launch_pair_offsets_sorted).
//
// The reference walks every initial block, finds the final cells it overlaps
// through the grid cover and adds the overlap sizes to a hash map keyed by the
// rank pair (transform.cpp:9-39, utils.hpp:91-139). Every overlap is a cell of
// the union of the two split vectors, so here one device thread per union cell
// emits (pair key, element count); a radix sort by key and a reduction by key
// give the per-pair totals in ascending (src, dest) order.
#include "grid2grid_amd.h"
#include "g2g_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

using namespace g2g;

namespace {

// Key of union cell k = (column interval) * n_ur + (row interval): the pair
// {initial owner, final owner} as lo * P + hi; weight: the cell's elements.
__global__ __launch_bounds__(256) void k_volume_keys(const span *ur, int32_t n_ur, const span *uc,
                                                     int32_t n_uc, const int32_t *iown, int32_t i_nbc,
                                                     const int32_t *fown, int32_t f_nbc, uint32_t P,
                                                     uint64_t *keys, int64_t *weights) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n_ur * n_uc) return;
    const span r = ur[k % n_ur];
    const span c = uc[k / n_ur];
    const uint32_t a = (uint32_t)iown[(int64_t)r.iblk * i_nbc + c.iblk];
    const uint32_t b = (uint32_t)fown[(int64_t)r.fblk * f_nbc + c.fblk];
    keys[k] = (uint64_t)min(a, b) * P + max(a, b);
    weights[k] = (int64_t)(r.end - r.start) * (c.end - c.start);
}

// Device buffers of one call, freed on return.
struct scratch_set {
    std::vector<void *> ptrs;
    ~scratch_set() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *get(size_t count) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

#define VOL_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return internal_fail(G2G_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                 __FILE__, __LINE__);                                                 \
    } while (0)

int comm_volume(const g2g_layout *init, const g2g_layout *fin, std::vector<g2g_edge> &out,
                hipStream_t s) {
    int rc = internal_check_grid("initial", init);
    if (rc != G2G_OK) return rc;
    rc = internal_check_grid("final", fin);
    if (rc != G2G_OK) return rc;
    std::vector<int32_t> irs(init->rows_split, init->rows_split + init->n_rows_split);
    std::vector<int32_t> ics(init->cols_split, init->cols_split + init->n_cols_split);
    std::vector<int32_t> frs(fin->rows_split, fin->rows_split + fin->n_rows_split);
    std::vector<int32_t> fcs(fin->cols_split, fin->cols_split + fin->n_cols_split);
    std::vector<span> ur, uc;
    rc = internal_union_spans(frs, irs, "rows", ur);
    if (rc != G2G_OK) return rc;
    rc = internal_union_spans(fcs, ics, "columns", uc);
    if (rc != G2G_OK) return rc;
    const int64_t n = (int64_t)ur.size() * (int64_t)uc.size();
    out.clear();
    if (n == 0) return G2G_OK;  // an empty matrix exchanges nothing
    if (n > 0x7fffffffll) return internal_fail(G2G_ERR_UNSUPPORTED, "comm_volume: %lld union cells", (long long)n);
    const uint32_t P = (uint32_t)std::max(init->n_ranks, fin->n_ranks);
    const int i_cells = (init->n_rows_split - 1) * (init->n_cols_split - 1);
    const int f_cells = (fin->n_rows_split - 1) * (fin->n_cols_split - 1);

    scratch_set sc;
    span *d_ur = sc.get<span>(ur.size()), *d_uc = sc.get<span>(uc.size());
    int32_t *d_io = sc.get<int32_t>(i_cells), *d_fo = sc.get<int32_t>(f_cells);
    uint64_t *keys = sc.get<uint64_t>(n), *keys_sorted = sc.get<uint64_t>(n), *ukeys = sc.get<uint64_t>(n);
    int64_t *w = sc.get<int64_t>(n), *w_sorted = sc.get<int64_t>(n), *sums = sc.get<int64_t>(n);
    int *d_runs = sc.get<int>(1);
    if (!d_ur || !d_uc || !d_io || !d_fo || !keys || !keys_sorted || !ukeys || !w || !w_sorted || !sums || !d_runs)
        return internal_fail(G2G_ERR_NOMEM, "comm_volume: device allocation failed");
    VOL_HIP(hipMemcpyAsync(d_ur, ur.data(), ur.size() * sizeof(span), hipMemcpyHostToDevice, s));
    VOL_HIP(hipMemcpyAsync(d_uc, uc.data(), uc.size() * sizeof(span), hipMemcpyHostToDevice, s));
    VOL_HIP(hipMemcpyAsync(d_io, init->owners, (size_t)i_cells * sizeof(int32_t), hipMemcpyHostToDevice, s));
    VOL_HIP(hipMemcpyAsync(d_fo, fin->owners, (size_t)f_cells * sizeof(int32_t), hipMemcpyHostToDevice, s));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_volume_keys, dim3(blocks), dim3(256), 0, s, d_ur, (int32_t)ur.size(), d_uc,
                       (int32_t)uc.size(), d_io, init->n_cols_split - 1, d_fo, fin->n_cols_split - 1, P, keys, w);
    VOL_HIP(hipGetLastError());
    // keys < P*P <= 2^32: sort only the bits that can be set
    int end_bit = 1;
    while (end_bit < 64 && ((uint64_t)1 << end_bit) < (uint64_t)P * P) ++end_bit;
    size_t sort_bytes = 0, red_bytes = 0;
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_sorted, w, w_sorted, (int)n, 0,
                                               end_bit, s));
    VOL_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, red_bytes, keys_sorted, ukeys, w_sorted, sums, d_runs,
                                              hipcub::Sum(), (int)n, s));
    void *tmp = sc.get<unsigned char>(std::max(sort_bytes, red_bytes));
    if (!tmp) return internal_fail(G2G_ERR_NOMEM, "comm_volume: device allocation failed");
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, keys_sorted, w, w_sorted, (int)n, 0, end_bit,
                                               s));
    VOL_HIP(hipcub::DeviceReduce::ReduceByKey(tmp, red_bytes, keys_sorted, ukeys, w_sorted, sums, d_runs,
                                              hipcub::Sum(), (int)n, s));
    int runs = 0;
    VOL_HIP(hipMemcpyAsync(&runs, d_runs, sizeof(int), hipMemcpyDeviceToHost, s));
    VOL_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> hk(runs);
    std::vector<int64_t> hw(runs);
    VOL_HIP(hipMemcpyAsync(hk.data(), ukeys, sizeof(uint64_t) * runs, hipMemcpyDeviceToHost, s));
    VOL_HIP(hipMemcpyAsync(hw.data(), sums, sizeof(int64_t) * runs, hipMemcpyDeviceToHost, s));
    VOL_HIP(hipStreamSynchronize(s));
    out.resize(runs);
    for (int k = 0; k < runs; ++k) {
        out[k].src = (int32_t)(hk[k] / P);
        out[k].dest = (int32_t)(hk[k] % P);
        out[k].weight = hw[k];
    }
    return G2G_OK;
}

struct wedge {
    int32_t a, b;
    int64_t w;
};

// ---- wire order by sorting (many ranks)

constexpr uint32_t kLocalKey = 0xffffffffu;  // pieces whose source and destination rank agree

// keys[p] = source * P + destination (kLocalKey for local pieces), idx[p] = p.
__global__ __launch_bounds__(256) void k_wire_keys(const uint32_t *ranks, int64_t n, uint32_t P, uint32_t *keys,
                                                   uint32_t *idx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t r = ranks[p], s = r >> 16, d = r & 0xffffu;
    keys[p] = s == d ? kLocalKey : s * P + d;
    idx[p] = (uint32_t)p;
}

// In sorted order: element count of each piece (0 for local ones) and, at the
// first piece of every key, its own position (0 elsewhere; a max-scan turns
// this into each piece's segment start).
__global__ __launch_bounds__(256) void k_wire_sizes(const uint32_t *skeys, const uint32_t *sidx, const piece *pieces,
                                                    int64_t n, uint64_t *sz, uint32_t *head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const piece pc = pieces[sidx[i]];
    sz[i] = skeys[i] == kLocalKey ? 0ull : (uint64_t)pc.rows * pc.cols;
    head[i] = (i == 0 || skeys[i] != skeys[i - 1]) ? (uint32_t)i : 0u;
}

// within[p]: offset of piece p in its pair's segment (~0 for local pieces);
// pair_counts[key] at the last piece of each key's segment.
__global__ __launch_bounds__(256) void k_wire_within(const uint32_t *skeys, const uint32_t *sidx, const uint64_t *sz,
                                                     const uint64_t *ex, const uint32_t *start, int64_t n,
                                                     uint64_t *within, uint64_t *pair_counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = skeys[i];
    if (key == kLocalKey) {
        within[sidx[i]] = ~0ull;
        return;
    }
    const uint64_t off = ex[i] - ex[start[i]];
    within[sidx[i]] = off;
    if (i + 1 == n || skeys[i + 1] != key) pair_counts[key] = off + sz[i];
}

__global__ __launch_bounds__(256) void k_zero_u64(uint64_t *x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

}  // namespace

namespace g2g {

// Same results as launch_pair_offsets (reference communication_data.cpp:51-99
// with the message order of utils.hpp:141-161): a stable radix sort of the
// pieces by (source, destination) keeps, inside every pair, the pieces in
// table order, which is the reference's (columns, rows) order; an exclusive
// scan of the sorted piece sizes, less the scan at each pair's first piece,
// gives the offsets. Synchronous on `s` (scratch is freed on return).
int launch_pair_offsets_sorted(const uint32_t *ranks, const piece *pieces, int64_t n, int32_t P,
                               uint64_t *within, uint64_t *pair_counts, hipStream_t s) {
    const int64_t K = (int64_t)P * P;
    if (K >= (int64_t)kLocalKey) return internal_fail(G2G_ERR_UNSUPPORTED, "wire order: %d ranks", P);
    if (n > 0x7fffffffll) return internal_fail(G2G_ERR_UNSUPPORTED, "wire order: %lld pieces", (long long)n);
    const unsigned gk = (unsigned)((K + 255) / 256);
    hipLaunchKernelGGL(k_zero_u64, dim3(gk), dim3(256), 0, s, pair_counts, K);
    VOL_HIP(hipGetLastError());
    if (n == 0) return G2G_OK;
    scratch_set sc;
    uint32_t *keys = sc.get<uint32_t>(n), *skeys = sc.get<uint32_t>(n), *idx = sc.get<uint32_t>(n),
             *sidx = sc.get<uint32_t>(n), *head = sc.get<uint32_t>(n), *start = sc.get<uint32_t>(n);
    uint64_t *sz = sc.get<uint64_t>(n), *ex = sc.get<uint64_t>(n);
    if (!keys || !skeys || !idx || !sidx || !head || !start || !sz || !ex)
        return internal_fail(G2G_ERR_NOMEM, "wire order: device allocation failed");
    const unsigned gn = (unsigned)((n + 255) / 256);
    const int ni = (int)n;
    hipLaunchKernelGGL(k_wire_keys, dim3(gn), dim3(256), 0, s, ranks, n, (uint32_t)P, keys, idx);
    VOL_HIP(hipGetLastError());
    size_t b1 = 0, b2 = 0, b3 = 0;
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, keys, skeys, idx, sidx, ni, 0, 32, s));
    VOL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, sz, ex, ni, s));
    VOL_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, b3, head, start, hipcub::Max(), ni, s));
    void *tmp = sc.get<unsigned char>(std::max(b1, std::max(b2, b3)));
    if (!tmp) return internal_fail(G2G_ERR_NOMEM, "wire order: device allocation failed");
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b1, keys, skeys, idx, sidx, ni, 0, 32, s));
    hipLaunchKernelGGL(k_wire_sizes, dim3(gn), dim3(256), 0, s, skeys, sidx, pieces, n, sz, head);
    VOL_HIP(hipGetLastError());
    VOL_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b2, sz, ex, ni, s));
    VOL_HIP(hipcub::DeviceScan::InclusiveScan(tmp, b3, head, start, hipcub::Max(), ni, s));
    hipLaunchKernelGGL(k_wire_within, dim3(gn), dim3(256), 0, s, skeys, sidx, sz, ex, start, n, within,
                       pair_counts);
    VOL_HIP(hipGetLastError());
    VOL_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}

namespace {

__global__ __launch_bounds__(256) void k_dst_keys(const piece *in, int64_t n, uint64_t *keys, uint32_t *idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    keys[k] = in[k].dst;
    idx[k] = (uint32_t)k;
}

__global__ __launch_bounds__(256) void k_gather_pieces(const piece *in, const uint32_t *sidx, int64_t n,
                                                       uint32_t es, piece *out, uint64_t *bytes) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const piece pc = in[sidx[k]];
    out[k] = pc;
    bytes[k] = (uint64_t)pc.rows * pc.cols * es;
}

}  // namespace

// out = the n pieces of `in` ordered by destination address (a stable radix
// sort of the 64-bit addresses), bytes[k] = out[k]'s bytes. Synchronous on
// `s` (scratch is freed on return).
int sort_pieces_by_dst(const piece *in, int64_t n, uint32_t es, piece *out, uint64_t *bytes, hipStream_t s) {
    if (n == 0) return G2G_OK;
    if (n > 0x7fffffffll) return internal_fail(G2G_ERR_UNSUPPORTED, "destination order: %lld pieces", (long long)n);
    scratch_set sc;
    uint64_t *keys = sc.get<uint64_t>(n), *skeys = sc.get<uint64_t>(n);
    uint32_t *idx = sc.get<uint32_t>(n), *sidx = sc.get<uint32_t>(n);
    if (!keys || !skeys || !idx || !sidx) return internal_fail(G2G_ERR_NOMEM, "destination order: device allocation failed");
    const unsigned gn = (unsigned)((n + 255) / 256);
    const int ni = (int)n;
    hipLaunchKernelGGL(k_dst_keys, dim3(gn), dim3(256), 0, s, in, n, keys, idx);
    VOL_HIP(hipGetLastError());
    size_t b = 0;
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, keys, skeys, idx, sidx, ni, 0, 64, s));
    void *tmp = sc.get<unsigned char>(b);
    if (!tmp) return internal_fail(G2G_ERR_NOMEM, "destination order: device allocation failed");
    VOL_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, skeys, idx, sidx, ni, 0, 64, s));
    hipLaunchKernelGGL(k_gather_pieces, dim3(gn), dim3(256), 0, s, in, sidx, n, es, out, bytes);
    VOL_HIP(hipGetLastError());
    VOL_HIP(hipStreamSynchronize(s));
    return G2G_OK;
}

}  // namespace g2g

extern "C" {

int g2g_comm_volume(const g2g_layout *initial, const g2g_layout *final, g2g_edge *edges, int64_t capacity,
                    int64_t *n_edges, void *stream) {
    if (!n_edges) return internal_fail(G2G_ERR_INVALID, "comm_volume: n_edges is NULL");
    if (capacity < 0) return internal_fail(G2G_ERR_INVALID, "comm_volume: negative capacity");
    std::vector<g2g_edge> out;
    const int rc = comm_volume(initial, final, out, (hipStream_t)stream);
    if (rc != G2G_OK) return rc;
    *n_edges = (int64_t)out.size();
    if (edges) std::copy(out.begin(), out.begin() + std::min<int64_t>(capacity, (int64_t)out.size()), edges);
    return G2G_OK;
}

int g2g_optimal_reordering(const g2g_edge *edges, int64_t n_edges, int32_t n_ranks, int32_t *permutation,
                           int32_t *reordered) {
    if (n_ranks < 1) return internal_fail(G2G_ERR_INVALID, "optimal_reordering: n_ranks must be positive");
    if (!permutation || !reordered) return internal_fail(G2G_ERR_INVALID, "optimal_reordering: NULL output");
    if (n_edges < 0 || (n_edges > 0 && !edges))
        return internal_fail(G2G_ERR_INVALID, "optimal_reordering: bad edge array");
    // pairs as unordered {lo, hi}, repeats summed (comm_volume::operator+=)
    std::vector<wedge> e;
    e.reserve((size_t)n_edges);
    for (int64_t k = 0; k < n_edges; ++k) {
        const int32_t a = edges[k].src, b = edges[k].dest;
        if (a < 0 || a >= n_ranks || b < 0 || b >= n_ranks)
            return internal_fail(G2G_ERR_INVALID, "optimal_reordering: edge %lld (%d, %d) outside [0, %d)",
                                 (long long)k, a, b, n_ranks);
        e.push_back(wedge{std::min(a, b), std::max(a, b), edges[k].weight});
    }
    std::sort(e.begin(), e.end(), [](const wedge &x, const wedge &y) {
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    });
    std::vector<wedge> pairs;
    for (const wedge &x : e) {
        if (!pairs.empty() && pairs.back().a == x.a && pairs.back().b == x.b)
            pairs.back().w += x.w;
        else
            pairs.push_back(x);
    }
    std::vector<int64_t> self(n_ranks, 0);
    std::vector<char> occurs(n_ranks, 0), visited(n_ranks, 0);
    for (const wedge &x : pairs) {
        if (x.w == 0) continue;  // a pair that is not in the volume
        if (x.a == x.b) self[x.a] = x.w;
        occurs[x.a] = occurs[x.b] = 1;
    }
    // adjusted weights (ranks_reordering.cpp:17-34)
    std::vector<wedge> cand;
    for (const wedge &x : pairs)
        if (x.a != x.b && x.w != 0) {
            const int64_t w = x.w - self[x.a] - self[x.b];
            if (w) cand.push_back(wedge{x.a, x.b, w});
        }
    for (int32_t a = 0; a < n_ranks; ++a)
        if (occurs[a]) cand.push_back(wedge{a, a, 1});
    // decreasing weight, then ascending (a, b)
    std::sort(cand.begin(), cand.end(), [](const wedge &x, const wedge &y) {
        if (x.w != y.w) return x.w > y.w;
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    });
    for (int32_t a = 0; a < n_ranks; ++a) permutation[a] = a;
    *reordered = 0;
    for (const wedge &x : cand) {  // greedy matching (ranks_reordering.cpp:39-58)
        if (visited[x.a] || visited[x.b]) continue;
        permutation[x.a] = x.b;
        permutation[x.b] = x.a;
        if (x.a != x.b) *reordered = 1;
        visited[x.a] = visited[x.b] = 1;
    }
    return G2G_OK;
}

}  // extern "C"
