// ref_driver.cpp -- TEST INFRASTRUCTURE: drives the reference's own code.
//
// oracle/Makefile compiles this file together with the reference's MPI-free
// sources, where they lie in the reference tree (src/grid2grid/interval.cpp,
// grid2D.cpp, block.cpp, grid_cover.cpp, communication_data.cpp), into
// oracle/_ref/libg2g_ref.so. Nothing of the reference is copied here.
//
// The reference's transform.cpp, utils.hpp, grid_layout.hpp and
// scalapack_layout.cpp include <mpi.h>, which this image does not have, so they
// are not built. What they do on the hot path is reproduced here with the
// reference's own building blocks:
//   * decompose_block (utils.hpp:10-65): the loop is restated below, but the
//     cover (grid_cover::decompose_block), owners (assigned_grid2D::owner),
//     piece pointers (block::subblock) and messages are the reference's;
//   * merge_messages (utils.hpp:141-144): std::sort with the reference's
//     message::operator<;
//   * prepare_to_send/recv (utils.hpp:146-192): the reference's
//     communication_data constructor;
//   * exchange_async (transform.cpp:41-124): the reference's copy_to_buffer,
//     copy_local_blocks and copy_from_buffer, with MPI_Isend/Irecv replaced by
//     a memcpy of each (source, destination) segment at the reference's
//     counts/dspls;
//   * grid_layout::transpose_or_conjugate (grid_layout.hpp:18-25): the
//     reference's assigned_grid2D::transpose and
//     local_blocks::transpose_or_conjugate;
//   * communication_volume (transform.cpp:9-39): the per-block loop of
//     utils.hpp:91-139 restated on the reference's cover and owners, summed
//     into the reference's comm_volume;
//   * optimal_reordering: the reference's ranks_reordering.cpp itself.
#include <grid2grid/block.hpp>
#include <grid2grid/communication_data.hpp>
#include <grid2grid/grid2D.hpp>
#include <grid2grid/grid_cover.hpp>
#include <grid2grid/ranks_reordering.hpp>

#include "g2g_oracle.h"

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <vector>

using namespace grid2grid;

namespace {

assigned_grid2D make_grid(const orc_layout *l) {
    std::vector<int> rs(l->rows_split, l->rows_split + l->n_rows_split);
    std::vector<int> cs(l->cols_split, l->cols_split + l->n_cols_split);
    int nr = l->n_rows_split - 1, nc = l->n_cols_split - 1;
    std::vector<std::vector<int>> owners(nr, std::vector<int>(nc));
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nc; ++j) owners[i][j] = l->owners[(size_t)i * nc + j];
    grid2D g(std::move(rs), std::move(cs));
    return assigned_grid2D(std::move(g), std::move(owners), l->n_ranks);
}

template <typename T>
local_blocks<T> make_blocks(const orc_layout *l, const assigned_grid2D &g, int rank) {
    std::vector<block<T>> v;
    for (int b = 0; b < l->n_blocks; ++b) {
        const orc_block &d = l->blocks[b];
        if (g.owner(d.row, d.col) != rank) continue;
        v.emplace_back(g, block_coordinates(d.row, d.col), (T *)d.data, (int)d.stride);
    }
    return local_blocks<T>(std::move(v));
}

// utils.hpp:10-65, on the reference's grid_cover / owner / subblock.
template <typename T>
void decompose(const local_blocks<T> &blocks, const assigned_grid2D &from,
               const assigned_grid2D &to, T alpha, T beta, int tag,
               std::vector<message<T>> &out) {
    grid_cover cover(from.grid(), to.grid());
    for (int k = 0; k < blocks.num_blocks(); ++k) {
        block<T> b = blocks.get_block(k);
        b.tag = tag;
        block_cover bc = cover.decompose_block(b);
        int row_start = b.rows_interval.start;
        for (int i = bc.rows_cover.start_index; i < bc.rows_cover.end_index; ++i) {
            int row_end = std::min(to.grid().rows_split[i + 1], b.rows_interval.end);
            int col_start = b.cols_interval.start;
            for (int j = bc.cols_cover.start_index; j < bc.cols_cover.end_index; ++j) {
                int rank = to.owner(i, j);
                int col_end = std::min(to.grid().cols_split[j + 1], b.cols_interval.end);
                block<T> sb = b.subblock({row_start, row_end}, {col_start, col_end});
                if (sb.non_empty()) {
                    out.push_back(message<T>(sb, rank));
                    out.back().alpha = alpha;
                    out.back().beta = beta;
                }
                col_start = col_end;
            }
            row_start = row_end;
        }
    }
}

template <typename T>
T load_scalar(const void *p, T dflt) {
    if (!p) return dflt;
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

// One virtual rank's send and recv communication_data for a set of pairs.
template <typename T>
struct rank_data {
    communication_data<T> send, recv;
};

template <typename T>
int run(int n_pairs, const orc_layout *inits, const orc_layout *fins, const char *trans,
        const void *alphas, const void *betas, int only_rank, int side, int64_t *counts,
        int64_t *dspls, void *pack_out, int64_t pack_cap, int64_t *total_out, double *secs) {
    int P = 0;
    for (int k = 0; k < n_pairs; ++k)
        P = std::max(P, std::max(inits[k].n_ranks, fins[k].n_ranks));
    std::vector<assigned_grid2D> gi(n_pairs), gf(n_pairs);
    std::vector<T> al(n_pairs), be(n_pairs);
    for (int k = 0; k < n_pairs; ++k) {
        gi[k] = make_grid(&inits[k]);
        gf[k] = make_grid(&fins[k]);
        al[k] = load_scalar<T>(alphas ? (const char *)alphas + k * sizeof(T) : nullptr, T{1});
        be[k] = load_scalar<T>(betas ? (const char *)betas + k * sizeof(T) : nullptr, T{0});
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<rank_data<T>> rd(P);
    for (int r = 0; r < P; ++r) {
        if (only_rank >= 0 && r != only_rank) continue;
        std::vector<message<T>> smsg, rmsg;
        for (int k = 0; k < n_pairs; ++k) {
            char t = trans ? trans[k] : 'N';
            // transform.cpp:151: init.transpose_or_conjugate(trans)
            assigned_grid2D g_init = gi[k];
            local_blocks<T> b_init = make_blocks<T>(&inits[k], g_init, r);
            if (t == 'T' || t == 'C') {
                g_init.transpose();
                b_init.transpose_or_conjugate(t);
            }
            local_blocks<T> b_fin = make_blocks<T>(&fins[k], gf[k], r);
            decompose(b_init, g_init, gf[k], al[k], be[k], k, smsg);
            decompose(b_fin, gf[k], g_init, al[k], be[k], k, rmsg);
        }
        std::sort(smsg.begin(), smsg.end());
        std::sort(rmsg.begin(), rmsg.end());
        rd[r].send = communication_data<T>(smsg, r, P);
        rd[r].recv = communication_data<T>(rmsg, r, P);
    }
    if (only_rank >= 0) {
        communication_data<T> &cd = side == 0 ? rd[only_rank].send : rd[only_rank].recv;
        for (int i = 0; i < P; ++i) {
            counts[i] = cd.counts[i];
            dspls[i] = cd.dspls[i];
        }
        *total_out = cd.total_size;
        if (pack_out && side == 0) {
            if (cd.total_size > pack_cap) return -1;
            cd.copy_to_buffer();
            std::memcpy(pack_out, cd.data(), sizeof(T) * (size_t)cd.total_size);
        }
        return 0;
    }
    for (int r = 0; r < P; ++r) rd[r].send.copy_to_buffer();
    for (int s = 0; s < P; ++s)
        for (int d = 0; d < P; ++d) {
            int cnt = rd[s].send.counts[d];
            if (!cnt) continue;
            if (rd[d].recv.counts[s] != cnt) return -1;
            std::memcpy(rd[d].recv.data() + rd[d].recv.dspls[s], rd[s].send.data() + rd[s].send.dspls[d],
                        sizeof(T) * (size_t)cnt);
        }
    for (int r = 0; r < P; ++r) copy_local_blocks(rd[r].send.local_messages, rd[r].recv.local_messages);
    for (int r = 0; r < P; ++r) rd[r].recv.copy_from_buffer();
    auto t1 = std::chrono::steady_clock::now();
    if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

template <typename F>
int dispatch(int dtype, F &&f) {
    try {
        switch (dtype) {
        case ORC_F32: return f(float{});
        case ORC_F64: return f(double{});
        case ORC_C64: return f(std::complex<float>{});
        case ORC_C128: return f(std::complex<double>{});
        }
    } catch (const std::exception &) {
        return -2;
    }
    return -1;
}

} // namespace

extern "C" {

// grid_cover.cpp:54-121, the reference's own function.
int ref_decomp_cover(const int32_t *blue, int nb, const int32_t *red, int nr, int32_t *out) {
    std::vector<int> b(blue, blue + nb), r(red, red + nr);
    std::vector<interval_cover> c = get_decomp_cover(b, r);
    for (size_t k = 0; k < c.size(); ++k) {
        out[2 * k] = c[k].start_index;
        out[2 * k + 1] = c[k].end_index;
    }
    return (int)c.size();
}

// Full pipeline for all virtual ranks (batched form when n_pairs > 1).
// seconds (may be NULL) receives the wall time of planning + copies.
int ref_transform(int n_pairs, const orc_layout *inits, const orc_layout *fins, int dtype,
                  const char *trans, const void *alphas, const void *betas, double *seconds) {
    return dispatch(dtype, [&](auto tag) {
        using T = decltype(tag);
        int64_t total = 0;
        return run<T>(n_pairs, inits, fins, trans, alphas, betas, -1, 0, nullptr, nullptr, nullptr, 0,
                      &total, seconds);
    });
}

// counts/dspls (elements) of one rank's send (side 0) or recv (side 1)
// communication_data, and for side 0 the packed send buffer.
int ref_plan_rank(const orc_layout *init, const orc_layout *fin, int dtype, char trans, int rank,
                  int side, int64_t *counts, int64_t *dspls, void *pack_out, int64_t pack_cap,
                  int64_t *total) {
    return dispatch(dtype, [&](auto tag) {
        using T = decltype(tag);
        return run<T>(1, init, fin, &trans, nullptr, nullptr, rank, side, counts, dspls, pack_out,
                      pack_cap, total, nullptr);
    });
}

// transform.cpp:9-39 (communication_volume): the loop of
// utils.hpp:91-139 (rank_to_comm_vol_for_block) restated on the reference's
// grid_cover::decompose_block and assigned_grid2D::owner / rows_interval /
// cols_interval, summed into the reference's comm_volume (edge_t keys), then
// written out densely (vol[a*n+b], a <= b).
int ref_comm_volume(const orc_layout *init, const orc_layout *fin, int64_t *vol) {
    try {
        assigned_grid2D gi = make_grid(init), gf = make_grid(fin);
        grid_cover cover(gi.grid(), gf.grid());
        std::unordered_map<edge_t, int> weights;
        for (int i = 0; i < gi.grid().n_rows; ++i) {
            for (int j = 0; j < gi.grid().n_cols; ++j) {
                block_cover bc = cover.decompose_block(block_coordinates{i, j});
                interval rows = gi.rows_interval(i), cols = gi.cols_interval(j);
                std::unordered_map<int, int> per_rank;
                int row_start = rows.start;
                for (int a = bc.rows_cover.start_index; a < bc.rows_cover.end_index; ++a) {
                    int row_end = std::min(gf.grid().rows_split[a + 1], rows.end);
                    int col_start = cols.start;
                    for (int b = bc.cols_cover.start_index; b < bc.cols_cover.end_index; ++b) {
                        int col_end = std::min(gf.grid().cols_split[b + 1], cols.end);
                        int size = (row_end - row_start) * (col_end - col_start);
                        if (size) per_rank[gf.owner(a, b)] += size;
                        col_start = col_end;
                    }
                    row_start = row_end;
                }
                const int rank = gi.owner(i, j);
                for (const auto &kv : per_rank)
                    weights[edge_t{rank, kv.first}.sorted()] += kv.second;
            }
        }
        comm_volume cv(std::move(weights));
        const int n = std::max(init->n_ranks, fin->n_ranks);
        std::memset(vol, 0, sizeof(int64_t) * (size_t)n * n);
        for (const auto &kv : cv.volume) vol[(size_t)kv.first.src * n + kv.first.dest] = kv.second;
        return 0;
    } catch (...) {
        return -1;
    }
}

// grid_layout::scale_by (grid_layout.hpp:48-56, restated: that header
// includes <mpi.h>): unless beta == 1, the reference's own block::scale_by
// (block.cpp:291-306) on every block of the layout.
int ref_scale_layout(const orc_layout *l, int dtype, const void *beta) {
    return dispatch(dtype, [&](auto tag) {
        using T = decltype(tag);
        const T b = load_scalar<T>(beta, T{1});
        if (b == T{1}) return 0;
        assigned_grid2D g = make_grid(l);
        for (int r = 0; r < l->n_ranks; ++r) {
            local_blocks<T> blocks = make_blocks<T>(l, g, r);
            for (int k = 0; k < blocks.num_blocks(); ++k) blocks.get_block(k).scale_by(b);
        }
        return 0;
    });
}

// ranks_reordering.cpp:4-61, the reference's own optimal_reordering, on a
// comm_volume holding the nonzero entries of the dense upper triangle.
int ref_optimal_reordering(const int64_t *vol, int n, int32_t *perm, int *reordered) {
    comm_volume cv;
    for (int a = 0; a < n; ++a)
        for (int b = a; b < n; ++b)
            if (vol[(size_t)a * n + b] > 0) cv.volume[edge_t{a, b}] = (int)vol[(size_t)a * n + b];
    bool r = false;
    std::vector<int> p = optimal_reordering(cv, n, r);
    for (int a = 0; a < n; ++a) perm[a] = p[a];
    *reordered = r ? 1 : 0;
    return 0;
}

} // extern "C"
