// grid2grid::grid_layout<T>: an assigned grid plus the blocks of its virtual
// ranks (reference src/grid2grid/grid_layout.hpp:7-66). With P virtual ranks
// on one GPU, one layout may hold the blocks of every rank; layouts built per
// rank (as each MPI rank builds its own in the reference) are merged with
// merge_ranks().
#pragma once

#include <grid2grid/block.hpp>
#include <grid2grid/grid2D.hpp>

#include <cctype>
#include <functional>
#include <stdexcept>
#include <vector>

namespace grid2grid {

template <typename T>
class grid_layout {
  public:
    grid_layout() = default;
    grid_layout(assigned_grid2D &&g, local_blocks<T> &&b) : grid(std::move(g)), blocks(std::move(b)) {}

    int num_ranks() const { return grid.num_ranks(); }

    // transpose ('T') or conjugate-transpose ('C') the description; no data moves
    void transpose_or_conjugate(char flag) {
        flag = char(std::toupper((unsigned char)flag));
        if (flag != 'N' && flag != 'T' && flag != 'C')
            throw std::runtime_error("transpose_or_conjugate: flag must be 'N', 'T' or 'C'");
        if (flag == 'T' || flag == 'C') {
            grid.transpose();
            blocks.transpose_or_conjugate(flag);
        }
    }

    void reorder_ranks(std::vector<int> &reordering) { grid.reorder_ranks(reordering); }
    int reordered_rank(int rank) const { return grid.reordered_rank(rank); }
    bool ranks_reordered() { return grid.ranks_reordered(); }

    int num_cols() const noexcept { return grid.num_cols(); }
    int num_rows() const noexcept { return grid.num_rows(); }
    int num_blocks_col() const noexcept { return grid.num_blocks_col(); }
    int num_blocks_row() const noexcept { return grid.num_blocks_row(); }

    // Every element of every block times beta on the device, one launch for
    // the whole layout (grid_layout.hpp:48-56); nothing when beta == 1.
    // Blocks flagged for transposition are refused (block.cpp:294).
    void scale_by(const T beta, void *stream = nullptr) {
        if (beta == T{1}) return;
        const grid2D &g = grid.grid();
        std::vector<int32_t> rs(g.rows_split.begin(), g.rows_split.end());
        std::vector<int32_t> cs(g.cols_split.begin(), g.cols_split.end());
        std::vector<int32_t> owners(size_t(g.n_rows) * g.n_cols, 0);
        std::vector<g2g_block> bl;
        for (int k = 0; k < blocks.num_blocks(); ++k) {
            const block<T> &b = blocks.get_block(k);
            if (b.transpose_on_copy)
                throw std::runtime_error("grid2grid: scale_by of a block flagged for transposition");
            const int r = b.coordinates.row, c = b.coordinates.col;
            if (r < 0 || r >= g.n_rows || c < 0 || c >= g.n_cols || b.rows_interval != g.row_interval(r) ||
                b.cols_interval != g.col_interval(c))
                throw std::runtime_error("grid2grid: a block does not match its grid cell");
            bl.push_back(g2g_block{r, c, b.stride, (void *)b.data});
        }
        const g2g_layout l{int32_t(rs.size()), int32_t(cs.size()), rs.data(), cs.data(), owners.data(), 1,
                           int32_t(bl.size()), bl.data()};
        detail::check(g2g_scale_layout(&l, detail::dtype_of<T>::value, &beta, stream));
    }

    assigned_grid2D grid;
    local_blocks<T> blocks;
};

template <typename T>
using layout_ref = std::reference_wrapper<grid_layout<T>>;

// One layout holding the blocks of every per-rank layout (same grid).
template <typename T>
grid_layout<T> merge_ranks(const std::vector<grid_layout<T>> &per_rank) {
    if (per_rank.empty()) throw std::runtime_error("merge_ranks: no layouts");
    grid_layout<T> out;
    out.grid = per_rank[0].grid;
    for (const auto &l : per_rank) {
        if (!(l.grid == out.grid)) throw std::runtime_error("merge_ranks: layouts describe different grids");
        out.blocks.append(l.blocks);
    }
    return out;
}

}  // namespace grid2grid
