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
