// grid2grid::grid2D and assigned_grid2D: a grid of row/column splits and the
// virtual rank owning each cell. Names and semantics follow the reference
// (src/grid2grid/grid2D.hpp:17-125, grid2D.cpp:5-126).
#pragma once

#include <grid2grid/interval.hpp>

#include <cassert>
#include <ostream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace grid2grid {

struct grid2D {
    int n_rows = 0;  // blocks along rows
    int n_cols = 0;  // blocks along columns
    std::vector<int> rows_split;
    std::vector<int> cols_split;

    grid2D() = default;
    grid2D(std::vector<int> &&r, std::vector<int> &&c)
        : n_rows(r.empty() ? 0 : int(r.size()) - 1),
          n_cols(c.empty() ? 0 : int(c.size()) - 1),
          rows_split(std::move(r)),
          cols_split(std::move(c)) {}

    interval row_interval(int index) const {
        if (index < 0 || size_t(index) + 1 >= rows_split.size())
            throw std::runtime_error("ERROR: in class grid2D, row index out of range.");
        return {rows_split[index], rows_split[index + 1]};
    }
    interval col_interval(int index) const {
        if (index < 0 || size_t(index) + 1 >= cols_split.size())
            throw std::runtime_error("ERROR: in class grid2D, col index out of range.");
        return {cols_split[index], cols_split[index + 1]};
    }
    void transpose() {
        std::swap(rows_split, cols_split);
        std::swap(n_rows, n_cols);
    }
};

class assigned_grid2D {
  public:
    assigned_grid2D() = default;
    assigned_grid2D(grid2D &&g, std::vector<std::vector<int>> &&owners, int n_ranks)
        : g_(std::move(g)), ranks_(std::move(owners)), n_ranks_(n_ranks) {}

    // owner of cell (i, j), after the optional rank relabelling
    int owner(int i, int j) const { return reordered_rank(ranks_[i][j]); }
    const grid2D &grid() const { return g_; }
    int num_ranks() const { return n_ranks_; }
    interval rows_interval(int index) const { return g_.row_interval(index); }
    interval cols_interval(int index) const { return g_.col_interval(index); }
    int block_size(int row_index, int col_index) {
        return rows_interval(row_index).length() * cols_interval(col_index).length();
    }

    void transpose() {
        g_.transpose();
        const size_t m = ranks_.size(), n = m ? ranks_[0].size() : 0;
        std::vector<std::vector<int>> t(n, std::vector<int>(m));
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) t[j][i] = ranks_[i][j];
        ranks_ = std::move(t);
    }

    void reorder_ranks(std::vector<int> &reordering) { reordering_ = reordering; }
    int reordered_rank(int rank) const { return reordering_.empty() ? rank : reordering_[rank]; }
    bool ranks_reordered() const { return !reordering_.empty(); }

    int num_blocks_row() const noexcept { return g_.n_rows; }
    int num_blocks_col() const noexcept { return g_.n_cols; }
    int num_rows() const noexcept { return g_.rows_split.back(); }
    int num_cols() const noexcept { return g_.cols_split.back(); }

    friend bool operator==(const assigned_grid2D &a, const assigned_grid2D &b) noexcept {
        return a.g_.rows_split == b.g_.rows_split && a.g_.cols_split == b.g_.cols_split && a.ranks_ == b.ranks_;
    }
    friend std::ostream &operator<<(std::ostream &os, const assigned_grid2D &g) {
        for (int i = 0; i < g.num_blocks_row(); ++i)
            for (int j = 0; j < g.num_blocks_col(); ++j)
                os << "block (" << i << ", " << j << ") owned by " << g.owner(i, j) << std::endl;
        return os;
    }

  private:
    grid2D g_;
    std::vector<std::vector<int>> ranks_;
    int n_ranks_ = 0;
    std::vector<int> reordering_;
};

}  // namespace grid2grid
