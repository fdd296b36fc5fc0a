// Block-cyclic (ScaLAPACK) layouts: the reference's get_scalapack_grid
// overloads and helper types (src/grid2grid/scalapack_layout.hpp:11-431,
// scalapack_layout.cpp:10-346), building the layout of one virtual rank whose
// local buffer `ptr` is device memory. Merge the P per-rank layouts with
// merge_ranks() (grid_layout.hpp) or pass them to transform() as a vector.
#pragma once

#include <grid2grid/grid_layout.hpp>

#include <cmath>
#include <initializer_list>
#include <stdexcept>
#include <vector>

namespace grid2grid {
namespace scalapack {

enum ordering { row_major, column_major };

struct int_pair {
    int row = 0;
    int col = 0;
    int_pair() = default;
    int_pair(int r, int c) : row(r), col(c) {}
    int_pair(std::initializer_list<int> l) {
        if (l.size() != 2) throw std::runtime_error("int_pair needs two values");
        row = *l.begin();
        col = *(l.begin() + 1);
    }
    void transpose() { std::swap(row, col); }
};

struct rank_grid_coord : int_pair { using int_pair::int_pair; rank_grid_coord() = default; };
struct rank_decomposition : int_pair {
    using int_pair::int_pair;
    rank_decomposition() = default;
    int n_total() const { return row * col; }
};
struct elem_grid_coord : int_pair { using int_pair::int_pair; elem_grid_coord() = default; };
struct block_dim : int_pair {
    using int_pair::int_pair;
    block_dim() = default;
    int size() const { return row * col; }
};
struct matrix_dim : int_pair {
    using int_pair::int_pair;
    matrix_dim() = default;
    int size() const { return row * col; }
};

struct data_layout {
    matrix_dim matrix_dimension;
    block_dim block_dimension;
    rank_decomposition rank_grid;
    ordering rank_grid_ordering = column_major;
    data_layout() = default;
    data_layout(matrix_dim m, block_dim b, rank_decomposition r, ordering o)
        : matrix_dimension(m), block_dimension(b), rank_grid(r), rank_grid_ordering(o) {}
};

inline rank_grid_coord rank_to_grid(int rank, rank_decomposition g, ordering o) {
    if (rank < 0 || rank >= g.n_total())
        throw std::runtime_error("Error in rank_to_grid: rank does not belong to the grid.");
    return o == column_major ? rank_grid_coord{rank % g.row, rank / g.row}
                             : rank_grid_coord{rank / g.col, rank % g.col};
}

inline int rank_from_grid(rank_grid_coord c, rank_decomposition g, ordering o) {
    if (c.row < 0 || c.row >= g.row || c.col < 0 || c.col >= g.col)
        throw std::runtime_error("Error in rank_from_grid: rank coordinates do not belong to the rank grid.");
    return o == column_major ? c.col * g.row + c.row : c.row * g.col + c.col;
}

// Splits of [begin, end) at multiples of blk_len, relative to begin.
inline std::vector<int> line_split(int begin, int end, int blk_len) {
    const int len = end - begin;
    const int first = blk_len - begin % blk_len;  // rows up to the first block border
    std::vector<int> s{0};
    if (first >= len) {
        s.push_back(len);
        return s;
    }
    for (int pos = first; pos < len; pos += blk_len) s.push_back(pos);
    s.push_back(len);
    return s;
}

// Elements of rank `rank`'s local buffer in the padded layout (whole blocks).
inline size_t local_size(int rank, data_layout &l) {
    const rank_grid_coord rc = rank_to_grid(rank, l.rank_grid, l.rank_grid_ordering);
    const int nbr = (l.matrix_dimension.row + l.block_dimension.row - 1) / l.block_dimension.row;
    const int nbc = (l.matrix_dimension.col + l.block_dimension.col - 1) / l.block_dimension.col;
    const int own_r = nbr / l.rank_grid.row + (rc.row < nbr % l.rank_grid.row ? 1 : 0);
    const int own_c = nbc / l.rank_grid.col + (rc.col < nbc % l.rank_grid.col ? 1 : 0);
    return size_t(own_r) * own_c * l.block_dimension.size();
}

}  // namespace scalapack

// Layout of the submatrix A(ia:ia+subm-1, ja:ja+subn-1) (1-based) of an m x n
// block-cyclic matrix as seen by `rank`, whose local buffer is `ptr` with
// leading dimension lld.
template <typename T>
grid_layout<T> get_scalapack_grid(int lld, scalapack::matrix_dim m_dim, scalapack::elem_grid_coord ij,
                                  scalapack::matrix_dim subm_dim, scalapack::block_dim b_dim,
                                  scalapack::rank_decomposition r_grid, scalapack::ordering ord,
                                  scalapack::rank_grid_coord src, T *ptr, const int rank) {
    (void)m_dim;
    if (ij.row < 1 || ij.col < 1) throw std::runtime_error("get_scalapack_grid: ia, ja are 1-based");
    const int i0 = ij.row - 1, j0 = ij.col - 1;
    std::vector<int> rs = scalapack::line_split(i0, i0 + subm_dim.row, b_dim.row);
    std::vector<int> cs = scalapack::line_split(j0, j0 + subm_dim.col, b_dim.col);
    const int gr = int(rs.size()) - 1, gc = int(cs.size()) - 1;
    const int br0 = i0 / b_dim.row, bc0 = j0 / b_dim.col;  // first (border) block of the submatrix
    const int sr = (br0 % r_grid.row + src.row) % r_grid.row;
    const int sc = (bc0 % r_grid.col + src.col) % r_grid.col;
    std::vector<std::vector<int>> owners(gr, std::vector<int>(gc));
    std::vector<block<T>> mine;
    for (int j = 0; j < gc; ++j) {
        const int pc = (j % r_grid.col + sc) % r_grid.col;
        for (int i = 0; i < gr; ++i) {
            const int pr = (i % r_grid.row + sr) % r_grid.row;
            owners[i][j] = scalapack::rank_from_grid({pr, pc}, r_grid, ord);
            if (owners[i][j] != rank) continue;
            // local position: whole local blocks before it plus the offset of
            // the submatrix start inside its (border) block
            const long lrow = long((br0 + i) / r_grid.row) * b_dim.row + (i0 + rs[i] - (br0 + i) * b_dim.row);
            const long lcol = long((bc0 + j) / r_grid.col) * b_dim.col + (j0 + cs[j] - (bc0 + j) * b_dim.col);
            mine.push_back(block<T>({rs[i], rs[i + 1]}, {cs[j], cs[j + 1]}, {i, j}, ptr + lrow + long(lld) * lcol, lld));
        }
    }
    grid2D g(std::move(rs), std::move(cs));
    return grid_layout<T>(assigned_grid2D(std::move(g), std::move(owners), r_grid.n_total()),
                          local_blocks<T>(std::move(mine)));
}

template <typename T>
grid_layout<T> get_scalapack_grid(int lld, scalapack::matrix_dim m_dim, scalapack::elem_grid_coord ij,
                                  scalapack::matrix_dim subm_dim, scalapack::block_dim b_dim,
                                  scalapack::rank_decomposition r_grid, scalapack::ordering ord,
                                  scalapack::rank_grid_coord src, const T *ptr, const int rank) {
    return get_scalapack_grid<T>(lld, m_dim, ij, subm_dim, b_dim, r_grid, ord, src, const_cast<T *>(ptr), rank);
}

// Whole matrix, no submatrix; lld padded to whole owned blocks.
template <typename T>
grid_layout<T> get_scalapack_grid(scalapack::matrix_dim m_dim, scalapack::block_dim b_dim,
                                  scalapack::rank_decomposition r_grid, scalapack::ordering ord, T *ptr,
                                  int rank) {
    const int nbr = (m_dim.row + b_dim.row - 1) / b_dim.row;
    const scalapack::rank_grid_coord rc = scalapack::rank_to_grid(rank, r_grid, ord);
    const int own_r = nbr / r_grid.row + (rc.row < nbr % r_grid.row ? 1 : 0);
    return get_scalapack_grid<T>(own_r * b_dim.row, m_dim, {1, 1}, m_dim, b_dim, r_grid, ord, {0, 0}, ptr, rank);
}

template <typename T>
grid_layout<T> get_scalapack_grid(scalapack::data_layout &l, T *ptr, int rank) {
    return get_scalapack_grid<T>(l.matrix_dimension, l.block_dimension, l.rank_grid, l.rank_grid_ordering, ptr,
                                 rank);
}

}  // namespace grid2grid
