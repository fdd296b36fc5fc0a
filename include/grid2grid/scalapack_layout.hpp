// Block-cyclic (ScaLAPACK) layouts: the reference's get_scalapack_grid
// overloads and helper types (src/grid2grid/scalapack_layout.hpp:11-431,
// scalapack_layout.cpp:10-346), building the layout of one virtual rank whose
// local buffer `ptr` is device memory. Merge the P per-rank layouts with
// merge_ranks() (grid_layout.hpp) or pass them to transform() as a vector.
#pragma once

#include <grid2grid/grid_layout.hpp>

#include <cmath>
#include <cstdlib>
#include <initializer_list>
#include <iostream>
#include <ostream>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace grid2grid {
namespace scalapack {

enum ordering { row_major, column_major };

// Pair of ints with the reference's elementwise arithmetic and its
// componentwise comparisons (both components must compare).
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
    int_pair operator+(const int_pair &o) const { return {row + o.row, col + o.col}; }
    int_pair operator-(const int_pair &o) const { return {row - o.row, col - o.col}; }
    int_pair operator*(const int_pair &o) const { return {row * o.row, col * o.col}; }
    int_pair operator/(const int_pair &o) const { return {row / o.row, col / o.col}; }
    int_pair operator%(const int_pair &o) const { return {row % o.row, col % o.col}; }
    int_pair &operator+=(const int_pair &o) { return *this = *this + o; }
    int_pair &operator-=(const int_pair &o) { return *this = *this - o; }
    int_pair &operator*=(const int_pair &o) { return *this = *this * o; }
    int_pair &operator/=(const int_pair &o) { return *this = *this / o; }
    int_pair &operator%=(const int_pair &o) { return *this = *this % o; }
    bool operator<(const int_pair &o) const { return row < o.row && col < o.col; }
    bool operator<=(const int_pair &o) const { return row <= o.row && col <= o.col; }
    int &operator[](int i) { return i == 0 ? row : col; }
    int_pair &operator()(const int_pair &o) {
        row = o.row;
        col = o.col;
        return *this;
    }
    void transpose() { std::swap(row, col); }
    friend std::ostream &operator<<(std::ostream &os, const int_pair &p) {
        return os << "[" << p.row << ", " << p.col << "]";
    }
};

struct rank_grid_coord : int_pair {
    using int_pair::int_pair;
    rank_grid_coord() = default;
    rank_grid_coord(const int_pair &p) : int_pair(p) {}
};
struct rank_decomposition : int_pair {
    using int_pair::int_pair;
    rank_decomposition() = default;
    rank_decomposition(const int_pair &p) : int_pair(p) {}
    int n_total() const { return row * col; }
};
struct elem_grid_coord : int_pair {
    using int_pair::int_pair;
    elem_grid_coord() = default;
    elem_grid_coord(const int_pair &p) : int_pair(p) {}
};
struct block_dim : int_pair {
    using int_pair::int_pair;
    block_dim() = default;
    block_dim(const int_pair &p) : int_pair(p) {}
    int size() const { return row * col; }
};
struct matrix_dim : int_pair {
    using int_pair::int_pair;
    matrix_dim() = default;
    matrix_dim(const int_pair &p) : int_pair(p) {}
    int size() const { return row * col; }
};

// An element's local coordinates in its owner's buffer, with that owner's
// rank-grid coordinates.
struct local_grid_coord {
    elem_grid_coord el_coord;
    rank_grid_coord rank_coord;
    local_grid_coord() = default;
    local_grid_coord(elem_grid_coord e, rank_grid_coord r) : el_coord(e), rank_coord(r) {}
};

struct matrix_grid {
    matrix_dim matrix_dimension;
    block_dim block_dimension;
    matrix_grid() = default;
    matrix_grid(matrix_dim m, block_dim b) : matrix_dimension(m), block_dimension(b) {}
    void transpose() {
        matrix_dimension.transpose();
        block_dimension.transpose();
    }
};

// The whole blocks one rank owns, stored as one column-major buffer of
// n_blocks_row*mb rows (the padded local layout).
struct local_blocks {
    int n_blocks_row = 0;
    int n_blocks_col = 0;
    block_dim block_dimension;
    rank_grid_coord rank_coord;
    int stride = 0;
    size_t size_no_padding = 0;

    local_blocks() = default;
    local_blocks(int nbr, int nbc, block_dim b, rank_grid_coord r)
        : n_blocks_row(nbr), n_blocks_col(nbc), block_dimension(b), rank_coord(r) {
        size_no_padding = size_t(nbr) * nbc * b.size();
        stride = nbr * b.row;
    }
    // buffer index -> local coordinates
    local_grid_coord get_local_coordinates(int index) const {
        return {{index % stride, index / stride}, rank_coord};
    }
    size_t size_with_padding() const { return size_t(n_blocks_row) * n_blocks_col * block_dimension.size(); }
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

// With the process grid starting at `src` (scalapack_layout.cpp:26-38).
inline rank_grid_coord rank_to_grid(int rank, rank_decomposition g, ordering o, rank_grid_coord src) {
    return rank_grid_coord((rank_to_grid(rank, g, o) + src) % g);
}

// Inverse of the overload above. The reference declares this overload
// (scalapack_layout.hpp:273-276) without defining it.
inline int rank_from_grid(rank_grid_coord c, rank_decomposition g, ordering o, rank_grid_coord src) {
    return rank_from_grid(rank_grid_coord(((c - src) % g + g) % g), g, o);
}

// Global index -> (index in the owner's local buffer, owner's grid
// coordinate) along one dimension (scalapack_layout.cpp:59-70).
inline std::tuple<int, int> local_coordinate(int glob, int block, int p_block, int /*mat_dim*/) {
    const int blk = glob / block;
    return std::make_tuple((blk / p_block) * block + glob % block, blk % p_block);
}

// global -> local coordinates (scalapack_layout.cpp:72-89).
inline local_grid_coord local_coordinates(matrix_grid mg, rank_decomposition rg, elem_grid_coord g) {
    int row, p_row, col, p_col;
    std::tie(row, p_row) = local_coordinate(g.row, mg.block_dimension.row, rg.row, mg.matrix_dimension.row);
    std::tie(col, p_col) = local_coordinate(g.col, mg.block_dimension.col, rg.col, mg.matrix_dimension.row);
    return {{row, col}, {p_row, p_col}};
}

// local -> global coordinates (scalapack_layout.cpp:92-121); {-1, -1}
// outside the matrix. As in the reference, the row test is gi > m, so the
// padding row gi == m of a partial last block still maps to (m, gj).
inline elem_grid_coord global_coordinates(matrix_grid mg, rank_decomposition rg, local_grid_coord lc) {
    const int mb = mg.block_dimension.row, nb = mg.block_dimension.col;
    const int li = lc.el_coord.row, lj = lc.el_coord.col;
    const int gi = ((li / mb) * rg.row + lc.rank_coord.row) * mb + li % mb;
    const int gj = ((lj / nb) * rg.col + lc.rank_coord.col) * nb + lj % nb;
    if (gi < 0 || gi > mg.matrix_dimension.row || gj < 0 || gj >= mg.matrix_dimension.col) return {-1, -1};
    return {gi, gj};
}

// Whole blocks owned by the rank at `rc` (scalapack_layout.cpp:123-139).
inline local_blocks get_local_blocks(matrix_grid mg, rank_decomposition rg, rank_grid_coord rc) {
    const int nbr = (mg.matrix_dimension.row + mg.block_dimension.row - 1) / mg.block_dimension.row;
    const int nbc = (mg.matrix_dimension.col + mg.block_dimension.col - 1) / mg.block_dimension.col;
    const int own_r = nbr / rg.row + (rc.row < nbr % rg.row ? 1 : 0);
    const int own_c = nbc / rg.col + (rc.col < nbc % rg.col ? 1 : 0);
    return local_blocks(own_r, own_c, mg.block_dimension, rc);
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
    return get_local_blocks(matrix_grid(l.matrix_dimension, l.block_dimension), l.rank_grid, rc)
        .size_with_padding();
}

// Host buffer of `rank` filled with f(gi, gj) (value-initialised where the
// padded buffer lies outside the matrix), scalapack_layout.hpp:299-331.
template <typename T, typename Function>
void initialize_locally(T *buffer, Function f, int rank, data_layout &l) {
    const matrix_grid mg(l.matrix_dimension, l.block_dimension);
    const local_blocks lb =
        get_local_blocks(mg, l.rank_grid, rank_to_grid(rank, l.rank_grid, l.rank_grid_ordering));
    const size_t n = lb.size_with_padding();
    for (size_t i = 0; i < n; ++i) {
        const elem_grid_coord g = global_coordinates(mg, l.rank_grid, lb.get_local_coordinates((int)i));
        buffer[i] = (g.row > -1 && g.col > -1) ? (T)f(g.row, g.col) : T();
    }
}

// Checks a host buffer of `rank` against f within eps, reporting the first
// mismatches on stdout (scalapack_layout.hpp:333-386).
template <typename Function>
bool validate(Function f, std::vector<decltype(f(0, 0))> &buffer, int rank, data_layout &l, double eps = 1e-6) {
    using elem_type = decltype(f(0, 0));
    const matrix_grid mg(l.matrix_dimension, l.block_dimension);
    const local_blocks lb =
        get_local_blocks(mg, l.rank_grid, rank_to_grid(rank, l.rank_grid, l.rank_grid_ordering));
    bool ok = true;
    if (buffer.size() != lb.size_with_padding()) {
        std::cout << "Size of the buffer is " << buffer.size() << " and should be " << lb.size_with_padding()
                  << std::endl;
        ok = false;
    }
    for (size_t i = 0; i < buffer.size(); ++i) {
        const elem_grid_coord g = global_coordinates(mg, l.rank_grid, lb.get_local_coordinates((int)i));
        if (g.row < 0 || g.col < 0 || g.row >= l.matrix_dimension.row || g.col >= l.matrix_dimension.col) continue;
        const elem_type want = (elem_type)f(g.row, g.col);
        if (std::abs(buffer[i] - want) > eps) {
            std::cout << "(" << g.row << ", " << g.col << ") = " << buffer[i] << " instead of " << want << std::endl;
            ok = false;
        }
    }
    return ok;
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
