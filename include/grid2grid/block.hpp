// grid2grid::block<T>, local_blocks<T>: a column-major block of device memory
// and the blocks of a layout. Names, constructors and flags follow the
// reference (src/grid2grid/block.hpp:23-190, block.cpp:23-348); `data` is a
// device pointer here, so local_element returns a proxy that reads and
// writes the element in device memory, and scale_by runs on the device.
#pragma once

#include <grid2grid_amd.h>

#include <grid2grid/grid2D.hpp>
#include <grid2grid/interval.hpp>

#include <algorithm>
#include <cctype>
#include <complex>
#include <cstddef>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace grid2grid {

namespace detail {

template <typename T> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = G2G_FLOAT32; };
template <> struct dtype_of<double> { static constexpr int value = G2G_FLOAT64; };
template <> struct dtype_of<std::complex<float>> { static constexpr int value = G2G_COMPLEX64; };
template <> struct dtype_of<std::complex<double>> { static constexpr int value = G2G_COMPLEX128; };

inline void check(int rc) {
    if (rc != G2G_OK) throw std::runtime_error(std::string("grid2grid: ") + g2g_last_error());
}

}  // namespace detail

// One element of device memory: converts to T by reading it and takes a T by
// writing it (each a synchronous copy), as block::local_element's reference
// does on host memory.
template <typename T>
class device_ref {
  public:
    explicit device_ref(T *p) : p_(p) {}
    operator T() const {
        T v;
        detail::check(g2g_copy_bytes(&v, p_, sizeof(T)));
        return v;
    }
    device_ref &operator=(const T &v) {
        detail::check(g2g_copy_bytes(p_, &v, sizeof(T)));
        return *this;
    }
    device_ref &operator=(const device_ref &o) { return *this = T(o); }
    T *address() const { return p_; }

  private:
    T *p_;
};

struct block_coordinates {
    int row = 0;
    int col = 0;
    block_coordinates() = default;
    block_coordinates(int r, int c) : row(r), col(c) {}
    void transpose() { std::swap(row, col); }
};

struct block_range {
    interval rows_interval;
    interval cols_interval;
    block_range() = default;
    block_range(interval r, interval c) : rows_interval(r), cols_interval(c) {}
    block_range intersection(const block_range &o) const {
        return {rows_interval.intersection(o.rows_interval), cols_interval.intersection(o.cols_interval)};
    }
    // As in the reference (block.cpp:35-51), including its column test, which
    // compares cols_interval.end <= range.cols_interval.start twice.
    bool outside_of(const block_range &range) const {
        return (rows_interval.end <= range.rows_interval.start ||
                rows_interval.start >= range.rows_interval.end) &&
               (cols_interval.end <= range.cols_interval.start || cols_interval.end <= range.cols_interval.start);
    }
    bool inside(const block_range &range) const {
        return range.rows_interval.start < rows_interval.start && range.rows_interval.end > rows_interval.end &&
               range.cols_interval.start < cols_interval.start && range.cols_interval.end > cols_interval.end;
    }
    bool intersects(const block_range &range) const { return !outside_of(range) && !inside(range); }
    bool non_empty() const { return rows_interval.non_empty() && cols_interval.non_empty(); }
    bool empty() const { return rows_interval.empty() || cols_interval.empty(); }
    bool operator==(const block_range &o) const {
        return empty() ? o.empty() : (rows_interval == o.rows_interval && cols_interval == o.cols_interval);
    }
    bool operator!=(const block_range &o) const { return !(*this == o); }
};

template <typename T>
struct block {
    int tag = 0;
    interval rows_interval;
    interval cols_interval;
    bool transpose_on_copy = false;
    bool conjugate_on_copy = false;
    block_coordinates coordinates;
    T *data = nullptr;  // device pointer to element (0, 0)
    int stride = 0;     // leading dimension (elements)

    block() = default;
    block(const assigned_grid2D &g, block_coordinates c, T *ptr, int ld)
        : rows_interval(g.rows_interval(c.row)), cols_interval(g.cols_interval(c.col)), coordinates(c),
          data(ptr), stride(ld) {}
    block(const assigned_grid2D &g, block_coordinates c, T *ptr)
        : block(g, c, ptr, g.rows_interval(c.row).length()) {}
    block(interval r, interval c, block_coordinates coord, T *ptr, int ld)
        : rows_interval(r), cols_interval(c), coordinates(coord), data(ptr), stride(ld) {}
    block(interval r, interval c, block_coordinates coord, T *ptr) : block(r, c, coord, ptr, r.length()) {}
    block(block_range &rg, block_coordinates coord, T *ptr, int ld)
        : block(rg.rows_interval, rg.cols_interval, coord, ptr, ld) {}
    block(block_range &rg, block_coordinates coord, T *ptr) : block(rg.rows_interval, rg.cols_interval, coord, ptr) {}
    // coordinates looked up in the grid from the intervals' starts
    block(const assigned_grid2D &g, interval r, interval c, T *ptr, int ld)
        : rows_interval(r), cols_interval(c), data(ptr), stride(ld) {
        coordinates = block_coordinates(interval_index(g.grid().rows_split, r), interval_index(g.grid().cols_split, c));
    }
    block(const assigned_grid2D &g, interval r, interval c, T *ptr) : block(g, r, c, ptr, r.length()) {}
    block(const assigned_grid2D &g, block_range &rg, T *ptr, int ld)
        : block(g, rg.rows_interval, rg.cols_interval, ptr, ld) {}
    block(const assigned_grid2D &g, block_range &rg, T *ptr) : block(g, rg.rows_interval, rg.cols_interval, ptr) {}

    static int interval_index(const std::vector<int> &splits, interval inter) {
        return int(std::lower_bound(splits.begin(), splits.end(), inter.start) - splits.begin());
    }

    // The piece [r_range) x [c_range) of this block; pointer arithmetic in
    // storage orientation as in the reference's subblock (block.cpp:165-196).
    block<T> subblock(interval r_range, interval c_range) const {
        if (!rows_interval.contains(r_range) || !cols_interval.contains(c_range))
            throw std::runtime_error("ERROR: current block does not contain requested subblock.");
        interval ri = rows_interval, ci = cols_interval;
        block_coordinates coord = coordinates;
        if (transpose_on_copy) {
            std::swap(r_range, c_range);
            std::swap(ri, ci);
            coord.transpose();
        }
        T *ptr = data + std::ptrdiff_t(c_range.start - ci.start) * stride + (r_range.start - ri.start);
        block<T> b(r_range, c_range, coord, ptr, stride);
        b.transpose_or_conjugate(conjugate_on_copy ? 'C' : (transpose_on_copy ? 'T' : 'N'));
        b.tag = tag;
        return b;
    }

    bool non_empty() const { return rows_interval.non_empty() && cols_interval.non_empty(); }

    bool operator<(const block &o) const {
        const bool cols_less = cols_interval < o.cols_interval, cols_eq = cols_interval == o.cols_interval;
        const bool rows_less = rows_interval < o.rows_interval, rows_eq = rows_interval == o.rows_interval;
        return cols_less || (cols_eq && rows_less) || (cols_eq && rows_eq && tag < o.tag);
    }

    int n_rows() const { return rows_interval.length(); }
    int n_cols() const { return cols_interval.length(); }
    std::pair<int, int> size() const { return {n_rows(), n_cols()}; }
    size_t total_size() const { return size_t(n_rows()) * size_t(n_cols()); }

    // Element (li, lj) of the block, local coordinates, swapped first for a
    // transposed block (block.cpp:228-247): its device address, and a proxy
    // that reads or writes it.
    T *local_element_address(int li, int lj) const {
        if (transpose_on_copy) std::swap(li, lj);
        return data + std::ptrdiff_t(stride) * lj + li;
    }
    device_ref<T> local_element(int li, int lj) const { return device_ref<T>(local_element_address(li, lj)); }

    // Every element times beta on the device (block.cpp:291-306); nothing
    // when beta == 1. A block flagged for transposition is refused (the
    // reference asserts it is not).
    void scale_by(T beta, void *stream = nullptr) {
        if (beta == T{1}) return;
        if (transpose_on_copy) throw std::runtime_error("grid2grid: scale_by of a block flagged for transposition");
        const int32_t rs[2] = {rows_interval.start, rows_interval.end};
        const int32_t cs[2] = {cols_interval.start, cols_interval.end};
        const int32_t owner = 0;
        const g2g_block b{0, 0, stride, (void *)data};
        const g2g_layout l{2, 2, rs, cs, &owner, 1, 1, &b};
        detail::check(g2g_scale_layout(&l, detail::dtype_of<T>::value, &beta, stream));
    }

    // flags only: the data moves during the transform
    void transpose_or_conjugate(char flag) {
        flag = char(std::toupper((unsigned char)flag));
        if (flag == 'N') return;
        std::swap(rows_interval, cols_interval);
        coordinates.transpose();
        if (flag == 'T' || flag == 'C') transpose_on_copy = true;
        if (flag == 'C') conjugate_on_copy = true;
    }
};

template <typename T>
bool operator==(const block<T> &a, const block<T> &b) noexcept {
    return a.rows_interval.start == b.rows_interval.start && a.rows_interval.end == b.rows_interval.end &&
           a.cols_interval.start == b.cols_interval.start && a.cols_interval.end == b.cols_interval.end &&
           a.coordinates.row == b.coordinates.row && a.coordinates.col == b.coordinates.col &&
           a.stride == b.stride && a.data == b.data && a.transpose_on_copy == b.transpose_on_copy &&
           a.conjugate_on_copy == b.conjugate_on_copy;
}

template <typename T>
std::ostream &operator<<(std::ostream &os, const block<T> &b) {
    return os << "rows: " << b.rows_interval << "cols: " << b.cols_interval << std::endl;
}

template <typename T>
class local_blocks {
  public:
    local_blocks() = default;
    local_blocks(std::vector<block<T>> &&b) : blocks_(std::move(b)) {
        for (const auto &x : blocks_) total_size_ += x.total_size();
    }
    block<T> &get_block(int i) { return blocks_[i]; }
    const block<T> &get_block(int i) const { return blocks_[i]; }
    int num_blocks() const { return int(blocks_.size()); }
    size_t size() const { return total_size_; }
    void transpose_or_conjugate(char flag) {
        for (auto &b : blocks_) b.transpose_or_conjugate(flag);
    }
    // appends another rank's blocks (all virtual ranks share one process)
    void append(const local_blocks &o) {
        blocks_.insert(blocks_.end(), o.blocks_.begin(), o.blocks_.end());
        total_size_ += o.total_size_;
    }
    friend bool operator==(const local_blocks &a, const local_blocks &b) noexcept {
        return a.blocks_ == b.blocks_ && a.total_size_ == b.total_size_;
    }

  private:
    std::vector<block<T>> blocks_;
    size_t total_size_ = 0;
};

}  // namespace grid2grid
