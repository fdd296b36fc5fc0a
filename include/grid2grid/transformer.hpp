// grid2grid::transformer<T>: queue several (from, to, trans, alpha, beta)
// redistributions and run them together (reference
// src/grid2grid/transformer.hpp:8-62); here they run in one persistent launch.
#pragma once

#include <grid2grid/transform.hpp>

#include <vector>

namespace grid2grid {

template <typename T>
struct transformer {
    std::vector<layout_ref<T>> from;
    std::vector<layout_ref<T>> to;
    std::vector<T> alpha;
    std::vector<T> beta;
    std::vector<char> transpose;
    virtual_comm comm;

    transformer() = default;
    explicit transformer(const virtual_comm &c) : comm(c) {}

    void schedule(grid_layout<T> &from_layout, grid_layout<T> &to_layout) {
        schedule(from_layout, to_layout, 'N', T{1}, T{0});
    }
    void schedule(grid_layout<T> &from_layout, grid_layout<T> &to_layout, const char trans, const T a,
                  const T b) {
        from.push_back(from_layout);
        to.push_back(to_layout);
        transpose.push_back(trans);
        alpha.push_back(a);
        beta.push_back(b);
    }
    void transform() {
        if (!from.empty())
            grid2grid::transform<T>(from, to, transpose.data(), alpha.data(), beta.data(), comm);
        clear();
    }
    void clear() {
        from.clear();
        to.clear();
        alpha.clear();
        beta.clear();
        transpose.clear();
    }
};

}  // namespace grid2grid
