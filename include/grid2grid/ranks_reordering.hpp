// grid2grid::optimal_reordering (src/grid2grid/ranks_reordering.hpp:7,
// ranks_reordering.cpp:4-61) through g2g_optimal_reordering: a greedy
// matching of ranks on the communication volume; the returned permutation is
// meant for assigned_grid2D::reorder_ranks. Equal weights are taken in
// ascending (src, dest) order, where the reference's order depends on its
// hash map.
#pragma once

#include <grid2grid_amd.h>

#include <grid2grid/comm_volume.hpp>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace grid2grid {

inline std::vector<int> optimal_reordering(comm_volume &volume, int n_ranks, bool &reordered) {
    std::vector<g2g_edge> edges;
    edges.reserve(volume.volume.size());
    for (const auto &kv : volume.volume) edges.push_back(g2g_edge{kv.first.src, kv.first.dest, kv.second});
    std::vector<int32_t> perm(n_ranks > 0 ? n_ranks : 0);
    int32_t r = 0;
    if (g2g_optimal_reordering(edges.data(), (int64_t)edges.size(), n_ranks, perm.data(), &r) != G2G_OK)
        throw std::runtime_error(std::string("grid2grid: ") + g2g_last_error());
    reordered = r != 0;
    return std::vector<int>(perm.begin(), perm.end());
}

}  // namespace grid2grid
