// Communication volume between ranks: the reference's edge_t,
// weighted_edge_t and comm_volume (src/grid2grid/comm_volume.hpp:6-141),
// same names, members and operators, so code written against the reference
// compiles unchanged.
#pragma once

#include <algorithm>
#include <cstddef>
#include <functional>
#include <ostream>
#include <unordered_map>

namespace grid2grid {

// A pair of ranks; sorted() gives the unordered pair as {min, max}.
struct edge_t {
    int src = 0;
    int dest = 0;

    edge_t() = default;
    edge_t(int s, int d) : src(s), dest(d) {}

    edge_t sorted() const { return edge_t{std::min(src, dest), std::max(src, dest)}; }
    bool operator==(const edge_t &o) const { return src == o.src && dest == o.dest; }
};

}  // namespace grid2grid

namespace std {
template <>
struct hash<grid2grid::edge_t> {
    size_t operator()(const grid2grid::edge_t &e) const noexcept {
        const unsigned long long k = ((unsigned long long)(unsigned)e.src << 32) | (unsigned)e.dest;
        return std::hash<unsigned long long>()(k);
    }
};
}  // namespace std

namespace grid2grid {

// An edge with its weight; ordered by weight alone, as in the reference.
struct weighted_edge_t {
    edge_t e;
    int w = 0;

    weighted_edge_t() = default;
    weighted_edge_t(int src, int dest, int weight) : e{src, dest}, w(weight) {}

    int weight() const { return w; }
    int src() const { return e.src; }
    int dest() const { return e.dest; }
    const edge_t &edge() const { return e; }

    bool operator==(const weighted_edge_t &o) const { return e == o.e && w == o.w; }
    bool operator<(const weighted_edge_t &o) const { return w < o.w; }
};

// Elements exchanged per unordered pair of ranks ({a, a}: kept by rank a).
struct comm_volume {
    using volume_t = std::unordered_map<edge_t, int>;
    volume_t volume;

    comm_volume() = default;
    comm_volume(volume_t &&v) : volume(std::move(v)) {}

    comm_volume &operator+=(const comm_volume &other) {
        for (const auto &kv : other.volume) volume[kv.first.sorted()] += kv.second;
        return *this;
    }

    comm_volume operator+(const comm_volume &other) const {
        comm_volume sum;
        for (const auto &kv : volume) sum.volume[kv.first.sorted()] += kv.second;
        sum += other;
        return sum;
    }

    // elements that move between different ranks
    size_t total_volume() const {
        size_t s = 0;
        for (const auto &kv : volume)
            if (kv.first.src != kv.first.dest) s += (size_t)kv.second;
        return s;
    }

    friend std::ostream &operator<<(std::ostream &os, const comm_volume &v) {
        os << "Communication volume consists of the following:" << std::endl;
        for (const auto &kv : v.volume) os << kv.first.src << "->" << kv.first.dest << ": " << kv.second << std::endl;
        return os;
    }
};

}  // namespace grid2grid
