// grid2grid::interval -- [start, end) on one matrix dimension.
// Same names, construction checks and ordering as the reference's interval
// (src/grid2grid/interval.hpp:10-41, interval.cpp:5-84).
#pragma once

#include <algorithm>
#include <ostream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace grid2grid {

struct interval {
    int start = 0;
    int end = 0;

    interval() = default;
    interval(int s, int e) : start(s), end(e) {
        if (s < 0 || e < 0)
            throw std::runtime_error("ERROR: in class interval (grid2grid): start, end > 0 must be satisfied.");
        if (s > e) throw std::runtime_error("ERROR: in class interval (grid2grid): start<=end must be satisfied.");
    }

    int length() const { return end - start; }
    bool contains(interval o) const { return start <= o.start && end >= o.end; }
    bool non_empty() const { return end > start; }
    bool empty() const { return end == start; }

    interval intersection(const interval &o) const {
        if (!non_empty() || !o.non_empty() || start >= o.end || end <= o.start) return {};
        return {std::max(start, o.start), std::min(end, o.end)};
    }

    // indices [s, e) of the intervals of split vector v overlapping *this:
    // s = max i with v[i] <= start, e = min i with v[i] >= end; (-1, -1) if none.
    std::pair<int, int> overlapping_intervals(const std::vector<int> &v) const {
        if (v.empty() || start >= end || start >= v.back() || end <= v.front()) return {-1, -1};
        const int s = int(std::upper_bound(v.begin(), v.end(), start) - v.begin()) - 1;
        const int e = int(std::lower_bound(v.begin(), v.end(), end) - v.begin());
        return {s, e};
    }

    bool operator==(const interval &o) const { return empty() ? o.empty() : (start == o.start && end == o.end); }
    bool operator!=(const interval &o) const { return !(*this == o); }
    bool operator<(const interval &o) const { return start < o.start || (start == o.start && end < o.end); }
};

inline std::ostream &operator<<(std::ostream &os, const interval &i) {
    return os << "interval[" << i.start << ", " << i.end << ")" << std::endl;
}

}  // namespace grid2grid
