#pragma once
// Drop-in replacement of the reference's include/htrip/bfs_oracle.hpp: the
// parity ground truth the reference's own test programs check the index
// against (src/bfs_oracle.cpp).  Header-only, CPU.  Hop distances over the
// observed passable cells, 4-connected, capped at INF; INF for an impassable
// or out-of-grid endpoint, 0 from a passable cell to itself.
#include <algorithm>
#include <vector>

#include "grid_map.hpp"
#include "types.hpp"

namespace htrip {

class BfsOracle {
public:
    // src/bfs_oracle.cpp:11-54
    [[nodiscard]] static dist_t queryDistance(const GridMap& grid, Point s, Point t) {
        return queryOneToAll(grid, s, {t})[0];
    }

    // src/bfs_oracle.cpp:72-181: one BFS from s; each target's distance.
    [[nodiscard]] static std::vector<dist_t> queryOneToAll(const GridMap& grid, Point s,
                                                           const std::vector<Point>& targets) {
        std::vector<dist_t> out(targets.size(), INF);
        if (!grid.isPassable(s) || targets.empty()) return out;
        const std::vector<int> hops = distances_from(grid, s);
        for (size_t i = 0; i < targets.size(); ++i) {
            if (!grid.isPassable(targets[i])) continue;
            const int h = hops[size_t(grid.index(targets[i]))];
            if (h >= 0) out[i] = dist_t(std::min(h, int(INF)));
        }
        return out;
    }

    // src/bfs_oracle.cpp:187-202: row i is queryOneToAll from viewpoint i.
    [[nodiscard]] static std::vector<dist_t> queryAllPairs(const GridMap& grid, const std::vector<Point>& viewpoints) {
        const size_t k = viewpoints.size();
        std::vector<dist_t> m(k * k, INF);
        for (size_t i = 0; i < k; ++i) {
            const auto row = queryOneToAll(grid, viewpoints[i], viewpoints);
            std::copy(row.begin(), row.end(), m.begin() + std::ptrdiff_t(i * k));
        }
        return m;
    }

private:
    // Hop counts from s over passable cells (-1 where unreached).
    static std::vector<int> distances_from(const GridMap& grid, Point s) {
        std::vector<int> hops(size_t(grid.rows) * size_t(grid.cols), -1);
        std::vector<int> ring{grid.index(s)};
        hops[size_t(ring[0])] = 0;
        for (size_t head = 0; head < ring.size(); ++head) {
            const int u = ring[head], r = u / grid.cols, c = u % grid.cols;
            const Point nb[4] = {{r - 1, c}, {r + 1, c}, {r, c - 1}, {r, c + 1}};
            for (const Point& v : nb) {
                if (!grid.isPassable(v)) continue;
                const int vi = grid.index(v);
                if (hops[size_t(vi)] >= 0) continue;
                hops[size_t(vi)] = hops[size_t(u)] + 1;
                ring.push_back(vi);
            }
        }
        return hops;
    }
};

}  // namespace htrip
