#pragma once
// Drop-in replacement of the reference's include/htrip/grid_map.hpp for the
// parts the distance oracle reads: dimensions, the observed layer, passability
// and edge weights (grid_map.hpp:31-334), the MovingAI loader
// (src/grid_map.cpp:16-76), cell weights and clearance, and the two synthetic
// world producers the reference's test and benchmark programs build their
// grids with (src/grid_map.cpp:78-200).  Header-only.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <queue>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "types.hpp"

namespace htrip {

class GridMap {
public:
    int rows = 0;
    int cols = 0;
    std::vector<CellState> ground_truth;
    std::vector<CellState> observed;
    std::vector<dist_t> weights;

    GridMap() = default;
    GridMap(int r, int c, CellState initial_state = CellState::Unknown)
        : rows(r), cols(c), ground_truth(size_t(r) * c, CellState::Free), observed(size_t(r) * c, initial_state),
          weights(size_t(r) * c, 1) {}

    [[nodiscard]] bool inBounds(int r, int c) const noexcept { return r >= 0 && r < rows && c >= 0 && c < cols; }
    [[nodiscard]] bool inBounds(Point p) const noexcept { return inBounds(p.r, p.c); }
    [[nodiscard]] int index(int r, int c) const noexcept { return r * cols + c; }
    [[nodiscard]] int index(Point p) const noexcept { return index(p.r, p.c); }
    [[nodiscard]] Point point(int idx) const noexcept { return {idx / cols, idx % cols}; }
    [[nodiscard]] bool isPassable(int r, int c) const noexcept {
        return inBounds(r, c) && observed[size_t(index(r, c))] == CellState::Free;
    }
    [[nodiscard]] bool isPassable(Point p) const noexcept { return isPassable(p.r, p.c); }
    [[nodiscard]] CellState getObserved(int r, int c) const noexcept {
        return inBounds(r, c) ? observed[size_t(index(r, c))] : CellState::Obstacle;
    }
    [[nodiscard]] CellState getObserved(Point p) const noexcept { return getObserved(p.r, p.c); }
    void setObserved(int r, int c, CellState s) noexcept {
        if (inBounds(r, c)) observed[size_t(index(r, c))] = s;
    }
    void setObserved(Point p, CellState s) noexcept { setObserved(p.r, p.c, s); }
    [[nodiscard]] bool hasWeights() const noexcept { return !weights.empty(); }
    [[nodiscard]] dist_t getCellWeight(int r, int c) const noexcept {
        if (!inBounds(r, c)) return INF;
        if (observed[size_t(index(r, c))] == CellState::Obstacle) return INF;
        if (weights.empty()) return 1;
        return weights[size_t(index(r, c))];
    }
    [[nodiscard]] dist_t getCellWeight(Point p) const noexcept { return getCellWeight(p.r, p.c); }
    // grid_map.hpp:297-313: the weight array is created (all 1) on first use.
    void setCellWeight(int r, int c, dist_t w) noexcept {
        if (!inBounds(r, c)) return;
        const size_t n = size_t(rows) * size_t(cols);
        if (weights.size() < n) weights.assign(n, 1);
        weights[size_t(index(r, c))] = w;
    }
    void setCellWeight(Point p, dist_t w) noexcept { setCellWeight(p.r, p.c, w); }
    // grid_map.hpp:174-212: clearance tier of an observed cell.  0 when it is
    // not passable; 1 when one of its 8 neighbours is an obstacle or off the
    // grid; 2 when a cell at Chebyshev distance 2 is; 3 otherwise.  Unknown
    // cells do not count as obstacles.
    [[nodiscard]] uint8_t getClearance(int r, int c) const noexcept {
        if (!isPassable(r, c)) return 0;
        auto walled = [&](int y, int x) {
            return !inBounds(y, x) || observed[size_t(index(y, x))] == CellState::Obstacle;
        };
        for (int ring = 1; ring <= 2; ++ring)
            for (int dy = -ring; dy <= ring; ++dy)
                for (int dx = -ring; dx <= ring; ++dx)
                    if (std::max(std::abs(dy), std::abs(dx)) == ring && walled(r + dy, c + dx)) return uint8_t(ring);
        return 3;
    }
    [[nodiscard]] uint8_t getClearance(Point p) const noexcept { return getClearance(p.r, p.c); }
    [[nodiscard]] dist_t edgeWeight(Point u, Point v) const noexcept {
        const dist_t wu = getCellWeight(u), wv = getCellWeight(v);
        if (wu >= INF || wv >= INF) return INF;
        const dist_t avg = dist_t((uint32_t(wu) + uint32_t(wv)) / 2);
        return avg == 0 ? 1 : avg;
    }
    bool openDoorway(Point p) {
        if (!inBounds(p)) return false;
        observed[size_t(index(p))] = CellState::Free;
        return true;
    }
    bool blockCorridor(Point p) {
        if (!inBounds(p)) return false;
        observed[size_t(index(p))] = CellState::Obstacle;
        return true;
    }
    bool revealCell(Point p) {
        if (!inBounds(p)) return false;
        observed[size_t(index(p))] = ground_truth[size_t(index(p))];
        return true;
    }

    // src/grid_map.cpp:16-76.  MovingAI .map: header lines up to a line that is
    // exactly "map" (trailing CR, space and tab ignored); lines starting with
    // "height" / "width" give the size.  Then one line per row: '.', 'G' and
    // 'S' are Free, every other character is an Obstacle, and so are cells past
    // a short line or a missing row.  observed starts all Unknown.
    bool loadMovingAI(const std::string& filepath) {
        std::ifstream f(filepath);
        if (!f.is_open()) return false;
        std::string line;
        int h = 0, w = 0;
        while (std::getline(f, line)) {
            while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
            std::istringstream ss(line);
            std::string key;
            if (line.rfind("height", 0) == 0) {
                ss >> key >> h;
            } else if (line.rfind("width", 0) == 0) {
                ss >> key >> w;
            } else if (line == "map") {
                break;
            }
        }
        if (h <= 0 || w <= 0) return false;
        rows = h;
        cols = w;
        ground_truth.assign(size_t(h) * w, CellState::Obstacle);
        observed.assign(size_t(h) * w, CellState::Unknown);
        weights.assign(size_t(h) * w, 1);
        for (int r = 0; r < h && std::getline(f, line); ++r)
            for (int c = 0; c < w && c < int(line.size()); ++c) {
                const char ch = line[size_t(c)];
                ground_truth[size_t(r) * w + c] =
                    (ch == '.' || ch == 'G' || ch == 'S') ? CellState::Free : CellState::Obstacle;
            }
        return true;
    }

    // src/grid_map.cpp:78-141.  Rooms of room_size (at least 4) tile the grid;
    // a room's interior is its cells one step inside its square, so walls
    // between rooms are two cells thick.  Every room then gets a doorway
    // through the wall to its east neighbour and one to its south neighbour,
    // at a random offset inside the room, from one mt19937_64 stream drawn in
    // room order (row-major; east before south), so that a seed gives the
    // reference's world cell for cell.  observed starts equal to the truth.
    void generateProceduralRooms(int r, int c, int room_size = 16, uint64_t seed = 42) {
        const int rs = std::max(room_size, 4);
        rows = r;
        cols = c;
        const size_t n = size_t(rows) * size_t(cols);
        ground_truth.assign(n, CellState::Obstacle);
        observed.assign(n, CellState::Unknown);
        weights.assign(n, 1);
        auto free_cell = [&](int y, int x) {
            if (y < rows && x < cols) ground_truth[size_t(index(y, x))] = CellState::Free;
        };
        const int ny = rows / rs, nx = cols / rs;
        for (int ry = 0; ry < ny; ++ry)
            for (int rx = 0; rx < nx; ++rx) {
                const int y1 = std::min((ry + 1) * rs - 1, rows - 1), x1 = std::min((rx + 1) * rs - 1, cols - 1);
                for (int y = ry * rs + 1; y < y1; ++y)
                    for (int x = rx * rs + 1; x < x1; ++x) free_cell(y, x);
            }
        std::mt19937_64 rng(seed);
        auto offset = [&] { return 1 + int(rng() % uint64_t(rs - 2)); };
        for (int ry = 0; ry < ny; ++ry)
            for (int rx = 0; rx < nx; ++rx) {
                if (rx + 1 < nx) {  // east: both wall layers at one row
                    const int y = ry * rs + offset();
                    if (y < rows) free_cell(y, (rx + 1) * rs - 1), free_cell(y, (rx + 1) * rs);
                }
                if (ry + 1 < ny) {  // south: both wall layers at one column
                    const int x = rx * rs + offset();
                    if (x < cols) free_cell((ry + 1) * rs - 1, x), free_cell((ry + 1) * rs, x);
                }
            }
        observed = ground_truth;
    }

    // src/grid_map.cpp:143-200.  A partly explored snapshot: everything
    // Unknown, then a 4-neighbour BFS (north, south, west, east) from
    // start_point (or, when that is not free in the truth, the first free cell
    // in row-major order) over the truth until max(10, free * coverage) free
    // cells are revealed.  A reached obstacle is revealed as an obstacle.  A
    // reached free cell is skipped (left Unknown, not expanded) with
    // probability island_prob once more than 50 cells are revealed; the
    // uniform draw happens for every reached free cell while island_prob > 0,
    // as in the reference, so a seed gives the same snapshot.
    bool initExplorationSnapshot(Point start_point, double target_coverage = 0.45, double island_prob = 0.1,
                                 uint64_t seed = 12345) {
        auto truth_free = [&](int i) { return ground_truth[size_t(i)] == CellState::Free; };
        if (!inBounds(start_point) || !truth_free(index(start_point))) {
            const auto it = std::find(ground_truth.begin(), ground_truth.end(), CellState::Free);
            if (it == ground_truth.end()) return false;
            start_point = point(int(it - ground_truth.begin()));
        }
        const size_t n = size_t(rows) * size_t(cols);
        observed.assign(n, CellState::Unknown);
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<double> u01(0.0, 1.0);
        const int total = int(std::count(ground_truth.begin(), ground_truth.end(), CellState::Free));
        const int budget = std::max(10, int(total * target_coverage));
        std::vector<uint8_t> seen(n, 0);
        std::queue<Point> frontier;
        frontier.push(start_point);
        seen[size_t(index(start_point))] = 1;
        observed[size_t(index(start_point))] = CellState::Free;
        int revealed = 1;
        const Point step[4] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        while (!frontier.empty() && revealed < budget) {
            const Point at = frontier.front();
            frontier.pop();
            for (const Point& d : step) {
                const Point nb{at.r + d.r, at.c + d.c};
                if (!inBounds(nb)) continue;
                const size_t i = size_t(index(nb));
                if (seen[i]) continue;
                seen[i] = 1;
                if (ground_truth[i] == CellState::Obstacle) {
                    observed[i] = CellState::Obstacle;
                    continue;
                }
                if (ground_truth[i] != CellState::Free) continue;
                if (island_prob > 0.0 && u01(rng) < island_prob && revealed > 50) continue;
                observed[i] = CellState::Free;
                frontier.push(nb);
                if (++revealed >= budget) break;
            }
        }
        return true;
    }
};

}  // namespace htrip
