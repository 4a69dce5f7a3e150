#pragma once
// Drop-in replacement of the reference's include/htrip/grid_map.hpp for the
// parts the distance oracle reads: dimensions, the observed layer, passability
// and edge weights (grid_map.hpp:31-334), plus the MovingAI loader
// (src/grid_map.cpp:16-76).  Header-only.
#include <fstream>
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
};

}  // namespace htrip
