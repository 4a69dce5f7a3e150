#pragma once
// Drop-in replacement of the reference's include/htrip/types.hpp: the same
// names and semantics (distance type, INF, saturating semiring product, cell
// states, Point, decrease records, HierarchyConfig) so code written against
// the reference compiles unchanged against the MI355X build.
#include <compare>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace htrip {

using dist_t = uint16_t;                 // types.hpp:34
static constexpr dist_t INF = 60000;     // types.hpp:45
static constexpr dist_t UNIT_STEP = 1;   // types.hpp:50

// Saturating tropical product (types.hpp:66-70, :87-91).
inline constexpr dist_t add_dist(dist_t a, dist_t b) noexcept {
    if (a >= INF || b >= INF) return INF;
    const uint32_t s = uint32_t(a) + uint32_t(b);
    return dist_t(s >= INF ? INF : s);
}
inline constexpr dist_t add_dist(dist_t a, dist_t b, dist_t c) noexcept {
    if (a >= INF || b >= INF || c >= INF) return INF;
    const uint32_t s = uint32_t(a) + uint32_t(b) + uint32_t(c);
    return dist_t(s >= INF ? INF : s);
}

enum class CellState : uint8_t { Unknown = 0, Free = 1, Obstacle = 2 };  // types.hpp:101-105

struct Point {  // types.hpp:113-145
    int r = 0;
    int c = 0;
    constexpr Point() noexcept = default;
    constexpr Point(int r_, int c_) noexcept : r(r_), c(c_) {}
    constexpr bool operator==(const Point& o) const noexcept { return r == o.r && c == o.c; }
    constexpr bool operator!=(const Point& o) const noexcept { return !(*this == o); }
    constexpr auto operator<=>(const Point& o) const noexcept {
        if (r != o.r) return r <=> o.r;
        return c <=> o.c;
    }
    [[nodiscard]] constexpr int manhattan(const Point& o) const noexcept {
        return std::abs(r - o.r) + std::abs(c - o.c);
    }
};

struct PortDecrease {  // types.hpp:153-157
    int p_a = 0;
    int p_b = 0;
    dist_t w = INF;
};

struct ChildDecrease {  // types.hpp:165-170
    int quad = 0;
    int p_a = 0;
    int p_b = 0;
    dist_t w = INF;
};

struct SeamEdge {  // types.hpp:178-183
    Point p1;
    Point p2;
    int lca_lvl = 1;
    dist_t w = UNIT_STEP;
};

inline constexpr int kMaxActiveChildren = 256;  // types.hpp:191

struct HierarchyConfig {  // types.hpp:199-225
    int tile_size_b = 16;
    int group_w = 2;
    int group_h = 2;
    int hybrid_k_threshold = 0;
    [[nodiscard]] constexpr bool isQuadtree() const noexcept { return group_w == 2 && group_h == 2; }
    [[nodiscard]] constexpr bool isValid() const noexcept {
        return tile_size_b > 0 && group_w >= 2 && group_h >= 2 && group_w == group_h &&
               group_w * group_h <= kMaxActiveChildren;
    }
};

}  // namespace htrip
