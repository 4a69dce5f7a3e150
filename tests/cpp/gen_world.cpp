// CPU helper for tests/test_maps.py: the drop-in htrip::GridMap's
// generateProceduralRooms(rows, cols, room, seed) and, when coverage > 0,
// initExplorationSnapshot({sr, sc}, coverage, island, seed2).  Writes
// "ok rows cols\n", then the ground-truth codes, the observed codes and the
// observed grid's getClearance tiers as raw bytes to stdout.
#include <cstdio>
#include <cstdlib>

#include "htrip/grid_map.hpp"

int main(int argc, char** argv) {
    if (argc < 10) return 2;
    const int rows = std::atoi(argv[1]), cols = std::atoi(argv[2]), room = std::atoi(argv[3]);
    const unsigned long long seed = std::strtoull(argv[4], nullptr, 10), seed2 = std::strtoull(argv[9], nullptr, 10);
    const int sr = std::atoi(argv[5]), sc = std::atoi(argv[6]);
    const double coverage = std::atof(argv[7]), island = std::atof(argv[8]);
    htrip::GridMap g;
    g.generateProceduralRooms(rows, cols, room, seed);
    const bool ok = coverage > 0 ? g.initExplorationSnapshot({sr, sc}, coverage, island, seed2) : true;
    std::printf("%d %d %d\n", ok ? 1 : 0, g.rows, g.cols);
    for (auto v : g.ground_truth) std::putchar(int(v));
    for (auto v : g.observed) std::putchar(int(v));
    for (int r = 0; r < g.rows; ++r)
        for (int c = 0; c < g.cols; ++c) std::putchar(int(g.getClearance(r, c)));
    return 0;
}
