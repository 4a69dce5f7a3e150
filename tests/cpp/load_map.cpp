// CPU helper for tests/test_maps.py: loads a MovingAI map with the drop-in
// htrip::GridMap::loadMovingAI and writes "rows cols\n" followed by the
// rows*cols ground-truth codes as raw bytes to stdout (exit 1 on failure).
#include <cstdio>

#include "htrip/grid_map.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    htrip::GridMap g;
    if (!g.loadMovingAI(argv[1])) return 1;
    std::printf("%d %d\n", g.rows, g.cols);
    for (auto v : g.ground_truth) std::putchar(int(v));
    return 0;
}
