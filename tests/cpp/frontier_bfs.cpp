// CPU helper for tests/test_frontier.py: the drop-in htrip::FrontierManager and
// htrip::BfsOracle on an observed grid.  Reads a file of "rows cols k min max"
// on the first line, then rows*cols observed codes as raw bytes, then k
// viewpoints as "r c" pairs.  Writes the frontier flags as raw bytes, then
// "n" and per cluster "id r c size" and its cells as "r c" pairs, then the
// k x k BfsOracle::queryAllPairs matrix, all as text after the flags.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "htrip/bfs_oracle.hpp"
#include "htrip/frontier.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ifstream f(argv[1], std::ios::binary);
    int rows = 0, cols = 0, k = 0, mn = 2, mx = 16;
    f >> rows >> cols >> k >> mn >> mx;
    f.get();
    htrip::GridMap g(rows, cols);
    std::string buf(size_t(rows) * size_t(cols), '\0');
    f.read(buf.data(), std::streamsize(buf.size()));
    for (size_t i = 0; i < buf.size(); ++i) g.observed[i] = static_cast<htrip::CellState>(uint8_t(buf[i]));
    std::vector<htrip::Point> vps(static_cast<size_t>(k));
    for (auto& p : vps) f >> p.r >> p.c;
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) std::putchar(htrip::FrontierManager::isFrontierCell(g, {r, c}) ? 1 : 0);
    const auto cl = htrip::FrontierManager(mn, mx).detectAndCluster(g);
    std::printf("%zu\n", cl.size());
    for (const auto& c : cl) {
        std::printf("%d %d %d %d", c.id, c.representative.r, c.representative.c, c.size());
        for (const auto& p : c.cells) std::printf(" %d %d", p.r, p.c);
        std::printf("\n");
    }
    const auto m = htrip::BfsOracle::queryAllPairs(g, vps);
    for (auto v : m) std::printf("%d ", int(v));
    std::printf("\n");
    return 0;
}
