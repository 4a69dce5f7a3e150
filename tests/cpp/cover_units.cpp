// Host-side check of the cover-product work planner (htg::plan_cover_units in
// libhtrip_gpu.so; no GPU call): for random levels of nodes with cover counts
// c, uncovered-endpoint counts b and live octet counts l8, the units of either
// hop must tile every node's row x column tiles and split its B list into
// disjoint ranges that cover it exactly, stay within the stated unit budget,
// and use the tallest row tile that still gives 512 tiles.
// Usage: cover_units <seed> <levels>; prints "ok <units>" or the first failure.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <tuple>
#include <vector>

#include "kernels.hpp"

namespace {

int fail(const char* what, int level, int hop, int node) {
    std::printf("FAIL %s (level %d, hop %d, node %d)\n", what, level, hop, node);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const unsigned seed = argc > 1 ? unsigned(std::atoi(argv[1])) : 1u;
    const int levels = argc > 2 ? std::atoi(argv[2]) : 200;
    std::mt19937 rng(seed);
    long long total_units = 0;
    for (int lv = 0; lv < levels; ++lv) {
        // level shapes from one node with long lists to hundreds of small nodes
        const int n = 1 + int(rng() % (lv % 3 == 0 ? 4 : 300));
        const int cmax = 1 + int(rng() % (lv % 2 ? 64 : 400));
        const int bmax = 1 + int(rng() % (lv % 5 == 0 ? 3000 : 600));
        const int l8max = 1 + int(rng() % 1022);
        std::vector<int> c(static_cast<size_t>(n)), b(static_cast<size_t>(n)), l8(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            c[size_t(i)] = rng() % 7 == 0 ? 0 : int(rng() % unsigned(cmax + 1));
            b[size_t(i)] = rng() % 5 == 0 ? 0 : int(rng() % unsigned(bmax + 1));
            l8[size_t(i)] = std::max(1, int(rng() % unsigned(l8max + 1)));
            l8[size_t(i)] = std::max(l8[size_t(i)], (c[size_t(i)] + 7) / 8);  // cover rows are live
        }
        for (int hop = 0; hop < 2; ++hop) {
            auto rows = [&](int i, int rpt) { return (c[size_t(i)] + 16 * rpt - 1) / (16 * rpt); };
            auto colt = [&](int i) {
                return hop ? (c[size_t(i)] + 127) / 128 : (l8[size_t(i)] + 15) / 16;
            };
            auto active = [&](int i) { return c[size_t(i)] > 0 && b[size_t(i)] > 0; };
            std::vector<int4> u;
            const int rpt = htg::plan_cover_units(hop != 0, n, c.data(), b.data(), l8.data(), u);
            // the tallest tile with 512 or more tiles, else 1
            int want = 1;
            for (int r = 4; r > 1; r /= 2) {
                long long t = 0;
                for (int i = 0; i < n; ++i)
                    if (active(i)) t += (long long)rows(i, r) * colt(i);
                if (t >= 512) {
                    want = r;
                    break;
                }
            }
            if (rpt != want) return fail("rows per thread", lv, hop, -1);
            long long tiles = 0, work = 0;
            for (int i = 0; i < n; ++i)
                if (active(i)) {
                    tiles += (long long)rows(i, rpt) * colt(i);
                    work += (long long)rows(i, rpt) * colt(i) * b[size_t(i)];
                }
            if ((long long)u.size() > tiles + 512) return fail("unit budget", lv, hop, -1);
            // per (node, row tile, column tile): its B ranges
            std::map<std::tuple<int, int, int>, std::vector<std::pair<int, int>>> ranges;
            for (const int4& x : u) {
                const int i = x.x, I = x.y & 0xffff, J = x.y >> 16;
                if (i < 0 || i >= n || !active(i)) return fail("unit of an inactive node", lv, hop, i);
                if (I < 0 || I >= rows(i, rpt) || J < 0 || J >= colt(i)) return fail("tile out of range", lv, hop, i);
                if (x.z < 0 || x.z >= x.w || x.w > b[size_t(i)]) return fail("B range out of order", lv, hop, i);
                ranges[std::make_tuple(i, I, J)].push_back({x.z, x.w});
            }
            const long long part_cap = std::max<long long>(64, (work + 511) / 512);
            for (int i = 0; i < n; ++i) {
                if (!active(i)) continue;
                for (int I = 0; I < rows(i, rpt); ++I)
                    for (int J = 0; J < colt(i); ++J) {
                        auto it = ranges.find(std::make_tuple(i, I, J));
                        if (it == ranges.end()) return fail("tile without a unit", lv, hop, i);
                        auto& v = it->second;
                        std::sort(v.begin(), v.end());
                        int at = 0;
                        for (const auto& r : v) {
                            if (r.first != at) return fail("B ranges overlap or leave a gap", lv, hop, i);
                            // parts of at most the level's share (64 parts at most per node)
                            const long long cap = std::max(part_cap, (long long)(b[size_t(i)] + 63) / 64) + 1;
                            if (r.second - r.first > cap) return fail("B part too long", lv, hop, i);
                            at = r.second;
                        }
                        if (at != b[size_t(i)]) return fail("B ranges stop short", lv, hop, i);
                    }
            }
            total_units += (long long)u.size();
        }
    }
    std::printf("ok %lld\n", total_units);
    return 0;
}
