// C++20 drop-in test: code written against the reference's MinPlusOracle API,
// compiled against htrip_amd/include/htrip/*.hpp and linked with
// libhtrip_gpu.so.  Mirrors suites of the reference's apps/test_exactness.cpp
// (known answers 1 and 37, blocked viewpoint INF, batch openings, copy
// semantics, incremental refresh, reopened viewpoint, liftPoint, liftRoot,
// a non-square workspace, changing viewpoint lists, queryDistance against the
// matrix, on the reference's own procedural worlds),
// plus const queries from several threads each with its own scratch, as the
// reference's simulator issues them (src/robot_simulator.cpp:1541-1546), and
// checks every matrix against a BFS written here.
// Exit status 0 = all passed.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <thread>
#include <vector>

#include "htrip/grid_map.hpp"
#include "htrip/minplus_oracle.hpp"

using namespace htrip;

static std::vector<dist_t> bfs_all_pairs(const GridMap& g, const std::vector<Point>& vps) {
    const size_t k = vps.size();
    std::vector<dist_t> m(k * k, INF);
    std::vector<int> dist(size_t(g.rows) * g.cols);
    for (size_t i = 0; i < k; ++i) {
        if (!g.isPassable(vps[i])) continue;
        std::fill(dist.begin(), dist.end(), -1);
        std::queue<Point> q;
        q.push(vps[i]);
        dist[size_t(g.index(vps[i]))] = 0;
        while (!q.empty()) {
            Point u = q.front();
            q.pop();
            const Point nb[4] = {{u.r - 1, u.c}, {u.r + 1, u.c}, {u.r, u.c - 1}, {u.r, u.c + 1}};
            for (const Point& v : nb) {
                if (!g.isPassable(v) || dist[size_t(g.index(v))] >= 0) continue;
                dist[size_t(g.index(v))] = dist[size_t(g.index(u))] + 1;
                q.push(v);
            }
        }
        for (size_t j = 0; j < k; ++j)
            if (g.isPassable(vps[j]) && dist[size_t(g.index(vps[j]))] >= 0)
                m[i * k + j] = dist_t(std::min(dist[size_t(g.index(vps[j]))], int(INF)));
    }
    return m;
}

static int failures = 0;
static void expect(bool ok, const char* what) {
    std::printf("  %-58s %s\n", what, ok ? "PASSED" : "FAILED");
    if (!ok) ++failures;
}

static bool matches_bfs(const MinPlusOracle& o, const GridMap& g, const std::vector<Point>& vps) {
    return o.queryAllPairs(vps, g) == bfs_all_pairs(g, vps);
}

int main() {
    {  // Suite 1: seam crossing between leaves, d = 1 (test_exactness.cpp:97-146)
        GridMap g(32, 32, CellState::Obstacle);
        g.openDoorway({4, 16});
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        expect(o.build(g), "suite 1 build");
        g.openDoorway({4, 15});
        o.onCellOpened({4, 15}, g);
        expect(o.queryDistance({4, 15}, {4, 16}, g) == 1, "suite 1 seam (4,15)-(4,16) = 1");
        for (int c = 10; c <= 20; ++c) {
            g.openDoorway({4, c});
            o.onCellOpened({4, c}, g);
        }
        expect(matches_bfs(o, g, {{4, 10}, {4, 13}, {4, 15}, {4, 16}, {4, 18}, {4, 20}}), "suite 1 corridor == BFS");
    }
    {  // Suite 8: dense-meet tail, d = 37 (:594-667)
        GridMap g(32, 32, CellState::Obstacle);
        for (int c = 16; c <= 31; ++c) g.openDoorway({2, c});
        for (int r = 2; r <= 15; ++r) g.openDoorway({r, 24});
        for (int r = 16; r <= 20; ++r) g.openDoorway({r, 24});
        for (int c = 16; c <= 31; ++c) g.openDoorway({20, c});
        for (int r = 16; r <= 31; ++r)
            for (int c = 0; c <= 15; ++c) g.openDoorway({r, c});
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        HTripOracle o(cfg);
        o.build(g);
        expect(o.queryDistance({2, 24}, {25, 10}, g) == 37, "suite 8 d((2,24),(25,10)) = 37");
        expect(matches_bfs(o, g, {{2, 24}, {25, 10}, {20, 20}, {8, 24}}), "suite 8 matrix == BFS");
    }
    {  // Suite 12: blocked viewpoint INF, door opens a finite path (:842-893)
        GridMap g(32, 32, CellState::Free);
        for (int r = 0; r < 32; ++r) g.blockCorridor({r, 16});
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> vps = {{4, 4}, {8, 12}, {4, 20}, {8, 28}, {20, 16}};
        auto m = o.queryAllPairs(vps, g);
        expect(m[4 * 5 + 4] == INF && m[0 * 5 + 2] == INF, "suite 12 blocked viewpoint and split halves INF");
        g.openDoorway({8, 16});
        o.onCellOpened({8, 16}, g);
        expect(matches_bfs(o, g, vps) && o.queryAllPairs(vps, g)[2] < INF, "suite 12 door opened == BFS");
    }
    {  // Suite 16.3: one 64-cell batch, open grid -> Manhattan (:1132-1163)
        GridMap g(64, 64, CellState::Free);
        for (int r = 0; r < 64; ++r) g.setObserved(r, 31, CellState::Obstacle);
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> batch;
        for (int r = 0; r < 64; ++r) {
            g.setObserved(r, 31, CellState::Free);
            batch.push_back({r, 31});
        }
        o.onBatchCellsOpened(batch, g);
        expect(o.queryDistance({32, 10}, {32, 45}, g) == 35, "suite 16.3 batch column, d = 35");
        expect(o.last_batch_stats.n_touched_leaves == 4, "suite 16.3 last_batch_stats touched leaves");
    }
    {  // Random exploration with batches; copy semantics (bench_update_latency.cpp:178)
        std::mt19937_64 rng(7);
        GridMap g(96, 96, CellState::Unknown);
        std::vector<Point> hidden;
        for (int r = 0; r < 96; ++r)
            for (int c = 0; c < 96; ++c) {
                const bool wall = (r % 12 == 0 && c % 5 != 0) || (c % 17 == 0 && r % 7 != 0);
                if (wall) g.setObserved(r, c, CellState::Obstacle);
                else if (r < 40 && c < 40) g.setObserved(r, c, CellState::Free);
                else hidden.push_back({r, c});
            }
        std::shuffle(hidden.begin(), hidden.end(), rng);
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g);
        MinPlusOracle snapshot = o;  // deep copy of the device index
        GridMap g0 = g;
        bool ok = true;
        for (size_t s = 0; s < 10; ++s) {
            std::vector<Point> batch(hidden.begin() + long(s * 300), hidden.begin() + long((s + 1) * 300));
            for (const Point& p : batch) g.openDoorway(p);
            o.onBatchCellsOpened(batch, g);
            std::vector<Point> vps;
            for (int i = 0; i < 20; ++i) vps.push_back({int(rng() % 96), int(rng() % 96)});
            ok = ok && matches_bfs(o, g, vps);
        }
        expect(ok, "random batch reveals == BFS after every batch");
        std::vector<Point> vps = {{1, 1}, {30, 30}, {39, 2}, {60, 60}};
        expect(matches_bfs(snapshot, g0, vps), "copied oracle keeps the pre-update state");
        expect(o.totalMemoryBytes() > 0 && o.depth() == o.H, "totalMemoryBytes / depth");
    }
    {  // Suite 7: incremental refresh == full query after every opening (:520-580),
       // with an unrelated full query on the same scratch in between
        std::mt19937_64 rng(99);
        GridMap g(64, 64, CellState::Unknown);
        for (int r = 0; r < 64; ++r)
            for (int c = 0; c < 64; ++c) {
                const bool wall = (r % 10 == 0 && c % 7 != 3) || (c % 13 == 0 && r % 9 != 4);
                g.setObserved(r, c, wall ? CellState::Obstacle : (r < 30 ? CellState::Free : CellState::Unknown));
            }
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> vps;
        for (int i = 0; vps.size() < 24; ++i) {
            const Point p{int(rng() % 64), int(rng() % 64)};
            if (g.isPassable(p) || p.r >= 30) vps.push_back(p);
        }
        MinPlusOracle::HTripQueryScratch scratch;
        MinPlusOracle::HTripIncrementalState state;
        std::vector<dist_t> inc;
        o.queryAllPairsIncremental(vps, g, inc, scratch, state);
        bool ok = inc == o.queryAllPairs(vps, g);
        int opened = 0;
        for (int step = 0; step < 400 && opened < 40; ++step) {
            const Point p{30 + int(rng() % 34), int(rng() % 64)};
            if (g.getObserved(p) != CellState::Unknown) continue;
            g.openDoorway(p);
            o.onCellOpened(p, g);
            ++opened;
            o.queryAllPairsIncremental(vps, g, inc, scratch, state);
            if (opened % 5 == 0) {
                std::vector<dist_t> throwaway;
                std::vector<Point> other(vps.rbegin(), vps.rbegin() + 10);
                o.queryAllPairs(other, g, throwaway, scratch);
            }
            ok = ok && inc == o.queryAllPairs(vps, g) && inc == bfs_all_pairs(g, vps);
        }
        o.queryAllPairsIncremental(vps, g, inc, scratch, state);  // no update since: cached
        expect(ok && opened == 40, "suite 7 incremental == full == BFS after 40 openings");
        expect(state.last_recomputed == 0 && inc == bfs_all_pairs(g, vps), "suite 7 cached refresh");
        std::vector<dist_t> empty_out{1};
        o.queryAllPairsIncremental({}, g, empty_out, scratch, state);
        expect(empty_out.empty() && !state.initialized, "suite 7 k = 0 clears the state");
    }
    {  // Suite 9b: a viewpoint blocked during a reset and reopened regains 0 (:717-748)
        GridMap g(32, 32, CellState::Free);
        for (int r = 0; r < 32; ++r) g.blockCorridor({r, 9});
        g.openDoorway({5, 9});
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> vps = {{1, 1}, {3, 3}, {20, 20}, {30, 2}};
        MinPlusOracle::HTripQueryScratch scratch;
        MinPlusOracle::HTripIncrementalState state;
        std::vector<dist_t> m;
        o.queryAllPairsIncremental(vps, g, m, scratch, state);
        g.blockCorridor(vps[1]);
        o.onCellBlocked(vps[1], g);
        o.queryAllPairsIncremental(vps, g, m, scratch, state);
        const bool blocked_inf = m[1 * 4 + 1] == INF && m[0 * 4 + 1] == INF;
        g.openDoorway(vps[1]);
        o.onCellOpened(vps[1], g);
        o.queryAllPairsIncremental(vps, g, m, scratch, state);
        expect(blocked_inf && m[1 * 4 + 1] == 0 && m == bfs_all_pairs(g, vps), "suite 9b reopened viewpoint");
    }
    {  // liftPoint (src/minplus_oracle.cpp:731-823): level 0 = BFS inside the
       // leaf, level H = BFS over the whole map (the root tile), both at the
       // tile's ports; consistency with the matrix of a port pair (suite 15 style)
        GridMap g(64, 64, CellState::Free);
        for (int r = 0; r < 64; ++r)
            if (r != 40) g.blockCorridor({r, 20});
        for (int c = 0; c < 64; ++c)
            if (c != 50) g.blockCorridor({33, c});
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g);
        const Point p{10, 12};
        MinPlusOracle::LiftedViewpoint lv;
        o.liftPoint(p, lv);
        bool ok = int(lv.X.size()) == o.H + 1 && lv.parent_node_idx[0] == -1 && lv.child_quad[0] == -1;
        // level 0 inside leaf (1, 1): rows 8..15, cols 8..15
        const int b = 8;
        auto ports = [](int R0, int C0, int s) {
            std::vector<Point> v;
            for (int c = 0; c < s; ++c) v.push_back({R0, C0 + c});
            for (int c = 0; c < s; ++c) v.push_back({R0 + s - 1, C0 + c});
            for (int r = 1; r < s - 1; ++r) v.push_back({R0 + r, C0});
            for (int r = 1; r < s - 1; ++r) v.push_back({R0 + r, C0 + s - 1});
            return v;
        };
        const auto lp = ports(8, 8, b);
        for (size_t i = 0; i < lp.size(); ++i)  // an open leaf: Manhattan distance
            ok = ok && lv.X[0][i] == dist_t(std::abs(lp[i].r - p.r) + std::abs(lp[i].c - p.c));
        const auto rp = ports(0, 0, 64);
        std::vector<Point> pair = {p};
        pair.insert(pair.end(), rp.begin(), rp.end());
        const auto d = bfs_all_pairs(g, pair);
        for (size_t e = 0; e < rp.size(); ++e) ok = ok && lv.X[size_t(o.H)][e] == d[1 + e];
        for (size_t l = 0; l < lv.X.size(); ++l)
            for (size_t i = 1; i < lv.active[l].size(); ++i)
                ok = ok && lv.X[l][size_t(lv.active[l][i - 1])] <= lv.X[l][size_t(lv.active[l][i])];
        expect(ok, "liftPoint levels 0 and H == BFS, active sorted");
    }
    {  // Suite 3: liftRoot 16x16 (H=1) -> 32x32 and 8x8 (H=0) -> 16x16 (:309-366)
        GridMap g16(16, 16), g32(32, 32), g8(8, 8);
        for (int r = 0; r < 32; ++r)
            for (int c = 0; c < 32; ++c) {
                g32.openDoorway({r, c});
                if (r < 16 && c < 16) g16.openDoorway({r, c});
                if (r < 8 && c < 8) g8.openDoorway({r, c});
            }
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g16);
        o.liftRoot(g32);
        expect(o.S == 32 && o.H == 2 && matches_bfs(o, g32, {{2, 2}, {5, 5}, {2, 20}, {20, 20}, {25, 10}}),
               "suite 3 liftRoot 16 -> 32 == BFS");
        MinPlusOracle o0(cfg);
        o0.build(g8);
        const bool h0 = o0.depth() == 0;
        o0.liftRoot(g16);
        expect(h0 && o0.depth() == 1 && matches_bfs(o0, g16, {{1, 1}, {3, 3}, {6, 6}, {10, 10}, {14, 14}}),
               "suite 3 liftRoot from H = 0 == BFS");
    }
    {  // Suite 16.4: liftRoot after batch openings (:1165-1208)
        GridMap g(64, 64);
        for (int r = 0; r < 64; ++r)
            for (int c = 0; c < 64; ++c) g.setObserved(r, c, r == 0 ? CellState::Obstacle : CellState::Free);
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> row0;
        for (int c = 0; c < 64; ++c) {
            g.setObserved(0, c, CellState::Free);
            row0.push_back({0, c});
        }
        o.onBatchCellsOpened(row0, g);
        GridMap g128(128, 128);
        for (int r = 0; r < 128; ++r)
            for (int c = 0; c < 128; ++c)
                g128.setObserved(r, c, (r < 64 && c < 64 && !g.isPassable({r, c})) ? CellState::Obstacle : CellState::Free);
        o.liftRoot(g128);
        const dist_t d = o.queryDistance({0, 0}, {0, 100}, g128);
        expect(d == bfs_all_pairs(g128, {{0, 0}, {0, 100}})[1] && d == 100, "suite 16.4 liftRoot after a batch, d = 100");
    }
    {  // Suite 11: non-square workspace, 20 x 44 rooms (b = 8 pads S to 64), openings (:798-841)
        GridMap g;
        g.generateProceduralRooms(20, 44, 8, 453);
        g.initExplorationSnapshot({2, 2}, 0.65, 0.10, 54);
        std::vector<Point> vps;
        for (int i = 0; i < 20 * 44 && vps.size() < 10; ++i)
            if (g.observed[size_t(i)] == CellState::Free) vps.push_back(g.point(i));
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g);
        MinPlusOracle::HTripQueryScratch scratch;
        MinPlusOracle::HTripIncrementalState state;
        std::vector<dist_t> inc;
        o.queryAllPairsIncremental(vps, g, inc, scratch, state);
        bool ok = o.S == 64 && matches_bfs(o, g, vps) && inc == bfs_all_pairs(g, vps);
        std::mt19937_64 rng(55);
        int opened = 0;
        for (int step = 0; step < 400 && opened < 15; ++step) {
            const Point p{int(rng() % 20), int(rng() % 44)};
            if (g.ground_truth[size_t(g.index(p))] != CellState::Free || g.getObserved(p) == CellState::Free) continue;
            g.openDoorway(p);
            o.onCellOpened(p, g);
            ++opened;
        }
        o.queryAllPairsIncremental(vps, g, inc, scratch, state);
        ok = ok && opened == 15 && matches_bfs(o, g, vps) && inc == bfs_all_pairs(g, vps);
        expect(ok, "suite 11 20x44 rooms, 15 openings == BFS");
    }
    {  // Suite 14: the viewpoint list changes (k and order) between incremental calls (:945-985)
        GridMap g;
        g.generateProceduralRooms(32, 32, 16, 856);
        g.initExplorationSnapshot({1, 1}, 0.70, 0.10, 58);
        std::vector<Point> cells;
        for (int i = 0; i < 32 * 32; ++i)
            if (g.observed[size_t(i)] == CellState::Free) cells.push_back(g.point(i));
        std::mt19937_64 rng(59);
        std::shuffle(cells.begin(), cells.end(), rng);
        const std::vector<Point> a(cells.begin(), cells.begin() + 5);
        std::vector<Point> bl(cells.begin(), cells.begin() + 8);
        std::reverse(bl.begin(), bl.end());
        std::vector<Point> ar(a.rbegin(), a.rend());
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        MinPlusOracle::HTripQueryScratch scratch;
        MinPlusOracle::HTripIncrementalState state;
        bool ok = true;
        const std::vector<Point>* lists[] = {&a, &bl, &a, &ar};
        for (const auto* l : lists) {
            std::vector<dist_t> inc;
            o.queryAllPairsIncremental(*l, g, inc, scratch, state);
            ok = ok && inc == bfs_all_pairs(g, *l) && inc == o.queryAllPairs(*l, g);
        }
        expect(ok, "suite 14 lists A, B, A, A reversed: incremental == BFS");
    }
    {  // Suite 15: queryDistance equals the matrix for every pair (:987-1024)
        GridMap g;
        g.generateProceduralRooms(40, 40, 16, 957);
        g.initExplorationSnapshot({2, 2}, 0.68, 0.10, 60);
        std::vector<Point> vps;
        for (int i = 0; i < 40 * 40 && vps.size() < 10; ++i)
            if (g.observed[size_t(i)] == CellState::Free) vps.push_back(g.point(i));
        HierarchyConfig cfg;
        cfg.tile_size_b = 8;
        MinPlusOracle o(cfg);
        o.build(g);
        const auto m = o.queryAllPairs(vps, g);
        bool ok = m == bfs_all_pairs(g, vps);
        for (size_t i = 0; i < vps.size(); ++i)
            for (size_t j = 0; j < vps.size(); ++j) ok = ok && o.queryDistance(vps[i], vps[j], g) == m[i * vps.size() + j];
        expect(ok, "suite 15 queryDistance == matrix for 100 pairs");
    }
    {  // onBatchLeavesChanged: cells of two leaves opened and blocked in the grid (:507-527);
       // unit weights accepted, other weights refused with the grid unchanged
        GridMap g;
        g.generateProceduralRooms(64, 64, 16, 77);
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        std::vector<Point> vps = {{3, 3}, {20, 20}, {40, 8}, {60, 60}, {17, 40}, {33, 33}};
        bool ok = matches_bfs(o, g, vps);
        for (int r = 0; r < 16; ++r) g.openDoorway({r, 15}), g.openDoorway({r, 16});  // a wide passage, leaves 0 and 1
        for (int c = 18; c < 30; ++c) g.blockCorridor({20, c});                         // a wall across leaf 5
        o.onBatchLeavesChanged({0, 1, 5}, g);
        ok = ok && matches_bfs(o, g, vps);
        o.onCellWeightChanged({3, 3}, 1, g);
        bool refused = false;
        try {
            o.onCellWeightChanged({3, 4}, 5, g);
        } catch (const std::invalid_argument&) {
            refused = true;
        }
        ok = ok && refused && g.getCellWeight({3, 4}) == 1 && matches_bfs(o, g, vps);
        expect(ok, "onBatchLeavesChanged open + block == BFS; weight 5 refused");
    }
    {  // const queries from 4 threads, each with its own HTripQueryScratch, on one oracle
        GridMap g(96, 96, CellState::Free);
        std::mt19937 rng(11);
        for (int i = 0; i < 900; ++i) g.blockCorridor({int(rng() % 96), int(rng() % 96)});
        HierarchyConfig cfg;
        cfg.tile_size_b = 16;
        MinPlusOracle o(cfg);
        o.build(g);
        constexpr int kThreads = 4, kQueries = 12;
        std::vector<std::vector<std::vector<Point>>> sets(kThreads);
        std::vector<std::vector<std::vector<dist_t>>> want(kThreads);
        for (int t = 0; t < kThreads; ++t)
            for (int q = 0; q < kQueries; ++q) {
                std::vector<Point> vps;
                const size_t k = 8 + size_t(rng() % 40);
                while (vps.size() < k) {
                    const Point p{int(rng() % 96), int(rng() % 96)};
                    if (g.isPassable(p)) vps.push_back(p);
                }
                want[size_t(t)].push_back(bfs_all_pairs(g, vps));
                sets[size_t(t)].push_back(std::move(vps));
            }
        const MinPlusOracle& co = o;
        std::vector<int> good(kThreads, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < kThreads; ++t)
            th.emplace_back([&, t] {
                MinPlusOracle::HTripQueryScratch scratch;
                std::vector<dist_t> out;
                for (int q = 0; q < kQueries; ++q) {
                    co.queryAllPairs(sets[size_t(t)][size_t(q)], g, out, scratch);
                    good[size_t(t)] += out == want[size_t(t)][size_t(q)];
                }
            });
        for (auto& x : th) x.join();
        bool ok = true;
        for (int t = 0; t < kThreads; ++t) ok = ok && good[size_t(t)] == kQueries;
        expect(ok, "4 threads, own scratch each, 48 matrices == BFS");
    }
    std::printf("%s (%d failure%s)\n", failures ? "FAILED" : "ALL PASSED", failures, failures == 1 ? "" : "s");
    return failures ? 1 : 0;
}
