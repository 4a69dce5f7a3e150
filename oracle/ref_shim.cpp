// TEST INFRASTRUCTURE ONLY -- C-ABI wrapper around the compiled reference.
//
// oracle/Makefile compiles this file together with the reference's own
// sources (src/grid_map.cpp, leaf_tile.cpp, internal_node.cpp,
// minplus_oracle.cpp, bfs_oracle.cpp, dijkstra_oracle.cpp, frontier.cpp, read in place
// from the reference checkout) into oracle/_ref/libhtrip_ref.so.  The result
// is the reference's CPU AVX2 path itself, used by tests/ to pin the
// restatement in htrip_oracle.cpp and the GPU tables, and by bench.py as the
// "reference" CPU baseline.  Nothing here is shipped with the product.
//
// The exported functions mirror htrip_oracle.cpp's orc_* names with a ref_
// prefix so that tests can run the same checks against either library.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "htrip/bfs_oracle.hpp"
#include "htrip/frontier.hpp"
#include "htrip/grid_map.hpp"
#include "htrip/minplus_oracle.hpp"

using namespace htrip;

namespace {
struct RefHandle {
    GridMap grid;
    MinPlusOracle oracle;
};

std::vector<Point> to_pts(const int32_t* rc, int n) {
    std::vector<Point> v(size_t(n > 0 ? n : 0));
    for (int i = 0; i < n; ++i) v[i] = Point{rc[2 * i], rc[2 * i + 1]};
    return v;
}
}  // namespace

extern "C" {

void* ref_create(int b, int g, int rows, int cols, const uint8_t* cells) {
    auto* h = new RefHandle();
    h->grid = GridMap(rows, cols, CellState::Unknown);
    for (size_t i = 0; i < size_t(rows) * cols; ++i) h->grid.observed[i] = static_cast<CellState>(cells[i]);
    HierarchyConfig cfg;
    cfg.tile_size_b = b;
    cfg.group_w = g;
    cfg.group_h = g;
    h->oracle = MinPlusOracle(cfg);
    return h;
}
void ref_destroy(void* p) { delete static_cast<RefHandle*>(p); }
int ref_build(void* p) {
    auto* h = static_cast<RefHandle*>(p);
    return h->oracle.build(h->grid) ? 1 : 0;
}
void ref_open_cells(void* p, const int32_t* rc, int n) {
    auto* h = static_cast<RefHandle*>(p);
    auto pts = to_pts(rc, n);
    for (const Point& q : pts) h->grid.openDoorway(q);
    h->oracle.onBatchCellsOpened(pts, h->grid);
}
void ref_open_cell(void* p, int r, int c) {
    auto* h = static_cast<RefHandle*>(p);
    h->grid.openDoorway(Point{r, c});
    h->oracle.onCellOpened(Point{r, c}, h->grid);
}
void ref_query(void* p, const int32_t* rc, int k, uint16_t* out) {
    auto* h = static_cast<RefHandle*>(p);
    auto m = h->oracle.queryAllPairs(to_pts(rc, k), h->grid);
    std::memcpy(out, m.data(), m.size() * sizeof(dist_t));
}
void ref_bfs(void* p, const int32_t* rc, int k, uint16_t* out) {
    auto* h = static_cast<RefHandle*>(p);
    auto m = BfsOracle::queryAllPairs(h->grid, to_pts(rc, k));
    std::memcpy(out, m.data(), m.size() * sizeof(dist_t));
}
int ref_S(void* p) { return static_cast<RefHandle*>(p)->oracle.S; }
int ref_H(void* p) { return static_cast<RefHandle*>(p)->oracle.H; }
int ref_num_leaves(void* p) { return int(static_cast<RefHandle*>(p)->oracle.leaves.size()); }
int ref_num_nodes(void* p, int lvl) { return int(static_cast<RefHandle*>(p)->oracle.levels[lvl - 1].size()); }
int ref_node_nV(void* p, int lvl, int idx) {
    return static_cast<RefHandle*>(p)->oracle.levels[lvl - 1][idx].numInterfaceVertices();
}
int ref_node_nExt(void* p, int lvl, int idx) {
    return static_cast<RefHandle*>(p)->oracle.levels[lvl - 1][idx].numExtPorts();
}
void ref_leaf_D(void* p, int idx, uint16_t* out) {
    const auto& D = static_cast<RefHandle*>(p)->oracle.leaves[idx].D;
    std::memcpy(out, D.data(), D.size() * 2);
}
void ref_leaf_S(void* p, int idx, uint16_t* out) {
    const auto& S = static_cast<RefHandle*>(p)->oracle.leaves[idx].S;
    std::memcpy(out, S.data(), S.size() * 2);
}
void ref_node_DU(void* p, int lvl, int idx, uint16_t* out) {
    const auto& D = static_cast<RefHandle*>(p)->oracle.levels[lvl - 1][idx].D_U;
    std::memcpy(out, D.data(), D.size() * 2);
}
void ref_node_SU(void* p, int lvl, int idx, uint16_t* out) {
    const auto& S = static_cast<RefHandle*>(p)->oracle.levels[lvl - 1][idx].S_U;
    std::memcpy(out, S.data(), S.size() * 2);
}
void ref_last_batch_stats(void* p, int32_t* out3) {
    const auto& s = static_cast<RefHandle*>(p)->oracle.last_batch_stats;
    out3[0] = s.n_touched_leaves;
    out3[1] = s.n_dirty_ancestors;
    out3[2] = s.n_pivots;
}
uint64_t ref_total_memory_bytes(void* p) { return static_cast<RefHandle*>(p)->oracle.totalMemoryBytes(); }
// effectiveMemoryBytes over a caller-given grid (rows x cols cell codes).
uint64_t ref_effective_memory_bytes(void* p, int rows, int cols, const uint8_t* cells) {
    GridMap g(rows, cols, CellState::Unknown);
    for (size_t i = 0; i < size_t(rows) * cols; ++i) g.observed[i] = static_cast<CellState>(cells[i]);
    return static_cast<RefHandle*>(p)->oracle.effectiveMemoryBytes(g);
}

// ---- Table import: lay out the reference hierarchy from tables computed
// elsewhere (the GPU), so that the reference query path can be timed on
// large maps without paying its CPU build.  Geometry follows
// src/minplus_oracle.cpp:54-67 and the LeafTile / InternalNode constructors.
void ref_import_begin(void* p) {
    auto* h = static_cast<RefHandle*>(p);
    MinPlusOracle& o = h->oracle;
    const int b = o.config.tile_size_b, g = o.config.group_w;
    const int md = std::max(h->grid.rows, h->grid.cols);
    o.S = b;
    while (o.S < md) o.S *= g;
    o.H = 0;
    for (int t = o.S / b; t > 1; t /= g) ++o.H;
    o.num_leaf_tiles_x = o.num_leaf_tiles_y = o.S / b;
    o.leaves.clear();
    o.leaves.reserve(size_t(o.num_leaf_tiles_x) * o.num_leaf_tiles_y);
    for (int ty = 0; ty < o.num_leaf_tiles_y; ++ty)
        for (int tx = 0; tx < o.num_leaf_tiles_x; ++tx) o.leaves.emplace_back(tx, ty, b);
    o.levels.assign(size_t(o.H), {});
    for (int lvl = 1; lvl <= o.H; ++lvl) {
        int side = b;
        for (int l = 0; l < lvl; ++l) side *= g;
        const int nt = std::max(1, o.S / side);
        for (int ty = 0; ty < nt; ++ty)
            for (int tx = 0; tx < nt; ++tx) o.levels[lvl - 1].emplace_back(lvl, tx, ty, b, g, g);
    }
    o.epoch++;
}
void ref_import_leaf(void* p, int idx, const uint16_t* D) {
    LeafTile& L = static_cast<RefHandle*>(p)->oracle.leaves[idx];
    std::memcpy(L.D.data(), D, L.D.size() * 2);
    const int n = L.b * L.b, np = L.numPorts();
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j)
            L.S[size_t(i) * np + j] = L.D[size_t(L.localIdx(L.ports[i])) * n + L.localIdx(L.ports[j])];
}
void ref_import_node(void* p, int lvl, int idx, const uint16_t* DU) {
    InternalNode& N = static_cast<RefHandle*>(p)->oracle.levels[lvl - 1][idx];
    std::memcpy(N.D_U.data(), DU, N.D_U.size() * 2);
    const int ne = N.numExtPorts(), nV = N.numInterfaceVertices();
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j)
            N.S_U[size_t(i) * ne + j] = N.D_U[size_t(N.extPortIdx[i]) * nV + N.extPortIdx[j]];
    N.refreshChildToExt();
}

// Time nq queries (query q has ks[q] viewpoints, packed back to back in rc)
// over nthreads worker threads, each with its own HTripQueryScratch -- the
// reference supports concurrent const queries with distinct scratch
// (include/htrip/minplus_oracle.hpp:189-240).  Returns wall seconds; writes
// each result to out (packed k*k blocks) when out is not null.
double ref_time_queries(void* p, const int32_t* rc, const int32_t* ks, int nq, int nthreads, uint16_t* out) {
    auto* h = static_cast<RefHandle*>(p);
    std::vector<size_t> vp_off(size_t(nq) + 1, 0), out_off(size_t(nq) + 1, 0);
    for (int q = 0; q < nq; ++q) {
        vp_off[q + 1] = vp_off[q] + size_t(ks[q]);
        out_off[q + 1] = out_off[q] + size_t(ks[q]) * ks[q];
    }
    std::atomic<int> next{0};
    auto worker = [&]() {
        MinPlusOracle::HTripQueryScratch scratch;
        std::vector<dist_t> m;
        for (;;) {
            int q = next.fetch_add(1);
            if (q >= nq) break;
            auto pts = to_pts(rc + 2 * vp_off[q], ks[q]);
            h->oracle.queryAllPairs(pts, h->grid, m, scratch);
            if (out) std::memcpy(out + out_off[q], m.data(), m.size() * 2);
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    if (nthreads <= 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
        for (auto& t : th) t.join();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// FrontierManager::detectAndCluster on an observed grid.  Returns the number
// of clusters; the first cap_clusters get (id, representative r, c, size) in
// meta and their cells back to back in cells (up to cap_cells pairs).  clr
// (rows*cols) receives GridMap::getClearance and front the frontier flags.
int ref_frontiers(int rows, int cols, const uint8_t* obs, int min_size, int max_size, int32_t* meta,
                  int cap_clusters, int32_t* cells, int cap_cells, uint8_t* clr, uint8_t* front) {
    GridMap g(rows, cols, CellState::Unknown);
    for (size_t i = 0; i < size_t(rows) * cols; ++i) g.observed[i] = static_cast<CellState>(obs[i]);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            clr[size_t(r) * cols + c] = g.getClearance(r, c);
            front[size_t(r) * cols + c] = FrontierManager::isFrontierCell(g, Point{r, c}) ? 1 : 0;
        }
    FrontierManager fm(min_size, max_size);
    const auto cl = fm.detectAndCluster(g);
    int w = 0;
    for (size_t i = 0; i < cl.size() && int(i) < cap_clusters; ++i) {
        meta[4 * i] = cl[i].id;
        meta[4 * i + 1] = cl[i].representative.r;
        meta[4 * i + 2] = cl[i].representative.c;
        meta[4 * i + 3] = cl[i].size();
        for (const Point& p : cl[i].cells)
            if (w < cap_cells) {
                cells[2 * w] = p.r;
                cells[2 * w + 1] = p.c;
                ++w;
            }
    }
    return int(cl.size());
}
// MinPlusOracle::liftRoot on a grown grid given as rows x cols observed codes.
void ref_lift_root(void* p, int rows, int cols, const uint8_t* cells) {
    auto* h = static_cast<RefHandle*>(p);
    h->grid = GridMap(rows, cols, CellState::Unknown);
    for (size_t i = 0; i < size_t(rows) * cols; ++i) h->grid.observed[i] = static_cast<CellState>(cells[i]);
    h->oracle.liftRoot(h->grid);
}
// GridMap::loadMovingAI: 1 on success with rows/cols set and, when they fit
// in cap cells, the ground truth codes written to out; 0 where it returns false.
int ref_load_movingai(const char* path, int32_t* rows, int32_t* cols, uint8_t* out, int64_t cap) {
    GridMap g;
    if (!g.loadMovingAI(path)) return 0;
    *rows = g.rows;
    *cols = g.cols;
    if (int64_t(g.rows) * g.cols <= cap)
        for (size_t i = 0; i < g.ground_truth.size(); ++i) out[i] = uint8_t(g.ground_truth[i]);
    return 1;
}
// GridMap::generateProceduralRooms(rows, cols, room, seed), then, when
// coverage > 0, initExplorationSnapshot({sr, sc}, coverage, island, seed2):
// the ground truth and observed codes (rows * cols each) and the snapshot's
// return value (1 when it was not run).
int ref_world(int rows, int cols, int room, uint64_t seed, int sr, int sc, double coverage, double island,
              uint64_t seed2, uint8_t* truth, uint8_t* obs) {
    GridMap g;
    g.generateProceduralRooms(rows, cols, room, seed);
    const int ok = coverage > 0 ? int(g.initExplorationSnapshot(Point{sr, sc}, coverage, island, seed2)) : 1;
    for (size_t i = 0; i < g.ground_truth.size(); ++i) truth[i] = uint8_t(g.ground_truth[i]);
    for (size_t i = 0; i < g.observed.size(); ++i) obs[i] = uint8_t(g.observed[i]);
    return ok;
}
// MinPlusOracle::liftPoint: X levels back to back (lens[l] entries each),
// par/quad/off H+1 entries, active[l] in the reference's order padded with -1
// to lens[l].
void ref_lift_point(void* p, int r, int c, uint16_t* X, int32_t* lens, int32_t* par, int32_t* quad, int32_t* off,
                    int32_t* active) {
    auto* h = static_cast<RefHandle*>(p);
    MinPlusOracle::LiftedViewpoint lv;
    h->oracle.liftPoint(Point{r, c}, lv);
    size_t o = 0;
    for (size_t l = 0; l < lv.X.size(); ++l) {
        lens[l] = int32_t(lv.X[l].size());
        par[l] = lv.parent_node_idx[l];
        quad[l] = lv.child_quad[l];
        off[l] = lv.child_offset[l];
        std::memcpy(X + o, lv.X[l].data(), lv.X[l].size() * 2);
        for (size_t i = 0; i < lv.X[l].size(); ++i) active[o + i] = i < lv.active[l].size() ? lv.active[l][i] : -1;
        o += lv.X[l].size();
    }
}
// MinPlusOracle::queryAllPairsIncremental with a persistent state: the
// handle keeps the scratch, the state and the matrix between calls.
struct IncHandle {
    MinPlusOracle::HTripQueryScratch scratch;
    MinPlusOracle::HTripIncrementalState state;
    std::vector<dist_t> m;
};
void* ref_inc_create() { return new IncHandle(); }
void ref_inc_destroy(void* q) { delete static_cast<IncHandle*>(q); }
// Seconds for one incremental refresh of the k listed viewpoints; out gets k*k.
double ref_inc_query(void* p, void* q, const int32_t* rc, int k, uint16_t* out) {
    auto* h = static_cast<RefHandle*>(p);
    auto* inc = static_cast<IncHandle*>(q);
    const auto pts = to_pts(rc, k);
    auto t0 = std::chrono::steady_clock::now();
    h->oracle.queryAllPairsIncremental(pts, h->grid, inc->m, inc->scratch, inc->state);
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::memcpy(out, inc->m.data(), inc->m.size() * 2);
    return dt;
}
// Seconds for n liftPoint calls on the listed cells (one thread).
double ref_time_lifts(void* p, const int32_t* rc, int n) {
    auto* h = static_cast<RefHandle*>(p);
    MinPlusOracle::LiftedViewpoint lv;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) h->oracle.liftPoint(Point{rc[2 * i], rc[2 * i + 1]}, lv);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
