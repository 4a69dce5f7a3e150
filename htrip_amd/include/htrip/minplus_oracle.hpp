#pragma once
// Drop-in replacement of the reference's include/htrip/minplus_oracle.hpp:
// htrip::MinPlusOracle (alias HTripOracle) with the same hot-path signatures
// (build, onCellOpened, onBatchCellsOpened, onCellBlocked, liftRoot,
// onBatchLeavesChanged, onCellWeightChanged / onTileWeightsChanged for unit
// weights, queryAllPairs with HTripQueryScratch, queryAllPairsIncremental,
// queryDistance, liftPoint) and
// the public metadata callers read
// (config, S, H, num_leaf_tiles_x/y, leaves.size(), depth(), findLcaLevel,
// totalMemoryBytes, last_batch_stats).  The index itself lives in HBM behind
// the C ABI of include/htrip_gpu.h; link with -lhtrip_gpu.
//
// Differences from the reference, all outside the results:
//  * leaves / levels hold tile metadata only; tables are read with
//    readLeafD / readNodeDU (copies from the device).
//  * GPU failures throw std::runtime_error (the reference has no error channel
//    besides build()'s bool).
//  * hybrid_k_threshold is accepted and ignored: the GPU query is exact for
//    every k, as the reference's BFS fallback is (minplus_oracle.cpp:951-958).
//  * Cell weights other than 1 are refused (std::invalid_argument) by build,
//    liftRoot and the weight-change calls: the index is unit-weight.
#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "grid_map.hpp"
#include "htrip_gpu.h"
#include "types.hpp"

namespace htrip {

namespace detail {
inline htg_ctx* device_context() {
    static htg_ctx* ctx = [] {
        htg_ctx* c = nullptr;
        if (htg_ctx_create(0, &c) != HTG_OK) throw std::runtime_error(std::string("htrip: ") + htg_last_error());
        return c;
    }();
    return ctx;
}
inline void check(int rc) {
    if (rc != HTG_OK) throw std::runtime_error(std::string("htrip: ") + htg_last_error());
}
// The GPU index computes unit-weight distances: a grid whose passable cells
// carry another weight is refused rather than answered wrongly.
inline void require_unit_weights(const GridMap& g) {
    const size_t n = std::min(g.weights.size(), g.observed.size());
    for (size_t i = 0; i < n; ++i)
        if (g.observed[i] != CellState::Obstacle && g.weights[i] != 1)
            throw std::invalid_argument("htrip: the GPU index supports cell weight 1 only");
}
}  // namespace detail

// Tile metadata (include/htrip/leaf_tile.hpp:41-52, internal_node.hpp:35-46).
struct LeafTile {
    int tx = 0, ty = 0, R0 = 0, C0 = 0, b = 16;
};
struct InternalNode {
    int level = 1, tx = 0, ty = 0, R0 = 0, C0 = 0, s = 32;
};

class MinPlusOracle {
public:
    HierarchyConfig config;
    int S = 0;
    int H = 0;
    int num_leaf_tiles_x = 0;
    int num_leaf_tiles_y = 0;
    std::vector<LeafTile> leaves;
    std::vector<std::vector<InternalNode>> levels;
    uint32_t update_counter = 0;
    uint64_t epoch = 0;

    struct LastBatchStats {  // minplus_oracle.hpp:251-256
        int n_touched_leaves = 0;
        int n_dirty_ancestors = 0;
        int n_pivots = 0;
    };
    LastBatchStats last_batch_stats{};

    // Caller-owned query workspace (minplus_oracle.hpp:189-238).  It owns a
    // device workspace (htg_query_ws: stream, plan and lift buffers, result
    // staging; created by the first query through it and grown, never
    // shrunk), so that const queryAllPairs calls with distinct scratch may run
    // on one oracle from several threads at once, as the reference allows.
    // The public host fields keep the reference's names and types for source
    // compatibility and are sized by ensureCapacity as there; the GPU query
    // does not fill them.  A copy starts with a workspace of its own.
    struct HTripQueryScratch {
        struct ViewpointLift {
            Point pt;
            int leaf_tx = 0, leaf_ty = 0;
            std::vector<std::vector<dist_t>> X;    // [lvl][port]
            std::vector<std::vector<int>> active;  // [lvl][idx]
        };
        std::vector<ViewpointLift> vp_data;
        std::vector<std::vector<int>> leaf_vps;
        std::vector<std::vector<int>> node_vps[2];
        std::vector<dist_t> proj_buf;
        std::vector<dist_t> block_buf;
        std::vector<int> row_marks;
        std::vector<int> row_list;
        std::vector<uint8_t> closed;
        std::vector<uint8_t> vp_new;
        std::vector<int32_t> rc;  // host staging of the viewpoints

        // Grow-only: k viewpoints with H + 1 lift levels, num_leaves leaves.
        void ensureCapacity(size_t k, int H, size_t num_leaves) {
            const size_t levels = size_t(H + 1);
            if (vp_data.size() < k) vp_data.resize(k);
            for (size_t i = 0; i < k; ++i) {
                if (vp_data[i].X.size() < levels) vp_data[i].X.resize(levels);
                if (vp_data[i].active.size() < levels) vp_data[i].active.resize(levels);
            }
            if (leaf_vps.size() < num_leaves) leaf_vps.resize(num_leaves);
            for (auto& nv : node_vps)
                if (nv.size() < num_leaves) nv.resize(num_leaves);
            if (closed.size() < k * k) closed.resize(k * k);
            if (vp_new.size() < k) vp_new.resize(k);
            if (rc.capacity() < 2 * k) rc.reserve(2 * k);
        }

        HTripQueryScratch() = default;
        HTripQueryScratch(const HTripQueryScratch& o)
            : vp_data(o.vp_data), leaf_vps(o.leaf_vps), node_vps{o.node_vps[0], o.node_vps[1]},
              proj_buf(o.proj_buf), block_buf(o.block_buf), row_marks(o.row_marks), row_list(o.row_list),
              closed(o.closed), vp_new(o.vp_new), rc(o.rc) {}
        HTripQueryScratch& operator=(const HTripQueryScratch& o) {
            if (this != &o) *this = HTripQueryScratch(o);
            return *this;
        }
        HTripQueryScratch(HTripQueryScratch&&) noexcept = default;
        HTripQueryScratch& operator=(HTripQueryScratch&&) noexcept = default;

        // The device workspace (created on first use).
        htg_query_ws* workspace() {
            if (!ws_) {
                htg_query_ws* w = nullptr;
                detail::check(htg_query_ws_create(detail::device_context(), &w));
                ws_.reset(w);
            }
            return ws_.get();
        }
        // Device bytes the workspace holds (0 before the first query).
        [[nodiscard]] uint64_t deviceBytes() const {
            uint64_t n = 0;
            if (ws_) detail::check(htg_query_ws_bytes(ws_.get(), &n));
            return n;
        }

    private:
        struct WsDeleter {
            void operator()(htg_query_ws* w) const noexcept { htg_query_ws_destroy(w); }
        };
        std::unique_ptr<htg_query_ws, WsDeleter> ws_;
    };

    MinPlusOracle() = default;
    explicit MinPlusOracle(HierarchyConfig cfg) : config(cfg) {}
    MinPlusOracle(const MinPlusOracle& o) { *this = o; }
    MinPlusOracle& operator=(const MinPlusOracle& o) {
        if (this == &o) return *this;
        copyMeta(o);
        map_.reset();
        if (o.map_) {
            htg_map* m = nullptr;
            detail::check(htg_map_clone(o.map_.get(), 0, &m));
            map_.reset(m);
        }
        return *this;
    }
    MinPlusOracle(MinPlusOracle&&) noexcept = default;
    MinPlusOracle& operator=(MinPlusOracle&&) noexcept = default;

    // src/minplus_oracle.cpp:48-137
    bool build(const GridMap& grid) {
        if (!config.isValid() || grid.rows <= 0 || grid.cols <= 0) return false;
        detail::require_unit_weights(grid);
        std::vector<uint8_t> cells(grid.observed.size());
        for (size_t i = 0; i < cells.size(); ++i) cells[i] = uint8_t(grid.observed[i]);
        htg_config c{config.tile_size_b, config.group_w, config.group_h, config.hybrid_k_threshold};
        htg_map* m = nullptr;
        const int rc = htg_map_create(detail::device_context(), &c, grid.rows, grid.cols, cells.data(), 0, &m);
        if (rc == HTG_E_INVALID) return false;
        detail::check(rc);
        map_.reset(m);
        detail::check(htg_build(m));
        readMeta();
        epoch++;
        return true;
    }

    // src/minplus_oracle.cpp:2008-2114.  grid is the grown map (it contains
    // the old one, within g*S on each side); S doubles (x group_w) and H grows
    // by one.  Kept tables and the map's cell state inside the old square are
    // the device's; the rest is computed from grid.observed.
    void liftRoot(const GridMap& grid) {
        if (!map_) throw std::runtime_error("htrip: liftRoot before build");
        detail::require_unit_weights(grid);
        std::vector<uint8_t> cells(grid.observed.size());
        for (size_t i = 0; i < cells.size(); ++i) cells[i] = uint8_t(grid.observed[i]);
        htg_map* m = map_.get();
        detail::check(htg_lift_root(&m, grid.rows, grid.cols, cells.data()));
        (void)map_.release();  // destroyed by htg_lift_root
        map_.reset(m);
        readMeta();
        epoch++;
    }

    [[nodiscard]] int depth() const noexcept { return H; }
    [[nodiscard]] int leafIndex(int tx, int ty) const noexcept { return ty * num_leaf_tiles_x + tx; }
    // src/minplus_oracle.cpp:15-22
    [[nodiscard]] int tileIndexAtLevel(int lvl, int tx, int ty) const noexcept {
        int side = config.tile_size_b;
        for (int l = 0; l < lvl; ++l) side *= config.group_w;
        int nx = S / side;
        if (nx <= 0) nx = 1;
        return ty * nx + tx;
    }
    // src/minplus_oracle.cpp:24-46
    [[nodiscard]] int findLcaLevel(Point p1, Point p2) const noexcept {
        const int b = config.tile_size_b;
        int ax = p1.c / b, ay = p1.r / b, bx = p2.c / b, by = p2.r / b;
        if (ax == bx && ay == by) return 0;
        for (int lvl = 1; lvl <= H; ++lvl) {
            ax /= config.group_w;
            ay /= config.group_h;
            bx /= config.group_w;
            by /= config.group_h;
            if (ax == bx && ay == by) return lvl;
        }
        return H + 1;
    }

    // src/minplus_oracle.cpp:139-177.  The cell must already be Free in grid.
    void onCellOpened(Point p, const GridMap& grid) {
        if (!map_ || grid.getObserved(p) == CellState::Obstacle) return;
        detail::check(htg_open_cell(map_.get(), p.r, p.c));
        ++update_counter;
    }

    // src/minplus_oracle.cpp:304-362.  The GPU index computes unit-weight
    // distances only: setting a cell's weight to 1 changes no table, any other
    // weight is refused (std::invalid_argument, the grid left unchanged)
    // rather than answered wrongly.
    void onCellWeightChanged(Point p, dist_t new_weight, GridMap& grid) {
        if (!grid.inBounds(p) || grid.getCellWeight(p) == new_weight) return;
        if (new_weight != 1) throw std::invalid_argument("htrip: the GPU index supports cell weight 1 only");
        grid.setCellWeight(p, new_weight);
    }
    // src/minplus_oracle.cpp:364-425, with the same rule for the weights.
    void onTileWeightsChanged(int tx, int ty, const std::vector<std::pair<Point, dist_t>>& cell_weights,
                              GridMap& grid) {
        if (tx < 0 || tx >= num_leaf_tiles_x || ty < 0 || ty >= num_leaf_tiles_y) return;
        for (const auto& pw : cell_weights)
            if (grid.inBounds(pw.first) && pw.second != 1)
                throw std::invalid_argument("htrip: the GPU index supports cell weight 1 only");
        for (const auto& pw : cell_weights) grid.setCellWeight(pw.first, pw.second);
    }
    // src/minplus_oracle.cpp:507-527: the listed leaves' cells may have
    // changed in grid.  Cells that became impassable are blocked first (their
    // leaves and ancestors are rebuilt), then cells that became passable are
    // opened (the incremental reveal); both are exact for any change, where
    // the reference's upward propagation assumes decreases only.
    void onBatchLeavesChanged(const std::vector<int>& leaf_indices, const GridMap& grid) {
        if (!map_ || leaf_indices.empty()) return;
        const int b = config.tile_size_b, nleaves = num_leaf_tiles_x * num_leaf_tiles_y;
        std::vector<int32_t> open_rc, block_rc;
        for (int li : leaf_indices) {
            if (li < 0 || li >= nleaves) continue;
            const int r0 = (li / num_leaf_tiles_x) * b, c0 = (li % num_leaf_tiles_x) * b;
            for (int r = r0; r < std::min(r0 + b, grid.rows); ++r)
                for (int c = c0; c < std::min(c0 + b, grid.cols); ++c) {
                    auto& to = grid.isPassable(r, c) ? open_rc : block_rc;
                    to.push_back(r);
                    to.push_back(c);
                }
        }
        if (!block_rc.empty()) detail::check(htg_block_cells(map_.get(), block_rc.data(), int(block_rc.size() / 2)));
        if (!open_rc.empty()) detail::check(htg_open_cells(map_.get(), open_rc.data(), int(open_rc.size() / 2)));
        ++update_counter;
        epoch++;
    }

    // src/minplus_oracle.cpp:247-302.  The cell must already be blocked in
    // grid; the device map marks it Obstacle, recomputes its leaf and rebuilds
    // the ancestor nodes along its branch.
    void onCellBlocked(Point p, const GridMap& /*grid*/) {
        if (!map_) return;
        const int32_t rc[2] = {p.r, p.c};
        detail::check(htg_block_cells(map_.get(), rc, 1));
        ++update_counter;
        epoch++;
    }

    // src/minplus_oracle.cpp:530-729.  Cells must already be Free in grid
    // (README.md:71-91); cells the grid marks Obstacle carry no edges in the
    // reference (edgeWeight is INF) and are skipped here.
    void onBatchCellsOpened(const std::vector<Point>& newly_opened, const GridMap& grid) {
        last_batch_stats = {};
        if (!map_ || newly_opened.empty()) return;
        std::vector<int32_t> rc;
        rc.reserve(2 * newly_opened.size());
        for (const Point& p : newly_opened) {
            if (grid.getObserved(p) == CellState::Obstacle) continue;
            rc.push_back(p.r);
            rc.push_back(p.c);
        }
        detail::check(htg_open_cells(map_.get(), rc.data(), int32_t(rc.size() / 2)));
        htg_batch_stats s{};
        detail::check(htg_last_batch_stats(map_.get(), &s));
        last_batch_stats = {s.n_touched_leaves, s.n_dirty_ancestors, s.n_pivots};
        ++update_counter;
        epoch++;
    }

    // src/minplus_oracle.cpp:934-1419.  Runs on the scratch's own device
    // workspace and stream: distinct scratch objects may query one oracle
    // from several threads at once (no update may run meanwhile, as in the
    // reference).
    void queryAllPairs(const std::vector<Point>& viewpoints, const GridMap& /*grid*/, std::vector<dist_t>& out,
                       HTripQueryScratch& scratch) const {
        const size_t k = viewpoints.size();
        out.resize(k * k);
        if (k == 0) return;
        if (!map_) {
            std::fill(out.begin(), out.end(), INF);
            return;
        }
        scratch.rc.resize(2 * k);
        for (size_t i = 0; i < k; ++i) {
            scratch.rc[2 * i] = viewpoints[i].r;
            scratch.rc[2 * i + 1] = viewpoints[i].c;
        }
        detail::check(htg_query_all_pairs_ws(map_.get(), scratch.workspace(), scratch.rc.data(), int32_t(k),
                                             out.data()));
    }
    // src/minplus_oracle.cpp:1421-1426
    [[nodiscard]] std::vector<dist_t> queryAllPairs(const std::vector<Point>& viewpoints, const GridMap& grid) const {
        std::vector<dist_t> m;
        queryAllPairs(viewpoints, grid, m, query_scratch);
        return m;
    }
    // src/minplus_oracle.cpp:829-928.  The early answers are the reference's
    // (INF for an endpoint that grid does not hold passable or that lies
    // outside the leaf tiles, 0 for s == t); otherwise the 2-viewpoint
    // batched query, which lifts both points and meets them at every common
    // ancestor, as the reference's LCA meet does.
    [[nodiscard]] dist_t queryDistance(Point s, Point t, const GridMap& grid) const {
        if (!grid.isPassable(s) || !grid.isPassable(t)) return INF;
        if (s == t) return 0;
        if (!inLeafRange(s) || !inLeafRange(t)) return INF;
        auto m = queryAllPairs({s, t}, grid);
        return m[1];
    }

    // include/htrip/minplus_oracle.hpp:177-183
    struct LiftedViewpoint {
        std::vector<std::vector<dist_t>> X;    // [lvl][port]
        std::vector<std::vector<int>> active;  // [lvl]: ports with X < INF, by increasing X
        std::vector<int> parent_node_idx;      // [lvl]: ancestor node at lvl (-1 at 0)
        std::vector<int> child_quad;           // [lvl]: quadrant of the child in that node (-1 at 0)
        std::vector<int> child_offset;         // [lvl]: start of that child's ports in the node's V_U
    };
    // queryDistance that also returns both lifts (minplus_oracle.hpp:317-319).
    [[nodiscard]] dist_t queryDistance(Point s, Point t, const GridMap& grid, LiftedViewpoint& X_s,
                                       LiftedViewpoint& Y_t) const;

    // src/minplus_oracle.cpp:731-823.  X comes from the device; the active
    // lists hold the same ports as the reference's, ordered by X and, among
    // equal distances, by port index (the reference's std::sort leaves the
    // order of ties unspecified).
    void liftPoint(Point p, LiftedViewpoint& lifted) const {
        if (!map_) throw std::runtime_error("htrip: liftPoint before build");
        size_t cap = size_t(4 * config.tile_size_b - 4);
        for (int lvl = 1; lvl <= H; ++lvl) {
            int32_t nn = 0, nv = 0, ne = 0;
            detail::check(htg_node_dims(map_.get(), lvl, &nn, &nv, &ne));
            cap += size_t(ne);
        }
        const size_t n = size_t(H) + 1;
        std::vector<dist_t> x(cap);
        std::vector<int32_t> lens(n), par(n), quad(n), off(n);
        detail::check(htg_lift_point(map_.get(), p.r, p.c, x.data(), int64_t(cap), lens.data(), par.data(),
                                     quad.data(), off.data()));
        lifted.X.resize(n);
        lifted.active.resize(n);
        lifted.parent_node_idx.assign(par.begin(), par.end());
        lifted.child_quad.assign(quad.begin(), quad.end());
        lifted.child_offset.assign(off.begin(), off.end());
        size_t o = 0;
        for (size_t l = 0; l < n; ++l) {
            auto& X = lifted.X[l];
            X.assign(x.begin() + std::ptrdiff_t(o), x.begin() + std::ptrdiff_t(o + size_t(lens[l])));
            o += size_t(lens[l]);
            auto& a = lifted.active[l];
            a.clear();
            for (int i = 0; i < int(X.size()); ++i)
                if (X[size_t(i)] < INF) a.push_back(i);
            std::stable_sort(a.begin(), a.end(), [&X](int u, int v) { return X[size_t(u)] < X[size_t(v)]; });
        }
    }

private:
    [[nodiscard]] bool inLeafRange(Point p) const noexcept {
        const int b = config.tile_size_b;
        return p.r >= 0 && p.c >= 0 && p.c / b < num_leaf_tiles_x && p.r / b < num_leaf_tiles_y;
    }

public:
    // include/htrip/minplus_oracle.hpp:262-272.  The device-side state
    // (htg_inc) keeps the viewpoints, the matrix and the table clock it was
    // taken at; the public fields mirror the reference's for callers that
    // read them.  Copies start from an empty state (the next call recomputes).
    struct HTripIncrementalState {
        bool initialized = false;
        uint64_t epoch_seen = 0;
        uint32_t last_update_counter = 0;
        std::vector<Point> vp_pts;
        std::vector<dist_t> matrix;
        int last_recomputed = 0;  // viewpoints recomputed by the last call (k or 0)

        HTripIncrementalState() = default;
        HTripIncrementalState(const HTripIncrementalState& o)
            : initialized(false), epoch_seen(o.epoch_seen), last_update_counter(o.last_update_counter),
              vp_pts(o.vp_pts), matrix(o.matrix) {}
        HTripIncrementalState& operator=(const HTripIncrementalState& o) {
            if (this != &o) {
                *this = HTripIncrementalState(o);
            }
            return *this;
        }
        HTripIncrementalState(HTripIncrementalState&&) noexcept = default;
        HTripIncrementalState& operator=(HTripIncrementalState&&) noexcept = default;

        htg_inc* handle() {
            if (!h_) {
                htg_inc* p = nullptr;
                detail::check(htg_inc_create(&p));
                h_.reset(p);
            }
            return h_.get();
        }

    private:
        struct IncDeleter {
            void operator()(htg_inc* p) const noexcept { htg_inc_destroy(p); }
        };
        std::unique_ptr<htg_inc, IncDeleter> h_;
    };

    // src/minplus_oracle.cpp:1444-2002.  Same contract: the result equals
    // queryAllPairs on the current map.  k = 0 clears the state (:1450-1454).
    // The device keeps a version stamp per leaf and node table; the cached
    // matrix is returned without device work while no update has rewritten
    // the leaf or an ancestor of any viewpoint since it was computed (no
    // entry can have changed), otherwise one batched query recomputes it.
    void queryAllPairsIncremental(const std::vector<Point>& viewpoints, const GridMap& /*grid*/,
                                  std::vector<dist_t>& out_matrix, HTripQueryScratch& scratch,
                                  HTripIncrementalState& state) const {
        const size_t k = viewpoints.size();
        if (k == 0 || !map_) {
            if (map_) detail::check(htg_query_all_pairs_incremental(map_.get(), state.handle(), nullptr, 0,
                                                                    nullptr, nullptr));
            out_matrix.assign(k * k, INF);
            state.initialized = false;
            return;
        }
        scratch.rc.resize(2 * k);
        for (size_t i = 0; i < k; ++i) {
            scratch.rc[2 * i] = viewpoints[i].r;
            scratch.rc[2 * i + 1] = viewpoints[i].c;
        }
        out_matrix.resize(k * k);
        int32_t n = 0;
        detail::check(htg_query_all_pairs_incremental(map_.get(), state.handle(), scratch.rc.data(), int32_t(k),
                                                      out_matrix.data(), &n));
        state.last_recomputed = n;
        state.initialized = true;
        state.epoch_seen = epoch;
        state.last_update_counter = update_counter;
        state.vp_pts = viewpoints;
        state.matrix = out_matrix;
    }

    // src/minplus_oracle.cpp:2116-2134
    [[nodiscard]] size_t totalMemoryBytes() const {
        if (!map_) return 0;
        htg_map_info info{};
        detail::check(htg_map_info_get(map_.get(), &info));
        return size_t(info.total_memory_bytes);
    }

    // src/minplus_oracle.cpp:2136-2182: the leaf D + S and node D_U + S_U
    // bytes of the tiles that lie inside grid and hold at least one Free cell
    // of it (child_to_ext is not counted, as there).
    [[nodiscard]] size_t effectiveMemoryBytes(const GridMap& grid) const {
        if (!map_) return 0;
        auto has_free = [&grid](int R0, int C0, int side) {
            if (R0 >= grid.rows || C0 >= grid.cols) return false;
            const int r1 = std::min(grid.rows, R0 + side), c1 = std::min(grid.cols, C0 + side);
            for (int r = R0; r < r1; ++r)
                for (int c = C0; c < c1; ++c)
                    if (grid.getObserved({r, c}) == CellState::Free) return true;
            return false;
        };
        const size_t b = size_t(config.tile_size_b), pl = 4 * b - 4;
        size_t total = 0;
        for (const LeafTile& l : leaves)
            if (has_free(l.R0, l.C0, l.b)) total += (b * b * b * b + pl * pl) * sizeof(dist_t);
        for (int lvl = 1; lvl <= H; ++lvl) {
            int32_t nn = 0, nv = 0, ne = 0;
            detail::check(htg_node_dims(map_.get(), lvl, &nn, &nv, &ne));
            for (const InternalNode& n : levels[size_t(lvl - 1)])
                if (has_free(n.R0, n.C0, n.s))
                    total += (size_t(nv) * size_t(nv) + size_t(ne) * size_t(ne)) * sizeof(dist_t);
        }
        return total;
    }

    // Table copies (the reference exposes LeafTile::D and InternalNode::D_U).
    [[nodiscard]] std::vector<dist_t> readLeafD(int leaf) const {
        const size_t n = size_t(config.tile_size_b) * config.tile_size_b;
        std::vector<dist_t> d(n * n);
        detail::check(htg_read_leaf_D(map_.get(), leaf, d.data()));
        return d;
    }
    [[nodiscard]] std::vector<dist_t> readNodeDU(int lvl, int idx) const {
        int32_t nn = 0, nv = 0, ne = 0;
        detail::check(htg_node_dims(map_.get(), lvl, &nn, &nv, &ne));
        std::vector<dist_t> d(size_t(nv) * nv);
        detail::check(htg_read_node_DU(map_.get(), lvl, idx, d.data()));
        return d;
    }
    [[nodiscard]] htg_map* handle() const noexcept { return map_.get(); }

    mutable HTripQueryScratch query_scratch;

private:
    struct MapDeleter {
        void operator()(htg_map* m) const noexcept { htg_map_destroy(m); }
    };
    std::unique_ptr<htg_map, MapDeleter> map_;

    void readMeta() {
        htg_map_info info{};
        detail::check(htg_map_info_get(map_.get(), &info));
        S = info.S;
        H = info.H;
        num_leaf_tiles_x = num_leaf_tiles_y = info.num_leaf_tiles_x;
        leaves.clear();
        const int b = config.tile_size_b;
        for (int ty = 0; ty < num_leaf_tiles_y; ++ty)
            for (int tx = 0; tx < num_leaf_tiles_x; ++tx) leaves.push_back({tx, ty, ty * b, tx * b, b});
        levels.assign(size_t(H), {});
        for (int lvl = 1; lvl <= H; ++lvl) {
            int side = b;
            for (int l = 0; l < lvl; ++l) side *= config.group_w;
            const int nt = S / side > 0 ? S / side : 1;
            for (int ty = 0; ty < nt; ++ty)
                for (int tx = 0; tx < nt; ++tx) levels[size_t(lvl - 1)].push_back({lvl, tx, ty, ty * side, tx * side, side});
        }
    }

    void copyMeta(const MinPlusOracle& o) {
        config = o.config;
        S = o.S;
        H = o.H;
        num_leaf_tiles_x = o.num_leaf_tiles_x;
        num_leaf_tiles_y = o.num_leaf_tiles_y;
        leaves = o.leaves;
        levels = o.levels;
        update_counter = o.update_counter;
        epoch = o.epoch;
        last_batch_stats = o.last_batch_stats;
    }
};

inline dist_t MinPlusOracle::queryDistance(Point s, Point t, const GridMap& grid, LiftedViewpoint& X_s,
                                           LiftedViewpoint& Y_t) const {
    if (!grid.isPassable(s) || !grid.isPassable(t)) return INF;
    if (s == t) return 0;
    if (!inLeafRange(s) || !inLeafRange(t)) return INF;
    liftPoint(s, X_s);
    liftPoint(t, Y_t);
    return queryDistance(s, t, grid);
}

using HTripOracle = MinPlusOracle;  // minplus_oracle.hpp:411

}  // namespace htrip
