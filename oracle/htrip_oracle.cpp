// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// CPU restatement of HTRIP's hierarchical min-plus distance oracle (the
// MinPlusOracle hot path: build, incremental reveal, queryAllPairs) and of
// its BFS ground truth.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker.  The
// product path (htrip_amd/, libhtrip_gpu.so) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root).  Semantics reproduced exactly:
//   * dist_t = uint16, INF = 60000, saturating add (include/htrip/types.hpp:34,45,66-70)
//   * canonical port order (src/leaf_tile.cpp:11-30)
//   * leaf D = tile-restricted BFS, D[i][i] = 0 for every cell (src/leaf_tile.cpp:56-208)
//   * node D_U = closure over V_U with seam pivots (src/internal_node.cpp:94-245)
//   * incremental batch reveal with pivot closures (src/minplus_oracle.cpp:530-729)
//   * liftPoint (src/minplus_oracle.cpp:731-823)
//   * all-pairs query by lift + inter-child meet + same-child shortcut
//     (src/minplus_oracle.cpp:934-1419), without the data-dependent pruning
//     (sorted active lists, early breaks, Manhattan "closed" flags), which the
//     reference documents and which never changes a result.
//   * BFS all-pairs (src/bfs_oracle.cpp:79-202) with a visited bitmap and
//     min(d, INF) reporting.
//
// Pinned by: tests/test_oracle.py (against the compiled reference in
// oracle/_ref, against brute-force region BFS on every table, against the
// known answers of the reference's test_exactness suites, and against the
// committed golden digests under tests/golden/).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

using dist_t = uint16_t;
static constexpr dist_t INF = 60000;

// include/htrip/types.hpp:66-70
static inline dist_t add_dist(dist_t a, dist_t b) {
    if (a >= INF || b >= INF) return INF;
    uint32_t s = uint32_t(a) + uint32_t(b);
    return dist_t(s >= INF ? INF : s);
}

struct Pt {
    int r, c;
    bool operator==(const Pt& o) const { return r == o.r && c == o.c; }
};

// include/htrip/grid_map.hpp:142-145 (isPassable), :273-278 (getCellWeight),
// :327-334 (edgeWeight).  Cell codes follow types.hpp:101-105:
// 0 = Unknown, 1 = Free, 2 = Obstacle.  All weights are 1 (the reference
// constructor fills weights with 1, src/grid_map.cpp:10-14).
struct Grid {
    int rows = 0, cols = 0;
    std::vector<uint8_t> obs;
    bool in(int r, int c) const { return r >= 0 && r < rows && c >= 0 && c < cols; }
    bool passable(int r, int c) const { return in(r, c) && obs[size_t(r) * cols + c] == 1; }
    dist_t weight(int r, int c) const {
        if (!in(r, c)) return INF;
        if (obs[size_t(r) * cols + c] == 2) return INF;
        return 1;
    }
    dist_t edge(Pt u, Pt v) const {
        dist_t wu = weight(u.r, u.c), wv = weight(v.r, v.c);
        if (wu >= INF || wv >= INF) return INF;
        dist_t avg = dist_t((uint32_t(wu) + wv) / 2);
        return avg == 0 ? 1 : avg;
    }
};

// src/leaf_tile.cpp:11-30
static void tile_ports(int R0, int C0, int s, std::vector<Pt>& out) {
    out.clear();
    for (int c = 0; c < s; ++c) out.push_back({R0, C0 + c});
    for (int c = 0; c < s; ++c) out.push_back({R0 + s - 1, C0 + c});
    for (int r = 1; r < s - 1; ++r) out.push_back({R0 + r, C0});
    for (int r = 1; r < s - 1; ++r) out.push_back({R0 + r, C0 + s - 1});
}

// src/internal_node.cpp:290-304
static int port_index(int r, int c, int R0, int C0, int s) {
    if (r == R0 && c >= C0 && c < C0 + s) return c - C0;
    if (r == R0 + s - 1 && c >= C0 && c < C0 + s) return s + (c - C0);
    if (c == C0 && r > R0 && r < R0 + s - 1) return 2 * s + (r - R0 - 1);
    if (c == C0 + s - 1 && r > R0 && r < R0 + s - 1) return 3 * s - 2 + (r - R0 - 1);
    return -1;
}

struct Decrease {  // types.hpp:153-170 (PortDecrease / ChildDecrease)
    int quad, a, b;
    dist_t w;
};

struct Leaf {
    int tx, ty, R0, C0;
    std::vector<Pt> ports;
    std::vector<dist_t> D, S;
};

struct Node {
    int lvl, tx, ty, R0, C0, s, cs;
    std::vector<Pt> VU;
    std::vector<int> off;
    std::vector<Pt> ext;
    std::vector<int> extIdx;
    std::vector<std::pair<int, int>> seams;
    std::vector<dist_t> DU, SU;
    int nV() const { return int(VU.size()); }
    int nExt() const { return int(ext.size()); }
};

// Pivot closure shared by leaf and node updates: for each pivot k (in order),
// row_i = min(row_i, D[i][k] (+) row_k) for every i with D[i][k] < INF.
// src/leaf_tile.cpp:348-371, src/internal_node.cpp:199-231, :372-399.
static void pivot_closure(std::vector<dist_t>& D, int n, const std::vector<int>& K) {
    for (int k : K) {
        const dist_t* rk = &D[size_t(k) * n];
        for (int i = 0; i < n; ++i) {
            dist_t dik = rk[i];
            if (dik >= INF) continue;
            dist_t* ri = &D[size_t(i) * n];
            for (int j = 0; j < n; ++j) {
                dist_t alt = add_dist(dik, rk[j]);
                if (alt < ri[j]) ri[j] = alt;
            }
        }
    }
}

struct Oracle {
    int b = 16, g = 2;
    Grid grid;
    int S = 0, H = 0, nl = 0;
    std::vector<Leaf> leaves;
    std::vector<std::vector<Node>> levels;
    int last_touched = 0, last_anc = 0, last_pivots = 0;

    int side(int lvl) const {
        int s = b;
        for (int l = 0; l < lvl; ++l) s *= g;
        return s;
    }
    int tiles_at(int lvl) const { return std::max(1, S / side(lvl)); }

    // src/leaf_tile.cpp:56-208 (unit-weight BFS branch :96-153, projection :198-207)
    void leaf_recompute(Leaf& L) {
        const int n = b * b;
        L.D.assign(size_t(n) * n, INF);
        for (int i = 0; i < n; ++i) L.D[size_t(i) * n + i] = 0;
        std::vector<uint8_t> pass(n);
        for (int i = 0; i < n; ++i) pass[i] = grid.passable(L.R0 + i / b, L.C0 + i % b);
        std::vector<int> q;
        q.reserve(n);
        for (int i = 0; i < n; ++i) {
            if (!pass[i]) continue;
            dist_t* row = &L.D[size_t(i) * n];
            q.clear();
            q.push_back(i);
            for (size_t h = 0; h < q.size(); ++h) {
                int u = q[h];
                dist_t alt = add_dist(row[u], 1);
                int ur = u / b, uc = u % b;
                int nb[4] = {ur > 0 ? u - b : -1, ur < b - 1 ? u + b : -1, uc > 0 ? u - 1 : -1,
                             uc < b - 1 ? u + 1 : -1};
                for (int v : nb) {
                    if (v < 0 || !pass[v] || row[v] != INF || v == i) continue;
                    row[v] = alt;
                    q.push_back(v);
                }
            }
        }
        leaf_project(L);
    }
    void leaf_project(Leaf& L) {
        const int n = b * b, np = int(L.ports.size());
        L.S.assign(size_t(np) * np, INF);
        for (int i = 0; i < np; ++i) {
            int u = (L.ports[i].r - L.R0) * b + (L.ports[i].c - L.C0);
            for (int j = 0; j < np; ++j) {
                int v = (L.ports[j].r - L.R0) * b + (L.ports[j].c - L.C0);
                L.S[size_t(i) * np + j] = L.D[size_t(u) * n + v];
            }
        }
    }

    // src/internal_node.cpp:13-92 (V_U, childOffset, extPorts/extPortIdx, seam candidates)
    Node make_node(int lvl, int tx, int ty) {
        Node N;
        N.lvl = lvl;
        N.tx = tx;
        N.ty = ty;
        N.cs = side(lvl - 1);
        N.s = N.cs * g;
        N.R0 = ty * N.s;
        N.C0 = tx * N.s;
        const int nch = g * g;
        N.off.assign(nch + 1, 0);
        std::vector<Pt> cp;
        for (int k = 0; k < nch; ++k) {
            tile_ports(N.R0 + (k / g) * N.cs, N.C0 + (k % g) * N.cs, N.cs, cp);
            N.off[k] = int(N.VU.size());
            N.VU.insert(N.VU.end(), cp.begin(), cp.end());
        }
        N.off[nch] = int(N.VU.size());
        tile_ports(N.R0, N.C0, N.s, N.ext);
        for (const Pt& e : N.ext) N.extIdx.push_back(find_vertex(N, e));
        // Seam candidates: cross-child port pairs at Manhattan distance 1,
        // listed in (i ascending, j ascending) order like the O(|V_U|^2) scan
        // at internal_node.cpp:81-89.
        for (int i = 0; i < N.nV(); ++i) {
            const Pt p = N.VU[i];
            const int ci = child_of(N, p);
            std::vector<int> js;
            const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
            for (int d = 0; d < 4; ++d) {
                Pt q{p.r + dr[d], p.c + dc[d]};
                if (q.r < N.R0 || q.r >= N.R0 + N.s || q.c < N.C0 || q.c >= N.C0 + N.s) continue;
                if (child_of(N, q) == ci) continue;
                int j = find_vertex(N, q);
                if (j > i) js.push_back(j);
            }
            std::sort(js.begin(), js.end());
            for (int j : js) N.seams.push_back({i, j});
        }
        const int nV = N.nV();
        N.DU.assign(size_t(nV) * nV, INF);
        for (int i = 0; i < nV; ++i) N.DU[size_t(i) * nV + i] = 0;
        const int ne = N.nExt();
        N.SU.assign(size_t(ne) * ne, INF);
        for (int i = 0; i < ne; ++i) N.SU[size_t(i) * ne + i] = 0;
        return N;
    }
    int child_of(const Node& N, Pt p) const {
        return ((p.r - N.R0) / N.cs) * g + (p.c - N.C0) / N.cs;
    }
    // src/internal_node.cpp:306-319
    int find_vertex(const Node& N, Pt p) const {
        const int cy = (p.r - N.R0) / N.cs, cx = (p.c - N.C0) / N.cs;
        if (p.r < N.R0 || p.c < N.C0 || cx >= g || cy >= g) return -1;
        const int q = cy * g + cx;
        int pi = port_index(p.r, p.c, N.R0 + cy * N.cs, N.C0 + cx * N.cs, N.cs);
        return pi < 0 ? -1 : N.off[q] + pi;
    }

    const std::vector<dist_t>& child_summary(int lvl, int ctx, int cty) const {
        if (lvl == 1) return leaves[size_t(cty) * nl + ctx].S;
        return levels[lvl - 2][size_t(cty) * tiles_at(lvl - 1) + ctx].SU;
    }

    // src/internal_node.cpp:94-245
    void node_recompute(Node& N) {
        const int nV = N.nV(), nch = g * g;
        std::fill(N.DU.begin(), N.DU.end(), INF);
        for (int i = 0; i < nV; ++i) N.DU[size_t(i) * nV + i] = 0;
        bool any_child_edge = false;
        for (int k = 0; k < nch; ++k) {
            const auto& sum = child_summary(N.lvl, N.tx * g + k % g, N.ty * g + k / g);
            const int o = N.off[k], np = N.off[k + 1] - o;
            bool has = false;
            for (int a = 0; a < np && !has; ++a)
                for (int c = 0; c < np; ++c)
                    if (a != c && sum[size_t(a) * np + c] < INF) { has = true; break; }
            if (!has) continue;
            any_child_edge = true;
            for (int a = 0; a < np; ++a)
                for (int c = 0; c < np; ++c) {
                    dist_t w = sum[size_t(a) * np + c];
                    dist_t& dst = N.DU[size_t(o + a) * nV + o + c];
                    if (w < INF && w < dst) dst = w;
                }
        }
        std::vector<int> K;
        std::vector<uint8_t> inK(nV, 0);
        for (auto [u, v] : N.seams) {
            Pt pu = N.VU[u], pv = N.VU[v];
            if (!grid.passable(pu.r, pu.c) || !grid.passable(pv.r, pv.c)) continue;
            dist_t w = grid.edge(pu, pv);
            if (w >= INF) continue;
            N.DU[size_t(u) * nV + v] = std::min(N.DU[size_t(u) * nV + v], w);
            N.DU[size_t(v) * nV + u] = std::min(N.DU[size_t(v) * nV + u], w);
            if (!inK[u]) { inK[u] = 1; K.push_back(u); }
            if (!inK[v]) { inK[v] = 1; K.push_back(v); }
        }
        if (!any_child_edge && K.empty()) {  // internal_node.cpp:176-189
            const int ne = N.nExt();
            std::fill(N.SU.begin(), N.SU.end(), INF);
            for (int i = 0; i < ne; ++i) N.SU[size_t(i) * ne + i] = 0;
            return;
        }
        pivot_closure(N.DU, nV, K);
        project_su(N);
    }
    void project_su(Node& N) {  // internal_node.cpp:233-242
        const int ne = N.nExt(), nV = N.nV();
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j)
                N.SU[size_t(i) * ne + j] = N.DU[size_t(N.extIdx[i]) * nV + N.extIdx[j]];
    }

    // src/minplus_oracle.cpp:48-137
    bool build() {
        if (!(b > 0 && g >= 2 && g * g <= 256)) return false;
        const int md = std::max(grid.rows, grid.cols);
        if (md <= 0) return false;
        S = b;
        while (S < md) S *= g;
        H = 0;
        for (int t = S / b; t > 1; t /= g) ++H;
        nl = S / b;
        leaves.clear();
        for (int ty = 0; ty < nl; ++ty)
            for (int tx = 0; tx < nl; ++tx) {
                Leaf L;
                L.tx = tx;
                L.ty = ty;
                L.R0 = ty * b;
                L.C0 = tx * b;
                tile_ports(L.R0, L.C0, b, L.ports);
                leaf_recompute(L);
                leaves.push_back(std::move(L));
            }
        levels.assign(H, {});
        for (int lvl = 1; lvl <= H; ++lvl) {
            const int nt = tiles_at(lvl);
            for (int ty = 0; ty < nt; ++ty)
                for (int tx = 0; tx < nt; ++tx) {
                    Node N = make_node(lvl, tx, ty);
                    node_recompute(N);
                    levels[lvl - 1].push_back(std::move(N));
                }
        }
        return true;
    }

    // src/leaf_tile.cpp:297-374 (+ refreshBoundarySummary :253-273)
    std::vector<Decrease> leaf_open(Leaf& L, const std::vector<Pt>& cells, int& npiv) {
        const int n = b * b;
        std::vector<uint8_t> inK(n, 0);
        std::vector<int> K;
        const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
        for (Pt p : cells) {
            const int u = (p.r - L.R0) * b + (p.c - L.C0);
            for (int d = 0; d < 4; ++d) {
                int vr = p.r + dr[d], vc = p.c + dc[d];
                if (vr < L.R0 || vr >= L.R0 + b || vc < L.C0 || vc >= L.C0 + b) continue;
                if (!grid.passable(vr, vc)) continue;
                dist_t w = grid.edge(p, {vr, vc});
                if (w >= INF) continue;
                const int v = (vr - L.R0) * b + (vc - L.C0);
                if (w < L.D[size_t(u) * n + v]) {
                    L.D[size_t(u) * n + v] = w;
                    L.D[size_t(v) * n + u] = w;
                    if (!inK[u]) { inK[u] = 1; K.push_back(u); }
                    if (!inK[v]) { inK[v] = 1; K.push_back(v); }
                }
            }
        }
        npiv = int(K.size());
        if (K.empty()) return {};
        pivot_closure(L.D, n, K);
        std::vector<Decrease> dec;
        const int np = int(L.ports.size());
        for (int i = 0; i < np; ++i) {
            int u = (L.ports[i].r - L.R0) * b + (L.ports[i].c - L.C0);
            for (int j = 0; j < np; ++j) {
                int v = (L.ports[j].r - L.R0) * b + (L.ports[j].c - L.C0);
                dist_t nd = L.D[size_t(u) * n + v];
                if (nd < L.S[size_t(i) * np + j]) {
                    L.S[size_t(i) * np + j] = nd;
                    dec.push_back({0, i, j, nd});
                }
            }
        }
        return dec;
    }

    struct Seam {
        Pt p1, p2;
        dist_t w;
    };

    // src/internal_node.cpp:332-413 (+ refreshBoundarySummary :415-433)
    std::vector<Decrease> node_perturb(Node& N, const std::vector<Decrease>& cd,
                                       const std::vector<Seam>& se, bool is_root) {
        const int nV = N.nV();
        std::vector<uint8_t> inK(nV, 0);
        std::vector<int> K;
        auto relax = [&](int u, int v, dist_t w) {
            if (w < N.DU[size_t(u) * nV + v]) {
                N.DU[size_t(u) * nV + v] = w;
                N.DU[size_t(v) * nV + u] = w;
                if (!inK[u]) { inK[u] = 1; K.push_back(u); }
                if (!inK[v]) { inK[v] = 1; K.push_back(v); }
            }
        };
        for (const auto& d : cd) relax(N.off[d.quad] + d.a, N.off[d.quad] + d.b, d.w);
        for (const auto& s : se) {
            int u = find_vertex(N, s.p1), v = find_vertex(N, s.p2);
            if (u != -1 && v != -1) relax(u, v, s.w);
        }
        if (K.empty()) return {};
        pivot_closure(N.DU, nV, K);
        const int ne = N.nExt();
        std::vector<Decrease> dec;
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) {
                dist_t nd = N.DU[size_t(N.extIdx[i]) * nV + N.extIdx[j]];
                if (nd < N.SU[size_t(i) * ne + j]) {
                    N.SU[size_t(i) * ne + j] = nd;
                    dec.push_back({0, i, j, nd});
                }
            }
        if (is_root) return {};
        return dec;
    }

    // src/minplus_oracle.cpp:15-22
    int tile_index_at(int lvl, int tx, int ty) const { return ty * tiles_at(lvl) + tx; }
    // src/minplus_oracle.cpp:24-46
    int lca_level(Pt p1, Pt p2) const {
        int ax = p1.c / b, ay = p1.r / b, bx = p2.c / b, by = p2.r / b;
        if (ax == bx && ay == by) return 0;
        for (int lvl = 1; lvl <= H; ++lvl) {
            ax /= g; ay /= g; bx /= g; by /= g;
            if (ax == bx && ay == by) return lvl;
        }
        return H + 1;
    }

    // src/minplus_oracle.cpp:530-729.  onCellOpened (:139-245) is the same
    // batch with one cell: the rank-one path of the reference computes the
    // same exact closure, so every table ends bit-identical.
    void open_cells(const std::vector<Pt>& cells) {
        last_touched = last_anc = last_pivots = 0;
        if (cells.empty()) return;
        std::vector<std::vector<Pt>> by_leaf(leaves.size());
        std::vector<int> dirty;
        for (Pt p : cells) {
            int tx = p.c / b, ty = p.r / b;
            if (p.c < 0 || p.r < 0 || tx >= nl || ty >= nl) continue;
            int li = ty * nl + tx;
            if (by_leaf[li].empty()) dirty.push_back(li);
            by_leaf[li].push_back(p);
        }
        if (dirty.empty()) return;
        std::vector<std::vector<std::vector<Seam>>> seams(H + 1);
        for (int l = 1; l <= H; ++l) seams[l].resize(levels[l - 1].size());
        const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
        for (Pt p : cells) {
            int tx = p.c / b, ty = p.r / b;
            for (int d = 0; d < 4; ++d) {
                int nr = p.r + dr[d], nc = p.c + dc[d];
                if (!grid.passable(nr, nc)) continue;
                if (nc / b == tx && nr / b == ty) continue;
                dist_t w = grid.edge(p, {nr, nc});
                if (w >= INF) continue;
                int lca = lca_level(p, {nr, nc});
                if (lca < 1 || lca > H) continue;
                int px = tx, py = ty;
                for (int l = 1; l <= lca; ++l) { px /= g; py /= g; }
                int pi = tile_index_at(lca, px, py);
                if (pi >= 0 && pi < int(seams[lca].size())) seams[lca][pi].push_back({p, {nr, nc}, w});
            }
        }
        int piv = 0;
        if (H == 0) {
            for (int li : dirty) {
                int np_ = 0;
                leaf_open(leaves[li], by_leaf[li], np_);
                piv += np_;
            }
            last_touched = int(dirty.size());
            last_pivots = piv;
            return;
        }
        std::vector<std::vector<Decrease>> cur(levels[0].size());
        std::vector<int> active;
        for (int li : dirty) {
            int np_ = 0;
            auto decs = leaf_open(leaves[li], by_leaf[li], np_);
            piv += np_;
            if (decs.empty()) continue;
            int tx = li % nl, ty = li / nl;
            int quad = (ty % g) * g + (tx % g);
            int pi = tile_index_at(1, tx / g, ty / g);
            if (cur[pi].empty()) active.push_back(pi);
            for (auto d : decs) cur[pi].push_back({quad, d.a, d.b, d.w});
        }
        int anc = 0;
        for (int lvl = 1; lvl <= H; ++lvl) {
            for (size_t pi = 0; pi < seams[lvl].size(); ++pi)
                if (!seams[lvl][pi].empty() && cur[pi].empty()) active.push_back(int(pi));
            if (active.empty()) {
                bool higher = false;
                for (int hl = lvl + 1; hl <= H && !higher; ++hl)
                    for (auto& s : seams[hl])
                        if (!s.empty()) { higher = true; break; }
                if (!higher) break;
            }
            anc += int(active.size());
            const int next_n = lvl < H ? int(levels[lvl].size()) : 0;
            std::vector<std::vector<Decrease>> next(next_n);
            std::vector<int> next_active;
            const int ntx = tiles_at(lvl);
            for (int pi : active) {
                auto ext = node_perturb(levels[lvl - 1][pi], cur[pi], seams[lvl][pi], lvl == H);
                if (ext.empty() || lvl >= H) continue;
                int ptx = pi % ntx, pty = pi / ntx;
                int quad = (pty % g) * g + (ptx % g);
                int gpi = tile_index_at(lvl + 1, ptx / g, pty / g);
                if (gpi < 0 || gpi >= next_n) continue;
                if (next[gpi].empty()) next_active.push_back(gpi);
                for (auto d : ext) next[gpi].push_back({quad, d.a, d.b, d.w});
            }
            cur = std::move(next);
            active = std::move(next_active);
        }
        last_touched = int(dirty.size());
        last_anc = anc;
        last_pivots = piv;
    }

    // src/minplus_oracle.cpp:731-823 (liftPoint) restated: level 0 is the
    // cell's row of its leaf D at the leaf ports; level l takes level l-1
    // through the child's rows of the ancestor's D_U at the exterior ports
    // (the child_to_ext block, src/internal_node.cpp:435-455).
    void lift_point(Pt p, std::vector<std::vector<dist_t>>& X, int* par, int* quad, int* off) const {
        X.assign(size_t(H) + 1, {});
        int tx = p.c / b, ty = p.r / b;
        const Leaf& L = leaves[size_t(ty) * nl + tx];
        const int n = b * b, u = (p.r - L.R0) * b + (p.c - L.C0);
        for (const Pt& q : L.ports) X[0].push_back(L.D[size_t(u) * n + (q.r - L.R0) * b + (q.c - L.C0)]);
        par[0] = quad[0] = -1;
        off[0] = 0;
        for (int lvl = 1; lvl <= H; ++lvl) {
            const int ptx = tx / g, pty = ty / g, q = (ty % g) * g + tx % g;
            const int pi = tile_index_at(lvl, ptx, pty);
            const Node& N = levels[size_t(lvl) - 1][size_t(pi)];
            par[lvl] = pi;
            quad[lvl] = q;
            off[lvl] = N.off[size_t(q)];
            const int nV = N.nV(), ne = N.nExt();
            auto& xo = X[size_t(lvl)];
            xo.assign(size_t(ne), INF);
            const auto& xi = X[size_t(lvl) - 1];
            for (size_t cp = 0; cp < xi.size(); ++cp) {
                if (xi[cp] >= INF) continue;
                const dist_t* row = &N.DU[(size_t(N.off[size_t(q)]) + cp) * nV];
                for (int e = 0; e < ne; ++e) {
                    const dist_t c = add_dist(xi[cp], row[N.extIdx[size_t(e)]]);
                    if (c < xo[size_t(e)]) xo[size_t(e)] = c;
                }
            }
            tx = ptx;
            ty = pty;
        }
    }

    // src/minplus_oracle.cpp:934-1419 restated without pruning.
    void query(const std::vector<Pt>& vp, dist_t* out) const {
        const size_t k = vp.size();
        std::fill(out, out + k * k, INF);
        if (k == 0) return;
        for (size_t i = 0; i < k; ++i)
            if (grid.passable(vp[i].r, vp[i].c)) out[i * k + i] = 0;
        // X[i][lvl] : distance from viewpoint i to the ports of its level-lvl tile
        std::vector<std::vector<std::vector<dist_t>>> X(k, std::vector<std::vector<dist_t>>(H + 1));
        std::vector<int> valid;
        const int n = b * b;
        for (size_t i = 0; i < k; ++i) {
            Pt p = vp[i];
            if (!grid.passable(p.r, p.c)) continue;
            int tx = p.c / b, ty = p.r / b;
            if (tx >= nl || ty >= nl) continue;
            valid.push_back(int(i));
            const Leaf& L = leaves[size_t(ty) * nl + tx];
            int u = (p.r - L.R0) * b + (p.c - L.C0);
            for (const Pt& q : L.ports) X[i][0].push_back(L.D[size_t(u) * n + (q.r - L.R0) * b + (q.c - L.C0)]);
        }
        // same-leaf pairs (:1027-1048)
        for (int i : valid)
            for (int j : valid) {
                if (i >= j) continue;
                Pt a = vp[i], c = vp[j];
                if (a.c / b != c.c / b || a.r / b != c.r / b) continue;
                const Leaf& L = leaves[size_t(a.r / b) * nl + a.c / b];
                dist_t d = L.D[size_t((a.r - L.R0) * b + (a.c - L.C0)) * n + (c.r - L.R0) * b + (c.c - L.C0)];
                out[size_t(i) * k + j] = d;
                out[size_t(j) * k + i] = d;
            }
        for (int lvl = 1; lvl <= H; ++lvl) {
            const int cside = side(lvl - 1), pside = side(lvl), ntx = tiles_at(lvl);
            // group viewpoints by node, then by child quad (quad order like :1082-1094)
            std::vector<std::vector<std::vector<int>>> groups(levels[lvl - 1].size(),
                                                              std::vector<std::vector<int>>(g * g));
            for (int i : valid) {
                Pt p = vp[i];
                int pi = (p.r / pside) * ntx + p.c / pside;
                int quad = ((p.r / cside) % g) * g + (p.c / cside) % g;
                groups[pi][quad].push_back(i);
            }
            for (size_t pi = 0; pi < groups.size(); ++pi) {
                const Node& N = levels[lvl - 1][pi];
                const int nV = N.nV(), ne = N.nExt();
                // lift (:1113-1154): X_lvl[e] = min_p X_{lvl-1}[p] (+) D_U[off+p][extIdx[e]]
                if (lvl < H)
                    for (int q = 0; q < g * g; ++q)
                        for (int i : groups[pi][q]) {
                            auto& xo = X[i][lvl];
                            xo.assign(ne, INF);
                            const auto& xi = X[i][lvl - 1];
                            for (size_t p = 0; p < xi.size(); ++p) {
                                if (xi[p] >= INF) continue;
                                const dist_t* row = &N.DU[size_t(N.off[q] + p) * nV];
                                for (int e = 0; e < ne; ++e) {
                                    dist_t c = add_dist(xi[p], row[N.extIdx[e]]);
                                    if (c < xo[e]) xo[e] = c;
                                }
                            }
                        }
                // inter-child meets (:1161-1326) and same-child shortcut (:1334-1411)
                for (int q1 = 0; q1 < g * g; ++q1) {
                    if (groups[pi][q1].empty()) continue;
                    const int o1 = N.off[q1], n1 = N.off[q1 + 1] - o1;
                    for (int q2 = q1; q2 < g * g; ++q2) {
                        if (groups[pi][q2].empty()) continue;
                        const int o2 = N.off[q2], n2 = N.off[q2 + 1] - o2;
                        std::vector<dist_t> proj(n2);
                        for (int i : groups[pi][q1]) {
                            const auto& xi = X[i][lvl - 1];
                            std::fill(proj.begin(), proj.end(), INF);
                            for (int p1 = 0; p1 < n1; ++p1) {
                                if (xi[p1] >= INF) continue;
                                const dist_t* row = &N.DU[size_t(o1 + p1) * nV + o2];
                                for (int p2 = 0; p2 < n2; ++p2) {
                                    dist_t c = add_dist(xi[p1], row[p2]);
                                    if (c < proj[p2]) proj[p2] = c;
                                }
                            }
                            for (int j : groups[pi][q2]) {
                                if (q1 == q2 && j <= i) continue;
                                const auto& yj = X[j][lvl - 1];
                                dist_t best = out[size_t(i) * k + j];
                                for (int p2 = 0; p2 < n2; ++p2) {
                                    dist_t c = add_dist(proj[p2], yj[p2]);
                                    if (c < best) best = c;
                                }
                                out[size_t(i) * k + j] = best;
                                out[size_t(j) * k + i] = best;
                            }
                        }
                    }
                }
            }
        }
    }
};

// src/bfs_oracle.cpp:79-202, with a visited bitmap so that a distance that
// reaches INF cannot be mistaken for "unvisited"; reports min(d, INF).
// Row i of out (k entries) from source i; the buffers belong to the caller.
static void bfs_row(const Grid& G, const std::vector<Pt>& vp, size_t i, dist_t* out, std::vector<uint32_t>& dist,
                    std::vector<uint8_t>& seen, std::vector<int>& q) {
    const size_t k = vp.size();
    std::fill(out + i * k, out + (i + 1) * k, INF);
    if (!G.passable(vp[i].r, vp[i].c)) return;
    std::fill(seen.begin(), seen.end(), 0);
    q.clear();
    int s = vp[i].r * G.cols + vp[i].c;
    seen[s] = 1;
    dist[s] = 0;
    q.push_back(s);
    for (size_t h = 0; h < q.size(); ++h) {
        int u = q[h], ur = u / G.cols, uc = u % G.cols;
        const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
        for (int d = 0; d < 4; ++d) {
            int vr = ur + dr[d], vc = uc + dc[d];
            if (!G.passable(vr, vc)) continue;
            int v = vr * G.cols + vc;
            if (seen[v]) continue;
            seen[v] = 1;
            dist[v] = dist[u] + 1;
            q.push_back(v);
        }
    }
    for (size_t j = 0; j < k; ++j) {
        if (!G.passable(vp[j].r, vp[j].c)) continue;
        int t = vp[j].r * G.cols + vp[j].c;
        if (seen[t]) out[i * k + j] = dist[t] >= INF ? INF : dist_t(dist[t]);
    }
}

// All pairs, sources spread over `threads` threads (rows are independent).
static void bfs_all_pairs(const Grid& G, const std::vector<Pt>& vp, dist_t* out, int threads = 1) {
    const size_t k = vp.size();
    const size_t cells = size_t(G.rows) * G.cols;
    const int nt = int(std::max<size_t>(1, std::min<size_t>(size_t(std::max(threads, 1)), k)));
    auto work = [&](int t) {
        std::vector<uint32_t> dist(cells);
        std::vector<uint8_t> seen(cells);
        std::vector<int> q;
        q.reserve(cells);
        for (size_t i = size_t(t); i < k; i += size_t(nt)) bfs_row(G, vp, i, out, dist, seen, q);
    };
    if (nt == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(work, t);
    for (auto& th : pool) th.join();
}

}  // namespace orc

using namespace orc;

static std::vector<Pt> to_pts(const int32_t* rc, int n) {
    std::vector<Pt> v(size_t(n > 0 ? n : 0));
    for (int i = 0; i < n; ++i) v[i] = {rc[2 * i], rc[2 * i + 1]};
    return v;
}

extern "C" {

// cells: rows*cols bytes, 0 = Unknown, 1 = Free, 2 = Obstacle (types.hpp:101-105)
void* orc_create(int b, int g, int rows, int cols, const uint8_t* cells) {
    auto* o = new Oracle();
    o->b = b;
    o->g = g;
    o->grid.rows = rows;
    o->grid.cols = cols;
    o->grid.obs.assign(cells, cells + size_t(rows) * cols);
    return o;
}
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
int orc_build(void* h) { return static_cast<Oracle*>(h)->build() ? 1 : 0; }
// Callers set the cells Free before the update (README.md:71-91); the shim does it here.
void orc_open_cells(void* h, const int32_t* rc, int n) {
    auto* o = static_cast<Oracle*>(h);
    auto pts = to_pts(rc, n);
    for (Pt p : pts)
        if (o->grid.in(p.r, p.c)) o->grid.obs[size_t(p.r) * o->grid.cols + p.c] = 1;
    o->open_cells(pts);
}
// onCellOpened leaves last_batch_stats untouched (src/minplus_oracle.cpp:139-177).
void orc_open_cell(void* h, int r, int c) {
    auto* o = static_cast<Oracle*>(h);
    const int t = o->last_touched, a = o->last_anc, p = o->last_pivots;
    int32_t rc[2] = {r, c};
    orc_open_cells(h, rc, 1);
    o->last_touched = t;
    o->last_anc = a;
    o->last_pivots = p;
}
void orc_query(void* h, const int32_t* rc, int k, uint16_t* out) {
    static_cast<Oracle*>(h)->query(to_pts(rc, k), out);
}
void orc_bfs(void* h, const int32_t* rc, int k, uint16_t* out) {
    bfs_all_pairs(static_cast<Oracle*>(h)->grid, to_pts(rc, k), out);
}
// Stand-alone BFS on a cell array (no hierarchy needed).
void orc_grid_bfs(int rows, int cols, const uint8_t* cells, const int32_t* rc, int k, uint16_t* out) {
    Grid G;
    G.rows = rows;
    G.cols = cols;
    G.obs.assign(cells, cells + size_t(rows) * cols);
    bfs_all_pairs(G, to_pts(rc, k), out);
}
// Same, with the sources spread over `threads` threads.
void orc_grid_bfs_mt(int rows, int cols, const uint8_t* cells, const int32_t* rc, int k, uint16_t* out, int threads) {
    Grid G;
    G.rows = rows;
    G.cols = cols;
    G.obs.assign(cells, cells + size_t(rows) * cols);
    bfs_all_pairs(G, to_pts(rc, k), out, threads);
}
// BFS restricted to the rectangle [R0,R0+h) x [C0,C0+w) from (sr,sc); out is h*w.
// Source distance 0 even when the source is impassable (leaf_tile.cpp:63-65).
void orc_region_bfs(void* h, int R0, int C0, int hh, int ww, int sr, int sc, uint16_t* out) {
    const Grid& G = static_cast<Oracle*>(h)->grid;
    std::fill(out, out + size_t(hh) * ww, INF);
    const int s = (sr - R0) * ww + (sc - C0);
    out[s] = 0;
    if (!G.passable(sr, sc)) return;
    std::vector<int> q{s};
    std::vector<uint8_t> seen(size_t(hh) * ww, 0);
    seen[s] = 1;
    std::vector<uint32_t> d(size_t(hh) * ww, 0);
    for (size_t i = 0; i < q.size(); ++i) {
        int u = q[i], ur = u / ww, uc = u % ww;
        const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
        for (int t = 0; t < 4; ++t) {
            int vr = ur + dr[t], vc = uc + dc[t];
            if (vr < 0 || vr >= hh || vc < 0 || vc >= ww) continue;
            if (!G.passable(R0 + vr, C0 + vc)) continue;
            int v = vr * ww + vc;
            if (seen[v]) continue;
            seen[v] = 1;
            d[v] = d[u] + 1;
            out[v] = d[v] >= INF ? INF : uint16_t(d[v]);
            q.push_back(v);
        }
    }
}
// liftPoint: X levels back to back (lens[l] entries each), par/quad/off H+1
// entries, active[l] = ports with X < INF by increasing X (ties by index),
// padded with -1 to lens[l].
void orc_lift_point(void* h, int r, int c, uint16_t* X, int32_t* lens, int32_t* par, int32_t* quad, int32_t* off,
                    int32_t* active) {
    const Oracle* o = static_cast<Oracle*>(h);
    std::vector<std::vector<dist_t>> x;
    std::vector<int> pv(size_t(o->H) + 1), qv(size_t(o->H) + 1), ov(size_t(o->H) + 1);
    o->lift_point(Pt{r, c}, x, pv.data(), qv.data(), ov.data());
    size_t w = 0;
    for (size_t l = 0; l < x.size(); ++l) {
        lens[l] = int32_t(x[l].size());
        par[l] = pv[l];
        quad[l] = qv[l];
        off[l] = ov[l];
        std::vector<int> a;
        for (int i = 0; i < int(x[l].size()); ++i)
            if (x[l][size_t(i)] < INF) a.push_back(i);
        std::stable_sort(a.begin(), a.end(), [&](int u, int v) { return x[l][size_t(u)] < x[l][size_t(v)]; });
        for (size_t i = 0; i < x[l].size(); ++i) {
            X[w + i] = x[l][i];
            active[w + i] = i < a.size() ? a[i] : -1;
        }
        w += x[l].size();
    }
}
int orc_S(void* h) { return static_cast<Oracle*>(h)->S; }
int orc_H(void* h) { return static_cast<Oracle*>(h)->H; }
int orc_num_leaves(void* h) { return int(static_cast<Oracle*>(h)->leaves.size()); }
int orc_num_nodes(void* h, int lvl) { return int(static_cast<Oracle*>(h)->levels[lvl - 1].size()); }
int orc_node_nV(void* h, int lvl, int idx) { return static_cast<Oracle*>(h)->levels[lvl - 1][idx].nV(); }
int orc_node_nExt(void* h, int lvl, int idx) { return static_cast<Oracle*>(h)->levels[lvl - 1][idx].nExt(); }
void orc_leaf_D(void* h, int idx, uint16_t* out) {
    const auto& D = static_cast<Oracle*>(h)->leaves[idx].D;
    std::memcpy(out, D.data(), D.size() * 2);
}
void orc_leaf_S(void* h, int idx, uint16_t* out) {
    const auto& S = static_cast<Oracle*>(h)->leaves[idx].S;
    std::memcpy(out, S.data(), S.size() * 2);
}
void orc_node_DU(void* h, int lvl, int idx, uint16_t* out) {
    const auto& D = static_cast<Oracle*>(h)->levels[lvl - 1][idx].DU;
    std::memcpy(out, D.data(), D.size() * 2);
}
void orc_node_SU(void* h, int lvl, int idx, uint16_t* out) {
    const auto& S = static_cast<Oracle*>(h)->levels[lvl - 1][idx].SU;
    std::memcpy(out, S.data(), S.size() * 2);
}
void orc_last_batch_stats(void* h, int32_t* out3) {
    auto* o = static_cast<Oracle*>(h);
    out3[0] = o->last_touched;
    out3[1] = o->last_anc;
    out3[2] = o->last_pivots;
}
// src/minplus_oracle.cpp:2116-2134: leaf D+S, node D_U+S_U and child_to_ext tables.
uint64_t orc_total_memory_bytes(void* h) {
    auto* o = static_cast<Oracle*>(h);
    uint64_t t = 0;
    for (auto& L : o->leaves) t += (L.D.size() + L.S.size()) * 2;
    for (auto& lv : o->levels)
        for (auto& N : lv) t += (N.DU.size() + N.SU.size() + size_t(N.nV()) * N.nExt()) * 2;
    return t;
}

}  // extern "C"
