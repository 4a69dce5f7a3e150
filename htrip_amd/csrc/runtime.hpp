// Internal types of the host runtime (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/htrip_gpu.h"
#include "geometry.hpp"
#include "kernels.hpp"

namespace htg {

extern thread_local std::string g_err;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    ~DevBuf() { release(); }
    int ensure(size_t n);
    void release();
};

// Packed update lists reserved per map (grown, doubling, past it).
constexpr size_t kUpdateListReserve = size_t(1) << 20;
// Cover-product work units reserved per map (16 bytes each; a level past it grows the buffers).
constexpr size_t kCoverUnitReserve = size_t(1) << 17;

// Grow-only pinned host buffer.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    HostBuf(HostBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    ~HostBuf() { release(); }
    int ensure(size_t n);
    void release();
};

// Packs several host vectors into one upload, each 16-byte aligned.
struct Packer {
    struct Part {
        const void* src;
        size_t off, bytes;
    };
    std::vector<Part> parts;
    size_t size = 0;
    template <typename T>
    size_t add(const std::vector<T>& v) {
        const size_t off = size;
        parts.push_back({v.data(), off, v.size() * sizeof(T)});
        size = (off + v.size() * sizeof(T) + 15) / 16 * 16 + (v.empty() ? 16 : 0);
        return off;
    }
    void write(char* dst) const {
        for (const Part& p : parts)
            if (p.bytes) std::memcpy(dst + p.off, p.src, p.bytes);
    }
    void clear() {
        parts.clear();
        size = 0;
    }
};

struct LevelGeo {
    LevelDims d;
    int nn = 0;              // nodes at this level
    size_t node_stride = 0;  // elements per node D_U (nV * ld)
    std::vector<short2> vu_rc;
    std::vector<int> ext_idx, ext_child, ext_cport;
    std::vector<int> grp_of4;        // projection group of each aligned quad of V_U ports, -1 if none
    std::vector<uint8_t> seam_flag;  // per V_U port: a seam port of its child
    std::vector<int2> seams;
    std::vector<int> seam_list, seam_off;  // per-child seam ports
    // Query projection: natural columns (V_U indices) the projection computes,
    // and the Z layout: per-child seam blocks (even-padded).  zsrc[pos] =
    // projection column of Z position pos or -1.
    std::vector<int> zcols, zsrc, zoff, zlen, ext_col;  // ext_col[e]: projection column of port e
    // per projection column: Z position (or -1) and exterior port index below
    // the root (or -1), for projections that fold in registers
    std::vector<int> colpos, colext;
    std::vector<int> vu_ext;  // [v]: exterior port index of V_U port v, or -1
    int ncols = 0, ldp = 0, nz = 0, ldz = 0, nsmax = 0;
    bool gather = false;
    // Live column groups: group gi = projection columns 4gi..4gi+3.  A group
    // is live in node W when one of its ports is passable; the others give
    // INF and the projection skips them.  live_list[W*ngrp ...] holds the
    // first live_cnt[W] live groups of W, live_map[W*ngrp + gi] the flags.
    // Live octets for the closure: octet o = V_U ports 8o..8o+7, live when
    // one of them is passable (only live rows and columns of D_U can change).
    // These arrays live in the map's pinned live block (h_*) and its device
    // copy (d_*), uploaded with one copy after each change.
    int ngrp = 0, ng8 = 0;
    int *h_live_cnt = nullptr, *h_live_list = nullptr, *h_live8_cnt = nullptr, *h_live8_list = nullptr;
    uint8_t* h_live_map = nullptr;
    const int *d_live_cnt = nullptr, *d_live_list = nullptr, *d_live8_cnt = nullptr, *d_live8_list = nullptr;
    const uint8_t* d_live_map = nullptr;
    // At the root the queries read only the seam x seam block of D_U, and a
    // closure whose pivots are seam ports keeps that block exact by itself:
    // the root's live octets (h_live8_*, used by every build and update
    // kernel) count passable seam ports only, and its full octet lists
    // (h_all8_*, every passable port) serve the on-demand closure of the
    // whole table.  Below the root both name the same arrays.
    bool seam_only = false;
    int *h_all8_cnt = nullptr, *h_all8_list = nullptr;
    const int *d_all8_cnt = nullptr, *d_all8_list = nullptr;
    std::vector<int> seam_live;  // [W * g^2 + q]: passable seam ports of child q (rows a projection may use)
    DevBuf d_vu_rc, d_ext_idx, d_ext_child, d_ext_cport, d_seams, d_seam_list, d_seam_off, d_zcols, d_zsrc, d_zoff,
        d_zlen, d_ext_col, d_colpos, d_colext, d_vu_ext, d_all;  // d_all: 0..nn-1
    DevBuf DU, SU;
};

struct LevelPlan {
    int M = 1, ksplit = 1, col_tiles = 0;
    std::vector<int4> items;  // (W, q | first-of-run << 16, first slot, count)
};

struct SlotKey {
    unsigned long long key;  // tile path, base-g^2 digits from the root down
    int q, i, vp;            // query, index in the query, viewpoint index
};

struct QueryPlan {
    std::vector<SlotKey> sk;          // slot order
    std::vector<SlotKey> sk_tmp;      // counting-sort output
    std::vector<int> key_count;       // counting-sort buckets
    std::vector<unsigned long long> kc, kw;  // per-level keys of the slots: (W, q) and W
    std::vector<int> by_q, qstart;    // meet order: slots grouped by query
    std::vector<int> q_k, q_vp;
    std::vector<long long> q_out;
    std::vector<int2> vp_rc;
    std::vector<uint8_t> vp_valid;
    std::vector<int2> slot_rc;
    std::vector<int> slot_q, slot_i;
    std::vector<LevelPlan> levels;
    // meet, every level in one launch
    std::vector<int> mtiles;   // 8 ints each: group size, q2, i0, j0, j_end, level | 16x16 << 16, glist offset, query
    std::vector<int> glist;    // slot indices
    std::vector<int> zgeo;     // per level: Z offset lo, hi, ldz, 0, zoff[g2], zlen[g2]
    std::vector<int4> fills;   // INF fills of Q: (start lo, start hi, n8, first block)
    // the batch's kernel arguments (device addresses), from bind_batch
    QFrontArgs front{};
    std::vector<QZArgs> zargs;  // per level
    QMeetArgs meet{};
    double m_bytes = 0, m_updates = 0;
    void clear() {
        q_k.clear();
        q_vp.clear();
        q_out.clear();
        vp_rc.clear();
        vp_valid.clear();
        slot_rc.clear();
        slot_q.clear();
        slot_i.clear();
        for (auto& l : levels) {
            l.col_tiles = 0;
            l.items.clear();
        }
        mtiles.clear();
        glist.clear();
        zgeo.clear();
        fills.clear();
        m_bytes = m_updates = 0;
    }
};

// Where the parts of one planned query batch go: offsets of each list in the
// packed plan, the X / Z regions of Q, and the sizes of the device regions
// (Q, row-chunk partials Zp, result).
struct BatchLayout {
    long long total_vp = 0, total_out = 0;
    int kmax = 0, ns = 0, fill_blocks = 0;
    size_t nsl = 1, qsz = 0, maxpart = 0;
    std::vector<size_t> xoff, zoffq, o_items;  // X of levels 0..H-1, Z of levels 1..H; per-level items
    size_t o_qk = 0, o_qout = 0, o_qvp = 0, o_vprc = 0, o_vpvalid = 0, o_src = 0, o_si = 0, o_tiles = 0,
           o_glist = 0, o_zgeo = 0, o_fills = 0;
    Packer pk;
    size_t q_bytes() const { return (qsz > 8 ? qsz : 8) * 2; }
    size_t zp_bytes() const { return nsl * (maxpart > 8 ? maxpart : 8) * 2; }
    // the meet's 16-bit atomics address whole 32-bit words
    size_t out_bytes() const { return (size_t(total_out) * 2 + 4 + 15) / 16 * 16; }
};

// Per-stream query workspace (grow-only, like HTripQueryScratch,
// include/htrip/minplus_oracle.hpp:189-238).
struct QueryWS {
    hipStream_t stream = nullptr;
    bool own_stream = true;
    hipEvent_t done = nullptr;  // recorded after the batch's last operation
    long long total_out = 0;
    QueryPlan plan;
    BatchLayout lay;
    HostBuf h_plan, h_out, h_wstat;
    // Q: X of levels 0..H-1 then Z of levels 1..H, each region 8-aligned
    DevBuf d_plan, Q, Zp, d_out, d_wstat;  // d_wstat: k_q_z work counters (bytes, updates)
    bool counted = false;                        // this batch's counters go to the profiler
    double host_s[3] = {0, 0, 0};                // host time: planning, packing + upload, launches
    double host_max[3] = {0, 0, 0};              // longest of each phase
    long long host_slow[3] = {0, 0, 0};          // phases over 0.5 ms
    double host_sub[4] = {0, 0, 0, 0};           // planning: keys, sort, meet tiles, layout + items
    long long host_batches = 0;
    // Runs on `ext` when given (not owned), else on a stream of its own.
    int init(hipStream_t ext = nullptr);
    ~QueryWS();
};

// Kernel timing (see htg_profile_enable): start/stop events around each
// tracked dispatch.  Thread-safe; a session's two snapshot maps share one
// instance.
struct Profiler {
    std::mutex mu;
    bool on = false;
    int every = 1;                // time one operation (query batch, update, build) in `every`
    std::atomic<long long> tick{0};
    // Whether the next operation carries timing events (counters and
    // durations then come from the same sampled operations).
    bool take() { return on && (every <= 1 || tick.fetch_add(1) % every == 0); }
    struct Pending {
        int cls;
        hipEvent_t a, b;
        double bytes, updates;
        hipStream_t st;
    };
    // Events recorded on the stream right before and after the dispatch
    // (default), or attached to the dispatch itself (HTG_PROFILE_ATTACH=1):
    // an attached pair slows the kernel it times about twofold on gfx950 and
    // holds other queues back.
    bool attach = false;
    std::vector<Pending> pending;          // ended, not yet harvested
    std::unordered_map<int, Pending> open_;  // reserved, not yet closed
    int next_token = 0;
    std::vector<hipEvent_t> pool;
    htg_kernel_stats acc[HTG_K_COUNT]{};
    hipEvent_t get();
    // Reserves the start/stop events of one kernel dispatch (kt, for the
    // launcher); returns a token for close(), or -1 with null events when
    // timing is off.
    int open(int cls, double bytes, double updates, KTimer& kt, hipStream_t st);
    // launched: the dispatch was made (its stop event is recorded now, unless
    // attached); otherwise the events are recycled.
    void close(int token, bool launched);
    // Accumulate every completed pair; pairs still running stay pending.
    void harvest();
    // Work counted on the device (bytes, updates), added without a launch.
    void add_work(int cls, double bytes, double updates);
    void reset();
    ~Profiler();
};

// Tables an update rewrote: recomputed leaves and, per level, touched nodes.
struct DirtySet {
    std::vector<int> leaves;
    std::vector<std::vector<int>> nodes;  // nodes[l-1]
    std::vector<int2> cells;              // cells that became passable
    void clear() {
        leaves.clear();
        nodes.clear();
        cells.clear();
    }
};

struct Map {
    htg_ctx* ctx = nullptr;
    htg_config cfg{};
    int b = 16, g = 2, rows = 0, cols = 0, S = 0, H = 0, nl = 0;
    int min_S = 0;  // side forced by a lifted root (0: the smallest square over the grid)
    std::vector<uint8_t> cells_code;  // rows*cols cell codes as last known
    std::vector<uint8_t> pass;        // S*S passable flags (host mirror of d_pass)
    DevBuf d_pass, leafD, leafS;
    std::vector<LevelGeo> lv;  // lv[l-1] is level l
    // Build/update stream; also the plain query stream.  A server's maps
    // share the server's update stream (own_stream false), so that a process
    // needs few hardware queues.
    hipStream_t stream = nullptr;
    bool own_stream = true;
    bool built = false;
    uint64_t epoch = 0;
    htg_batch_stats stats{0, 0, 0};
    uint64_t budget = 0, device_bytes = 0;
    htg_tuning tune{32, 4, 0, 0, 0};
    QueryWS qws;
    // Held by the calls that use the map's own workspaces (qws, lift_*), so
    // that concurrent queries on one map serialise instead of sharing them.
    std::mutex qmu;
    // An update failed after it had changed the cell state: tables may not
    // match it, so every call but build answers HTG_E_STATE.
    bool broken = false;
    // Fault injection (htg_inject_fault): the fault_count-th arrival at
    // fault_site fails as a HIP error would.
    std::atomic<int> fault_site{0}, fault_count{0};
    bool fault_hit(int site) {
        if (fault_site.load() != site) return false;
        if (fault_count.fetch_sub(1) != 1) return false;
        fault_site = 0;
        return true;
    }
    std::shared_ptr<Profiler> profp = std::make_shared<Profiler>();
    Profiler& prof_ref() { return *profp; }
    DirtySet dirty;  // filled by open_cells
    // Build captured as one hipGraph (tune.use_graph); replayed while the
    // launch signature (per-level pivot blocks and grid sizes) is unchanged.
    hipGraphExec_t build_graph = nullptr;
    std::vector<int> build_sig;
    bool capturing = false;   // no timing events inside a capture
    bool timing_now = false;  // the current build / update was sampled for timing
    int prof_open(int cls, double bytes, double updates, KTimer& kt) {
        kt = KTimer{};
        return capturing || !timing_now ? -1 : prof_ref().open(cls, bytes, updates, kt, stream);
    }
    // Build pivots per node of a level (host): the endpoints of seam pairs
    // whose cells are both passable.
    // (every node of the level, or the listed ones in that order)
    std::vector<int> build_pivot_counts(int lvl, const std::vector<int>* nodes = nullptr) const;
    int enqueue_build(const std::vector<std::vector<int>>& hc);
    HostBuf live_h;  // live arrays of every level (pinned)
    DevBuf live_d;
    hipEvent_t live_evt = nullptr;  // recorded after the last upload of live_h
    bool live_pending = false;
    ~Map();
    DevBuf ws_nodes, ws_flags, ws_piv, ws_bpiv, ws_counts, ws_R, ws_changed, ws_mask, ws_edges, ws_cells, ws_leaves,
        ws_leaf_changed;
    DevBuf ws_upd, ws_chg;  // update: packed per-level lists; changed flags (leaves, then every level)
    DevBuf ws_leafP;        // update: BFS rows of the newly opened cells of incrementally revealed leaves
    DevBuf ws_leafl;        // update: packed leaf-phase lists (new cells, leaf lists, pivot items)
    HostBuf ws_leafl_h;
    HostBuf ws_upd_h, ws_chg_h, ws_counts_h;  // pinned staging of the lists, changed flags, pivot counts
    DevBuf ws_cov;  // update: work units of a level's cover products
    HostBuf ws_cov_h;
    std::vector<int4> cov_units;
    // snapshot catch-up lists: staging of their own, since the update that
    // follows refills ws_upd_h while the catch-up's upload may still read it
    DevBuf ws_cu;
    HostBuf ws_cu_h;

    int build();
    int open_cells(const int32_t* rc, int n);
    int block_cells(const int32_t* rc, int n);
    // One level of a rebuild: assemble, seam edges, closure over the seam
    // endpoints, S_U (changed[i] set when listed node i's S_U decreased).
    int rebuild_level(int lvl, const int* dnodes, const std::vector<int>& nodes, int* changed);
    // Update planning: estimated closure time of the listed nodes, and whether
    // rebuilding a level's candidates beats updating them (hc: cover counts,
    // then uncovered-endpoint counts; pb: build_pivot_counts of the same nodes).
    double closure_us(int lvl, const std::vector<int>& nodes, const int* cnt) const;
    bool rebuild_is_cheaper(int lvl, const std::vector<int>& nodes, const std::vector<int>& hc, int na,
                            const std::vector<int>& pb) const;
    long long rebuilt_levels = 0;  // update levels done by a rebuild (printed at HTG_DEBUG >= 2)
    int rebuild(const std::vector<int2>& closed, const std::vector<int>& leaves,
                const std::vector<std::vector<int>>& anc);
    void reset_versions();
    int query(const int32_t* rc, const int32_t* ks, int nq, uint16_t* out_host, QueryWS& ws);
    // liftPoint of cell (r, c): x receives levels 0..H back to back (lens[l]
    // entries each), meta the ancestor node, child quad and child offset.
    int lift_point(int r, int c, uint16_t* x, long long x_cap, int32_t* lens, int32_t* parent, int32_t* quad,
                   int32_t* offset);
    DevBuf lift_x, lift_part;  // liftPoint: the lifts of every level; row-chunk partial minima
    HostBuf lift_h;
    // Table versions for incremental queries.  uid names this map for the
    // process's lifetime, gen counts full builds, and every update that
    // rewrites tables advances clock and stamps those leaves and nodes with
    // it.  A query result depends only on the tables of its viewpoints'
    // leaves and ancestor chains, so it is unchanged while none of them is
    // stamped after the result was taken.
    uint64_t uid = 0, gen = 0, clock = 0;
    std::vector<uint64_t> leaf_ver;
    std::vector<std::vector<uint64_t>> node_ver;  // node_ver[l-1][node]
    void stamp(const std::vector<int>& leaves, const std::vector<std::vector<int>>& nodes, int first_level);
    // Latest stamp on the leaf and ancestor chain of cell (r, c).
    uint64_t chain_version(int r, int c) const;
    int query_incremental(htg_inc& st, const int32_t* rc, int k, uint16_t* out, int32_t* n_recomputed);
    // Asynchronous halves of query(): plan + launch on ws.stream (results
    // copied to ws.h_out when to_host), then wait for ws.done.
    int query_launch(const int32_t* rc, const int32_t* ks, int nq, QueryWS& ws, bool to_host);
    // The two halves of a launch: host planning (P and B are reused and keep
    // their capacity), and the kernels of a planned batch on given buffers.
    int plan_batch(const int32_t* rc, const int32_t* ks, int nq, QueryPlan& P, BatchLayout& B, bool counting,
                   double* host_sub) const;
    // Device addresses of a planned batch (plan at d_plan, X / Z regions Q,
    // row-chunk partials Zp, result d_out, counters wstat) into the kernels'
    // argument blocks in P.
    void bind_batch(QueryPlan& P, const BatchLayout& B, const char* d_plan, uint16_t* Q, uint16_t* Zp,
                    uint16_t* d_out, unsigned long long* wstat) const;
    // The kernels of a bound batch on stream st (plan already on the device).
    int enqueue_batch(const QueryPlan& P, const BatchLayout& B, bool counted, hipStream_t st);
    int query_finish(QueryWS& ws, uint16_t* out_host);
    // Grow ws ahead of time for batches of up to max_vps viewpoints and
    // max_out result entries, so that serving does not allocate (a device
    // free synchronises the whole device).
    int reserve_query_ws(QueryWS& ws, int max_vps, long long max_out) const;
    int ensure_build_ws(int lvl);
    int reserve_update_ws();
    std::vector<uint8_t> cell_mark;            // per grid cell, all zero between updates (deduplication)
    std::vector<int> leaf_idx;                 // per leaf, all zero between updates (dirty-leaf positions)
    double upd_time[12] = {};  // update phases (printed at HTG_DEBUG >= 2)
    long long upd_count = 0, catchup_count = 0;
    // Lay out the live block of every level (once, at creation).
    int init_live();
    // Recompute the live column groups and octets of the nodes that contain
    // one of `cells` (every node when cells is null) and upload the changed
    // byte range with one asynchronous copy on the map stream.
    int refresh_live(const std::vector<int2>* cells);
    // nodes: the listed nodes (host copy of ws_nodes)
    // full: close over every live octet (the root's full lists) instead of
    // the maintained ones.
    int run_closure(int lvl, const int* dnodes, const std::vector<int>& nodes, const std::vector<int>& hcounts,
                    bool full = false);
    // The root's entries outside its seam x seam block (outer-boundary
    // ports) are only kept exact on demand: root_full says they are, and
    // materialize_root() rebuilds the root over its full octet lists when
    // they are not (liftPoint, table reads, liftRoot).
    bool root_full = false;
    int materialize_root();
    const uint16_t* child_summaries(int lvl, size_t& stride) const;
    // Make this map equal to src after src applied one update with dirty set
    // ds (this map equal to src before it): copies only the rewritten tables.
    int catch_up_from(const Map& src, const DirtySet& ds);
    // Make this map equal to src (same geometry) by copying every table and
    // the cell state; clears `broken`.
    int copy_all_from(const Map& src);
};

}  // namespace htg

struct htg_map : htg::Map {};

// Caller-owned state of MinPlusOracle::queryAllPairsIncremental
// (include/htrip/minplus_oracle.hpp:262-272): the map and build it was taken
// on, the clock at that time, the viewpoints and the matrix.
struct htg_inc {
    bool valid = false;
    uint64_t map_uid = 0, gen = 0, clock = 0;
    std::vector<int32_t> rc;
    std::vector<uint16_t> m;
};

struct htg_ctx {
    int device = 0;
};

// Caller-owned query workspace (HTripQueryScratch,
// include/htrip/minplus_oracle.hpp:189-242): its own stream and buffers;
// `after` orders its queries behind the map's update stream.
struct htg_query_ws {
    htg_ctx* ctx = nullptr;
    htg::QueryWS ws;
    hipEvent_t after = nullptr;
    ~htg_query_ws() {
        if (after) (void)hipEventDestroy(after);
    }
};

namespace htg {
// htg_map_create / htg_map_clone on a given stream (nullptr: a new one).
// min_S: the hierarchy square is at least this side (a lifted root).
int map_create_on(htg_ctx* ctx, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                  uint64_t hbm_budget, hipStream_t shared, htg_map** out, int min_S = 0);
// liftRoot: a new map of side g S over the grown grid, holding src's tables
// in its old square and building the rest; src is left unchanged.
int lift_root(htg_map* src, int32_t rows, int32_t cols, const uint8_t* cells, htg_map** out);
int map_clone_on(const htg_map* src, uint64_t hbm_budget, hipStream_t shared, htg_map** out);
}  // namespace htg
