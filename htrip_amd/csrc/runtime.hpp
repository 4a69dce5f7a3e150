// Internal types of the host runtime (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/htrip_gpu.h"
#include "geometry.hpp"

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
};

struct LevelGeo {
    LevelDims d;
    int nn = 0;              // nodes at this level
    size_t node_stride = 0;  // elements per node D_U (nV * ld)
    std::vector<short2> vu_rc;
    std::vector<int> ext_idx, ext_child, ext_cport;
    std::vector<int2> seams;
    std::vector<int> seam_list, seam_off;  // per-child seam ports
    // Query projection: natural columns (V_U indices) the projection computes,
    // and the Z layout: per-child seam blocks (even-padded) then the other
    // exterior ports.  zsrc[pos] = natural column of Z position pos or -1.
    std::vector<int> zcols, zsrc, zoff, zlen, ext_zpos;
    int ncols = 0, ldp = 0, nz = 0, ldz = 0, nsmax = 0;
    bool gather = false;
    DevBuf d_vu_rc, d_ext_idx, d_ext_child, d_ext_cport, d_seams, d_seam_list, d_seam_off, d_zcols, d_zsrc, d_zoff,
        d_zlen, d_ext_zpos;
    DevBuf DU, SU;
};

struct LevelPlan {
    int M = 1, ksplit = 1;
    std::vector<int> mtiles;   // meet tiles, 8 ints each
    double z_bytes = 0, z_updates = 0, m_bytes = 0, m_updates = 0;
    std::vector<int4> items;   // (W, q, first slot, count)
    std::vector<int4> groups;  // (query, 0, glist offset, count)
    std::vector<int> glist;    // slot indices
};

struct QueryPlan {
    std::vector<int> q_k, q_vp;
    std::vector<long long> q_out;
    std::vector<int2> vp_rc;
    std::vector<uint8_t> vp_valid;
    std::vector<int2> slot_rc;
    std::vector<int> slot_q, slot_i;
    std::vector<LevelPlan> levels;
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
            l.z_bytes = l.z_updates = l.m_bytes = l.m_updates = 0;
            l.items.clear();
            l.mtiles.clear();
            l.groups.clear();
            l.glist.clear();
        }
    }
};

// Per-stream query workspace (grow-only, like HTripQueryScratch,
// include/htrip/minplus_oracle.hpp:189-238).
struct QueryWS {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // recorded after the batch's last operation
    long long total_out = 0;
    QueryPlan plan;
    HostBuf h_plan, h_out;
    DevBuf d_plan, X[2], Z, Zp, d_out;
    int init();
    ~QueryWS();
};

// Event-bracketed launch timing (see htg_profile_enable).  Thread-safe; a
// session's two snapshot maps share one instance.
struct Profiler {
    std::mutex mu;
    bool on = false;
    struct Pending {
        int cls;
        hipEvent_t a, b;
        double bytes, updates;
    };
    std::vector<Pending> pending;          // ended, not yet harvested
    std::unordered_map<int, Pending> open;  // begun, not yet ended
    int next_token = 0;
    std::vector<hipEvent_t> pool;
    htg_kernel_stats acc[HTG_K_COUNT]{};
    hipEvent_t get();
    // Records the start event; returns a token to pass to end().
    int begin(int cls, hipStream_t st, double bytes, double updates);
    void end(int idx, hipStream_t st);
    // Accumulate every completed pair; pairs still running stay pending.
    void harvest();
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
    std::vector<uint8_t> cells_code;  // rows*cols cell codes as last known
    std::vector<uint8_t> pass;        // S*S passable flags (host mirror of d_pass)
    DevBuf d_pass, leafD, leafS;
    std::vector<LevelGeo> lv;  // lv[l-1] is level l
    hipStream_t stream = nullptr;
    bool built = false;
    uint64_t epoch = 0;
    htg_batch_stats stats{0, 0, 0};
    uint64_t budget = 0, device_bytes = 0;
    htg_tuning tune{64, 16, 0, 0};
    QueryWS qws;
    std::shared_ptr<Profiler> profp = std::make_shared<Profiler>();
    Profiler& prof_ref() { return *profp; }
    DirtySet dirty;  // filled by open_cells
    DevBuf ws_nodes, ws_flags, ws_piv, ws_counts, ws_R, ws_changed, ws_mask, ws_edges, ws_cells, ws_leaves,
        ws_leaf_changed;

    int build();
    int open_cells(const int32_t* rc, int n);
    int query(const int32_t* rc, const int32_t* ks, int nq, uint16_t* out_host, QueryWS& ws);
    // Asynchronous halves of query(): plan + launch on ws.stream (results
    // copied to ws.h_out when to_host), then wait for ws.done.
    int query_launch(const int32_t* rc, const int32_t* ks, int nq, QueryWS& ws, bool to_host);
    int query_finish(QueryWS& ws, uint16_t* out_host);
    int ensure_build_ws(int lvl);
    int run_closure(int lvl, int n_list, const std::vector<int>& hcounts);
    const uint16_t* child_summaries(int lvl, size_t& stride) const;
    // Make this map equal to src after src applied one update with dirty set
    // ds (this map equal to src before it): copies only the rewritten tables.
    int catch_up_from(const Map& src, const DirtySet& ds);
};

}  // namespace htg

struct htg_map : htg::Map {};

struct htg_ctx {
    int device = 0;
};
