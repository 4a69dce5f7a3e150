// Host runtime of libhtrip_gpu.so: the C ABI of include/htrip_gpu.h.
//
// A map (htg_map) holds one hierarchy resident in HBM:
//   pass   S x S u8           passable cells, padded to the root side
//   leafD  L x b^2 x b^2 u16  tile-restricted all-pairs distances
//   leafS  L x (4b-4)^2 u16   leaf port summaries
//   DU[l]  N_l x nV x ld u16  node interface closures (rows padded to 128 B)
//   SU[l]  N_l x ne x ne u16  node exterior summaries
// Geometry (port order, V_U order, seam pairs) is identical to the reference
// so every table can be compared element for element.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/htrip_gpu.h"
#include "geometry.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace htg;

namespace htg {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(HTG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Buffers that have to grow double: freeing device or pinned memory waits
// for the whole device, which stalls every other stream's work.
static int debug_level();

// Growing a buffer frees the old one, and freeing device or pinned memory
// waits for the whole device; HTG_DEBUG >= 3 reports every growth.
static void note_growth(const char* what, size_t from, size_t to) {
    if (debug_level() >= 3) std::fprintf(stderr, "[htg] %s grows %zu -> %zu bytes\n", what, from, to);
}

int DevBuf::ensure(size_t n) {
    if (n <= bytes) return HTG_OK;
    if (bytes > 0) n = std::max(n, 2 * bytes);
    note_growth("device buffer", bytes, n);
    release();
    if (n == 0) return HTG_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return fail(HTG_E_NOMEM, "hipMalloc(" + std::to_string(n) + "): " + hipGetErrorString(e));
    }
    bytes = n;
    return HTG_OK;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int HostBuf::ensure(size_t n) {
    if (n <= bytes) return HTG_OK;
    if (bytes > 0) n = std::max(n, 2 * bytes);
    note_growth("pinned buffer", bytes, n);
    release();
    if (n == 0) return HTG_OK;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return fail(HTG_E_NOMEM, "hipHostMalloc(" + std::to_string(n) + "): " + hipGetErrorString(e));
    }
    bytes = n;
    return HTG_OK;
}
void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

hipEvent_t Profiler::get() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    // no system-scope fence when the event completes: that fence writes back
    // and invalidates the L2, so the next kernel would start cold
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}
int Profiler::open(int cls, double bytes, double updates, KTimer& kt, hipStream_t st) {
    kt = KTimer{};
    if (!on) return -1;
    std::lock_guard<std::mutex> lk(mu);
    Pending p{cls, get(), get(), bytes, updates, st};
    if (!p.a || !p.b) {
        if (p.a) pool.push_back(p.a);
        if (p.b) pool.push_back(p.b);
        return -1;
    }
    if (attach) {
        kt.a = p.a;
        kt.b = p.b;
    } else if (hipEventRecord(p.a, st) != hipSuccess) {
        (void)hipGetLastError();
        pool.push_back(p.a);
        pool.push_back(p.b);
        return -1;
    }
    const int token = next_token++ & 0x3fffffff;
    this->open_[token] = p;
    return token;
}
void Profiler::close(int token, bool launched) {
    if (token < 0) return;
    std::lock_guard<std::mutex> lk(mu);
    auto it = open_.find(token);
    if (it == open_.end()) return;
    if (launched && !attach && hipEventRecord(it->second.b, it->second.st) != hipSuccess) {
        (void)hipGetLastError();
        launched = false;
    }
    if (launched) {
        pending.push_back(it->second);
    } else {
        pool.push_back(it->second.a);
        pool.push_back(it->second.b);
    }
    open_.erase(it);
}
void Profiler::harvest() {
    std::lock_guard<std::mutex> lk(mu);
    std::vector<Pending> keep;
    for (auto& p : pending) {
        float ms = 0.f;
        hipError_t q = hipEventQuery(p.b);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            keep.push_back(p);
            continue;
        }
        if (q == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            htg_kernel_stats& a = acc[p.cls];
            a.launches++;
            a.total_ms += ms;
            a.bytes += p.bytes;
            a.updates += p.updates;
        }
        (void)hipGetLastError();
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.swap(keep);
}
void Profiler::add_work(int cls, double bytes, double updates) {
    std::lock_guard<std::mutex> lk(mu);
    acc[cls].bytes += bytes;
    acc[cls].updates += updates;
}
void Profiler::reset() {
    harvest();
    std::lock_guard<std::mutex> lk(mu);
    for (auto& a : acc) a = htg_kernel_stats{0, 0.0, 0.0, 0.0};
}
Profiler::~Profiler() {
    for (auto& kv : open_) pending.push_back(kv.second);
    for (auto& p : pending) {
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// ---------------------------------------------------------------------------
// Level geometry templates (relative to the node origin), shared by every
// node of a level.  Follows src/internal_node.cpp:13-92.
static void make_level(LevelGeo& L, int lvl, int b, int g, int S, int H) {
    L.d = level_dims(lvl, b, g, S);
    const LevelDims& d = L.d;
    L.nn = d.nt * d.nt;
    L.node_stride = size_t(d.nV) * d.ld;
    L.vu_rc.resize(d.nV);
    for (int v = 0; v < d.nV; ++v) {
        const int q = v / d.pc, p = v % d.pc;
        int r, c;
        port_rc(p, d.cs, r, c);
        L.vu_rc[v] = make_short2(short((q / g) * d.cs + r), short((q % g) * d.cs + c));
    }
    auto vertex = [&](int r, int c) -> int {  // src/internal_node.cpp:306-319
        if (r < 0 || c < 0 || r >= d.s || c >= d.s) return -1;
        const int q = (r / d.cs) * g + c / d.cs;
        const int p = port_index(r % d.cs, c % d.cs, d.cs);
        return p < 0 ? -1 : q * d.pc + p;
    };
    L.ext_idx.resize(d.ne);
    L.ext_child.resize(d.ne);
    L.ext_cport.resize(d.ne);
    for (int e = 0; e < d.ne; ++e) {
        int r, c;
        port_rc(e, d.s, r, c);
        const int v = vertex(r, c);
        L.ext_idx[e] = v;
        L.ext_child[e] = v / d.pc;
        L.ext_cport[e] = v % d.pc;
    }
    // Seam candidates: cross-child ports at Manhattan distance 1, listed once (u < v).
    L.seams.clear();
    L.seam_off.assign(g * g + 1, 0);
    L.seam_list.clear();
    const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
    for (int q = 0; q < g * g; ++q) {
        L.seam_off[q] = int(L.seam_list.size());
        for (int p = 0; p < d.pc; ++p) {
            const int u = q * d.pc + p;
            const short2 a = L.vu_rc[u];
            bool seam_port = false;
            for (int k = 0; k < 4; ++k) {
                const int r = a.x + dr[k], c = a.y + dc[k];
                if (r < 0 || c < 0 || r >= d.s || c >= d.s) continue;
                if ((r / d.cs) * g + c / d.cs == q) continue;
                seam_port = true;
                const int v = vertex(r, c);
                if (v > u) L.seams.push_back(make_int2(u, v));
            }
            if (seam_port) L.seam_list.push_back(p);
        }
    }
    L.seam_off[g * g] = int(L.seam_list.size());
    // Query projection columns: every V_U port below the root (seam ports
    // for the meet, exterior ports for the lift); only seam ports at the root.
    L.gather = (lvl == H);
    L.zcols.clear();
    if (!L.gather) {
        L.zcols.resize(d.nV);
        std::iota(L.zcols.begin(), L.zcols.end(), 0);
    } else {
        // seam ports, widened to aligned groups of 4 consecutive V_U indices
        std::vector<int> grp;
        for (int q = 0; q < g * g; ++q)
            for (int i = L.seam_off[q]; i < L.seam_off[q + 1]; ++i) grp.push_back((q * d.pc + L.seam_list[i]) / 4);
        std::sort(grp.begin(), grp.end());
        grp.erase(std::unique(grp.begin(), grp.end()), grp.end());
        for (int g4 : grp)
            for (int e = 0; e < 4; ++e) L.zcols.push_back(4 * g4 + e);
    }
    L.ncols = int(L.zcols.size());
    L.ldp = round_up(std::max(L.ncols, 8), 8);
    std::vector<int> natcol(d.nV, -1), zpos_of(d.nV, -1);
    for (int i = 0; i < L.ncols; ++i) natcol[L.zcols[i]] = i;
    L.zsrc.clear();
    L.zoff.assign(g * g, 0);
    L.zlen.assign(g * g, 0);
    L.nsmax = 0;
    for (int q = 0; q < g * g; ++q) {
        L.zoff[q] = int(L.zsrc.size());
        const int ns = L.seam_off[q + 1] - L.seam_off[q];
        L.nsmax = std::max(L.nsmax, ns);
        for (int i = L.seam_off[q]; i < L.seam_off[q + 1]; ++i) {
            const int v = q * d.pc + L.seam_list[i];
            zpos_of[v] = int(L.zsrc.size());
            L.zsrc.push_back(natcol[v]);
        }
        if (ns & 1) L.zsrc.push_back(-1);
        L.zlen[q] = int(L.zsrc.size()) - L.zoff[q];
    }
    // Z holds the seam blocks only: the meet reads nothing else, and the
    // lift to the exterior ports goes to X of the next level directly.
    L.nz = int(L.zsrc.size());
    L.ldz = round_up(std::max(L.nz, 8), 8);
    L.ext_col.assign(d.ne, -1);
    if (!L.gather)
        for (int e = 0; e < d.ne; ++e) L.ext_col[e] = natcol[size_t(L.ext_idx[e])];
    L.vu_ext.assign(d.nV, -1);
    for (int e = 0; e < d.ne; ++e) L.vu_ext[size_t(L.ext_idx[e])] = e;
    L.colpos.assign(L.ncols, -1);
    L.colext.assign(L.ncols, -1);
    for (int i = 0; i < L.ncols; ++i) L.colpos[i] = zpos_of[L.zcols[i]];
    if (!L.gather)
        for (int e = 0; e < d.ne; ++e) L.colext[size_t(L.ext_col[e])] = e;
    L.seam_live.assign(size_t(L.nn) * g * g, 0);
    L.ng8 = d.ld / 8;
    L.ngrp = L.ncols / 4;
    // inverse tables for incremental liveness: group of each aligned quad of
    // V_U ports (-1: not projected), seam flag of each V_U port
    L.grp_of4.assign(size_t(d.nV / 4), -1);
    for (int gi = 0; gi < L.ngrp; ++gi) L.grp_of4[size_t(L.zcols[size_t(4 * gi)] / 4)] = gi;
    L.seam_flag.assign(size_t(d.nV), 0);
    for (int q = 0; q < g * g; ++q)
        for (int i = L.seam_off[q]; i < L.seam_off[q + 1]; ++i) L.seam_flag[size_t(q * d.pc + L.seam_list[i])] = 1;
}

int Map::init_live() {
    // per level: cnt (nn), list (nn * ngrp), cnt8 (nn), list8 (nn * ng8) ints, map (nn * ngrp) bytes, then at
    // the root its full octet lists: cnt (nn), list (nn * ng8) ints
    auto a16 = [](size_t x) { return (x + 15) / 16 * 16; };
    std::vector<size_t> off(lv.size() * 7);
    size_t total = 0;
    for (size_t l = 0; l < lv.size(); ++l) {
        LevelGeo& L = lv[l];
        L.seam_only = l + 1 == lv.size();
        const size_t nn = size_t(L.nn);
        const size_t sz[7] = {nn * 4, nn * size_t(L.ngrp) * 4, nn * 4, nn * size_t(L.ng8) * 4, nn * size_t(L.ngrp),
                              L.seam_only ? nn * 4 : 0, L.seam_only ? nn * size_t(L.ng8) * 4 : 0};
        for (int k = 0; k < 7; ++k) {
            off[l * 7 + k] = total;
            total += a16(std::max<size_t>(sz[k], 16));
        }
    }
    int rc;
    if ((rc = live_h.ensure(total)) || (rc = live_d.ensure(total))) return rc;
    std::memset(live_h.p, 0, total);
    char* h = static_cast<char*>(live_h.p);
    char* d = static_cast<char*>(live_d.p);
    for (size_t l = 0; l < lv.size(); ++l) {
        LevelGeo& L = lv[l];
        L.h_live_cnt = reinterpret_cast<int*>(h + off[l * 7 + 0]);
        L.h_live_list = reinterpret_cast<int*>(h + off[l * 7 + 1]);
        L.h_live8_cnt = reinterpret_cast<int*>(h + off[l * 7 + 2]);
        L.h_live8_list = reinterpret_cast<int*>(h + off[l * 7 + 3]);
        L.h_live_map = reinterpret_cast<uint8_t*>(h + off[l * 7 + 4]);
        L.d_live_cnt = reinterpret_cast<const int*>(d + off[l * 7 + 0]);
        L.d_live_list = reinterpret_cast<const int*>(d + off[l * 7 + 1]);
        L.d_live8_cnt = reinterpret_cast<const int*>(d + off[l * 7 + 2]);
        L.d_live8_list = reinterpret_cast<const int*>(d + off[l * 7 + 3]);
        L.d_live_map = reinterpret_cast<const uint8_t*>(d + off[l * 7 + 4]);
        const int a = L.seam_only ? 5 : 2, b = L.seam_only ? 6 : 3;
        L.h_all8_cnt = reinterpret_cast<int*>(h + off[l * 7 + a]);
        L.h_all8_list = reinterpret_cast<int*>(h + off[l * 7 + b]);
        L.d_all8_cnt = reinterpret_cast<const int*>(d + off[l * 7 + a]);
        L.d_all8_list = reinterpret_cast<const int*>(d + off[l * 7 + b]);
    }
    hipError_t e;
    if (!live_evt && (e = hipEventCreateWithFlags(&live_evt, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    return HTG_OK;
}

static int debug_level();

Map::~Map() {
    if (debug_level() >= 2 && upd_count > 0)
        std::fprintf(stderr,
                     "[htg] update host phases per call (%lld calls): cells+live %.0f us, plan %.0f us, levels "
                     "host %.0f us (rebuild decisions %.0f us), levels device wait %.0f us, stats %.0f us; leaf phase "
                     "%.0f us of the first; plan: pivots %.0f, seams %.0f, candidates %.0f, new ports %.0f us; "
                     "catch-up %.0f us per call (%lld)\n",
                     upd_count, 1e6 * upd_time[0] / upd_count, 1e6 * upd_time[1] / upd_count,
                     1e6 * upd_time[2] / upd_count, 1e6 * upd_time[6] / upd_count, 1e6 * upd_time[3] / upd_count,
                     1e6 * upd_time[4] / upd_count, 1e6 * upd_time[7] / upd_count, 1e6 * upd_time[8] / upd_count,
                     1e6 * upd_time[9] / upd_count, 1e6 * upd_time[10] / upd_count, 1e6 * upd_time[11] / upd_count,
                     catchup_count ? 1e6 * upd_time[5] / catchup_count : 0.0, catchup_count);
    if (build_graph) (void)hipGraphExecDestroy(build_graph);
    if (live_evt) {
        (void)hipEventSynchronize(live_evt);
        (void)hipEventDestroy(live_evt);
    }
}

int Map::refresh_live(const std::vector<int2>* cells) {
    hipError_t e;
    if (live_pending) {  // the previous upload still reads the pinned block
        if ((e = hipEventSynchronize(live_evt)) != hipSuccess) return hip_fail(e, "live upload wait");
        live_pending = false;
    }
    const char* hbase = static_cast<const char*>(live_h.p);
    size_t lo = ~size_t(0), hi = 0;  // changed byte range of the block
    auto touch = [&](const void* p, size_t bytes) {
        const size_t o = size_t(static_cast<const char*>(p) - hbase);
        lo = std::min(lo, o);
        hi = std::max(hi, o + bytes);
    };
    // insert x into the ascending list a[0..n) (absent); returns the new size
    auto insert_sorted = [](int* a, int n, int x) {
        int i = n;
        while (i > 0 && a[i - 1] > x) {
            a[i] = a[i - 1];
            --i;
        }
        a[i] = x;
        return n + 1;
    };
    for (int l = 1; l <= H; ++l) {
        LevelGeo& L = lv[l - 1];
        const int s = L.d.s, nt = L.d.nt, ng = L.ngrp, cs = L.d.cs;
        int a = L.nn, z = -1;  // touched node range
        if (!cells) {  // every node from scratch
            for (int W = 0; W < L.nn; ++W) {
                const int R0 = (W / nt) * s, C0 = (W % nt) * s;
                auto live_port = [&](int v) {
                    const short2 p = L.vu_rc[size_t(v)];
                    return pass[size_t(R0 + p.x) * S + size_t(C0 + p.y)] != 0;
                };
                int cnt = 0;
                for (int gi = 0; gi < ng; ++gi) {
                    bool live = false;
                    for (int k = 0; k < 4 && !live; ++k) live = live_port(L.zcols[size_t(4 * gi + k)]);
                    L.h_live_map[size_t(W) * ng + gi] = live ? 1 : 0;
                    if (live) L.h_live_list[size_t(W) * ng + cnt++] = gi;
                }
                L.h_live_cnt[W] = cnt;
                for (int q = 0; q < g * g; ++q) {
                    int ns_live = 0;
                    for (int i = L.seam_off[q]; i < L.seam_off[q + 1]; ++i)
                        ns_live += live_port(q * L.d.pc + L.seam_list[i]);
                    L.seam_live[size_t(W) * g * g + q] = ns_live;
                }
                int c8 = 0, a8 = 0;
                for (int o = 0; o < L.ng8; ++o) {
                    bool live = false, seam = false;
                    for (int v = 8 * o; v < 8 * o + 8 && v < L.d.nV && !seam; ++v)
                        if (live_port(v)) {
                            live = true;
                            seam = L.seam_flag[size_t(v)] != 0;
                        }
                    if (L.seam_only ? seam : live) L.h_live8_list[size_t(W) * L.ng8 + c8++] = o;
                    if (L.seam_only && live) L.h_all8_list[size_t(W) * L.ng8 + a8++] = o;
                }
                L.h_live8_cnt[W] = c8;
                if (L.seam_only) L.h_all8_cnt[W] = a8;
            }
            a = 0;
            z = L.nn - 1;
        } else {  // cells that became passable: only their own ports change
            for (const int2& c : *cells) {
                const int ty = c.x / s, tx = c.y / s, W = ty * nt + tx;
                const int r = c.x - ty * s, cc = c.y - tx * s;
                const int pi = port_index(r % cs, cc % cs, cs);
                if (pi < 0) continue;  // not a port of its node at this level
                const int q = (r / cs) * g + cc / cs;
                const int v = q * L.d.pc + pi;
                a = std::min(a, W);
                z = std::max(z, W);
                const int gi = L.grp_of4[size_t(v / 4)];
                if (gi >= 0 && !L.h_live_map[size_t(W) * ng + gi]) {
                    L.h_live_map[size_t(W) * ng + gi] = 1;
                    L.h_live_cnt[W] = insert_sorted(L.h_live_list + size_t(W) * ng, L.h_live_cnt[W], gi);
                }
                if (L.seam_flag[size_t(v)]) L.seam_live[size_t(W) * g * g + q]++;
                const int o = v / 8;
                if (!L.seam_only || L.seam_flag[size_t(v)]) {
                    int* l8 = L.h_live8_list + size_t(W) * L.ng8;
                    const int n8 = L.h_live8_cnt[W];
                    if (!std::binary_search(l8, l8 + n8, o)) L.h_live8_cnt[W] = insert_sorted(l8, n8, o);
                }
                if (L.seam_only) {
                    int* a8 = L.h_all8_list + size_t(W) * L.ng8;
                    const int n8 = L.h_all8_cnt[W];
                    if (!std::binary_search(a8, a8 + n8, o)) L.h_all8_cnt[W] = insert_sorted(a8, n8, o);
                }
            }
        }
        if (z < a) continue;
        const size_t a0 = size_t(a), z0 = size_t(z) + 1;
        touch(L.h_live_cnt + a0, (z0 - a0) * 4);
        touch(L.h_live_list + a0 * ng, (z0 - a0) * ng * 4);
        touch(L.h_live8_cnt + a0, (z0 - a0) * 4);
        touch(L.h_live8_list + a0 * L.ng8, (z0 - a0) * L.ng8 * 4);
        if (L.seam_only) {
            touch(L.h_all8_cnt + a0, (z0 - a0) * 4);
            touch(L.h_all8_list + a0 * L.ng8, (z0 - a0) * L.ng8 * 4);
        }
        touch(L.h_live_map + a0 * ng, (z0 - a0) * ng);
    }
    if (hi <= lo) return HTG_OK;
    if ((e = hipMemcpyAsync(static_cast<char*>(live_d.p) + lo, hbase + lo, hi - lo, hipMemcpyHostToDevice, stream)) !=
            hipSuccess ||
        (e = hipEventRecord(live_evt, stream)) != hipSuccess)
        return hip_fail(e, "upload live groups");
    live_pending = true;
    return HTG_OK;
}

// Synchronous copies go through the map's stream: the null stream would take
// a hardware queue of its own.
template <typename T>
static int upload(DevBuf& dst, const std::vector<T>& src, hipStream_t st) {
    int rc = dst.ensure(std::max<size_t>(src.size() * sizeof(T), 16));
    if (rc) return rc;
    if (!src.empty()) {
        hipError_t e = hipMemcpyAsync(dst.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy geometry");
    }
    return HTG_OK;
}

// Reserve of the incremental leaf reveal: BFS rows of the new cells (up to
// 16 MiB) and the leaf-phase list staging.
static uint64_t leaf_ws_bytes(int b, int S) {
    const uint64_t nl = uint64_t(S / b), n = uint64_t(b) * b;
    return std::min<uint64_t>(16ull << 20, nl * nl * kLeafRelaxMax * n * 2) + kUpdateListReserve / 4;
}

static uint64_t tables_bytes(int b, int g, int S, int H) {
    const uint64_t nl = uint64_t(S / b), n = uint64_t(b) * b, pl = 4ull * b - 4;
    uint64_t t = uint64_t(S) * S + nl * nl * (n * n + pl * pl) * 2;
    for (int l = 1; l <= H; ++l) {
        LevelDims d = level_dims(l, b, g, S);
        t += uint64_t(d.nt) * d.nt * (uint64_t(d.nV) * d.ld + uint64_t(d.ne) * d.ne) * 2;
    }
    return t;
}

// Largest per-level build workspace: flag planes, pivot and B lists, counts,
// pivot rows.
static uint64_t build_ws_bytes(int b, int g, int S, int H, int P) {
    uint64_t m = 0;
    for (int l = 1; l <= H; ++l) {
        LevelDims d = level_dims(l, b, g, S);
        const uint64_t nn = uint64_t(d.nt) * d.nt;
        const uint64_t w = nn * (uint64_t(d.nV) * (3 + 4 + 4) + 24 + uint64_t(P) * d.ld * 2) + nn * 4 * 4;
        m = std::max(m, w);
    }
    return m + kUpdateListReserve + leaf_ws_bytes(b, S);
}

static void geometry(int b, int g, int rows, int cols, int& S, int& H, int min_S = 0) {
    S = b;
    while (S < std::max(rows, cols) || S < min_S) S *= g;
    H = 0;
    for (int t = S / b; t > 1; t /= g) ++H;
}

// Update lists and changed flags, and every level's closure workspace, at
// their working sizes before the first update (a map that is only ever
// updated, such as the second copy of a snapshot session, never builds).
int Map::reserve_update_ws() {
    size_t nchg = size_t(nl) * nl;
    for (const LevelGeo& L : lv) nchg += size_t(L.nn);
    int rc;
    size_t nn_max = 0;
    for (const LevelGeo& L : lv) nn_max = std::max(nn_max, size_t(L.nn));
    const uint64_t leafP = leaf_ws_bytes(b, S) - kUpdateListReserve / 4;
    if ((rc = ws_upd.ensure(kUpdateListReserve)) || (rc = ws_upd_h.ensure(kUpdateListReserve)) ||
        (rc = ws_leafl.ensure(kUpdateListReserve / 4)) || (rc = ws_leafl_h.ensure(kUpdateListReserve / 4)) ||
        (rc = ws_leafP.ensure(size_t(leafP))) ||
        (rc = ws_chg.ensure(nchg * 4 + 16)) || (rc = ws_chg_h.ensure(nchg * 4 + 16)) ||
        (rc = ws_counts_h.ensure(2 * nn_max * 4 + 16)) || (rc = ws_cov.ensure(kCoverUnitReserve)) ||
        (rc = ws_cov_h.ensure(kCoverUnitReserve)))
        return rc;
    for (int lvl = 1; lvl <= H; ++lvl)
        if ((rc = ensure_build_ws(lvl))) return rc;
    return HTG_OK;
}

int Map::ensure_build_ws(int lvl) {
    const LevelGeo& L = lv[lvl - 1];
    const size_t nn = size_t(L.nn);
    int rc;
    if ((rc = ws_nodes.ensure(nn * 4))) return rc;
    if ((rc = ws_flags.ensure(3 * nn * L.d.nV))) return rc;  // C marks, B marks, new ports
    if ((rc = ws_piv.ensure(nn * L.d.nV * 4))) return rc;
    if ((rc = ws_bpiv.ensure(nn * L.d.nV * 4))) return rc;
    if ((rc = ws_counts.ensure(2 * nn * 4 + 16))) return rc;
    // two pivot-row buffers: a block's update runs beside the next block's pivots
    if ((rc = ws_R.ensure(2 * nn * size_t(tune.pivot_block) * L.d.ld * 2))) return rc;
    if ((rc = ws_changed.ensure(nn * 4 + 16))) return rc;
    if ((rc = ws_mask.ensure(nn * 4 + 16))) return rc;
    return HTG_OK;
}

// Closure over each listed node's pivot list (already compacted into ws_piv
// with counts in ws_counts; host copy of counts in hcounts).
int Map::run_closure(int lvl, const int* nodes, const std::vector<int>& hnodes, const std::vector<int>& hcounts,
                     bool full) {
    LevelGeo& L = lv[lvl - 1];
    const int* h8 = full ? L.h_all8_cnt : L.h_live8_cnt;
    const int* d8c = full ? L.d_all8_cnt : L.d_live8_cnt;
    const int* d8l = full ? L.d_all8_list : L.d_live8_list;
    const int n_list = int(hnodes.size());
    int maxc = 0, max8 = 0;
    for (int i = 0; i < n_list; ++i) {
        maxc = std::max(maxc, hcounts[i]);
        if (hcounts[i] > 0) max8 = std::max(max8, h8[hnodes[i]]);
    }
    const int P = tune.pivot_block;
    const int blocks = (maxc + P - 1) / P;
    // Pivot blocks of up to 32: block 0's pivots, then per block t its update
    // together with block t + 1's pivots, then the last block's update
    // (blocks + 1 launches; the pivot rows alternate between the two halves
    // of ws_R).  Blocks of 64 run the pivots and the update of each block as
    // two launches: the next block's pivot rows then cost each pivot
    // workgroup a 64 x 64 x 64 correction that outweighs the saved launch.
    const bool fused = P <= 32;
    uint16_t* Rb[2] = {static_cast<uint16_t*>(ws_R.p), static_cast<uint16_t*>(ws_R.p) + size_t(L.nn) * P * L.d.ld};
    auto figures = [&](int t, double& upd_p, double& upd_u, double& by_p, double& by_u) {
        upd_p = upd_u = by_p = by_u = 0;
        if (!prof_ref().on || t >= blocks) return;
        for (int i = 0; i < n_list; ++i) {
            const double p = std::max(0, std::min(P, hcounts[i] - t * P));
            if (p <= 0) continue;
            const double nV = 8.0 * h8[hnodes[i]];  // live rows / columns
            upd_p += p * p * nV + p * p * p;
            by_p += (p * nV + p * nV) * 2;  // pivot rows in, R out
            upd_u += p * nV * nV;
            by_u += (2 * nV * nV + 2 * p * nV) * 2;  // D_U tile in + out, R in
        }
    };
    for (int t = 0; t < blocks; ++t) {
        double upd_p, upd_u, by_p, by_u, up2 = 0, uu2 = 0, bp2 = 0, bu2 = 0;
        figures(t, upd_p, upd_u, by_p, by_u);
        uint16_t* R = fused ? Rb[t & 1] : Rb[0];
        if (t == 0 || !fused) {
            KTimer kt;
            const int pi = prof_open(HTG_K_CLOSURE_PIVOTS, by_p, upd_p, kt);
            const bool la = launch_closure_pivots(P, L.d.nV, L.d.ld, nodes, n_list, static_cast<const int*>(ws_piv.p),
                                                  L.d.nV, static_cast<const int*>(ws_counts.p), t,
                                                  static_cast<const uint16_t*>(L.DU.p), L.node_stride, d8c, d8l,
                                                  L.ng8, max8, R, stream, kt);
            prof_ref().close(pi, la);
        }
        // the update of block t (fused: with block t + 1's pivots, counted with it)
        const bool step = fused && t + 1 < blocks;
        if (step) figures(t + 1, up2, uu2, bp2, bu2);
        KTimer kt;
        const int pi = prof_open(HTG_K_CLOSURE_UPDATE, by_u + bp2, upd_u + up2, kt);
        const bool la =
            step ? launch_closure_step(P, L.d.nV, L.d.ld, nodes, n_list, static_cast<const int*>(ws_piv.p), L.d.nV,
                                       static_cast<const int*>(ws_counts.p), t, static_cast<uint16_t*>(L.DU.p),
                                       L.node_stride, d8c, d8l, L.ng8, max8, R, Rb[(t + 1) & 1], stream, kt)
                 : launch_closure_update(P, L.d.nV, L.d.ld, nodes, n_list, static_cast<const int*>(ws_counts.p), t,
                                         static_cast<uint16_t*>(L.DU.p), L.node_stride, d8c, d8l, L.ng8, max8, R,
                                         stream, kt);
        prof_ref().close(pi, la);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "closure launch");
    return HTG_OK;
}

const uint16_t* Map::child_summaries(int lvl, size_t& stride) const {
    if (lvl == 1) {
        stride = size_t(4 * b - 4) * (4 * b - 4);
        return static_cast<const uint16_t*>(leafS.p);
    }
    const LevelGeo& C = lv[lvl - 2];
    stride = size_t(C.d.ne) * C.d.ne;
    return static_cast<const uint16_t*>(C.SU.p);
}

std::vector<int> Map::build_pivot_counts(int lvl, const std::vector<int>* nodes) const {
    const LevelGeo& L = lv[size_t(lvl) - 1];
    const int n_out = nodes ? int(nodes->size()) : L.nn;
    std::vector<int> cnt(size_t(n_out), 0);
    std::vector<uint8_t> mark(size_t(L.d.nV));
    for (int i = 0; i < n_out; ++i) {
        const int W = nodes ? (*nodes)[size_t(i)] : i;
        const int R0 = (W / L.d.nt) * L.d.s, C0 = (W % L.d.nt) * L.d.s;
        std::fill(mark.begin(), mark.end(), 0);
        int c = 0;
        for (const int2& e : L.seams) {
            const short2 a = L.vu_rc[size_t(e.x)], d = L.vu_rc[size_t(e.y)];
            if (!pass[size_t(R0 + a.x) * S + size_t(C0 + a.y)] || !pass[size_t(R0 + d.x) * S + size_t(C0 + d.y)])
                continue;
            c += !mark[size_t(e.x)] + !mark[size_t(e.y)];
            mark[size_t(e.x)] = mark[size_t(e.y)] = 1;
        }
        cnt[size_t(i)] = c;
    }
    return cnt;
}

// onCellBlocked (src/minplus_oracle.cpp:247-302), for a list of cells: they
// become impassable (Obstacle); their leaves are recomputed by the BFS kernel
// and every ancestor node is rebuilt from its children with the build's
// kernels on those nodes only (assemble, seam edges, closure, S_U).  Other
// nodes' tables do not depend on these cells.  The live lists stay a
// superset: a port that turned impassable keeps an INF row and column, which
// the live-row kernels and the snapshot copy rely on.
int Map::block_cells(const int32_t* rcs, int n) {
    timing_now = prof_ref().take();
    std::vector<int2> closed;
    std::vector<int> leaves;
    for (int i = 0; i < n; ++i) {
        const int r = rcs[2 * i], c = rcs[2 * i + 1];
        if (r < 0 || c < 0 || r >= rows || c >= cols) continue;
        const size_t key = size_t(r) * S + c;
        if (!pass[key]) continue;  // already impassable: no table changes
        pass[key] = 0;
        cells_code[size_t(r) * cols + c] = 2;
        closed.push_back(make_int2(r, c));
        leaves.push_back((r / b) * nl + c / b);
    }
    epoch++;  // the reference bumps it on every call (src/minplus_oracle.cpp:300)
    if (closed.empty()) return HTG_OK;
    std::sort(leaves.begin(), leaves.end());
    leaves.erase(std::unique(leaves.begin(), leaves.end()), leaves.end());
    // ancestors per level, ascending
    std::vector<std::vector<int>> anc(size_t(H) + 1);
    {
        std::vector<int> below = leaves;
        int below_nt = nl;
        for (int lvl = 1; lvl <= H; ++lvl) {
            const LevelGeo& L = lv[size_t(lvl) - 1];
            std::vector<int>& a = anc[size_t(lvl)];
            for (int ch : below) a.push_back(((ch / below_nt) / g) * L.d.nt + (ch % below_nt) / g);
            std::sort(a.begin(), a.end());
            a.erase(std::unique(a.begin(), a.end()), a.end());
            below = a;
            below_nt = L.d.nt;
        }
    }
    return rebuild(closed, leaves, anc);
}

// Estimated device time (us) of the blocked closure of the listed nodes over
// cnt[i] pivots each: blocks of P pivots, each block at least the latency of
// its two launches, else bound by the packed-u16 VALU rate (P x live^2 updates
// per node and block).
static constexpr double kUpdatesPerUs = 23.6e6;  // ~60 % of the 39.3 T/s packed min-plus peak (measured)
static constexpr double kBlockUs = 55.0;         // a pivot block's two launches at small sizes (measured)
double Map::closure_us(int lvl, const std::vector<int>& nodes, const int* cnt) const {
    const LevelGeo& L = lv[size_t(lvl) - 1];
    const int P = tune.pivot_block;
    int maxc = 0;
    for (size_t i = 0; i < nodes.size(); ++i) maxc = std::max(maxc, cnt[i]);
    double t = 0;
    for (int b0 = 0; b0 < maxc; b0 += P) {
        double upd = 0;
        for (size_t i = 0; i < nodes.size(); ++i) {
            const double p = std::max(0, std::min(P, cnt[i] - b0));
            const double live = 8.0 * L.h_live8_cnt[nodes[i]];
            upd += p * live * live;
        }
        t += std::max(kBlockUs, upd / kUpdatesPerUs);
    }
    return t;
}

// The update of a level after its counts are known (cover C and uncovered
// endpoints B per node, hc = C then B) against a rebuild of the same nodes.
bool Map::rebuild_is_cheaper(int lvl, const std::vector<int>& nodes, const std::vector<int>& hc, int na,
                             const std::vector<int>& pb) const {
    if (tune.update_rebuild < 0) return false;
    if (tune.update_rebuild > 0) return true;
    const LevelGeo& L = lv[size_t(lvl) - 1];
    double cover = 0;
    bool any_b = false;
    for (int i = 0; i < na; ++i) {
        const double c = hc[size_t(i)], b = hc[size_t(na + i)], live = 8.0 * L.h_live8_cnt[nodes[size_t(i)]];
        if (c <= 0 || b <= 0) continue;
        cover += c * b * live + c * c * b;
        any_b = true;
    }
    const double upd = closure_us(lvl, nodes, hc.data()) + cover / kUpdatesPerUs + (any_b ? 2 * kBlockUs : 0);
    double assemble = 0;
    for (size_t i = 0; i < nodes.size(); ++i) assemble += double(L.d.nV) * L.d.ld * 2 / 4.0e6;  // ~4 TB/s
    const double reb = closure_us(lvl, nodes, pb.data()) + assemble + 3 * kBlockUs / 2;
    if (debug_level())
        std::fprintf(stderr, "[htg] update lvl %d: update ~%.0f us, rebuild ~%.0f us\n", lvl, upd, reb);
    return reb < upd;
}

// Rebuild the listed nodes of one level from their children with the build
// kernels (assemble, seam edges, closure over the seam endpoints, S_U);
// changed[i] (nullable) is set when listed node i's S_U decreased.
int Map::rebuild_level(int lvl, const int* dnodes, const std::vector<int>& hn, int* changed) {
    LevelGeo& L = lv[size_t(lvl) - 1];
    const int na = int(hn.size());
    if (na == 0) return HTG_OK;
    if (L.seam_only) root_full = false;
    int rc;
    if ((rc = ensure_build_ws(lvl))) return rc;
    launch_fill_u8(static_cast<uint8_t*>(ws_flags.p), size_t(na) * L.d.nV, 0, stream);
    size_t cstride;
    const uint16_t* csum = child_summaries(lvl, cstride);
    launch_node_assemble(lvl, g, L.d.pc, L.d.nV, L.d.ld, L.d.nt, csum, cstride, dnodes, na,
                         static_cast<uint16_t*>(L.DU.p), L.node_stride, stream);
    launch_seam_relax(L.d.nV, L.d.ld, L.d.s, L.d.nt, static_cast<const int2*>(L.d_seams.p), int(L.seams.size()),
                      static_cast<const short2*>(L.d_vu_rc.p), static_cast<const uint8_t*>(d_pass.p), S, dnodes, na,
                      static_cast<uint16_t*>(L.DU.p), L.node_stride, static_cast<uint8_t*>(ws_flags.p), stream);
    launch_compact_flags(L.d.nV, dnodes, na, static_cast<uint8_t*>(ws_flags.p), static_cast<int*>(ws_piv.p), L.d.nV,
                         static_cast<int*>(ws_counts.p), stream);
    if ((rc = run_closure(lvl, dnodes, hn, build_pivot_counts(lvl, &hn)))) return rc;
    launch_project_su(L.d.nV, L.d.ld, L.d.ne, static_cast<const int*>(L.d_ext_idx.p), dnodes, na,
                      static_cast<const uint16_t*>(L.DU.p), L.node_stride, static_cast<uint16_t*>(L.SU.p), changed,
                      stream);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HTG_OK : hip_fail(e, "rebuild launch");
}

// The root's full table from the maintained one: entries with a row or a
// column outside the maintained octets are assembled afresh from the
// children, then one closure over the seam endpoints runs over every live
// octet.  The maintained block is exact already: the closure only rewrites
// its entries with the values they hold, so queries reading it from other
// streams meanwhile see no change.  S_U follows from the full table.
int Map::materialize_root() {
    if (H == 0 || root_full || !built || broken) return HTG_OK;
    LevelGeo& L = lv[size_t(H) - 1];
    int rc;
    if ((rc = ensure_build_ws(H))) return rc;
    const int* nodes = static_cast<const int*>(L.d_all.p);
    const std::vector<int> root(size_t(L.nn), 0);  // one node
    launch_fill_u8(static_cast<uint8_t*>(ws_flags.p), size_t(L.nn) * L.d.nV, 0, stream);
    size_t cstride;
    const uint16_t* csum = child_summaries(H, cstride);
    launch_node_assemble(H, g, L.d.pc, L.d.nV, L.d.ld, L.d.nt, csum, cstride, nodes, L.nn,
                         static_cast<uint16_t*>(L.DU.p), L.node_stride, stream, L.d_live8_cnt, L.d_live8_list, L.ng8);
    launch_seam_relax(L.d.nV, L.d.ld, L.d.s, L.d.nt, static_cast<const int2*>(L.d_seams.p), int(L.seams.size()),
                      static_cast<const short2*>(L.d_vu_rc.p), static_cast<const uint8_t*>(d_pass.p), S, nodes, L.nn,
                      static_cast<uint16_t*>(L.DU.p), L.node_stride, static_cast<uint8_t*>(ws_flags.p), stream);
    launch_compact_flags(L.d.nV, nodes, L.nn, static_cast<uint8_t*>(ws_flags.p), static_cast<int*>(ws_piv.p), L.d.nV,
                         static_cast<int*>(ws_counts.p), stream);
    if ((rc = run_closure(H, nodes, root, build_pivot_counts(H, &root), true))) return rc;
    launch_project_su(L.d.nV, L.d.ld, L.d.ne, static_cast<const int*>(L.d_ext_idx.p), nodes, L.nn,
                      static_cast<const uint16_t*>(L.DU.p), L.node_stride, static_cast<uint16_t*>(L.SU.p), nullptr,
                      stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "root closure");
    root_full = true;
    return HTG_OK;
}

// Recompute the listed leaves (BFS from the passability) and rebuild the
// listed nodes of every level from their children with the build kernels
// (assemble, seam edges, closure over the seam endpoints, S_U); closed cells
// are cleared in the device passability first.  anc[lvl] lists level lvl's
// nodes (ascending; anc[0] unused).  Stamps every listed table.
int Map::rebuild(const std::vector<int2>& closed, const std::vector<int>& leaves,
                 const std::vector<std::vector<int>>& anc) {
    hipError_t e;
    int rc;
    Packer pk;
    const size_t o_cells = pk.add(closed), o_leaves = pk.add(leaves);
    std::vector<size_t> o_anc(size_t(H) + 1);
    for (int lvl = 1; lvl <= H; ++lvl) o_anc[size_t(lvl)] = pk.add(anc[size_t(lvl)]);
    if ((rc = ws_upd_h.ensure(pk.size)) || (rc = ws_upd.ensure(pk.size))) return rc;
    pk.write(static_cast<char*>(ws_upd_h.p));
    if ((e = hipMemcpyAsync(ws_upd.p, ws_upd_h.p, pk.size, hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hip_fail(e, "upload block lists");
    const char* up = static_cast<const char*>(ws_upd.p);
    if (!closed.empty())
        launch_set_pass(static_cast<uint8_t*>(d_pass.p), S, reinterpret_cast<const int2*>(up + o_cells),
                        int(closed.size()), stream, 0);
    const double leaf_bytes = (double(b) * b * b * b + double(4 * b - 4) * (4 * b - 4)) * 2 + double(b) * b;
    KTimer lkt;
    const int lpi = prof_open(HTG_K_LEAF_APSP, leaf_bytes * double(leaves.size()), 0, lkt);
    const bool lla = launch_leaf_apsp(static_cast<const uint8_t*>(d_pass.p), S, b, nl,
                                      reinterpret_cast<const int*>(up + o_leaves), int(leaves.size()),
                                      static_cast<uint16_t*>(leafD.p), static_cast<uint16_t*>(leafS.p), nullptr, stream,
                                      lkt, kBackgroundWG);
    prof_ref().close(lpi, lla);
    for (int lvl = 1; lvl <= H; ++lvl) {
        if (anc[size_t(lvl)].empty()) continue;
        if ((rc = rebuild_level(lvl, reinterpret_cast<const int*>(up + o_anc[size_t(lvl)]), anc[size_t(lvl)],
                                nullptr)))
            return rc;
    }
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "rebuild launch");
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "rebuild sync");
    prof_ref().harvest();
    stamp(leaves, anc, 1);
    return HTG_OK;
}

// Build (src/minplus_oracle.cpp:48-137): leaf APSP, then per level assemble
// -> seam edges -> pivot compaction -> blocked closure -> S_U, every node of
// a level in the same launches.  The pivot counts come from the host, so the
// sequence has no host sync and can run as one captured graph.
int Map::enqueue_build(const std::vector<std::vector<int>>& hc) {
    hipError_t e;
    const double leaf_bytes = (double(b) * b * b * b + double(4 * b - 4) * (4 * b - 4)) * 2 + double(b) * b;
    KTimer kt;
    int pi = prof_open(HTG_K_LEAF_APSP, leaf_bytes * nl * nl, 0, kt);
    const bool la = launch_leaf_apsp(static_cast<const uint8_t*>(d_pass.p), S, b, nl, nullptr, nl * nl,
                                     static_cast<uint16_t*>(leafD.p), static_cast<uint16_t*>(leafS.p), nullptr, stream,
                                     kt);
    prof_ref().close(pi, la);
    for (int lvl = 1; lvl <= H; ++lvl) {
        LevelGeo& L = lv[size_t(lvl) - 1];
        const int* nodes = static_cast<const int*>(L.d_all.p);
        launch_fill_u8(static_cast<uint8_t*>(ws_flags.p), size_t(L.nn) * L.d.nV, 0, stream);
        size_t cstride;
        const uint16_t* csum = child_summaries(lvl, cstride);
        launch_node_assemble(lvl, g, L.d.pc, L.d.nV, L.d.ld, L.d.nt, csum, cstride, nodes, L.nn,
                             static_cast<uint16_t*>(L.DU.p), L.node_stride, stream);
        launch_seam_relax(L.d.nV, L.d.ld, L.d.s, L.d.nt, static_cast<const int2*>(L.d_seams.p), int(L.seams.size()),
                          static_cast<const short2*>(L.d_vu_rc.p), static_cast<const uint8_t*>(d_pass.p), S, nodes,
                          L.nn, static_cast<uint16_t*>(L.DU.p), L.node_stride, static_cast<uint8_t*>(ws_flags.p),
                          stream);
        launch_compact_flags(L.d.nV, nodes, L.nn, static_cast<uint8_t*>(ws_flags.p), static_cast<int*>(ws_piv.p),
                             L.d.nV, static_cast<int*>(ws_counts.p), stream);
        std::vector<int> all(size_t(L.nn));
        std::iota(all.begin(), all.end(), 0);
        int rc;
        if ((rc = run_closure(lvl, nodes, all, hc[size_t(lvl)]))) return rc;
        launch_project_su(L.d.nV, L.d.ld, L.d.ne, static_cast<const int*>(L.d_ext_idx.p), nodes, L.nn,
                          static_cast<const uint16_t*>(L.DU.p), L.node_stride, static_cast<uint16_t*>(L.SU.p), nullptr,
                          stream);
    }
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "build launch");
    return HTG_OK;
}

int Map::build() {
    hipError_t e;
    int rc;
    timing_now = prof_ref().take();
    if (broken) {
        // a failed update left the device passability and live lists behind
        // the host's cell state: bring them level before rebuilding
        if (live_pending) {
            if ((e = hipEventSynchronize(live_evt)) != hipSuccess) return hip_fail(e, "live upload wait");
            live_pending = false;
        }
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "recovery sync");
        if ((e = hipMemcpyAsync(d_pass.p, pass.data(), pass.size(), hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return hip_fail(e, "recovery: upload grid");
        if ((rc = refresh_live(nullptr))) return rc;
    }
    if (fault_hit(HTG_FAULT_BUILD)) return fail(HTG_E_HIP, "injected fault (build)");
    std::vector<std::vector<int>> hc(size_t(H) + 1);
    std::vector<int> sig;
    int ws_lvl = 0;
    size_t ws_need = 0;
    for (int lvl = 1; lvl <= H; ++lvl) {
        const LevelGeo& L = lv[size_t(lvl) - 1];
        const size_t need = size_t(L.nn) * L.d.nV;
        if (need > ws_need) ws_need = need, ws_lvl = lvl;
        hc[size_t(lvl)] = build_pivot_counts(lvl);
        int mx = 0, m8 = 0;
        for (int W = 0; W < L.nn; ++W) {
            mx = std::max(mx, hc[size_t(lvl)][size_t(W)]);
            if (hc[size_t(lvl)][size_t(W)] > 0) m8 = std::max(m8, L.h_live8_cnt[W]);
        }
        sig.push_back(mx);
        sig.push_back(m8);
    }
    sig.push_back(tune.pivot_block);
    for (int lvl = 1; lvl <= H; ++lvl)  // every level's workspace before any launch (and any capture)
        if ((rc = ensure_build_ws(lvl))) return rc;
    (void)ws_lvl;
    if (!tune.use_graph) {
        if ((rc = enqueue_build(hc))) return rc;
    } else {
        if (build_graph && sig != build_sig) {
            (void)hipGraphExecDestroy(build_graph);
            build_graph = nullptr;
        }
        if (!build_graph) {
            hipGraph_t graph = nullptr;
            if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
                return hip_fail(e, "begin capture");
            capturing = true;
            rc = enqueue_build(hc);
            capturing = false;
            e = hipStreamEndCapture(stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(e, "end capture");
            e = hipGraphInstantiate(&build_graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                build_graph = nullptr;
                return hip_fail(e, "graph instantiate");
            }
            build_sig = sig;
        }
        if ((e = hipGraphLaunch(build_graph, stream)) != hipSuccess) return hip_fail(e, "graph launch");
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "build sync");
    prof_ref().harvest();
    built = true;
    root_full = false;  // the build closes the root over its maintained octets only
    broken = false;
    epoch++;
    reset_versions();
    return HTG_OK;
}

void Map::reset_versions() {
    gen++;
    clock++;
    leaf_ver.assign(size_t(nl) * nl, 0);
    node_ver.assign(size_t(H), {});
    for (int l = 0; l < H; ++l) node_ver[size_t(l)].assign(size_t(lv[size_t(l)].nn), 0);
}

void Map::stamp(const std::vector<int>& leaves, const std::vector<std::vector<int>>& nodes, int first_level) {
    if (leaf_ver.size() != size_t(nl) * nl || node_ver.size() != size_t(H)) reset_versions();  // never built
    ++clock;
    for (int x : leaves) leaf_ver[size_t(x)] = clock;
    for (size_t i = size_t(first_level); i < nodes.size(); ++i) {
        const size_t l = i - size_t(first_level);  // level l + 1
        if (l >= node_ver.size()) break;
        for (int w : nodes[i]) node_ver[l][size_t(w)] = clock;
    }
}

uint64_t Map::chain_version(int r, int c) const {
    int tx = c / b, ty = r / b;
    uint64_t v = leaf_ver[size_t(ty) * nl + tx];
    for (int l = 1; l <= H; ++l) {
        tx /= g;
        ty /= g;
        v = std::max(v, node_ver[size_t(l) - 1][size_t(ty) * lv[size_t(l) - 1].d.nt + tx]);
    }
    return v;
}

int Map::query_incremental(htg_inc& st, const int32_t* rc, int k, uint16_t* out, int32_t* n_recomputed) {
    if (n_recomputed) *n_recomputed = 0;
    if (k == 0) {
        st.valid = false;
        return HTG_OK;
    }
    bool fresh = st.valid && st.map_uid == uid && st.gen == gen && st.rc.size() == size_t(2 * k) &&
                 std::memcmp(st.rc.data(), rc, size_t(2 * k) * 4) == 0;
    for (int i = 0; fresh && i < k; ++i) {
        const int r = rc[2 * i], c = rc[2 * i + 1];
        if (r < 0 || c < 0 || r >= S || c >= S) continue;  // outside the hierarchy: INF for good
        if (chain_version(r, c) > st.clock) fresh = false;
    }
    if (!fresh) {
        st.valid = false;
        st.m.resize(size_t(k) * k);
        const int32_t ks = k;
        const int e = query(rc, &ks, 1, st.m.data(), qws);
        if (e) return e;
        st.rc.assign(rc, rc + 2 * k);
        st.map_uid = uid;
        st.gen = gen;
        st.clock = clock;
        st.valid = true;
        if (n_recomputed) *n_recomputed = k;
    }
    std::memcpy(out, st.m.data(), size_t(k) * k * 2);
    return HTG_OK;
}

static int debug_level() {
    static const int lvl = [] {
        const char* e = std::getenv("HTG_DEBUG");
        return e ? std::atoi(e) : 0;
    }();
    return lvl;
}

// ---------------------------------------------------------------------------
// Incremental reveal (src/minplus_oracle.cpp:530-729).  Dirty leaves are
// recomputed by the leaf BFS kernel (the reference's pivot relaxation
// computes the same tile closure, src/leaf_tile.cpp:297-374); each dirty
// ancestor takes min(D_U block, new child summary) plus new seam edges and
// closes over the rows that decreased (src/internal_node.cpp:332-413).  By
// default the closure runs over a vertex cover C of the decreased entries
// (ports opened by this batch, plus both endpoints of decreases between old
// ports) instead of every decreased row: the cover rows are first relaxed
// through the uncovered endpoints B (k_cover_relax, k_cover_pairs), which
// leaves the same closed table with far fewer pivots.
int Map::open_cells(const int32_t* rcs, int n) {
    using clk = std::chrono::steady_clock;
    const auto u0 = clk::now();
    double sync_s = 0;
    auto sync = [&]() {
        const auto a = clk::now();
        const hipError_t r = hipStreamSynchronize(stream);
        sync_s += std::chrono::duration<double>(clk::now() - a).count();
        return r;
    };
    timing_now = prof_ref().take();
    stats = htg_batch_stats{0, 0, 0};
    dirty.clear();
    dirty.nodes.resize(H);
    if (n <= 0) return HTG_OK;
    hipError_t e;
    int rc;
    const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
    auto passable = [&](int r, int c) { return r >= 0 && c >= 0 && r < rows && c < cols && pass[size_t(r) * S + c]; };
    // 1. leaves touched by the batch (ascending); cells that become passable
    std::vector<int> dirty;
    std::vector<int2> newly;  // cells that became passable
    std::vector<int2> cells;
    cells.reserve(size_t(n));
    dirty.reserve(size_t(n));
    // leaf_idx: per leaf, its position in dirty + 1 (0 = not dirty); all zero
    // between updates
    if (leaf_idx.size() != size_t(nl) * nl) leaf_idx.assign(size_t(nl) * nl, 0);
    for (int i = 0; i < n; ++i) {
        const int r = rcs[2 * i], c = rcs[2 * i + 1];
        if (r < 0 || c < 0 || r >= S || c >= S) continue;
        cells.push_back(make_int2(r, c));
        const int leaf = (r / b) * nl + c / b;
        if (!leaf_idx[size_t(leaf)]) {
            leaf_idx[size_t(leaf)] = 1;
            dirty.push_back(leaf);
        }
    }
    std::sort(dirty.begin(), dirty.end());
    for (size_t i = 0; i < dirty.size(); ++i) leaf_idx[size_t(dirty[i])] = int(i) + 1;
    struct ClearLeafIdx {  // every return path leaves leaf_idx all zero
        std::vector<int>& idx;
        const std::vector<int>& d;
        ~ClearLeafIdx() {
            for (int leaf : d) idx[size_t(leaf)] = 0;
        }
    } clear_leaf_idx{leaf_idx, dirty};
    if (dirty.empty()) return HTG_OK;
    for (const int2& p : cells) {
        if (p.x >= rows || p.y >= cols) continue;  // padding cells stay impassable
        const size_t key = size_t(p.x) * S + p.y;
        if (!pass[key]) {
            pass[key] = 1;
            cells_code[size_t(p.x) * cols + p.y] = 1;
            newly.push_back(p);
        }
    }
    // Only cells on a leaf's border can be ports at some level or end a seam
    // edge (every tile above the leaves is a union of leaves).
    auto on_leaf_border = [&](const int2& p) {
        const int r = p.x % b, c = p.y % b;
        return r == 0 || c == 0 || r == b - 1 || c == b - 1;
    };
    std::vector<int2> bnew, bcells;
    for (const int2& p : newly)
        if (on_leaf_border(p)) bnew.push_back(p);
    for (const int2& p : cells)
        if (on_leaf_border(p)) bcells.push_back(p);
    if (!bnew.empty() && (rc = refresh_live(&bnew))) return rc;
    const auto u0b = clk::now();
    const int nd = int(dirty.size());
    // 5. dirty leaves: a leaf with at most kLeafRelaxMax new cells is revealed
    // incrementally (BFS rows of its new cells, then the rank-k relaxation of
    // its table through them), one with more is recomputed whole; a leaf
    // whose cells were all passable already keeps its tables
    const int nsq = b * b;
    const bool relax_ok = b >= 8 && nsq % 8 == 0;
    std::vector<int> full_leaves, full_idx, pcell;
    std::vector<int4> rleaves, pitems;
    {
        // new cells bucketed by dirty leaf (counting sort), in the order seen
        std::vector<int> start(size_t(nd) + 1, 0), cell_of(newly.size());
        for (const int2& p : newly) start[size_t(leaf_idx[size_t((p.x / b) * nl + p.y / b)])]++;
        for (int i = 0; i < nd; ++i) start[size_t(i) + 1] += start[size_t(i)];
        {
            std::vector<int> at(start.begin(), start.end() - 1);
            for (const int2& p : newly)
                cell_of[size_t(at[size_t(leaf_idx[size_t((p.x / b) * nl + p.y / b)] - 1)]++)] =
                    (p.x % b) * b + p.y % b;
        }
        const int swg = leaf_sources_per_wg(b);
        for (int i = 0; i < nd; ++i) {
            const int k = start[size_t(i) + 1] - start[size_t(i)];
            if (k == 0) continue;
            if (!relax_ok || k > kLeafRelaxMax) {
                full_leaves.push_back(dirty[size_t(i)]);
                full_idx.push_back(i);
                continue;
            }
            const int off = int(pcell.size());
            rleaves.push_back(make_int4(dirty[size_t(i)], off, k, i));
            for (int c0 = 0; c0 < k; c0 += swg)
                pitems.push_back(make_int4(dirty[size_t(i)], off + c0, std::min(swg, k - c0), i));
            pcell.insert(pcell.end(), cell_of.begin() + start[size_t(i)], cell_of.begin() + start[size_t(i) + 1]);
        }
    }
    if (!pcell.empty() && (rc = ws_leafP.ensure(pcell.size() * size_t(nsq) * 2))) return rc;
    // The leaf phase goes first, with lists of its own: its kernels run while
    // the host plans the levels.  Changed flags: leaves, then each level's
    // candidates (at most every node of the level).
    size_t nchg_max = size_t(nd);
    for (const LevelGeo& L : lv) nchg_max += size_t(L.nn);
    Packer lpk;
    const size_t o_cells = lpk.add(newly), o_full = lpk.add(full_leaves), o_fidx = lpk.add(full_idx),
                 o_rl = lpk.add(rleaves), o_pit = lpk.add(pitems), o_pcell = lpk.add(pcell);
    if ((rc = ws_leafl_h.ensure(lpk.size)) || (rc = ws_leafl.ensure(lpk.size)) ||
        (rc = ws_chg.ensure(nchg_max * 4 + 16)) || (rc = ws_chg_h.ensure(nchg_max * 4 + 16)))
        return rc;
    lpk.write(static_cast<char*>(ws_leafl_h.p));
    if ((e = hipMemcpyAsync(ws_leafl.p, ws_leafl_h.p, lpk.size, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemsetAsync(ws_chg.p, 0, nchg_max * 4, stream)) != hipSuccess)
        return hip_fail(e, "upload leaf lists");
    const char* lup = static_cast<const char*>(ws_leafl.p);
    int* chg = static_cast<int*>(ws_chg.p);
    if (!newly.empty())
        launch_set_pass(static_cast<uint8_t*>(d_pass.p), S, reinterpret_cast<const int2*>(lup + o_cells),
                        int(newly.size()), stream);
    const double leaf_bytes = (double(b) * b * b * b + double(4 * b - 4) * (4 * b - 4)) * 2 + double(b) * b;
    {
        KTimer lkt;
        int lpi = prof_open(HTG_K_LEAF_APSP, leaf_bytes * double(full_leaves.size()), 0, lkt);
        bool lla = launch_leaf_apsp(static_cast<const uint8_t*>(d_pass.p), S, b, nl,
                                    reinterpret_cast<const int*>(lup + o_full), int(full_leaves.size()),
                                    static_cast<uint16_t*>(leafD.p), static_cast<uint16_t*>(leafS.p), chg, stream, lkt,
                                    kBackgroundWG, reinterpret_cast<const int*>(lup + o_fidx));
        prof_ref().close(lpi, lla);
        // pivot rows: D row + P row per new cell; relaxation: D read and
        // written once, P rows of the tile's rows and columns per pivot
        const double np = double(pcell.size()), nsqd = double(nsq);
        lpi = prof_open(HTG_K_LEAF_APSP, np * nsqd * 4 + double(b) * b * double(pitems.size()), 0, lkt);
        lla = launch_leaf_pivot_rows(static_cast<const uint8_t*>(d_pass.p), S, b, nl,
                                     reinterpret_cast<const int4*>(lup + o_pit), int(pitems.size()),
                                     reinterpret_cast<const int*>(lup + o_pcell), static_cast<uint16_t*>(leafD.p),
                                     static_cast<uint16_t*>(ws_leafP.p), static_cast<uint16_t*>(leafS.p), chg, stream,
                                     lkt, kBackgroundWG);
        prof_ref().close(lpi, lla);
        double upd = 0;
        for (const int4& r : rleaves) upd += double(r.z) * nsqd * nsqd;
        lpi = prof_open(HTG_K_LEAF_APSP, double(rleaves.size()) * nsqd * nsqd * 4 + np * nsqd * 2 * 3, upd, lkt);
        lla = launch_leaf_relax(b, reinterpret_cast<const int4*>(lup + o_rl), int(rleaves.size()),
                                static_cast<const uint16_t*>(ws_leafP.p), static_cast<uint16_t*>(leafD.p),
                                static_cast<uint16_t*>(leafS.p), chg, stream, lkt);
        prof_ref().close(lpi, lla);
    }
    const auto u1 = clk::now();
    // 2. leaf pivot count, as the reference's leaf update would collect it:
    // the endpoints of intra-leaf edges with an endpoint opened by this batch
    int piv = 0;
    {
        if (cell_mark.size() != size_t(S) * S) cell_mark.assign(size_t(S) * S, 0);
        std::vector<size_t> marked;  // distinct endpoints, cleared below
        marked.reserve(newly.size() * 2);
        auto add = [&](size_t key) {
            if (!cell_mark[key]) {
                cell_mark[key] = 1;
                marked.push_back(key);
            }
        };
        for (const int2& p : newly) {
            const int R0 = (p.x / b) * b, C0 = (p.y / b) * b;
            for (int k = 0; k < 4; ++k) {
                const int vr = p.x + dr[k], vc = p.y + dc[k];
                if (vr < R0 || vr >= R0 + b || vc < C0 || vc >= C0 + b || !passable(vr, vc)) continue;
                add(size_t(p.x) * S + p.y);
                add(size_t(vr) * S + vc);
            }
        }
        piv = int(marked.size());
        for (size_t key : marked) cell_mark[key] = 0;
    }
    const auto p1 = clk::now();
    // 3. seam edges by LCA level and node (src/minplus_oracle.cpp:563-598),
    // as (level, node, u, v) sorted and unique
    struct SeamEdge {
        int lvl, node, u, v;
        bool operator<(const SeamEdge& o) const {
            return lvl != o.lvl ? lvl < o.lvl : node != o.node ? node < o.node : u != o.u ? u < o.u : v < o.v;
        }
        bool operator==(const SeamEdge& o) const { return lvl == o.lvl && node == o.node && u == o.u && v == o.v; }
    };
    std::vector<SeamEdge> seams;
    for (const int2& p : bcells) {
        if (!passable(p.x, p.y)) continue;
        for (int k = 0; k < 4; ++k) {
            const int nr = p.x + dr[k], nc = p.y + dc[k];
            if (!passable(nr, nc)) continue;
            int ax = p.y / b, ay = p.x / b, bx = nc / b, by = nr / b;
            if (ax == bx && ay == by) continue;
            int lca = 0;
            for (int l = 1; l <= H; ++l) {
                ax /= g; ay /= g; bx /= g; by /= g;
                if (ax == bx && ay == by) { lca = l; break; }
            }
            if (lca < 1) continue;
            const LevelGeo& L = lv[lca - 1];
            const int node = ay * L.d.nt + ax;
            const int R0 = ay * L.d.s, C0 = ax * L.d.s;
            auto vert = [&](int r, int c) {
                r -= R0;
                c -= C0;
                const int q = (r / L.d.cs) * g + c / L.d.cs;
                return q * L.d.pc + port_index(r % L.d.cs, c % L.d.cs, L.d.cs);
            };
            int u = vert(p.x, p.y), v = vert(nr, nc);
            if (u > v) std::swap(u, v);
            seams.push_back({lca, node, u, v});
        }
    }
    std::sort(seams.begin(), seams.end());
    seams.erase(std::unique(seams.begin(), seams.end()), seams.end());
    const auto p2 = clk::now();
    // 4. candidate nodes per level: every ancestor of a dirty leaf (plus, to
    // be safe, every node with new seam edges).  The child masks of the
    // reference's routing (a parent is dirty when a child summary changed,
    // src/minplus_oracle.cpp:650-722) are resolved on the device from the
    // changed flags of the level below, so a level needs one host sync.
    this->dirty.leaves = dirty;
    this->dirty.cells = newly;
    const int g2 = g * g;
    std::vector<std::vector<int>> cand(size_t(H) + 1), cslot(size_t(H) + 1);
    std::vector<std::vector<int4>> edges(size_t(H) + 1);
    {
        std::vector<int> below = dirty;  // ascending, as every candidate list
        int below_nt = nl;
        size_t se = 0;  // seam edges of this level start here
        for (int lvl = 1; lvl <= H; ++lvl) {
            const LevelGeo& L = lv[lvl - 1];
            size_t se_end = se;
            while (se_end < seams.size() && seams[se_end].lvl == lvl) ++se_end;
            std::vector<int>& c = cand[size_t(lvl)];
            c.reserve(below.size() + (se_end - se));
            for (int ch : below) c.push_back(((ch / below_nt) / g) * L.d.nt + (ch % below_nt) / g);
            for (size_t k = se; k < se_end; ++k) c.push_back(seams[k].node);
            std::sort(c.begin(), c.end());
            c.erase(std::unique(c.begin(), c.end()), c.end());
            std::vector<int>& cs = cslot[size_t(lvl)];
            cs.assign(c.size() * size_t(g2), -1);
            size_t k = se;
            for (size_t i = 0; i < c.size(); ++i) {
                const int tx = c[i] % L.d.nt, ty = c[i] / L.d.nt;
                for (int q = 0; q < g2; ++q) {
                    const int child = (ty * g + q / g) * below_nt + tx * g + q % g;
                    auto it = std::lower_bound(below.begin(), below.end(), child);
                    if (it != below.end() && *it == child) cs[i * size_t(g2) + size_t(q)] = int(it - below.begin());
                }
                while (k < se_end && seams[k].node < c[i]) ++k;
                for (; k < se_end && seams[k].node == c[i]; ++k)
                    edges[size_t(lvl)].push_back(make_int4(int(i), seams[k].u, seams[k].v, 1));
            }
            below.swap(c);
            c = below;
            below_nt = L.d.nt;
            se = se_end;
        }
    }
    const auto p3 = clk::now();
    // new ports per level as (candidate slot, V_U index): the cover of the
    // decreased entries is built around them
    const bool cover = tune.update_pivots == 0;
    std::vector<std::vector<int2>> newp(size_t(H) + 1);
    if (cover) {
        for (const int2& p : bnew) {
            for (int lvl = 1; lvl <= H; ++lvl) {
                const LevelGeo& L = lv[lvl - 1];
                const int ty = p.x / L.d.s, tx = p.y / L.d.s;
                const std::vector<int>& c = cand[size_t(lvl)];
                auto it = std::lower_bound(c.begin(), c.end(), ty * L.d.nt + tx);
                if (it == c.end() || *it != ty * L.d.nt + tx) continue;
                const int r = p.x - ty * L.d.s, cc = p.y - tx * L.d.s;
                const int pi = port_index(r % L.d.cs, cc % L.d.cs, L.d.cs);
                if (pi < 0) continue;
                const int q = (r / L.d.cs) * g + cc / L.d.cs;
                newp[size_t(lvl)].push_back(make_int2(int(it - c.begin()), q * L.d.pc + pi));
            }
        }
    }
    const auto u2 = clk::now();
    // one packed upload of the level lists; changed flags: leaves, then each
    // level's candidates
    Packer pk;
    std::vector<size_t> o_cand(size_t(H) + 1), o_cslot(size_t(H) + 1), o_edges(size_t(H) + 1), o_chg(size_t(H) + 1),
        o_newp(size_t(H) + 1);
    size_t nchg = size_t(nd);
    for (int lvl = 1; lvl <= H; ++lvl) {
        o_cand[size_t(lvl)] = pk.add(cand[size_t(lvl)]);
        o_cslot[size_t(lvl)] = pk.add(cslot[size_t(lvl)]);
        o_edges[size_t(lvl)] = pk.add(edges[size_t(lvl)]);
        o_newp[size_t(lvl)] = pk.add(newp[size_t(lvl)]);
        o_chg[size_t(lvl)] = nchg;
        nchg += cand[size_t(lvl)].size();
    }
    if ((rc = ws_upd_h.ensure(pk.size)) || (rc = ws_upd.ensure(pk.size))) return rc;
    pk.write(static_cast<char*>(ws_upd_h.p));
    if ((e = hipMemcpyAsync(ws_upd.p, ws_upd_h.p, pk.size, hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hip_fail(e, "upload update lists");
    const char* up = static_cast<const char*>(ws_upd.p);
    stats.n_touched_leaves = nd;
    stats.n_pivots = piv;
    if (H == 0) {
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "update sync");
        stamp(this->dirty.leaves, this->dirty.nodes, 0);
        return HTG_OK;  // the reference returns before bumping epoch here
    }
    bool below_any = true;  // the level below may have changed flags set
    const int* chg_below = chg;
    int top = 0;        // highest level processed
    int top_first = 0;  // lowest level processed
    for (int lvl = 1; lvl <= H; ++lvl) {
        LevelGeo& L = lv[lvl - 1];
        const std::vector<int>& cl = cand[size_t(lvl)];
        const int na = int(cl.size());
        int* chg_here = chg + o_chg[size_t(lvl)];
        if (!below_any && edges[size_t(lvl)].empty()) {  // nothing can change at this level
            bool higher = false;
            for (int hl = lvl + 1; hl <= H && !higher; ++hl) higher = !edges[size_t(hl)].empty();
            if (!higher) break;
            chg_below = chg_here;  // all clear
            continue;
        }
        top = lvl;
        if (top_first == 0) top_first = lvl;
        if (L.seam_only) root_full = false;
        if ((rc = ensure_build_ws(lvl))) return rc;
        const int* dnodes = reinterpret_cast<const int*>(up + o_cand[size_t(lvl)]);
        int max8 = 0;
        for (int node : cl) max8 = std::max(max8, L.h_live8_cnt[node]);
        const size_t plane = cover ? size_t(na) * L.d.nV : 0;
        uint8_t* flags = static_cast<uint8_t*>(ws_flags.p);
        if ((e = hipMemsetAsync(flags, 0, cover ? 3 * plane : size_t(na) * L.d.nV, stream)) != hipSuccess)
            return hip_fail(e, "clear flags");
        if (cover)
            launch_mark_u8(reinterpret_cast<const int2*>(up + o_newp[size_t(lvl)]), int(newp[size_t(lvl)].size()),
                           L.d.nV, flags + 2 * plane, stream);
        size_t cstride;
        const uint16_t* csum = child_summaries(lvl, cstride);
        launch_child_relax_live(g, L.d.pc, L.d.nV, L.d.ld, L.d.nt, csum, cstride, dnodes,
                                reinterpret_cast<const int*>(up + o_cslot[size_t(lvl)]), chg_below, na, L.d_live8_cnt,
                                L.d_live8_list, L.ng8, max8, static_cast<uint16_t*>(L.DU.p), L.node_stride, flags,
                                plane, stream);
        if (!edges[size_t(lvl)].empty())
            launch_edge_relax(L.d.nV, L.d.ld, reinterpret_cast<const int4*>(up + o_edges[size_t(lvl)]),
                              int(edges[size_t(lvl)].size()), dnodes, static_cast<uint16_t*>(L.DU.p), L.node_stride,
                              flags, plane, stream);
        int* cnt = static_cast<int*>(ws_counts.p);
        launch_compact_flags_live(L.d.nV, dnodes, na, L.d_live8_cnt, L.d_live8_list, L.ng8, flags, plane,
                                  static_cast<int*>(ws_piv.p), static_cast<int*>(ws_bpiv.p), L.d.nV, cnt, stream);
        std::vector<int> hc(static_cast<size_t>(cover ? 2 * na : na));
        if ((rc = ws_counts_h.ensure(hc.size() * 4))) return rc;  // pinned: a pageable copy would block
        if ((e = hipMemcpyAsync(ws_counts_h.p, cnt, hc.size() * 4, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return hip_fail(e, "download counts");
        // the rebuild estimate's pivot counts (passable seam endpoints) depend
        // on host state only: counted while the level's kernels run
        const auto r0 = clk::now();
        std::vector<int> pb;
        if (cover && tune.update_rebuild == 0) pb = build_pivot_counts(lvl, &cl);
        upd_time[6] += std::chrono::duration<double>(clk::now() - r0).count();
        if ((e = sync()) != hipSuccess) return hip_fail(e, "update sync");
        std::memcpy(hc.data(), ws_counts_h.p, hc.size() * 4);
        long long sum = 0, bsum = 0;
        int mx = 0, mxb = 0;
        for (int i = 0; i < na; ++i) {
            sum += hc[size_t(i)];
            mx = std::max(mx, hc[size_t(i)]);
            if (cover && hc[size_t(i)] > 0) {
                bsum += hc[size_t(na + i)];
                mxb = std::max(mxb, hc[size_t(na + i)]);
            }
            if (hc[size_t(i)] > 0) this->dirty.nodes[size_t(lvl) - 1].push_back(cl[size_t(i)]);
        }
        if (debug_level())
            std::fprintf(stderr, "[htg] update lvl %d: %d nodes (nV %d), pivots sum %lld max %d, B sum %lld max %d\n",
                         lvl, na, L.d.nV, sum, mx, bsum, mxb);
        // A level whose cover makes the update dearer than rebuilding its
        // candidate nodes from their children (the build's closure over the
        // passable seam endpoints) is rebuilt instead.  After any reveal every
        // table equals a fresh build of the revealed grid, so both paths give
        // the same tables, S_U and changed flags.
        const auto r1 = clk::now();
        const bool reb = sum > 0 && cover && rebuild_is_cheaper(lvl, cl, hc, na, pb);
        upd_time[6] += std::chrono::duration<double>(clk::now() - r1).count();
        if (reb) {
            if ((rc = rebuild_level(lvl, dnodes, cl, chg_here))) return rc;
            this->dirty.nodes[size_t(lvl) - 1] = cl;  // every candidate's table was rewritten
            rebuilt_levels++;
            below_any = true;
            chg_below = chg_here;
            if (lvl == top_first && fault_hit(HTG_FAULT_UPDATE)) return fail(HTG_E_HIP, "injected fault (update)");
            continue;
        }
        if (bsum > 0) {  // fold the uncovered endpoints into the cover rows
            double by_r = 0, upd_r = 0, by_p = 0, upd_p = 0;
            if (prof_ref().on) {
                for (int i = 0; i < na; ++i) {
                    const double c = hc[size_t(i)], bb = hc[size_t(na + i)];
                    if (c <= 0 || bb <= 0) continue;
                    const double nl = 8.0 * L.h_live8_cnt[cl[size_t(i)]];
                    upd_r += c * bb * nl;
                    by_r += (2 * c * nl + bb * nl) * 2;  // cover rows in + out, B rows in
                    upd_p += c * c * bb;
                    by_p += (2 * c * bb + c * c) * 2;
                }
            }
            // work units of both hops from the counts just read, in one upload
            // (the level's sync above has drained the previous level's upload)
            std::vector<int> l8(static_cast<size_t>(na));
            for (int i = 0; i < na; ++i) l8[size_t(i)] = L.h_live8_cnt[cl[size_t(i)]];
            cov_units.clear();
            const int rpt1 = plan_cover_units(false, na, hc.data(), hc.data() + na, l8.data(), cov_units);
            const int n1 = int(cov_units.size());
            const int rpt2 = plan_cover_units(true, na, hc.data(), hc.data() + na, l8.data(), cov_units);
            const int n2 = int(cov_units.size()) - n1;
            const size_t ub = cov_units.size() * sizeof(int4);
            if ((rc = ws_cov_h.ensure(ub)) || (rc = ws_cov.ensure(ub))) return rc;
            std::memcpy(ws_cov_h.p, cov_units.data(), ub);
            if ((e = hipMemcpyAsync(ws_cov.p, ws_cov_h.p, ub, hipMemcpyHostToDevice, stream)) != hipSuccess)
                return hip_fail(e, "upload cover units");
            const int4* units = static_cast<const int4*>(ws_cov.p);
            KTimer kt;
            int pi = prof_open(HTG_K_COVER_RELAX, by_r, upd_r, kt);
            bool la = launch_cover_relax(L.d.nV, L.d.ld, dnodes, na, cnt, static_cast<const int*>(ws_piv.p),
                                         static_cast<const int*>(ws_bpiv.p), L.d.nV, units, n1, rpt1,
                                         L.d_live8_cnt, L.d_live8_list, L.ng8, static_cast<uint16_t*>(L.DU.p),
                                         L.node_stride, stream, kt);
            prof_ref().close(pi, la);
            pi = prof_open(HTG_K_COVER_PAIRS, by_p, upd_p, kt);
            la = launch_cover_pairs(L.d.nV, L.d.ld, dnodes, na, cnt, static_cast<const int*>(ws_piv.p),
                                    static_cast<const int*>(ws_bpiv.p), L.d.nV, units + n1, n2, rpt2,
                                    static_cast<uint16_t*>(L.DU.p), L.node_stride, stream, kt);
            prof_ref().close(pi, la);
        }
        hc.resize(size_t(na));
        if (sum > 0) {
            if ((rc = run_closure(lvl, dnodes, cl, hc))) return rc;
            // S_U refresh for nodes that closed over something (the reference
            // skips nodes with an empty pivot set, src/internal_node.cpp:368-370)
            launch_project_su_live(L.d.nV, L.d.ld, L.d.ne, static_cast<const int*>(L.d_vu_ext.p), dnodes,
                                   static_cast<const int*>(ws_counts.p), na, L.d_live8_cnt, L.d_live8_list, L.ng8,
                                   max8, static_cast<const uint16_t*>(L.DU.p), L.node_stride,
                                   static_cast<uint16_t*>(L.SU.p), chg_here, stream);
            if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "update launch");
        }
        below_any = sum > 0;
        chg_below = chg_here;
        if (lvl == top_first && fault_hit(HTG_FAULT_UPDATE)) return fail(HTG_E_HIP, "injected fault (update)");
    }
    const auto u3 = clk::now();
    // stats: a node is a dirty ancestor when a child summary changed or it
    // got new seam edges (the reference's routing)
    if ((e = hipMemcpyAsync(ws_chg_h.p, ws_chg.p, nchg * 4, hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return hip_fail(e, "download changed");
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "update sync");
    const int* hchg = static_cast<const int*>(ws_chg_h.p);
    int anc = 0;
    for (int lvl = 1; lvl <= top; ++lvl) {
        const std::vector<int>& cl = cand[size_t(lvl)];
        const size_t below_off = lvl == 1 ? 0 : o_chg[size_t(lvl) - 1];
        std::vector<uint8_t> has_edge(cl.size(), 0);
        for (const int4& ed : edges[size_t(lvl)]) has_edge[size_t(ed.x)] = 1;
        for (size_t i = 0; i < cl.size(); ++i) {
            bool act = has_edge[i] != 0;
            for (int q = 0; q < g2 && !act; ++q) {
                const int sl = cslot[size_t(lvl)][i * size_t(g2) + size_t(q)];
                act = sl >= 0 && hchg[below_off + size_t(sl)] != 0;
            }
            anc += act;
        }
    }
    stats.n_dirty_ancestors = anc;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "update sync");
    prof_ref().harvest();
    stamp(this->dirty.leaves, this->dirty.nodes, 0);
    epoch++;
    const auto u4 = clk::now();
    auto dt = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    upd_time[0] += dt(u0, u1);  // cell bookkeeping and live groups, leaf phase
    upd_time[7] += dt(u0b, u1);  // leaf phase
    upd_time[8] += dt(u1, p1);   // leaf pivot count
    upd_time[9] += dt(p1, p2);   // seam edges
    upd_time[10] += dt(p2, p3);  // candidates
    upd_time[11] += dt(p3, u2);  // new ports
    upd_time[1] += dt(u1, u2);  // planning: leaf pivots, seams, candidates
    upd_time[2] += dt(u2, u3) - sync_s;  // level loop, host side
    upd_time[3] += sync_s;               // level loop, waiting for the device
    upd_time[4] += dt(u3, u4);           // changed flags, stats
    upd_count++;
    return HTG_OK;
}

int Map::catch_up_from(const Map& src, const DirtySet& ds) {
    const auto c0 = std::chrono::steady_clock::now();
    struct Acc {
        Map* m;
        std::chrono::steady_clock::time_point t;
        ~Acc() {
            m->upd_time[5] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
            m->catchup_count++;
        }
    } acc{this, c0};
    hipError_t e;
    const size_t n = size_t(b) * b, pl = size_t(4 * b - 4);
    for (const int2& c : ds.cells) {
        pass[size_t(c.x) * S + c.y] = 1;
        cells_code[size_t(c.x) * cols + c.y] = 1;
    }
    if (!ds.cells.empty() &&
        (e = hipMemcpyAsync(d_pass.p, src.d_pass.p, size_t(S) * S, hipMemcpyDeviceToDevice, stream)) != hipSuccess)
        return hip_fail(e, "catch-up pass");
    if (!ds.cells.empty()) {
        const int rc = refresh_live(&ds.cells);
        if (rc) return rc;
    }
    // one packed upload of the leaf and node lists, then one copy launch per
    // table kind and level
    Packer pk;
    const size_t o_leaves = pk.add(ds.leaves);
    std::vector<size_t> o_nodes(ds.nodes.size());
    for (size_t l = 0; l < ds.nodes.size(); ++l) o_nodes[l] = pk.add(ds.nodes[l]);
    int rc;
    if ((rc = ws_cu_h.ensure(pk.size)) || (rc = ws_cu.ensure(pk.size))) return rc;
    pk.write(static_cast<char*>(ws_cu_h.p));
    if ((e = hipMemcpyAsync(ws_cu.p, ws_cu_h.p, pk.size, hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hip_fail(e, "catch-up lists");
    const char* up = static_cast<const char*>(ws_cu.p);
    const int* dleaves = reinterpret_cast<const int*>(up + o_leaves);
    const int nlv = int(ds.leaves.size());
    launch_copy_items(dleaves, nlv, n * n, static_cast<const uint16_t*>(src.leafD.p), static_cast<uint16_t*>(leafD.p),
                      stream);
    launch_copy_items(dleaves, nlv, pl * pl, static_cast<const uint16_t*>(src.leafS.p),
                      static_cast<uint16_t*>(leafS.p), stream);
    for (size_t l = 0; l < ds.nodes.size() && l < lv.size(); ++l) {
        LevelGeo& L = lv[l];
        const LevelGeo& Ls = src.lv[l];
        int max8 = 0;
        for (int node : ds.nodes[l]) max8 = std::max(max8, L.h_live8_cnt[node]);
        launch_copy_live_rows(reinterpret_cast<const int*>(up + o_nodes[l]), int(ds.nodes[l].size()), L.d.nV, L.d.ld,
                              L.node_stride, L.d_live8_cnt, L.d_live8_list, L.ng8, max8, L.d.ne,
                              static_cast<const int*>(L.d_vu_ext.p), static_cast<const uint16_t*>(Ls.DU.p),
                              static_cast<uint16_t*>(L.DU.p), static_cast<const uint16_t*>(Ls.SU.p),
                              static_cast<uint16_t*>(L.SU.p), stream);
    }
    // No host wait: the update that the catch-up prepares runs behind the
    // copies on the map's stream, the map is not queried before that update
    // has finished (its own final wait), and the next catch-up into ws_cu_h
    // comes after that wait too.
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "catch-up copy");
    stats = src.stats;
    epoch = src.epoch;
    stamp(ds.leaves, ds.nodes, 0);
    return HTG_OK;
}

int Map::copy_all_from(const Map& src) {
    if (src.S != S || src.H != H || src.b != b || src.g != g || src.rows != rows || src.cols != cols)
        return fail(HTG_E_INVALID, "copy between maps of different geometry");
    hipError_t e;
    if ((e = hipStreamSynchronize(src.stream)) != hipSuccess || (e = hipStreamSynchronize(stream)) != hipSuccess)
        return hip_fail(e, "full copy sync");
    pass = src.pass;
    cells_code = src.cells_code;
    const auto d2d = hipMemcpyDeviceToDevice;
    if ((e = hipMemcpyAsync(d_pass.p, src.d_pass.p, size_t(S) * S, d2d, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(leafD.p, src.leafD.p, leafD.bytes, d2d, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(leafS.p, src.leafS.p, leafS.bytes, d2d, stream)) != hipSuccess)
        return hip_fail(e, "full copy");
    for (size_t l = 0; l < lv.size(); ++l) {
        const size_t du = size_t(lv[l].nn) * lv[l].node_stride * 2, su = size_t(lv[l].nn) * lv[l].d.ne * lv[l].d.ne * 2;
        if ((e = hipMemcpyAsync(lv[l].DU.p, src.lv[l].DU.p, du, d2d, stream)) != hipSuccess ||
            (e = hipMemcpyAsync(lv[l].SU.p, src.lv[l].SU.p, su, d2d, stream)) != hipSuccess)
            return hip_fail(e, "full copy");
    }
    int rc;
    if ((rc = refresh_live(nullptr))) return rc;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "full copy sync");
    stats = src.stats;
    epoch = src.epoch;
    built = src.built;
    root_full = src.root_full;
    broken = false;
    reset_versions();
    return HTG_OK;
}

// ---------------------------------------------------------------------------
// Batched all-pairs query.  Host planning: viewpoints are sorted by their
// tile key (base g^2 digits from the root down), so at every level the
// viewpoints that share (W, q) and those that share W are contiguous runs.
int Map::query(const int32_t* rcs, const int32_t* ks, int nq, uint16_t* out_host, QueryWS& ws) {
    int rc = query_launch(rcs, ks, nq, ws, out_host != nullptr);
    if (rc) return rc;
    return query_finish(ws, out_host);
}

// liftPoint (src/minplus_oracle.cpp:731-823).  Level 0 is the point's row of
// its leaf table at the leaf's ports; level l is the vector-matrix product of
// level l-1 with the rows of the point's child in the ancestor's D_U, read at
// the ancestor's exterior ports (the reference's child_to_ext submatrix,
// src/internal_node.cpp:435-455).  One stream, synchronous.
int Map::lift_point(int r, int c, uint16_t* x, long long x_cap, int32_t* lens, int32_t* parent, int32_t* quad,
                    int32_t* offset) {
    if (r < 0 || c < 0 || r >= S || c >= S) return fail(HTG_E_RANGE, "point outside the hierarchy");
    if (int mr = materialize_root()) return mr;  // the root's exterior rows
    std::vector<long long> off(size_t(H) + 2, 0);
    lens[0] = 4 * b - 4;
    for (int l = 1; l <= H; ++l) lens[l] = lv[size_t(l) - 1].d.ne;
    for (int l = 0; l <= H; ++l) off[size_t(l) + 1] = off[size_t(l)] + lens[l];
    const long long total = off[size_t(H) + 1];
    if (x_cap < total) return fail(HTG_E_RANGE, "lift buffer too small");
    parent[0] = quad[0] = -1;
    offset[0] = 0;
    int tx = c / b, ty = r / b;
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[size_t(l) - 1];
        parent[l] = (ty / g) * L.d.nt + tx / g;
        quad[l] = (ty % g) * g + tx % g;
        offset[l] = quad[l] * L.d.pc;
        tx /= g;
        ty /= g;
    }
    // row-chunk scratch: enough workgroups to spread the rows over the chip
    std::vector<int> rpc_l(size_t(H) + 1, 0), nch_l(size_t(H) + 1, 0);
    size_t part = 8;
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[size_t(l) - 1];
        const int col_blocks = (L.d.ld / 2 + 255) / 256;
        int nch = std::max(1, std::min(512 / col_blocks, (L.d.pc + 15) / 16));
        const int rpc = std::min(kLiftRowsPerChunk, (L.d.pc + nch - 1) / nch);
        nch = (L.d.pc + rpc - 1) / rpc;
        rpc_l[size_t(l)] = rpc;
        nch_l[size_t(l)] = nch;
        part = std::max(part, size_t(nch) * L.d.ld);
    }
    int rc;
    if ((rc = lift_x.ensure(size_t(total) * 2 + 16)) || (rc = lift_part.ensure(part * 2)) ||
        (rc = lift_h.ensure(size_t(total) * 2 + 16)))
        return rc;
    uint16_t* X = static_cast<uint16_t*>(lift_x.p);
    const size_t n = size_t(b) * b;
    const int leaf = (r / b) * nl + c / b, u = (r % b) * b + c % b;
    launch_lift_leaf(static_cast<const uint16_t*>(leafD.p) + size_t(leaf) * n * n + size_t(u) * n, b, X, stream);
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[size_t(l) - 1];
        launch_lift_level(static_cast<const uint16_t*>(L.DU.p) + size_t(parent[l]) * L.node_stride, L.d.ld,
                          quad[l] * L.d.pc, L.d.pc, rpc_l[size_t(l)], nch_l[size_t(l)], X + off[size_t(l) - 1],
                          static_cast<uint16_t*>(lift_part.p), static_cast<const int*>(L.d_ext_child.p),
                          static_cast<const int*>(L.d_ext_cport.p), L.d.ne, X + off[size_t(l)], stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "lift launch");
    if ((e = hipMemcpyAsync(lift_h.p, X, size_t(total) * 2, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return hip_fail(e, "lift download");
    std::memcpy(x, lift_h.p, size_t(total) * 2);
    return HTG_OK;
}

int Map::query_finish(QueryWS& ws, uint16_t* out_host) {
    hipError_t e = hipEventSynchronize(ws.done);
    if (e != hipSuccess) return hip_fail(e, "query sync");
    if (out_host && ws.total_out) std::memcpy(out_host, ws.h_out.p, size_t(ws.total_out) * 2);
    if (ws.counted) {
        const unsigned long long* w = static_cast<const unsigned long long*>(ws.h_wstat.p);
        double by = 0, up = 0;
        for (int i = 0; i < kWorkSlots; ++i) by += double(w[2 * i]), up += double(w[2 * i + 1]);
        prof_ref().add_work(HTG_K_Q_PROJECT, by, up);
        ws.counted = false;
    }
    prof_ref().harvest();
    return HTG_OK;
}

// Host planning of one query batch into P, with the packed plan and the
// sizes of its device regions in B.  Both are reused from call to call
// (their vectors keep their capacity), so planning allocates nothing once
// they have grown to the batch sizes in use.  host_sub (nullable, 4 entries)
// accumulates the planning sub-phase times.
int Map::plan_batch(const int32_t* rcs, const int32_t* ks, int nq, QueryPlan& P, BatchLayout& B, bool counting,
                    double* host_sub) const {
    const auto th0 = std::chrono::steady_clock::now();
    long long total_vp = 0, total_out = 0;
    int kmax = 0;
    for (int q = 0; q < nq; ++q) {
        if (ks[q] < 0) return fail(HTG_E_INVALID, "negative k");
        total_vp += ks[q];
        total_out += (long long)ks[q] * ks[q];
        kmax = std::max(kmax, ks[q]);
    }
    B.total_vp = total_vp;
    B.total_out = total_out;
    if (total_out == 0) return HTG_OK;
    // ---- plan
    P.clear();
    P.q_k.assign(ks, ks + nq);
    P.q_out.resize(nq);
    P.q_vp.resize(nq);
    long long oo = 0, vv = 0;
    for (int q = 0; q < nq; ++q) {
        P.q_out[q] = oo;
        P.q_vp[q] = int(vv);
        oo += (long long)ks[q] * ks[q];
        vv += ks[q];
    }
    P.vp_rc.resize(total_vp);
    P.vp_valid.resize(total_vp);
    // Slot order: viewpoints sorted by (tile key, viewpoint index); the key's
    // base-g^2 digits run from the root down, so the viewpoints sharing
    // (W, q) at any level form contiguous runs.
    std::vector<SlotKey>& sk = P.sk;
    sk.clear();
    sk.reserve(size_t(total_vp));
    // b and g powers of two (the usual case): the key's digits are bit fields
    int lb = 0, lg = 0;
    while ((1 << lb) < b) ++lb;
    while ((1 << lg) < g) ++lg;
    const bool shifts = (1 << lb) == b && (1 << lg) == g;
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < ks[q]; ++i) {
            const int v = P.q_vp[q] + i;
            const int r = rcs[2 * v], c = rcs[2 * v + 1];
            P.vp_rc[v] = make_int2(r, c);
            const bool ok = r >= 0 && c >= 0 && r < rows && c < cols && pass[size_t(r) * S + c];
            P.vp_valid[v] = ok;
            if (!ok) continue;
            unsigned long long key = 0;
            if (shifts) {
                const unsigned gm = unsigned(g - 1);
                for (int l = H; l >= 1; --l) {  // child side of level l: b * g^(l-1)
                    const int sh = lb + lg * (l - 1);
                    key = (key << (2 * lg)) | ((((unsigned(r) >> sh) & gm) << lg) | ((unsigned(c) >> sh) & gm));
                }
            } else {
                for (int l = H; l >= 1; --l) {
                    const int cs = lv[l - 1].d.cs;
                    key = key * unsigned(g * g) + unsigned(((r / cs) % g) * g + (c / cs) % g);
                }
            }
            sk.push_back({key, q, i, v});
        }
    const auto tk0 = std::chrono::steady_clock::now();
    // order by (key, vp): sk is built in vp order, so a stable counting sort on
    // the key does it when the key is short (2 H log2(g) bits)
    const int key_bits = shifts ? 2 * lg * H : 64;
    if (key_bits <= 16 && sk.size() > 32 && (size_t(1) << key_bits) <= 32 * sk.size()) {
        std::vector<int>& cnt = P.key_count;
        cnt.assign((size_t(1) << key_bits) + 1, 0);
        for (const SlotKey& k : sk) cnt[size_t(k.key) + 1]++;
        for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
        std::vector<SlotKey>& tmp = P.sk_tmp;
        tmp.resize(sk.size());
        for (const SlotKey& k : sk) tmp[size_t(cnt[size_t(k.key)]++)] = k;
        sk.swap(tmp);
    } else {
        std::sort(sk.begin(), sk.end(), [](const SlotKey& a, const SlotKey& c) {
            return a.key != c.key ? a.key < c.key : a.vp < c.vp;  // vp orders (q, i)
        });
    }
    const auto tk1 = std::chrono::steady_clock::now();
    double t_meet = 0;
    const int ns = int(sk.size());
    P.slot_rc.resize(ns);
    P.slot_q.resize(ns);
    P.slot_i.resize(ns);
    for (int s = 0; s < ns; ++s) {
        P.slot_rc[s] = P.vp_rc[sk[s].vp];
        P.slot_q[s] = sk[s].q;
        P.slot_i[s] = sk[s].i;
    }
    // Meet order: slots grouped by query, in slot order within a query
    // (counting sort); the (query, W) groups of any level are then
    // contiguous runs, and within a group the target children q2 are too.
    std::vector<int>& by_q = P.by_q;
    by_q.assign(size_t(ns), 0);
    {
        std::vector<int>& qstart = P.qstart;
        qstart.assign(size_t(nq) + 1, 0);
        for (int s2 = 0; s2 < ns; ++s2) qstart[size_t(sk[s2].q) + 1]++;
        for (int q = 0; q < nq; ++q) qstart[size_t(q) + 1] += qstart[size_t(q)];
        for (int s2 = 0; s2 < ns; ++s2) by_q[size_t(qstart[size_t(sk[s2].q)]++)] = s2;
    }
    P.levels.resize(H);
    const int g2 = g * g;
    const size_t nsl = size_t(std::max(ns, 1));
    // Q layout: X_0 .. X_{H-1}, then Z_1 .. Z_H
    std::vector<size_t>& xoff = B.xoff;
    std::vector<size_t>& zoffq = B.zoffq;
    xoff.assign(size_t(H) + 1, 0);
    zoffq.assign(size_t(H) + 1, 0);
    size_t qsz = 0;
    auto a8 = [](size_t x) { return (x + 7) / 8 * 8; };
    for (int l = 0; l < H; ++l) {
        xoff[size_t(l)] = qsz;
        qsz += a8(nsl * size_t(l == 0 ? 4 * b - 4 : lv[size_t(l) - 1].d.ne));
    }
    for (int l = 1; l <= H; ++l) {
        zoffq[size_t(l)] = qsz;
        qsz += a8(nsl * size_t(lv[size_t(l) - 1].ldz));
    }
    size_t zbase = zoffq[1 <= H ? 1 : 0];
    int fill_blocks = 0;
    auto add_fill = [&](size_t start, size_t count) {  // start and count are multiples of 8
        const long long n8 = (long long)(count / 8);
        if (n8 == 0) return;
        P.fills.push_back(make_int4(int(start & 0xffffffffu), int(start >> 32), int(n8), fill_blocks));
        fill_blocks += int((n8 + 255) / 256);
    };
    // per-level slot keys: kc[s] = key / g2^(l-1) identifies (W, q) at level
    // l, kw[s] = kc[s] / g2 identifies W; shifts when g2 is a power of two
    std::vector<unsigned long long>& kc = P.kc;
    std::vector<unsigned long long>& kw = P.kw;
    kc.resize(size_t(ns));
    kw.resize(size_t(ns));
    for (int s2 = 0; s2 < ns; ++s2) kc[size_t(s2)] = sk[size_t(s2)].key;
    const bool pow2 = (g2 & (g2 - 1)) == 0;
    int g2shift = 0;
    while ((1 << g2shift) < g2) ++g2shift;
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[l - 1];
        LevelPlan& LP = P.levels[l - 1];
        for (int s2 = 0; s2 < ns; ++s2)
            kw[size_t(s2)] = pow2 ? kc[size_t(s2)] >> g2shift : kc[size_t(s2)] / unsigned(g2);
        auto qdig = [&](int s2) { return int(kc[size_t(s2)] - kw[size_t(s2)] * unsigned(g2)); };
        // (W, q) runs -> projection items
        int maxrun = 0;
        for (int s = 0; s < ns;) {
            int t = s + 1;
            while (t < ns && kc[size_t(t)] == kc[size_t(s)]) ++t;
            maxrun = std::max(maxrun, t - s);
            s = t;
        }
        int M = 1;
        while (M < maxrun && M < tune.zq_tile) M *= 2;
        LP.M = M;
        long long waves = 0;  // wavefronts per row chunk
        int rows_max = 0;     // passable seam rows of the largest item
        for (int s = 0; s < ns;) {
            int t = s + 1;
            while (t < ns && kc[size_t(t)] == kc[size_t(s)]) ++t;
            const int2 rc0 = P.slot_rc[s];
            const int W = (rc0.x / L.d.s) * L.d.nt + rc0.y / L.d.s;
            const int q = qdig(s);
            const int tiles = (L.h_live_cnt[W] + 63) / 64;
            LP.col_tiles = std::max(LP.col_tiles, tiles);
            rows_max = std::max(rows_max, L.seam_live[size_t(W) * g2 + q]);
            for (int a = s; a < t; a += M) {
                // two int4 per item: (node, child | first flag, first slot, slots) and the node's live
                // group count, the child's seam range and its first seam port (no dependent loads)
                LP.items.push_back(make_int4(W, q | (a == s ? 1 << 16 : 0), a, std::min(M, t - a)));
                const int sl0 = L.seam_off[size_t(q)], sl1 = L.seam_off[size_t(q) + 1];
                LP.items.push_back(make_int4(L.h_live_cnt[W], sl0, sl1, sl1 > sl0 ? L.seam_list[size_t(sl0)] : 0));
                waves += tiles;
            }
            s = t;
        }
        {  // Row chunks: one (fold in registers, no fold launch) while an item
           // has at most kRowsPerChunk passable seam rows; beyond that, enough
           // chunks for that bound and to give the launch ~2048 workgroups.
            constexpr int kRowsPerChunk = 512;  // 4 wavefronts x 128 rows
            int ks = (rows_max + kRowsPerChunk - 1) / kRowsPerChunk;
            if (ks > 1 && waves > 0) ks = std::max(ks, int((2048 + waves - 1) / waves));
            ks = std::min(ks, std::max(1, L.nsmax / 32));
            LP.ksplit = std::max(1, ks);
        }
        // Z region of this level and its layout for the meet
        zbase = zoffq[size_t(l)];
        P.zgeo.push_back(int(zbase & 0xffffffffu));
        P.zgeo.push_back(int(zbase >> 32));
        P.zgeo.push_back(L.ldz);
        P.zgeo.push_back(0);
        P.zgeo.insert(P.zgeo.end(), L.zoff.begin(), L.zoff.end());
        P.zgeo.insert(P.zgeo.end(), L.zlen.begin(), L.zlen.end());
        if (LP.ksplit == 1 && ns > 0) {  // the projection writes live columns only
            add_fill(zoffq[size_t(l)], a8(nsl * size_t(L.ldz)));
            if (l < H) add_fill(xoff[size_t(l)], a8(nsl * size_t(L.d.ne)));
        }
        // (query, W) groups -> meet tiles per target child q2
        const auto tm0 = std::chrono::steady_clock::now();
        for (int a0 = 0; a0 < ns;) {
            const int s0 = by_q[size_t(a0)];
            const int q = sk[size_t(s0)].q;
            const unsigned long long w = kw[size_t(s0)];
            int a1 = a0 + 1;
            while (a1 < ns && sk[size_t(by_q[size_t(a1)])].q == q && kw[size_t(by_q[size_t(a1)])] == w) ++a1;
            const int n = a1 - a0;
            if (n >= 2) {
                const int goff = int(P.glist.size());
                P.glist.insert(P.glist.end(), by_q.begin() + a0, by_q.begin() + a1);
                for (int j0 = 0; j0 < n;) {
                    const int q2 = qdig(by_q[size_t(a0 + j0)]);
                    int je = j0 + 1;
                    while (je < n && qdig(by_q[size_t(a0 + je)]) == q2) ++je;
                    // 16 x 16 tiles when both sides have 16 viewpoints, else 8 x 8
                    const bool big = n >= 16 && je - j0 >= 16;
                    const int T = big ? 16 : 8;
                    for (int i0 = 0; i0 < n; i0 += T)
                        for (int jj = j0; jj < je; jj += T) {
                            const size_t o = P.mtiles.size();
                            P.mtiles.resize(o + 8);
                            int* tt = P.mtiles.data() + o;
                            tt[0] = n, tt[1] = q2, tt[2] = i0, tt[3] = jj, tt[4] = je, tt[5] = (l - 1) | (big << 16);
                            tt[6] = goff, tt[7] = q;
                        }
                    if (counting) {  // the meet's work figures, for kernel timing only
                        // algorithmic: the group's n rows of the q2 seam block read
                        // once (q2's own rows are among them), n x |q2| results
                        const double zl = L.zlen[size_t(q2)], nj = double(je - j0);
                        P.m_updates += double(n) * nj * zl;
                        P.m_bytes += (double(n) * zl + double(n) * nj) * 2;
                    }
                    j0 = je;
                }
            }
            a0 = a1;
        }
        t_meet += std::chrono::duration<double>(std::chrono::steady_clock::now() - tm0).count();
        kc.swap(kw);
    }
    const auto th1 = std::chrono::steady_clock::now();
    if (host_sub) {
        host_sub[0] += std::chrono::duration<double>(tk0 - th0).count();  // validation, keys
        host_sub[1] += std::chrono::duration<double>(tk1 - tk0).count();  // sort
        host_sub[2] += t_meet;                                            // meet tiles
        host_sub[3] += std::chrono::duration<double>(th1 - tk1).count() - t_meet;  // layout, items
    }
    // ---- upload (one packed staging buffer)
    Packer& pk = B.pk;
    pk.clear();
    B.o_qk = pk.add(P.q_k), B.o_qout = pk.add(P.q_out), B.o_qvp = pk.add(P.q_vp), B.o_vprc = pk.add(P.vp_rc);
    B.o_vpvalid = pk.add(P.vp_valid), B.o_src = pk.add(P.slot_rc), B.o_si = pk.add(P.slot_i);
    B.o_tiles = pk.add(P.mtiles), B.o_glist = pk.add(P.glist), B.o_zgeo = pk.add(P.zgeo), B.o_fills = pk.add(P.fills);
    B.o_items.assign(size_t(H), 0);
    size_t maxpart = 0;
    for (int l = 0; l < H; ++l) {
        B.o_items[size_t(l)] = pk.add(P.levels[size_t(l)].items);
        maxpart = std::max(maxpart, size_t(P.levels[size_t(l)].ksplit) * lv[size_t(l)].ldp);
    }
    B.kmax = kmax;
    B.ns = ns;
    B.nsl = nsl;
    B.qsz = qsz;
    B.maxpart = maxpart;
    B.fill_blocks = fill_blocks;
    if (P.fills.size() > 64) return fail(HTG_E_INVALID, "query plan has more than 64 fill ranges");
    return HTG_OK;
}

// Device addresses of a planned batch into the kernels' argument blocks.
void Map::bind_batch(QueryPlan& P, const BatchLayout& B, const char* dp, uint16_t* Qb, uint16_t* Zp,
                     uint16_t* dout, unsigned long long* wstat) const {
    auto I = [&](size_t o) { return reinterpret_cast<const int*>(dp + o); };
    const int nq = int(P.q_k.size());
    const int ns = H > 0 ? B.ns : 0;
    QFrontArgs& F = P.front;
    F = QFrontArgs{};
    F.q_k = I(B.o_qk);
    F.q_out = reinterpret_cast<const long long*>(dp + B.o_qout);
    F.q_vp = I(B.o_qvp);
    F.vp_rc = reinterpret_cast<const int2*>(dp + B.o_vprc);
    F.vp_valid = reinterpret_cast<const uint8_t*>(dp + B.o_vpvalid);
    F.slot_rc = reinterpret_cast<const int2*>(dp + B.o_src);
    F.leafD = static_cast<const uint16_t*>(leafD.p);
    F.out = dout;
    F.X0 = Qb + B.xoff[0];
    F.fills = reinterpret_cast<const int4*>(dp + B.o_fills);
    F.fill_base = Qb;
    F.wstat = wstat;
    F.nq = nq;
    F.ib = int(((long long)B.kmax * B.kmax + 255) / 256);
    F.leaf_blocks = int(((long long)ns * (4 * b - 4) + 255) / 256);
    F.n_slots = ns;
    F.b = b;
    F.nl = nl;
    F.n_fill = int(P.fills.size());
    F.fill_blocks = B.fill_blocks;
    P.zargs.resize(size_t(H));
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[size_t(l) - 1];
        const LevelPlan& LP = P.levels[size_t(l) - 1];
        QZArgs& Z = P.zargs[size_t(l) - 1];
        Z = QZArgs{};
        Z.DU = static_cast<const uint16_t*>(L.DU.p);
        Z.seam_list = static_cast<const int*>(L.d_seam_list.p);
        Z.cols = static_cast<const int*>(L.d_zcols.p);
        Z.live_list = L.d_live_list;
        Z.items = reinterpret_cast<const int4*>(dp + B.o_items[size_t(l) - 1]);
        Z.X = Qb + B.xoff[size_t(l) - 1];                // X of this level (stride pc)
        Z.Zp = Zp;
        Z.colpos = static_cast<const int*>(L.d_colpos.p);
        Z.colext = static_cast<const int*>(L.d_colext.p);
        Z.ext_child = static_cast<const int*>(L.d_ext_child.p);
        Z.ext_cport = static_cast<const int*>(L.d_ext_cport.p);
        Z.Z = Qb + B.zoffq[size_t(l)];
        Z.Xn = l < H ? Qb + B.xoff[size_t(l)] : nullptr;  // lifted X (stride ne)
        Z.wstat = wstat;
        Z.zsrc = static_cast<const int*>(L.d_zsrc.p);
        Z.slot_rc = F.slot_rc;
        Z.live_map = L.d_live_map;
        Z.ext_col = static_cast<const int*>(L.d_ext_col.p);
        Z.node_stride = L.node_stride;
        Z.part_stride = B.nsl * L.ldp;
        Z.ld = L.d.ld;
        Z.pc = L.d.pc;
        Z.ngrp = L.ngrp;
        Z.ldp = L.ldp;
        Z.n_items = int(LP.items.size() / 2);
        Z.col_tiles = LP.col_tiles;
        Z.ksplit = LP.ksplit;
        Z.ldz = L.ldz;
        Z.ne = l < H ? L.d.ne : 0;  // no lift above the root
        Z.M = LP.M;
        Z.n_slots = ns;
        Z.nz = L.nz;
        Z.side = L.d.s;
        Z.nt = L.d.nt;
        Z.cs = L.d.cs;
        Z.g = g;
    }
    QMeetArgs& Mt = P.meet;
    Mt = QMeetArgs{};
    Mt.tiles = I(B.o_tiles);
    Mt.glist = I(B.o_glist);
    Mt.zgeo = I(B.o_zgeo);
    Mt.slot_i = I(B.o_si);
    Mt.q_k = F.q_k;
    Mt.q_out = F.q_out;
    Mt.Z = Qb;
    Mt.out = dout;
    Mt.n_tiles = ns > 0 ? int(P.mtiles.size() / 8) : 0;
    Mt.g2 = g * g;
}

// The launches of one bound batch: front (output init, level-0 lift, INF
// fills), per level the projection (and the fold of its row chunks), then
// the meet of every level.
int Map::enqueue_batch(const QueryPlan& P, const BatchLayout& B, bool counted, hipStream_t st) {
    if (!launch_q_front(P.front, st)) return fail(HTG_E_INVALID, "query front: more than 64 fill ranges");
    for (int l = 1; l <= H && B.ns > 0; ++l) {
        const LevelGeo& L = lv[l - 1];
        const QZArgs& Z = P.zargs[size_t(l) - 1];
        // bytes and updates of the projection are counted on the device
        KTimer kt;
        const int pi = counted ? prof_ref().open(HTG_K_Q_PROJECT, 0.0, 0.0, kt, st) : -1;
        const bool la = launch_q_z(Z, st, kt);
        prof_ref().close(pi, la);
        if (Z.ksplit > 1) {
            const int pf = counted ? prof_ref().open(HTG_K_Q_LIFT, double(B.ns) * (L.nz + 2.0 * Z.ne + L.d.pc) * 2,
                                                     0, kt, st)
                                   : -1;
            const bool lf = launch_q_fold(Z, st, kt);
            prof_ref().close(pf, lf);
        }
    }
    if (B.ns > 0 && P.meet.n_tiles > 0) {
        KTimer kt;
        const int pm = counted ? prof_ref().open(HTG_K_Q_MEET, P.m_bytes, P.m_updates, kt, st) : -1;
        const bool lm = launch_q_meet(P.meet, st, kt);
        prof_ref().close(pm, lm);
    }
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "query launch");
    return HTG_OK;
}

int Map::query_launch(const int32_t* rcs, const int32_t* ks, int nq, QueryWS& ws, bool to_host) {
    hipError_t e;
    int rc;
    const auto th0 = std::chrono::steady_clock::now();
    QueryPlan& P = ws.plan;
    BatchLayout& B = ws.lay;
    ws.counted = prof_ref().take();  // this batch carries kernel timing
    if ((rc = plan_batch(rcs, ks, nq, P, B, ws.counted, ws.host_sub))) return rc;
    ws.total_out = B.total_out;
    if (B.total_out == 0) {
        ws.counted = false;
        hipError_t e0 = hipEventRecord(ws.done, ws.stream);
        return e0 == hipSuccess ? HTG_OK : hip_fail(e0, "record");
    }
    const auto th1 = std::chrono::steady_clock::now();
    // ---- buffers, device addresses, upload (one packed staging buffer)
    const Packer& pk = B.pk;
    const long long total_out = B.total_out;
    if ((rc = ws.h_plan.ensure(pk.size)) || (rc = ws.d_plan.ensure(pk.size)) || (rc = ws.Q.ensure(B.q_bytes())) ||
        (rc = ws.Zp.ensure(B.zp_bytes())) || (rc = ws.d_out.ensure(B.out_bytes())) ||
        (rc = ws.h_out.ensure((size_t(total_out) * 2 + 15) / 16 * 16)))
        return rc;
    unsigned long long* wstat = nullptr;
    if (ws.counted) {
        if ((rc = ws.d_wstat.ensure(kWorkSlots * 16)) || (rc = ws.h_wstat.ensure(kWorkSlots * 16))) return rc;
        wstat = static_cast<unsigned long long*>(ws.d_wstat.p);
    }
    hipStream_t st = ws.stream;
    bind_batch(P, B, static_cast<const char*>(ws.d_plan.p), static_cast<uint16_t*>(ws.Q.p),
               static_cast<uint16_t*>(ws.Zp.p), static_cast<uint16_t*>(ws.d_out.p), wstat);
    pk.write(static_cast<char*>(ws.h_plan.p));
    if (fault_hit(HTG_FAULT_QUERY)) return fail(HTG_E_HIP, "injected fault (query launch)");
    static const bool kernel_copy = [] {
        const char* e = std::getenv("HTG_KERNEL_COPY");
        return !e || std::atoi(e) != 0;
    }();
    const auto th2 = std::chrono::steady_clock::now();
    if (kernel_copy) {
        launch_copy_bytes(ws.h_plan.p, ws.d_plan.p, pk.size, st);
    } else if ((e = hipMemcpyAsync(ws.d_plan.p, ws.h_plan.p, pk.size, hipMemcpyHostToDevice, st)) != hipSuccess) {
        return hip_fail(e, "upload plan");
    }
    if ((rc = enqueue_batch(P, B, ws.counted, st))) return rc;
    if (to_host) {
        if (kernel_copy)
            launch_copy_bytes(ws.d_out.p, ws.h_out.p, size_t(total_out) * 2, st);
        else if ((e = hipMemcpyAsync(ws.h_out.p, ws.d_out.p, size_t(total_out) * 2, hipMemcpyDeviceToHost, st)) !=
                 hipSuccess)
            return hip_fail(e, "download result");
    }
    if (wstat && (e = hipMemcpyAsync(ws.h_wstat.p, wstat, kWorkSlots * 16, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "download counters");
    if ((e = hipEventRecord(ws.done, st)) != hipSuccess) return hip_fail(e, "record");
    const auto th3 = std::chrono::steady_clock::now();
    const double ph[3] = {std::chrono::duration<double>(th1 - th0).count(),
                          std::chrono::duration<double>(th2 - th1).count(),
                          std::chrono::duration<double>(th3 - th2).count()};
    for (int i = 0; i < 3; ++i) {
        ws.host_s[i] += ph[i];
        ws.host_max[i] = std::max(ws.host_max[i], ph[i]);
        ws.host_slow[i] += ph[i] > 0.5e-3;
    }
    ws.host_batches++;
    return HTG_OK;
}

int Map::reserve_query_ws(QueryWS& ws, int max_vps, long long max_out) const {
    const size_t nsl = size_t(std::max(max_vps, 1));
    auto a8 = [](size_t x) { return (x + 7) / 8 * 8; };
    size_t qsz = a8(nsl * size_t(4 * b - 4)), zp = 8;
    for (int l = 1; l <= H; ++l) {
        const LevelGeo& L = lv[size_t(l) - 1];
        if (l < H) qsz += a8(nsl * size_t(L.d.ne));
        qsz += a8(nsl * size_t(L.ldz));
        zp = std::max(zp, size_t(4) * L.ldp);  // up to four row chunks without growing
    }
    const size_t plan = nsl * 64 + (size_t(1) << 20);
    int rc;
    if ((rc = ws.Q.ensure(qsz * 2)) || (rc = ws.Zp.ensure(nsl * zp * 2)) ||
        (rc = ws.d_out.ensure(size_t(max_out) * 2 + 4)) || (rc = ws.h_out.ensure(size_t(max_out) * 2)) ||
        (rc = ws.d_plan.ensure(plan)) || (rc = ws.h_plan.ensure(plan)) || (rc = ws.d_wstat.ensure(kWorkSlots * 16)) ||
        (rc = ws.h_wstat.ensure(kWorkSlots * 16)))
        return rc;
    return HTG_OK;
}

int QueryWS::init(hipStream_t ext) {
    if (stream) return HTG_OK;
    hipError_t e;
    if (ext) {
        stream = ext;
        own_stream = false;
    } else if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess) {
        return hip_fail(e, "hipStreamCreate");
    }
    if ((e = hipEventCreateWithFlags(&done, hipEventDisableTiming)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    return HTG_OK;
}
QueryWS::~QueryWS() {
    if (debug_level() >= 2 && host_batches > 0)
        std::fprintf(stderr,
                     "[htg] query host time per batch (%lld batches): plan %.1f us (max %.0f, %lld > 0.5 ms), "
                     "pack+upload %.1f us (max %.0f, %lld), launch %.1f us (max %.0f, %lld)\n",
                     host_batches, 1e6 * host_s[0] / host_batches, 1e6 * host_max[0], host_slow[0],
                     1e6 * host_s[1] / host_batches, 1e6 * host_max[1], host_slow[1], 1e6 * host_s[2] / host_batches,
                     1e6 * host_max[2], host_slow[2]);
    if (debug_level() >= 2 && host_batches > 0)
        std::fprintf(stderr, "[htg]   planning: keys %.1f us, sort %.1f us, meet tiles %.1f us, layout+items %.1f us\n",
                     1e6 * host_sub[0] / host_batches, 1e6 * host_sub[1] / host_batches,
                     1e6 * host_sub[2] / host_batches, 1e6 * host_sub[3] / host_batches);
    if (stream) (void)hipStreamSynchronize(stream);
    if (done) (void)hipEventDestroy(done);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
}

}  // namespace htg

// ===========================================================================
// C ABI
// ===========================================================================
#define HTG_GUARD_BEGIN try {
#define HTG_GUARD_END                                                   \
    }                                                                   \
    catch (const std::bad_alloc&) {                                     \
        return htg::fail(HTG_E_NOMEM, "host allocation failed");        \
    }                                                                   \
    catch (...) {                                                       \
        return htg::fail(HTG_E_INVALID, "unexpected exception");        \
    }

// ---------------------------------------------------------------------------
// Device scratch of htg_query_batch_noalloc: one file-scope array per device
// (each GPU the library runs on holds its own copy until the process ends),
// carved per call into the packed plan, the X / Z regions, the row-chunk
// partials and the result.  Host code never uses the array's name as a
// device address: htg_ctx_create looks its address up once per device
// (hipGetSymbolAddress, with that device current) and keeps it in
// g_noalloc[device].ptr.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
struct NoAllocSlot {
    std::once_flag once;
    hipError_t err = hipSuccess;
    void* ptr = nullptr;  // this device's copy of `scratch`
    // Calls on one device share its array, so they run one at a time; the
    // host planning state is reused from call to call.
    std::mutex mu;
    QueryPlan plan;
    BatchLayout lay;
    std::vector<char> host_plan;
};
NoAllocSlot g_noalloc[kMaxDevices];

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(HTG_E_INVALID, "device index beyond the scratch table");
    NoAllocSlot& sl = g_noalloc[device];
    std::call_once(sl.once, [&sl] { sl.err = hipGetSymbolAddress(&sl.ptr, HIP_SYMBOL(scratch)); });
    if (sl.err != hipSuccess) return hip_fail(sl.err, "hipGetSymbolAddress(scratch)");
    return HTG_OK;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
}  // namespace

namespace {
// Calls on a map need a built map whose last update did not fail.
int usable(const htg_map* m) {
    if (!m->built) return fail(HTG_E_STATE, "map not built");
    if (m->broken) return fail(HTG_E_STATE, "an update of this map failed: it needs htg_build");
    return HTG_OK;
}
}  // namespace

extern "C" {

const char* htg_last_error(void) { return htg::g_err.c_str(); }

int htg_ctx_create(int device, htg_ctx** out) {
    HTG_GUARD_BEGIN
    if (!out) return fail(HTG_E_INVALID, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(HTG_E_INVALID, "no such device");
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (int rc = noalloc_setup(device)) return rc;
    auto* c = new htg_ctx();
    c->device = device;
    *out = c;
    return HTG_OK;
    HTG_GUARD_END
}

int htg_ctx_destroy(htg_ctx* ctx) {
    delete ctx;
    return HTG_OK;
}

int htg_ctx_mem_info(htg_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx) return fail(HTG_E_INVALID, "null ctx");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    size_t f = 0, t = 0;
    if ((e = hipMemGetInfo(&f, &t)) != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return HTG_OK;
}

int htg_map_footprint(const htg_config* cfg, int32_t rows, int32_t cols, uint64_t* bytes) {
    if (!cfg || !bytes) return fail(HTG_E_INVALID, "null argument");
    const int b = cfg->tile_size_b, g = cfg->group_w;
    if (b < 2 || b > 64 || g < 2 || cfg->group_h != g || g * g > 32 || rows <= 0 || cols <= 0)
        return fail(HTG_E_INVALID, "unsupported configuration");
    int S, H;
    geometry(b, g, rows, cols, S, H);
    *bytes = tables_bytes(b, g, S, H) + build_ws_bytes(b, g, S, H, 64);
    return HTG_OK;
}

}  // extern "C"

namespace htg {
int map_create_on(htg_ctx* ctx, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                  uint64_t hbm_budget, hipStream_t shared, htg_map** out, int min_S) {
    HTG_GUARD_BEGIN
    if (!ctx || !cfg || !cells || !out) return fail(HTG_E_INVALID, "null argument");
    // include/htrip/types.hpp:221-224 plus this build's limits: 2 <= b <= 64
    // (one tile row per lane) and g^2 <= 32 (child masks are 32-bit).
    const int b = cfg->tile_size_b, g = cfg->group_w;
    if (b <= 0 || g < 2 || cfg->group_h != g || g * g > 256) return fail(HTG_E_INVALID, "invalid HierarchyConfig");
    if (b < 2 || b > 64 || g * g > 32) return fail(HTG_E_INVALID, "tile_size_b must be 2..64 and group <= 5");
    if (rows <= 0 || cols <= 0) return fail(HTG_E_INVALID, "empty grid");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (min_S > (1 << 16)) return fail(HTG_E_INVALID, "hierarchy side too large");
    int rc;
    int S0, H0;
    geometry(b, g, rows, cols, S0, H0, min_S);
    const uint64_t need = tables_bytes(b, g, S0, H0) + build_ws_bytes(b, g, S0, H0, 64);
    if (hbm_budget && need > hbm_budget)
        return fail(HTG_E_NOMEM, "map needs " + std::to_string(need) + " bytes, budget " + std::to_string(hbm_budget));
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && need > fr)
        return fail(HTG_E_NOMEM, "map needs " + std::to_string(need) + " bytes, device has " + std::to_string(fr));
    static std::atomic<uint64_t> next_uid{0};
    std::unique_ptr<htg_map> m(new htg_map());
    m->uid = ++next_uid;
    m->ctx = ctx;
    m->cfg = *cfg;
    m->b = b;
    m->g = g;
    m->rows = rows;
    m->cols = cols;
    m->budget = hbm_budget;
    m->min_S = min_S;
    geometry(b, g, rows, cols, m->S, m->H, min_S);
    m->nl = m->S / b;
    m->cells_code.assign(cells, cells + size_t(rows) * cols);
    m->pass.assign(size_t(m->S) * m->S, 0);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) m->pass[size_t(r) * m->S + c] = cells[size_t(r) * cols + c] == 1;
    if (shared) {
        m->stream = shared;
        m->own_stream = false;
    } else if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) {
        return hip_fail(e, "hipStreamCreate");
    }
    hipStream_t st = m->stream;
    if ((rc = m->qws.init(st))) return rc;  // plain queries run in order with updates
    const size_t n = size_t(b) * b, pl = size_t(4 * b - 4), nleaves = size_t(m->nl) * m->nl;
    if ((rc = m->d_pass.ensure(m->pass.size()))) return rc;
    if ((rc = m->leafD.ensure(nleaves * n * n * 2))) return rc;
    if ((rc = m->leafS.ensure(nleaves * pl * pl * 2))) return rc;
    if ((e = hipMemcpyAsync(m->d_pass.p, m->pass.data(), m->pass.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "upload grid");
    m->lv.resize(m->H);
    for (int l = 1; l <= m->H; ++l) {
        LevelGeo& L = m->lv[l - 1];
        make_level(L, l, b, g, m->S, m->H);
        if ((rc = L.DU.ensure(size_t(L.nn) * L.node_stride * 2))) return rc;
        if ((rc = L.SU.ensure(size_t(L.nn) * L.d.ne * L.d.ne * 2))) return rc;
        if ((rc = upload(L.d_vu_rc, L.vu_rc, st)) || (rc = upload(L.d_ext_idx, L.ext_idx, st)) ||
            (rc = upload(L.d_ext_child, L.ext_child, st)) || (rc = upload(L.d_ext_cport, L.ext_cport, st)) ||
            (rc = upload(L.d_seams, L.seams, st)) || (rc = upload(L.d_seam_list, L.seam_list, st)) ||
            (rc = upload(L.d_seam_off, L.seam_off, st)) || (rc = upload(L.d_zcols, L.zcols, st)) ||
            (rc = upload(L.d_zsrc, L.zsrc, st)) || (rc = upload(L.d_zoff, L.zoff, st)) || (rc = upload(L.d_zlen, L.zlen, st)) ||
            (rc = upload(L.d_ext_col, L.ext_col, st)) ||
            (rc = upload(L.d_colpos, L.colpos, st)) || (rc = upload(L.d_colext, L.colext, st)) ||
            (rc = upload(L.d_vu_ext, L.vu_ext, st)))
            return rc;
        std::vector<int> all(size_t(L.nn));
        std::iota(all.begin(), all.end(), 0);
        if ((rc = upload(L.d_all, all, st)))
            return rc;
    }
    if ((rc = m->init_live()) || (rc = m->refresh_live(nullptr))) return rc;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "upload live groups");
    if ((rc = m->reserve_update_ws())) return rc;
    m->device_bytes = tables_bytes(b, g, m->S, m->H);
    *out = m.release();
    return HTG_OK;
    HTG_GUARD_END
}

int map_clone_on(const htg_map* src, uint64_t hbm_budget, hipStream_t shared, htg_map** out) {
    HTG_GUARD_BEGIN
    if (!src || !out) return fail(HTG_E_INVALID, "null argument");
    if (src->broken) return fail(HTG_E_STATE, "an update of this map failed: it needs htg_build");
    htg_map* m = nullptr;
    int rc = map_create_on(src->ctx, &src->cfg, src->rows, src->cols, src->cells_code.data(), hbm_budget, shared, &m,
                           src->min_S);
    if (rc) return rc;
    std::unique_ptr<htg_map> hold(m);
    m->pass = src->pass;
    m->tune = src->tune;
    m->built = src->built;
    m->root_full = src->root_full;
    m->stats = src->stats;
    m->epoch = src->epoch;
    m->gen = src->gen;
    m->clock = src->clock;
    m->leaf_ver = src->leaf_ver;
    m->node_ver = src->node_ver;
    hipError_t e;
    const hipStream_t st = m->stream;
    (void)hipStreamSynchronize(src->stream);
    const auto d2d = hipMemcpyDeviceToDevice;
    if ((e = hipMemcpyAsync(m->d_pass.p, src->d_pass.p,
                            src->d_pass.bytes < m->d_pass.bytes ? src->d_pass.bytes : m->d_pass.bytes, d2d, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(m->leafD.p, src->leafD.p, m->leafD.bytes, d2d, st)) != hipSuccess ||
        (e = hipMemcpyAsync(m->leafS.p, src->leafS.p, m->leafS.bytes, d2d, st)) != hipSuccess)
        return hip_fail(e, "clone copy");
    for (int l = 0; l < m->H; ++l) {
        if ((e = hipMemcpyAsync(m->lv[l].DU.p, src->lv[l].DU.p, size_t(m->lv[l].nn) * m->lv[l].node_stride * 2, d2d,
                                st)) != hipSuccess ||
            (e = hipMemcpyAsync(m->lv[l].SU.p, src->lv[l].SU.p,
                                size_t(m->lv[l].nn) * m->lv[l].d.ne * m->lv[l].d.ne * 2, d2d, st)) != hipSuccess)
            return hip_fail(e, "clone copy");
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "clone copy");
    *out = hold.release();
    return HTG_OK;
    HTG_GUARD_END
}

// liftRoot (src/minplus_oracle.cpp:2008-2114).  The square grows from S to
// g S and gains a root level.  Leaves and nodes inside the old square keep
// their tables (copied into their places in the larger level arrays); the
// leaves and nodes outside it and the new root are computed from the grid,
// leaves by BFS and nodes from their children, as the reference computes
// them (LeafTile::recomputeFromGrid, InternalNode::recomputeFromChildren).
// Inside the old square the map keeps its own cell state, which its kept
// tables describe; cells there that were outside the old grid stay
// impassable until opened.
int lift_root(htg_map* src, int32_t rows, int32_t cols, const uint8_t* cells, htg_map** out) {
    HTG_GUARD_BEGIN
    if (!src || !cells || !out) return fail(HTG_E_INVALID, "null argument");
    if (!src->built) return fail(HTG_E_STATE, "map not built");
    const int b = src->b, g = src->g, S0 = src->S, nl0 = src->nl, H0 = src->H;
    const long long S1 = (long long)S0 * g;
    if (S1 > (1 << 16)) return fail(HTG_E_INVALID, "hierarchy side too large");
    if (rows < src->rows || cols < src->cols || rows > S1 || cols > S1)
        return fail(HTG_E_INVALID, "the grid must contain the old one and fit the lifted square");
    hipError_t e = hipSetDevice(src->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(src->stream)) != hipSuccess) return hip_fail(e, "sync");
    // the old root becomes an inner node: every entry of its table counts
    if (int mr = src->materialize_root()) return mr;
    std::vector<uint8_t> merged(size_t(rows) * cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            uint8_t v = cells[size_t(r) * cols + c];
            if (r < S0 && c < S0) {
                if (r < src->rows && c < src->cols) v = src->cells_code[size_t(r) * src->cols + c];
                if (!src->pass[size_t(r) * S0 + c] && v == 1) v = 0;  // kept tables treat it as impassable
            }
            merged[size_t(r) * cols + c] = v;
        }
    htg_map* m = nullptr;
    int rc = map_create_on(src->ctx, &src->cfg, rows, cols, merged.data(), src->budget, nullptr, &m, int(S1));
    if (rc) return rc;
    std::unique_ptr<htg_map> hold(m);
    if (m->S != S1 || m->H != H0 + 1) return fail(HTG_E_INVALID, "unexpected lifted geometry");
    m->tune = src->tune;
    m->stats = src->stats;
    m->epoch = src->epoch + 1;  // src/minplus_oracle.cpp:2113
    const hipStream_t st = m->stream;
    const auto d2d = hipMemcpyDeviceToDevice;
    const int nl1 = m->nl;
    const size_t n = size_t(b) * b, pl = size_t(4 * b - 4);
    // old leaf (tx, ty) moves from ty nl0 + tx to ty nl1 + tx: one 2-D copy per array
    if ((e = hipMemcpy2DAsync(m->leafD.p, size_t(nl1) * n * n * 2, src->leafD.p, size_t(nl0) * n * n * 2,
                              size_t(nl0) * n * n * 2, size_t(nl0), d2d, st)) != hipSuccess ||
        (e = hipMemcpy2DAsync(m->leafS.p, size_t(nl1) * pl * pl * 2, src->leafS.p, size_t(nl0) * pl * pl * 2,
                              size_t(nl0) * pl * pl * 2, size_t(nl0), d2d, st)) != hipSuccess)
        return hip_fail(e, "lift copy leaves");
    std::vector<int> leaves;
    for (int ty = 0; ty < nl1; ++ty)
        for (int tx = 0; tx < nl1; ++tx)
            if (tx >= nl0 || ty >= nl0) leaves.push_back(ty * nl1 + tx);
    std::vector<std::vector<int>> nodes(size_t(H0) + 2);
    for (int l = 1; l <= H0; ++l) {
        const LevelGeo& Lo = src->lv[size_t(l) - 1];
        const LevelGeo& Ln = m->lv[size_t(l) - 1];
        if (Lo.node_stride != Ln.node_stride || Lo.d.ne != Ln.d.ne) return fail(HTG_E_INVALID, "level layout differs");
        const int nt0 = Lo.d.nt, nt1 = Ln.d.nt;
        const size_t du = Lo.node_stride * 2, su = size_t(Lo.d.ne) * Lo.d.ne * 2;
        if ((e = hipMemcpy2DAsync(Ln.DU.p, size_t(nt1) * du, Lo.DU.p, size_t(nt0) * du, size_t(nt0) * du, size_t(nt0),
                                  d2d, st)) != hipSuccess ||
            (e = hipMemcpy2DAsync(Ln.SU.p, size_t(nt1) * su, Lo.SU.p, size_t(nt0) * su, size_t(nt0) * su, size_t(nt0),
                                  d2d, st)) != hipSuccess)
            return hip_fail(e, "lift copy nodes");
        for (int ty = 0; ty < nt1; ++ty)
            for (int tx = 0; tx < nt1; ++tx)
                if (tx >= nt0 || ty >= nt0) nodes[size_t(l)].push_back(ty * nt1 + tx);
    }
    nodes[size_t(H0) + 1].push_back(0);  // the new root
    m->reset_versions();
    m->timing_now = false;
    if ((rc = m->rebuild({}, leaves, nodes))) return rc;
    m->built = true;
    *out = hold.release();
    return HTG_OK;
    HTG_GUARD_END
}
}  // namespace htg

extern "C" {

int htg_map_create(htg_ctx* ctx, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                   uint64_t hbm_budget, htg_map** out) {
    return htg::map_create_on(ctx, cfg, rows, cols, cells, hbm_budget, nullptr, out);
}

int htg_map_destroy(htg_map* map) {
    if (!map) return HTG_OK;
    (void)hipSetDevice(map->ctx->device);
    (void)hipStreamSynchronize(map->stream);
    (void)hipStreamSynchronize(map->qws.stream);
    if (map->stream && map->own_stream) (void)hipStreamDestroy(map->stream);
    delete map;
    return HTG_OK;
}

int htg_map_clone(const htg_map* src, uint64_t hbm_budget, htg_map** out) {
    return htg::map_clone_on(src, hbm_budget, nullptr, out);
}

int htg_map_info_get(const htg_map* map, htg_map_info* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    out->rows = map->rows;
    out->cols = map->cols;
    out->S = map->S;
    out->H = map->H;
    out->num_leaf_tiles_x = map->nl;
    out->num_leaves = map->nl * map->nl;
    out->built = map->built ? 1 : 0;
    out->reserved = 0;
    // Reference accounting (src/minplus_oracle.cpp:2116-2134): leaf D + S,
    // node D_U + S_U + child_to_ext (nV x nExt).
    const uint64_t n = uint64_t(map->b) * map->b, pl = 4ull * map->b - 4;
    uint64_t t = uint64_t(map->nl) * map->nl * (n * n + pl * pl) * 2;
    for (const auto& L : map->lv)
        t += uint64_t(L.nn) * (uint64_t(L.d.nV) * L.d.nV + uint64_t(L.d.ne) * L.d.ne + uint64_t(L.d.nV) * L.d.ne) * 2;
    out->total_memory_bytes = t;
    out->device_bytes = map->device_bytes;
    return HTG_OK;
}

int htg_build(htg_map* map) {
    HTG_GUARD_BEGIN
    if (!map) return fail(HTG_E_INVALID, "null map");
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return map->build();
    HTG_GUARD_END
}

int htg_open_cells(htg_map* map, const int32_t* rc, int32_t n) {
    HTG_GUARD_BEGIN
    if (!map || (n > 0 && !rc)) return fail(HTG_E_INVALID, "null argument");
    if (int u = usable(map)) return u;
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int r = map->open_cells(rc, n);
    if (r != HTG_OK) map->broken = true;  // the cell state already includes the batch
    return r;
    HTG_GUARD_END
}

int htg_open_cell(htg_map* map, int32_t r, int32_t c) {
    HTG_GUARD_BEGIN
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (int u = usable(map)) return u;
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // onCellOpened does not touch last_batch_stats (src/minplus_oracle.cpp:139-177)
    const htg_batch_stats keep = map->stats;
    const int32_t rcv[2] = {r, c};
    int rcode = map->open_cells(rcv, 1);
    map->stats = keep;
    if (rcode != HTG_OK) map->broken = true;
    return rcode;
    HTG_GUARD_END
}

int htg_block_cells(htg_map* map, const int32_t* rc, int32_t n) {
    HTG_GUARD_BEGIN
    if (!map || (n > 0 && !rc)) return fail(HTG_E_INVALID, "null argument");
    if (n < 0) return fail(HTG_E_INVALID, "negative count");
    if (int u = usable(map)) return u;
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int r = map->block_cells(rc, n);
    if (r != HTG_OK) map->broken = true;
    return r;
    HTG_GUARD_END
}

// Viewpoint count of a batch; HTG_E_INVALID for a negative k or missing rc.
static int check_batch(const int32_t* rc, const int32_t* ks, int32_t nq, uint16_t* out) {
    if (nq < 0) return fail(HTG_E_INVALID, "negative query count");
    if (nq > 0 && (!ks || !out)) return fail(HTG_E_INVALID, "null argument");
    long long total = 0;
    for (int q = 0; q < nq; ++q) {
        if (ks[q] < 0) return fail(HTG_E_INVALID, "negative k");
        total += ks[q];
    }
    if (total > 0 && !rc) return fail(HTG_E_INVALID, "null viewpoints");
    return HTG_OK;
}

int htg_query_batch(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, uint16_t* out) {
    HTG_GUARD_BEGIN
    if (int c = check_batch(rc, ks, nq, out)) return c;
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (int u = usable(map)) return u;
    std::lock_guard<std::mutex> lk(map->qmu);
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // queries read what the map's update stream wrote
    if ((e = hipStreamSynchronize(map->stream)) != hipSuccess) return hip_fail(e, "update sync");
    return map->query(rc, ks, nq, out, map->qws);
    HTG_GUARD_END
}

int htg_query_ws_create(htg_ctx* ctx, htg_query_ws** out) {
    HTG_GUARD_BEGIN
    if (!ctx || !out) return fail(HTG_E_INVALID, "null argument");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::unique_ptr<htg_query_ws> w(new htg_query_ws());
    w->ctx = ctx;
    if (int rc = w->ws.init()) return rc;  // a stream of its own
    if ((e = hipEventCreateWithFlags(&w->after, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    *out = w.release();
    return HTG_OK;
    HTG_GUARD_END
}

int htg_query_ws_destroy(htg_query_ws* ws) {
    if (!ws) return HTG_OK;
    (void)hipSetDevice(ws->ctx->device);
    if (ws->ws.stream) (void)hipStreamSynchronize(ws->ws.stream);
    delete ws;
    return HTG_OK;
}

int htg_query_ws_bytes(const htg_query_ws* ws, uint64_t* bytes) {
    if (!ws || !bytes) return fail(HTG_E_INVALID, "null argument");
    const QueryWS& w = ws->ws;
    *bytes = uint64_t(w.d_plan.bytes + w.Q.bytes + w.Zp.bytes + w.d_out.bytes + w.d_wstat.bytes);
    return HTG_OK;
}

int htg_query_batch_ws(htg_map* map, htg_query_ws* ws, const int32_t* rc, const int32_t* ks, int32_t nq,
                       uint16_t* out) {
    HTG_GUARD_BEGIN
    if (int c = check_batch(rc, ks, nq, out)) return c;
    if (!map || !ws) return fail(HTG_E_INVALID, "null argument");
    if (ws->ctx->device != map->ctx->device) return fail(HTG_E_INVALID, "workspace of another device");
    if (int u = usable(map)) return u;
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // the workspace's stream runs after every update enqueued on the map so far
    if ((e = hipEventRecord(ws->after, map->stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(ws->ws.stream, ws->after, 0)) != hipSuccess)
        return hip_fail(e, "order after updates");
    return map->query(rc, ks, nq, out, ws->ws);
    HTG_GUARD_END
}

int htg_query_all_pairs_ws(htg_map* map, htg_query_ws* ws, const int32_t* rc, int32_t k, uint16_t* out) {
    return htg_query_batch_ws(map, ws, rc, &k, 1, out);
}

int htg_query_batch_noalloc(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, uint16_t* out) {
    HTG_GUARD_BEGIN
    if (int c = check_batch(rc, ks, nq, out)) return c;
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (int u = usable(map)) return u;
    const int dev = map->ctx->device;
    if (dev < 0 || dev >= kMaxDevices || !g_noalloc[dev].ptr)
        return fail(HTG_E_STATE, "no-allocation scratch not set up for this device");
    NoAllocSlot& sl = g_noalloc[dev];
    std::lock_guard<std::mutex> dl(sl.mu);    // calls on one device share its array
    std::lock_guard<std::mutex> ml(map->qmu);  // and the map's stream with its own queries
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    QueryPlan& P = sl.plan;
    BatchLayout& B = sl.lay;
    if (int r = map->plan_batch(rc, ks, nq, P, B, false, nullptr)) return r;
    if (B.total_out == 0) return HTG_OK;
    const size_t o_plan = 0, o_q = o_plan + align256(B.pk.size), o_zp = o_q + align256(B.q_bytes()),
                 o_out = o_zp + align256(B.zp_bytes()), need = o_out + align256(B.out_bytes());
    if (need > kNoAllocScratchBytes)
        return fail(HTG_E_RANGE, "batch needs " + std::to_string(need) + " bytes of device scratch; "
                                 "htg_query_batch_noalloc holds " + std::to_string(kNoAllocScratchBytes));
    char* base = static_cast<char*>(sl.ptr);
    map->bind_batch(P, B, base + o_plan, reinterpret_cast<uint16_t*>(base + o_q),
                    reinterpret_cast<uint16_t*>(base + o_zp), reinterpret_cast<uint16_t*>(base + o_out), nullptr);
    if (sl.host_plan.size() < B.pk.size) sl.host_plan.resize(B.pk.size);  // host staging, grow-only
    B.pk.write(sl.host_plan.data());
    if (map->fault_hit(HTG_FAULT_QUERY)) return fail(HTG_E_HIP, "injected fault (query launch)");
    const hipStream_t st = map->stream;  // in order behind the map's updates
    if ((e = hipMemcpyAsync(base + o_plan, sl.host_plan.data(), B.pk.size, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "upload plan");
    // Nothing here needs to start at zero: the front kernel writes every
    // result entry and every INF fill of the X / Z regions, the projection
    // writes every partial the fold reads.
    if (int r = map->enqueue_batch(P, B, false, st)) return r;
    if ((e = hipMemcpyAsync(out, base + o_out, size_t(B.total_out) * 2, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "download result");
    return HTG_OK;
    HTG_GUARD_END
}

int htg_inject_fault(htg_map* map, int32_t site, int32_t countdown) {
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (countdown < 0 || (countdown > 0 && (site < HTG_FAULT_UPDATE || site > HTG_FAULT_BUILD)))
        return fail(HTG_E_INVALID, "bad fault site or countdown");
    map->fault_site = 0;
    map->fault_count = countdown;
    map->fault_site = countdown > 0 ? site : 0;
    return HTG_OK;
}

int htg_lift_point(htg_map* map, int32_t r, int32_t c, uint16_t* x, int64_t x_cap, int32_t* lens,
                   int32_t* parent_node_idx, int32_t* child_quad, int32_t* child_offset) {
    HTG_GUARD_BEGIN
    if (!map || !x || !lens || !parent_node_idx || !child_quad || !child_offset)
        return fail(HTG_E_INVALID, "null argument");
    if (int u = usable(map)) return u;
    std::lock_guard<std::mutex> lk(map->qmu);
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return map->lift_point(r, c, x, x_cap, lens, parent_node_idx, child_quad, child_offset);
    HTG_GUARD_END
}

int htg_lift_root(htg_map** map, int32_t rows, int32_t cols, const uint8_t* cells) {
    HTG_GUARD_BEGIN
    if (!map || !*map) return fail(HTG_E_INVALID, "null map");
    if ((*map)->broken) return fail(HTG_E_STATE, "an update of this map failed: it needs htg_build");
    htg_map* m = nullptr;
    const int rc = htg::lift_root(*map, rows, cols, cells, &m);
    if (rc) return rc;
    htg_map_destroy(*map);
    *map = m;
    return HTG_OK;
    HTG_GUARD_END
}

int htg_inc_create(htg_inc** out) {
    HTG_GUARD_BEGIN
    if (!out) return fail(HTG_E_INVALID, "null argument");
    *out = new htg_inc();
    return HTG_OK;
    HTG_GUARD_END
}

int htg_inc_destroy(htg_inc* state) {
    delete state;
    return HTG_OK;
}

int htg_query_all_pairs_incremental(htg_map* map, htg_inc* state, const int32_t* rc, int32_t k, uint16_t* out,
                                    int32_t* n_recomputed) {
    HTG_GUARD_BEGIN
    if (!map || !state || (k > 0 && (!rc || !out))) return fail(HTG_E_INVALID, "null argument");
    if (k < 0) return fail(HTG_E_INVALID, "negative k");
    if (int u = usable(map)) return u;
    std::lock_guard<std::mutex> lk(map->qmu);
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(map->stream)) != hipSuccess) return hip_fail(e, "update sync");
    return map->query_incremental(*state, rc, k, out, n_recomputed);
    HTG_GUARD_END
}

int htg_query_all_pairs(htg_map* map, const int32_t* rc, int32_t k, uint16_t* out) {
    return htg_query_batch(map, rc, &k, 1, out);
}

int htg_last_batch_stats(const htg_map* map, htg_batch_stats* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    *out = map->stats;
    return HTG_OK;
}

int htg_sync(htg_map* map) {
    if (!map) return fail(HTG_E_INVALID, "null map");
    hipError_t e = hipStreamSynchronize(map->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(map->qws.stream);
    if (e != hipSuccess) return hip_fail(e, "sync");
    return HTG_OK;
}

// Table reads at the root see its full table (under the map's query lock,
// as htg_lift_point does).
static int materialize_locked(htg_map* map, int lvl) {
    if (lvl != map->H || map->root_full || !map->built || map->broken) return HTG_OK;
    std::lock_guard<std::mutex> lk(map->qmu);
    const hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return map->materialize_root();
}

static int read_rows(htg_map* map, const void* src, size_t rows, size_t cols, size_t ld, uint16_t* out) {
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(map->stream)) != hipSuccess) return hip_fail(e, "sync");
    if ((e = hipMemcpy2DAsync(out, cols * 2, src, ld * 2, cols * 2, rows, hipMemcpyDeviceToHost, map->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(map->stream)) != hipSuccess)
        return hip_fail(e, "read table");
    return HTG_OK;
}

int htg_read_leaf_D(htg_map* map, int32_t leaf, uint16_t* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    if (leaf < 0 || leaf >= map->nl * map->nl) return fail(HTG_E_RANGE, "leaf index");
    const size_t n = size_t(map->b) * map->b;
    return read_rows(map, static_cast<const uint16_t*>(map->leafD.p) + size_t(leaf) * n * n, n, n, n, out);
}

int htg_read_leaf_S(htg_map* map, int32_t leaf, uint16_t* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    if (leaf < 0 || leaf >= map->nl * map->nl) return fail(HTG_E_RANGE, "leaf index");
    const size_t p = size_t(4 * map->b - 4);
    return read_rows(map, static_cast<const uint16_t*>(map->leafS.p) + size_t(leaf) * p * p, p, p, p, out);
}

int htg_node_dims(const htg_map* map, int32_t lvl, int32_t* num_nodes, int32_t* nV, int32_t* nExt) {
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (lvl < 1 || lvl > map->H) return fail(HTG_E_RANGE, "level");
    const LevelGeo& L = map->lv[lvl - 1];
    if (num_nodes) *num_nodes = L.nn;
    if (nV) *nV = L.d.nV;
    if (nExt) *nExt = L.d.ne;
    return HTG_OK;
}

int htg_read_node_DU(htg_map* map, int32_t lvl, int32_t idx, uint16_t* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    if (lvl < 1 || lvl > map->H || idx < 0 || idx >= map->lv[lvl - 1].nn) return fail(HTG_E_RANGE, "node index");
    const LevelGeo& L = map->lv[lvl - 1];
    if (int rc = materialize_locked(map, lvl)) return rc;
    return read_rows(map, static_cast<const uint16_t*>(L.DU.p) + size_t(idx) * L.node_stride, L.d.nV, L.d.nV, L.d.ld,
                     out);
}

int htg_read_node_SU(htg_map* map, int32_t lvl, int32_t idx, uint16_t* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    if (lvl < 1 || lvl > map->H || idx < 0 || idx >= map->lv[lvl - 1].nn) return fail(HTG_E_RANGE, "node index");
    const LevelGeo& L = map->lv[lvl - 1];
    const size_t ne = size_t(L.d.ne);
    if (int rc = materialize_locked(map, lvl)) return rc;
    return read_rows(map, static_cast<const uint16_t*>(L.SU.p) + size_t(idx) * ne * ne, ne, ne, ne, out);
}

int htg_profile_enable(htg_map* map, int on) {
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (on < 0) return fail(HTG_E_INVALID, "sampling period must be >= 0");
    map->prof_ref().reset();
    static const bool attach = [] {
        const char* e = std::getenv("HTG_PROFILE_ATTACH");
        return e && std::atoi(e) != 0;
    }();
    map->prof_ref().attach = attach;
    map->prof_ref().every = on > 1 ? on : 1;
    map->prof_ref().tick = 0;
    map->prof_ref().on = on != 0;
    return HTG_OK;
}

int htg_profile_get(htg_map* map, int kernel_class, htg_kernel_stats* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    if (kernel_class < 0 || kernel_class >= HTG_K_COUNT) return fail(HTG_E_RANGE, "kernel class");
    map->prof_ref().harvest();
    *out = map->prof_ref().acc[kernel_class];
    return HTG_OK;
}

int htg_get_tuning(const htg_map* map, htg_tuning* out) {
    if (!map || !out) return fail(HTG_E_INVALID, "null argument");
    *out = map->tune;
    return HTG_OK;
}

// Autotuner: times a rebuild for every pivot block with and without graph
// capture, and the sample query batch for every projection tile, and keeps
// the fastest of each (ties keep the smaller choice).  The tables after a
// rebuild are the same, so tuning never changes results.
int htg_autotune(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, int32_t reps, htg_tuning* out) {
    HTG_GUARD_BEGIN
    if (!map) return fail(HTG_E_INVALID, "null map");
    if (int u = usable(map)) return u;
    if (nq > 0 && (!rc || !ks)) return fail(HTG_E_INVALID, "null sample");
    std::lock_guard<std::mutex> lk(map->qmu);
    hipError_t e = hipSetDevice(map->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(map->stream)) != hipSuccess) return hip_fail(e, "sync");
    reps = std::max(1, reps);
    using clk = std::chrono::steady_clock;
    htg_tuning best = map->tune;
    double best_t = 1e30;
    for (int graph = 0; graph <= 1; ++graph)
        for (int P : {16, 32, 64}) {
            map->tune.pivot_block = P;
            map->tune.use_graph = graph;
            int rcode = map->build();  // warm-up (and capture)
            if (rcode) return rcode;
            double t = 1e30;
            for (int r = 0; r < reps; ++r) {
                const auto t0 = clk::now();
                if ((rcode = map->build())) return rcode;
                t = std::min(t, std::chrono::duration<double>(clk::now() - t0).count());
            }
            if (t < best_t) {
                best_t = t;
                best.pivot_block = P;
                best.use_graph = graph;
            }
        }
    map->tune.pivot_block = best.pivot_block;
    map->tune.use_graph = best.use_graph;
    int rcode = map->build();  // leave the map built with the chosen settings
    if (rcode) return rcode;
    if (nq > 0) {
        long long total = 0;
        for (int q = 0; q < nq; ++q) total += (long long)ks[q] * ks[q];
        std::vector<uint16_t> tmp(size_t(std::max<long long>(total, 1)));
        double best_q = 1e30;
        for (int Z : {4, 8, 16}) {
            map->tune.zq_tile = Z;
            if ((rcode = map->query(rc, ks, nq, tmp.data(), map->qws))) return rcode;
            double t = 1e30;
            for (int r = 0; r < reps; ++r) {
                const auto t0 = clk::now();
                if ((rcode = map->query(rc, ks, nq, tmp.data(), map->qws))) return rcode;
                t = std::min(t, std::chrono::duration<double>(clk::now() - t0).count());
            }
            if (t < best_q) {
                best_q = t;
                best.zq_tile = Z;
            }
        }
    }
    map->tune = best;
    if (out) *out = best;
    return HTG_OK;
    HTG_GUARD_END
}

int htg_set_tuning(htg_map* map, const htg_tuning* t) {
    if (!map || !t) return fail(HTG_E_INVALID, "null argument");
    if (t->pivot_block != 16 && t->pivot_block != 32 && t->pivot_block != 64)
        return fail(HTG_E_INVALID, "pivot_block must be 16, 32 or 64");
    if (t->zq_tile < 1 || t->zq_tile > 16) return fail(HTG_E_INVALID, "zq_tile must be 1..16");
    if (t->update_pivots != 0 && t->update_pivots != 1) return fail(HTG_E_INVALID, "update_pivots must be 0 or 1");
    if (t->update_rebuild < -1 || t->update_rebuild > 1) return fail(HTG_E_INVALID, "update_rebuild must be -1, 0 or 1");
    map->tune = *t;
    return HTG_OK;
}

}  // extern "C"
