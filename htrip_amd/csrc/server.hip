// Serving loop: resident sessions, admission by HBM budget, continuous
// batching of update batches and mixed-k all-pairs queries.
//
// The reference's only serving-like caller is its robot simulator, which
// double-buffers the oracle and runs updates and queries on worker threads
// (src/robot_simulator.cpp:897-1058, :1483-1580).  Two consistency modes:
//
//  strict   -- requests of a session are served in submission order: a query
//              observes exactly the updates submitted before it.  Consecutive
//              queries are coalesced into one batched launch, consecutive
//              updates into one reveal batch (the union of monotone openings
//              equals applying them in turn); an update waits until the
//              session's in-flight query batches completed.
//  snapshot -- every session holds two copies of its hierarchy.  Queries run
//              on the front copy as soon as a slot frees and never wait for
//              updates; an update thread merges all pending updates, applies
//              them to the back copy (after bringing it level by copying the
//              tables the previous update rewrote) and swaps front and back.
//              A query observes a prefix of its session's updates; the
//              prefix length is reported with the result.
//
// In both modes up to n_slots query batches are in flight on their own
// streams, batches of different sessions overlap on the device, and requests
// that arrive while batches run join the next batch as soon as a slot frees.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/htrip_gpu.h"
#include "runtime.hpp"

using namespace htg;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int session = 0;
    int type = 0;  // 0 query, 1 update
    std::vector<int32_t> rc;
    int32_t count = 0;
    uint16_t* out = nullptr;
    Clock::time_point t_submit, t_done;
    int64_t version = 0;  // query: updates in its snapshot; update: its index in the session
    int status = HTG_OK;
    bool done = false;
};

// One query slot: a stream workspace and the thread that serves it.
struct Batch {
    QueryWS ws;
    std::mutex ws_mu;  // held by the slot thread while a batch uses ws
    int session = -1;
    int map_idx = 0;
    std::vector<Request*> reqs;
    std::thread th;
};

struct Session {
    htg_map* maps[2] = {nullptr, nullptr};
    int front = 0;
    uint64_t bytes = 0;
    std::deque<Request*> pending;      // strict: everything; snapshot: queries
    std::deque<Request*> upd_pending;  // snapshot: updates
    int inflight[2] = {0, 0};          // query batches on the device, per copy
    int updating = 0;
    int64_t submitted_updates = 0;
    int64_t applied = 0;  // update requests applied to the front copy
    bool back_stale = false;
    bool full_resync = false;  // the back copy must copy every table (its last update was recovered by a rebuild)
    DirtySet last_dirty;
    Clock::time_point last_update{};  // start of the last update batch (snapshot mode)
};

// Update stream, restricted to about pct % of the CUs when 0 < pct < 100.
// The mask takes whole blocks of 8 consecutive CU bits, spread evenly over
// the index range, so every XCD gets its share whether the CU numbering
// interleaves the XCDs or not.
hipError_t make_update_stream(int device, int pct, hipStream_t* out) {
    int ncu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    if (pct <= 0 || pct >= 100 || ncu < 16) return hipStreamCreateWithFlags(out, hipStreamNonBlocking);
    std::vector<uint32_t> mask(size_t(ncu + 31) / 32, 0u);
    const int blocks = ncu / 8;
    for (int t = 0; t < blocks; ++t)
        if ((t + 1) * pct / 100 > t * pct / 100)
            for (int i = t * 8; i < t * 8 + 8; ++i) mask[size_t(i) / 32] |= 1u << (i % 32);
    return hipExtStreamCreateWithCUMask(out, uint32_t(mask.size()), mask.data());
}

void mark_done(Request* r, int st, Clock::time_point now) {
    r->status = st;
    r->t_done = now;
    r->done = true;
}

}  // namespace

struct htg_server {
    htg_ctx* ctx = nullptr;
    htg_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_upd;
    std::map<int, Session> sessions;
    int next_session = 1;
    uint64_t used = 0;
    std::vector<std::unique_ptr<Batch>> pool;
    std::vector<hipStream_t> ustreams;  // builds and updates; session i runs on ustreams[i % size]
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_ticket = 1;
    std::vector<std::thread> updaters;  // snapshot mode: one per update stream
    int rr = 0;  // session where the next slot starts looking (round robin)
    bool stop = false;
    htg_server_stats stats{};

    bool snapshot() const { return cfg.consistency == HTG_CONSISTENCY_SNAPSHOT; }
    void slot_loop(Batch* b);
    void run_updates();
    bool pending_work() const;
    bool update_one(std::unique_lock<std::mutex>& lk);
    // Applies a merged update batch to m; a failure is recovered by
    // rebuilding m from its cell state and applying the batch again.
    int apply_update(htg_map* m, const std::vector<int32_t>& cells, bool& recovered);
};

int htg_server::apply_update(htg_map* m, const std::vector<int32_t>& cells, bool& recovered) {
    recovered = false;
    int st = m->open_cells(cells.data(), int(cells.size() / 2));
    if (st == HTG_OK) return st;
    // The cell state already holds the batch (or part of it): rebuild from it,
    // then open the batch again (cells already open are skipped).
    m->broken = true;
    if (m->build() != HTG_OK) return st;
    const int again = m->open_cells(cells.data(), int(cells.size() / 2));
    if (again != HTG_OK) {
        m->broken = true;
        return again;
    }
    recovered = true;
    return HTG_OK;
}

bool htg_server::pending_work() const {
    for (auto& kv : sessions)
        if (!kv.second.pending.empty()) return true;
    return false;
}

// A slot thread takes the next dispatchable work of any session (round
// robin), outside the lock plans and launches it on its own stream, waits
// for it, and completes its requests.  Query batches of several slots (and
// sessions) overlap on the device, and their host-side planning runs on
// separate threads.  In strict mode a slot also applies the update at the
// head of a session once the session's query batches have completed.
void htg_server::slot_loop(Batch* b) {
    if (hipSetDevice(ctx->device) != hipSuccess) return;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        int sid = -1;
        bool is_update = false;
        std::vector<Request*> run;
        std::vector<int> ids;  // sessions, starting at the round-robin point
        for (auto it = sessions.lower_bound(rr); it != sessions.end(); ++it) ids.push_back(it->first);
        for (auto it = sessions.begin(); it != sessions.end() && it->first < rr; ++it) ids.push_back(it->first);
        for (int id : ids) {
            Session& ss = sessions.at(id);
            // strict: nothing passes an update being applied; snapshot:
            // updates go to the back copy and never hold queries back
            if (ss.pending.empty() || (!snapshot() && ss.updating)) continue;
            if (ss.pending.front()->type == 0) {
                long long vps = 0;
                while (!ss.pending.empty() && ss.pending.front()->type == 0 &&
                       int(run.size()) < cfg.max_batch_queries &&
                       (run.empty() || vps + ss.pending.front()->count <= cfg.max_batch_viewpoints)) {
                    vps += ss.pending.front()->count;
                    run.push_back(ss.pending.front());
                    ss.pending.pop_front();
                }
            } else if (ss.inflight[0] == 0) {  // strict: the queries before the update have completed
                while (!ss.pending.empty() && ss.pending.front()->type == 1) {
                    run.push_back(ss.pending.front());
                    ss.pending.pop_front();
                }
                is_update = true;
            } else {
                continue;
            }
            sid = id;
            rr = id + 1;
            break;
        }
        if (sid < 0) {
            if (stop && !pending_work()) return;
            cv_work.wait(lk);
            continue;
        }
        Session& ss = sessions.at(sid);
        if (is_update) {
            ss.updating++;
            std::vector<int32_t> cells;
            for (Request* r : run) cells.insert(cells.end(), r->rc.begin(), r->rc.end());
            htg_map* m = ss.maps[0];
            lk.unlock();
            const auto tu0 = Clock::now();
            bool recovered = false;
            const int st = apply_update(m, cells, recovered);
            const double tu = std::chrono::duration<double>(Clock::now() - tu0).count();
            lk.lock();
            stats.update_seconds += tu;
            stats.recovered_faults += recovered;
            ss.updating--;
            ss.applied += int64_t(run.size());
            stats.update_batches++;
            stats.updates += run.size();
            const auto now = Clock::now();
            for (Request* r : run) mark_done(r, st, now);
            cv_done.notify_all();
            cv_work.notify_all();  // queries queued behind the update may go now
            continue;
        }
        const int mi = snapshot() ? ss.front : 0;
        ss.inflight[mi]++;
        std::vector<int32_t> rc, ks;
        for (Request* r : run) {
            rc.insert(rc.end(), r->rc.begin(), r->rc.end());
            ks.push_back(r->count);
            r->version = ss.applied;
        }
        htg_map* m = ss.maps[mi];
        stats.query_batches++;
        stats.queries += run.size();
        if (pending_work()) cv_work.notify_one();  // more work queued: let another idle slot take it
        lk.unlock();
        std::unique_lock<std::mutex> wl(b->ws_mu);
        const auto tp0 = Clock::now();
        int st = m->query_launch(rc.data(), ks.data(), int(ks.size()), b->ws, true);
        const double tp = std::chrono::duration<double>(Clock::now() - tp0).count();
        const auto tg0 = Clock::now();
        if (st == HTG_OK) {
            hipError_t q;
            while ((q = hipEventQuery(b->ws.done)) == hipErrorNotReady) std::this_thread::yield();
            (void)hipGetLastError();
            st = m->query_finish(b->ws, nullptr);
        }
        const auto tg1 = Clock::now();
        if (st == HTG_OK) {
            const uint16_t* res = static_cast<const uint16_t*>(b->ws.h_out.p);
            long long off = 0;
            for (Request* r : run) {
                const size_t n = size_t(r->count) * r->count;
                if (r->out && n) std::memcpy(r->out, res + off, n * 2);
                off += (long long)n;
            }
        }
        wl.unlock();
        const auto tg2 = Clock::now();
        lk.lock();
        stats.plan_seconds += tp;
        stats.gpu_wait_seconds += std::chrono::duration<double>(tg1 - tg0).count();
        stats.finish_seconds += std::chrono::duration<double>(tg2 - tg1).count();
        const auto now = Clock::now();
        for (Request* r : run) mark_done(r, st, now);
        ss.inflight[mi]--;
        cv_done.notify_all();
        cv_upd.notify_all();
        if (!snapshot()) cv_work.notify_all();  // a strict update may wait for this batch
    }
}

// Snapshot mode: apply one session's pending updates to its back copy and
// swap.  Called with lk held; returns false when no session had work.  The
// session whose oldest pending update has waited longest goes first (a busy
// session would otherwise refill its queue during its own update and keep
// the others waiting).
bool htg_server::update_one(std::unique_lock<std::mutex>& lk) {
    const auto now = Clock::now();
    Session* pick = nullptr;
    Clock::time_point oldest{};
    for (auto& kv : sessions) {
        Session& ss = kv.second;
        if (ss.upd_pending.empty() || ss.updating) continue;
        if (!stop && cfg.update_tick_us > 0 && now - ss.last_update < std::chrono::microseconds(cfg.update_tick_us))
            continue;  // coalesce: the next tick takes everything pending then
        if (ss.inflight[1 - ss.front] > 0) continue;  // queries still read the back copy
        const Clock::time_point t = ss.upd_pending.front()->t_submit;
        if (!pick || t < oldest) pick = &ss, oldest = t;
    }
    if (pick) {
        Session& ss = *pick;
        const int back = 1 - ss.front;
        std::vector<Request*> run(ss.upd_pending.begin(), ss.upd_pending.end());
        ss.upd_pending.clear();
        ss.updating++;
        std::vector<int32_t> cells;
        for (Request* r : run) cells.insert(cells.end(), r->rc.begin(), r->rc.end());
        htg_map* fm = ss.maps[ss.front];
        htg_map* bm = ss.maps[back];
        const bool stale = ss.back_stale, full = ss.full_resync;
        DirtySet prev = ss.last_dirty;
        lk.unlock();
        const auto tu0 = Clock::now();
        ss.last_update = tu0;
        int st = !stale ? HTG_OK : full ? bm->copy_all_from(*fm) : bm->catch_up_from(*fm, prev);
        bool recovered = false;
        if (st != HTG_OK) {  // the back copy is part way through the catch-up: rebuild it from its cells
            bm->broken = true;
            if (bm->copy_all_from(*fm) == HTG_OK) {
                st = HTG_OK;
                recovered = true;
            }
        }
        if (st == HTG_OK) {
            bool r2 = false;
            st = apply_update(bm, cells, r2);
            recovered |= r2;
        }
        const double tu = std::chrono::duration<double>(Clock::now() - tu0).count();
        lk.lock();
        stats.update_seconds += tu;
        stats.recovered_faults += recovered;
        ss.updating--;
        if (st == HTG_OK) {
            ss.front = back;
            ss.applied += int64_t(run.size());
            ss.last_dirty = bm->dirty;
            ss.back_stale = true;
            // after a recovery the tables the batch changed are not all in
            // bm->dirty: the other copy catches up by copying everything
            ss.full_resync = recovered;
        }
        stats.update_batches++;
        stats.updates += run.size();
        const auto now = Clock::now();
        for (Request* r : run) mark_done(r, st, now);
        cv_done.notify_all();
        cv_work.notify_all();
        return true;
    }
    return false;
}

void htg_server::run_updates() {
    if (hipSetDevice(ctx->device) != hipSuccess) return;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (update_one(lk)) continue;
        bool pending = false;
        for (auto& kv : sessions) pending |= !kv.second.upd_pending.empty();
        if (stop && !pending) return;
        // woken by submissions and by retired batches (the back copy frees up)
        cv_upd.wait_for(lk, std::chrono::microseconds(200));
    }
}

// ===========================================================================
extern "C" {

int htg_server_create(htg_ctx* ctx, const htg_server_config* cfg, htg_server** out) {
    try {
        if (!ctx || !cfg || !out) return fail(HTG_E_INVALID, "null argument");
        if (cfg->n_slots < 1 || cfg->n_slots > 64 || cfg->max_batch_queries < 1 || cfg->max_batch_viewpoints < 1 ||
            cfg->update_cu_percent < 0 || cfg->update_cu_percent > 100 || cfg->update_tick_us < 0 ||
            cfg->update_streams < 0 || cfg->update_streams > 8 ||
            (cfg->consistency != HTG_CONSISTENCY_STRICT && cfg->consistency != HTG_CONSISTENCY_SNAPSHOT))
            return fail(HTG_E_INVALID, "bad server configuration");
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        std::unique_ptr<htg_server> s(new htg_server());
        s->ctx = ctx;
        s->cfg = *cfg;
        // A few update streams and n_slots query streams: a process has few
        // hardware queues (GPU_MAX_HW_QUEUES, 4 by default), and streams
        // beyond that share queues, so that a query would wait behind update
        // kernels.  Sessions are dealt to the update streams, so updates of
        // sessions on different streams overlap on the device.  With
        // update_cu_percent the update streams run on that share of the CUs
        // only, and queries always find free CUs.
        const int nus = cfg->update_streams > 0 ? cfg->update_streams : 2;
        for (int i = 0; i < nus; ++i) {
            hipStream_t us = nullptr;
            if ((e = make_update_stream(ctx->device, cfg->update_cu_percent, &us)) != hipSuccess)
                return hip_fail(e, "update stream");
            s->ustreams.push_back(us);
        }
        for (int i = 0; i < cfg->n_slots; ++i) {
            s->pool.emplace_back(new Batch());
            int rc = s->pool.back()->ws.init();
            if (rc) return rc;
        }
        htg_server* raw = s.get();
        for (auto& bp : s->pool) {
            Batch* bb = bp.get();
            bb->th = std::thread([raw, bb] { raw->slot_loop(bb); });
        }
        if (raw->snapshot())
            for (size_t i = 0; i < s->ustreams.size(); ++i) s->updaters.emplace_back([raw] { raw->run_updates(); });
        *out = s.release();
        return HTG_OK;
    } catch (...) {
        return fail(HTG_E_INVALID, "server create failed");
    }
}

int htg_server_destroy(htg_server* s) {
    if (!s) return HTG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_upd.notify_all();
    for (auto& bp : s->pool)
        if (bp->th.joinable()) bp->th.join();
    for (auto& t : s->updaters)
        if (t.joinable()) t.join();
    for (auto& kv : s->sessions)
        for (htg_map* m : kv.second.maps)
            if (m) htg_map_destroy(m);
    s->pool.clear();
    for (hipStream_t us : s->ustreams) (void)hipStreamDestroy(us);
    delete s;
    return HTG_OK;
}

int htg_server_add_map(htg_server* s, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                       int32_t* session) {
    if (!s || !cfg || !cells || !session) return fail(HTG_E_INVALID, "null argument");
    uint64_t need = 0;
    int rc = htg_map_footprint(cfg, rows, cols, &need);
    if (rc) return rc;
    const int copies = s->snapshot() ? 2 : 1;
    need *= uint64_t(copies);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->cfg.hbm_budget && s->used + need > s->cfg.hbm_budget)
            return fail(HTG_E_NOMEM, "admission refused: session needs " + std::to_string(need) + " bytes, " +
                                         std::to_string(s->cfg.hbm_budget - s->used) + " left in the budget");
        s->used += need;  // reserve before building
    }
    htg_map* m[2] = {nullptr, nullptr};
    hipStream_t us;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        us = s->ustreams[size_t(s->next_session) % s->ustreams.size()];  // the next session id's stream
    }
    rc = map_create_on(s->ctx, cfg, rows, cols, cells, 0, us, &m[0]);
    if (rc == HTG_OK) rc = htg_build(m[0]);
    if (rc == HTG_OK && copies == 2) {
        rc = map_clone_on(m[0], 0, us, &m[1]);
        if (rc == HTG_OK) m[1]->profp = m[0]->profp;  // one profiler per session
    }
    // size every slot's workspace for this map's largest batch up front
    const long long max_out = std::min<long long>((long long)s->cfg.max_batch_viewpoints * s->cfg.max_batch_viewpoints,
                                                  (long long)s->cfg.max_batch_queries * 256 * 256);
    for (size_t i = 0; rc == HTG_OK && i < s->pool.size(); ++i) {
        std::lock_guard<std::mutex> wl(s->pool[i]->ws_mu);
        rc = m[0]->reserve_query_ws(s->pool[i]->ws, s->cfg.max_batch_viewpoints, max_out);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != HTG_OK) {
        for (htg_map* x : m)
            if (x) htg_map_destroy(x);
        s->used -= need;
        return rc;
    }
    const int id = s->next_session++;
    Session& ss = s->sessions[id];
    ss.maps[0] = m[0];
    ss.maps[1] = m[1];
    ss.bytes = need;
    *session = id;
    return HTG_OK;
}

int htg_server_remove_map(htg_server* s, int32_t session) {
    if (!s) return fail(HTG_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return fail(HTG_E_RANGE, "no such session");
    Session& ss = it->second;
    s->cv_done.wait(lk, [&] {
        return ss.pending.empty() && ss.upd_pending.empty() && ss.inflight[0] == 0 && ss.inflight[1] == 0 &&
               ss.updating == 0;
    });
    htg_map* m[2] = {ss.maps[0], ss.maps[1]};
    s->used -= ss.bytes;
    s->sessions.erase(it);
    lk.unlock();
    for (htg_map* x : m)
        if (x) htg_map_destroy(x);
    return HTG_OK;
}

static int submit(htg_server* s, int32_t session, int type, const int32_t* rc, int32_t count, uint16_t* out,
                  uint64_t* ticket) {
    if (!s || !ticket || count < 0 || (count > 0 && !rc)) return fail(HTG_E_INVALID, "bad argument");
    std::unique_ptr<Request> r(new Request());
    r->type = type;
    r->session = session;
    r->rc.assign(rc, rc + 2 * size_t(count));
    r->count = count;
    r->out = out;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->sessions.find(session);
        if (it == s->sessions.end()) return fail(HTG_E_RANGE, "no such session");
        Session& ss = it->second;
        r->ticket = s->next_ticket++;
        *ticket = r->ticket;
        if (type == 1) r->version = ss.submitted_updates++;
        if (type == 1 && s->snapshot())
            ss.upd_pending.push_back(r.get());
        else
            ss.pending.push_back(r.get());
        s->reqs[r->ticket] = std::move(r);
    }
    if (type == 1 && s->snapshot())
        s->cv_upd.notify_one();
    else
        s->cv_work.notify_one();
    return HTG_OK;
}

int htg_server_submit_query(htg_server* s, int32_t session, const int32_t* rc, int32_t k, uint16_t* out,
                            uint64_t* ticket) {
    return submit(s, session, 0, rc, k, out, ticket);
}

int htg_server_submit_update(htg_server* s, int32_t session, const int32_t* rc, int32_t n, uint64_t* ticket) {
    return submit(s, session, 1, rc, n, nullptr, ticket);
}

int htg_server_wait_ex(htg_server* s, uint64_t ticket, double* latency_s, int64_t* version) {
    if (!s) return fail(HTG_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(HTG_E_RANGE, "unknown ticket");
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    const int st = r->status;
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (version) *version = r->version;
    s->reqs.erase(it);
    if (st != HTG_OK) return fail(st, "request failed");
    return HTG_OK;
}

int htg_server_wait(htg_server* s, uint64_t ticket, double* latency_s) {
    return htg_server_wait_ex(s, ticket, latency_s, nullptr);
}

int htg_server_stats_get(htg_server* s, htg_server_stats* out) {
    if (!s || !out) return fail(HTG_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->hbm_used = s->used;
    out->sessions = int32_t(s->sessions.size());
    return HTG_OK;
}

int htg_server_cancel(htg_server* s, uint64_t ticket) {
    if (!s) return fail(HTG_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(HTG_E_RANGE, "unknown ticket");
    Request* r = it->second.get();
    if (r->done) return fail(HTG_E_STATE, "request already completed");
    auto si = s->sessions.find(r->session);
    if (si == s->sessions.end()) return fail(HTG_E_STATE, "request already completed");
    Session& ss = si->second;
    for (std::deque<Request*>* q : {&ss.pending, &ss.upd_pending}) {
        for (auto qi = q->begin(); qi != q->end(); ++qi) {
            if (*qi != r) continue;
            q->erase(qi);
            mark_done(r, HTG_E_CANCELLED, Clock::now());
            s->cv_done.notify_all();
            s->cv_work.notify_all();  // a strict update may have waited behind it
            return HTG_OK;
        }
    }
    return fail(HTG_E_STATE, "request already running");
}

int htg_server_inject_fault(htg_server* s, int32_t session, int32_t site, int32_t countdown) {
    if (!s) return fail(HTG_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return fail(HTG_E_RANGE, "no such session");
    Session& ss = it->second;
    // updates go to copy 0 (strict) or the back copy (snapshot); queries to
    // copy 0 or the front copy
    const int mi = !s->snapshot() ? 0 : site == HTG_FAULT_QUERY ? ss.front : 1 - ss.front;
    return htg_inject_fault(ss.maps[mi], site, countdown);
}

int htg_server_map(htg_server* s, int32_t session, htg_map** out) {
    if (!s || !out) return fail(HTG_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return fail(HTG_E_RANGE, "no such session");
    *out = it->second.maps[it->second.front];
    return HTG_OK;
}

int htg_server_replay(htg_server* s, int32_t n, const double* t_arrival, const int32_t* session, const int32_t* type,
                      const int32_t* count, const int64_t* rc_off, const int32_t* rc, const int64_t* out_off,
                      uint16_t* out, double* t_done, int64_t* version) {
    if (!s || n < 0 || (n > 0 && (!t_arrival || !session || !type || !count || !rc_off || !rc || !t_done)))
        return fail(HTG_E_INVALID, "null argument");
    std::vector<uint64_t> tickets(size_t(n), 0);
    const auto t0 = Clock::now();
    int first_err = HTG_OK;
    for (int i = 0; i < n; ++i) {
        const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t_arrival[i]));
        while (Clock::now() < due) {
        }
        uint16_t* o = (type[i] == 0 && out && out_off && out_off[i] >= 0) ? out + out_off[i] : nullptr;
        int rcode = submit(s, session[i], type[i], rc + 2 * rc_off[i], count[i], o, &tickets[i]);
        if (rcode != HTG_OK && first_err == HTG_OK) first_err = rcode;
    }
    for (int i = 0; i < n; ++i) {
        if (!tickets[i]) {
            t_done[i] = -1;
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(tickets[i]);
        Request* r = it->second.get();
        s->cv_done.wait(lk, [&] { return r->done; });
        t_done[i] = std::chrono::duration<double>(r->t_done - t0).count();
        if (version) version[i] = r->version;
        if (r->status != HTG_OK && first_err == HTG_OK) first_err = r->status;
        s->reqs.erase(it);
    }
    if (first_err != HTG_OK) return fail(first_err, "replay: a request failed");
    return HTG_OK;
}

}  // extern "C"
