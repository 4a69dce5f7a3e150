/*
 * htrip_gpu.h -- C ABI of the MI355X (gfx950) min-plus distance oracle.
 *
 * This is the drop-in boundary beneath HTRIP's MinPlusOracle
 * (include/htrip/minplus_oracle.hpp in the reference).  Plain pointers and
 * sizes only; no C++ types, no exceptions.  Every call returns an int status:
 * HTG_OK (0) or a negative HTG_E_* code; htg_last_error() gives the message
 * of the last failure on the calling thread.
 *
 * Arithmetic is the reference's tropical semiring on uint16 distances:
 * INF = 60000, saturating add, min (include/htrip/types.hpp:34,45,66-70).
 * Cell codes follow include/htrip/types.hpp:101-105: 0 Unknown, 1 Free,
 * 2 Obstacle; only Free cells are passable (include/htrip/grid_map.hpp:142-145).
 *
 * Threading: updates (build, open, block, lift_root) of a map handle are used
 * by one host thread at a time, with no query running on the map, as in the
 * reference.  Queries that go through caller-owned workspaces
 * (htg_query_ws_*, the counterpart of HTripQueryScratch,
 * include/htrip/minplus_oracle.hpp:189-242, :337-341) may run on one map from
 * several threads at once, one workspace per thread.  The queries that use
 * the map's own workspace (htg_query_all_pairs, htg_query_batch,
 * htg_query_all_pairs_incremental, htg_lift_point) serialise on a per-map
 * lock.  Different handles may be used concurrently from different threads.
 * The server API owns its maps' ordering itself.
 */
#ifndef HTRIP_GPU_H
#define HTRIP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HTG_INF 60000

enum {
    HTG_OK = 0,
    HTG_E_INVALID = -1, /* bad argument or configuration */
    HTG_E_NOMEM = -2,   /* admission refused: HBM budget or device memory */
    HTG_E_HIP = -3,     /* HIP runtime error */
    HTG_E_STATE = -4,   /* call not valid in the handle's state (e.g. query before build, or a map
                           whose last update failed: it needs htg_build) */
    HTG_E_RANGE = -5,   /* index out of range */
    HTG_E_CANCELLED = -6 /* server: the request was cancelled before it ran */
};

typedef struct htg_ctx htg_ctx;
typedef struct htg_map htg_map;
typedef struct htg_inc htg_inc;
typedef struct htg_query_ws htg_query_ws;
typedef struct htg_server htg_server;

/* include/htrip/types.hpp:199-225 (HierarchyConfig): tile_size_b, group_w,
 * group_h, hybrid_k_threshold.  group_w must equal group_h. */
typedef struct htg_config {
    int32_t tile_size_b;
    int32_t group_w;
    int32_t group_h;
    int32_t hybrid_k_threshold;
} htg_config;

/* Geometry and footprint of a built map (include/htrip/minplus_oracle.hpp:41-47). */
typedef struct htg_map_info {
    int32_t rows, cols;
    int32_t S, H;              /* root side b*g^H >= max(rows, cols); hierarchy depth */
    int32_t num_leaf_tiles_x;  /* = num_leaf_tiles_y = S / b */
    int32_t num_leaves;
    int32_t built;
    int32_t reserved;
    uint64_t total_memory_bytes; /* reference accounting, src/minplus_oracle.cpp:2116-2134 */
    uint64_t device_bytes;       /* HBM actually held by this map */
} htg_map_info;

/* Counters of the last update call (include/htrip/minplus_oracle.hpp:251-256). */
typedef struct htg_batch_stats {
    int32_t n_touched_leaves;
    int32_t n_dirty_ancestors;
    int32_t n_pivots;
} htg_batch_stats;

/* ---- context --------------------------------------------------------- */
int htg_ctx_create(int device, htg_ctx** out);
int htg_ctx_destroy(htg_ctx* ctx);
const char* htg_last_error(void);
/* Free / total HBM of the context's device, in bytes. */
int htg_ctx_mem_info(htg_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- map lifecycle ---------------------------------------------------- */
/* Replaces MinPlusOracle(HierarchyConfig) (minplus_oracle.hpp:59) plus the
 * GridMap it will be built from.  cells: rows*cols bytes (0/1/2 codes).
 * hbm_budget: bytes the map may hold on the device, 0 = no limit beyond the
 * device's free memory; a map that would exceed it is refused with
 * HTG_E_NOMEM (admission by memory). */
int htg_map_create(htg_ctx* ctx, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                   uint64_t hbm_budget, htg_map** out);
int htg_map_destroy(htg_map* map);
/* Deep copy of device state: the reference oracle is copied by value by its
 * callers (apps/bench_update_latency.cpp:178, src/robot_simulator.cpp:160). */
int htg_map_clone(const htg_map* src, uint64_t hbm_budget, htg_map** out);
/* Bytes a map of this shape would hold on the device (for admission planning). */
int htg_map_footprint(const htg_config* cfg, int32_t rows, int32_t cols, uint64_t* bytes);
int htg_map_info_get(const htg_map* map, htg_map_info* out);

/* ---- hot path ----------------------------------------------------------- */
/* MinPlusOracle::build (src/minplus_oracle.cpp:48-137).  Returns
 * HTG_E_INVALID where the reference returns false (invalid config, empty grid). */
int htg_build(htg_map* map);
/* MinPlusOracle::onBatchCellsOpened (src/minplus_oracle.cpp:530-729).  rc holds
 * n (row, col) pairs.  The listed cells become Free (the reference's caller sets
 * them Free before the call, README.md:71-91); out-of-range cells are ignored
 * (src/minplus_oracle.cpp:546-547). */
int htg_open_cells(htg_map* map, const int32_t* rc, int32_t n);
/* MinPlusOracle::onCellOpened (src/minplus_oracle.cpp:139-177). */
int htg_open_cell(htg_map* map, int32_t r, int32_t c);
/* MinPlusOracle::onCellBlocked (src/minplus_oracle.cpp:247-302), for n cells
 * at once: the listed cells become impassable (Obstacle); their leaves are
 * recomputed and every ancestor node is rebuilt from its children.  Cells out
 * of range or already impassable are skipped; last_batch_stats is untouched.
 * Not supported inside an htg_server session (its maps are reveal-only). */
int htg_block_cells(htg_map* map, const int32_t* rc, int32_t n);
/* MinPlusOracle::liftRoot (src/minplus_oracle.cpp:2008-2114): the hierarchy's
 * square grows from S to g*S and gains a root level.  rows x cols cell codes
 * give the grown grid (it contains the old grid; rows, cols <= g*S).  Leaves
 * and nodes inside the old square keep their tables, and the map keeps its own
 * cell state there (cells of the old square outside the old grid stay
 * impassable until opened); the other leaves and nodes and the new root are
 * computed from the given cells.  On success *map is replaced by the grown
 * map's handle and the old handle is destroyed (incremental states taken on it
 * recompute, as the reference's epoch bump forces); on failure *map is kept.
 * Not for maps inside an htg_server session. */
int htg_lift_root(htg_map** map, int32_t rows, int32_t cols, const uint8_t* cells);
/* MinPlusOracle::queryAllPairs (src/minplus_oracle.cpp:934-1426).  rc holds k
 * (row, col) viewpoints; out receives k*k distances, row-major, host memory.
 * Impassable or out-of-range viewpoints give an INF row and column, INF
 * diagonal (src/minplus_oracle.cpp:961-965, :993-997). */
int htg_query_all_pairs(htg_map* map, const int32_t* rc, int32_t k, uint16_t* out);
/* MinPlusOracle::liftPoint (src/minplus_oracle.cpp:731-823,
 * include/htrip/minplus_oracle.hpp:177-183, :361).  Distances from cell (r, c)
 * to the ports of its leaf (level 0, 4b-4 entries in port order) and, for
 * l = 1..H, to the exterior ports of its level-l ancestor (nExt entries in
 * extPorts order), each restricted to that tile's region.  x receives the H+1
 * vectors back to back (x_cap entries available; lens[l] gets each length);
 * parent_node_idx, child_quad and child_offset get H+1 entries with entry 0 =
 * -1, -1, 0 as in the reference.  The active lists of the reference's
 * LiftedViewpoint are the finite entries of X (the drop-in derives them).
 * HTG_E_RANGE for a cell outside the S x S hierarchy (the reference indexes
 * out of bounds there) or x_cap too small. */
int htg_lift_point(htg_map* map, int32_t r, int32_t c, uint16_t* x, int64_t x_cap, int32_t* lens,
                   int32_t* parent_node_idx, int32_t* child_quad, int32_t* child_offset);
/* MinPlusOracle::queryAllPairsIncremental (src/minplus_oracle.cpp:1444-2002,
 * state include/htrip/minplus_oracle.hpp:262-272).  `state` (htg_inc_create)
 * keeps the last viewpoints and matrix; out receives k*k distances equal to
 * htg_query_all_pairs on the current map.  k = 0 clears the state.  The state
 * is refreshed by one batched query when it is new, was taken on another map
 * or before a build, lists other viewpoints, or when an update since then
 * rewrote the leaf or an ancestor node of one of the viewpoints; otherwise
 * the cached matrix is returned with no device work (no such table changed,
 * so no entry can have changed).  n_recomputed (may be null) gets k when the
 * matrix was recomputed, 0 when it was cached. */
int htg_inc_create(htg_inc** out);
int htg_inc_destroy(htg_inc* state);
int htg_query_all_pairs_incremental(htg_map* map, htg_inc* state, const int32_t* rc, int32_t k, uint16_t* out,
                                    int32_t* n_recomputed);
/* Same, for nq queries in one call: query q has ks[q] viewpoints (packed back to
 * back in rc) and its k*k result at out + sum_{p<q} ks[p]^2.  One coalesced set
 * of launches serves all of them. */
int htg_query_batch(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, uint16_t* out);
int htg_last_batch_stats(const htg_map* map, htg_batch_stats* out);
int htg_sync(htg_map* map);

/* ---- caller-owned query workspace --------------------------------------
 * MinPlusOracle::HTripQueryScratch (include/htrip/minplus_oracle.hpp:189-242)
 * and the const queryAllPairs that takes it (:337-341): a reusable,
 * grow-only workspace (plan, per-level lift and projection buffers, result
 * staging) with a stream of its own.  Queries through distinct workspaces
 * may run on one map from several threads at once (the reference's callers
 * do so with distinct scratch, src/robot_simulator.cpp:1541-1546, :1768); a
 * workspace serves one call at a time.  A query through a workspace sees
 * every update the map finished before the call (its stream waits for the
 * map's update stream).  A workspace may serve any map of its context's
 * device. */
int htg_query_ws_create(htg_ctx* ctx, htg_query_ws** out);
int htg_query_ws_destroy(htg_query_ws* ws);
/* Device bytes the workspace holds (it grows to the largest call so far). */
int htg_query_ws_bytes(const htg_query_ws* ws, uint64_t* bytes);
int htg_query_all_pairs_ws(htg_map* map, htg_query_ws* ws, const int32_t* rc, int32_t k, uint16_t* out);
int htg_query_batch_ws(htg_map* map, htg_query_ws* ws, const int32_t* rc, const int32_t* ks, int32_t nq,
                       uint16_t* out);
/* htg_query_batch that allocates no memory: no hipMalloc / hipFree, no pinned
 * or device buffer of any kind.  No counterpart in the reference; it is for
 * callers that cannot allocate (tight serving loops, code that must not
 * allocate or synchronise on a free).  Its device scratch (plan, lift and
 * projection buffers, result staging) is one fixed array of 64 MiB per GPU,
 * held from the library's first use on that GPU until the process ends;
 * a batch that needs more returns HTG_E_RANGE with the size it needs (on the
 * 1024^2, b = 32 serving map a single query of up to k ~ 1500 fits).  rc and out
 * are host memory as in htg_query_batch; the call runs on the map's stream
 * after its updates and returns when out is written.  Calls on one GPU share
 * its array: they are serialised by a per-GPU lock (unlocked, they would
 * race), and the array keeps the last call's contents, which no call reads
 * before writing. */
int htg_query_batch_noalloc(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, uint16_t* out);

/* ---- fault injection (test hook) ----------------------------------------
 * The countdown-th time (1 = the next time) the map reaches `site`, the call
 * fails there with HTG_E_HIP as a HIP error would.  countdown 0 disarms.
 * An update that fails this way (or by a real error) has already changed
 * the map's cell state, so the map answers HTG_E_STATE until htg_build
 * rebuilds it from that state.  No counterpart in the reference. */
enum {
    HTG_FAULT_UPDATE = 1, /* htg_open_cells / htg_open_cell, after its first level's kernels were enqueued */
    HTG_FAULT_QUERY = 2,  /* a query launch, before any kernel of the batch */
    HTG_FAULT_BUILD = 3   /* htg_build, before the first kernel */
};
int htg_inject_fault(htg_map* map, int32_t site, int32_t countdown);

/* ---- table access (the reference keeps these as public members:
 * LeafTile::D/S, InternalNode::D_U/S_U, include/htrip/leaf_tile.hpp,
 * include/htrip/internal_node.hpp).  Copies to host memory. --------------- */
int htg_read_leaf_D(htg_map* map, int32_t leaf, uint16_t* out);   /* b^2 x b^2 */
int htg_read_leaf_S(htg_map* map, int32_t leaf, uint16_t* out);   /* (4b-4)^2 */
int htg_node_dims(const htg_map* map, int32_t lvl, int32_t* num_nodes, int32_t* nV, int32_t* nExt);
int htg_read_node_DU(htg_map* map, int32_t lvl, int32_t idx, uint16_t* out); /* nV x nV */
int htg_read_node_SU(htg_map* map, int32_t lvl, int32_t idx, uint16_t* out); /* nExt x nExt */

/* ---- kernel timing ----------------------------------------------------------
 * When enabled, the runtime records start/stop hipEvents on the launching
 * stream immediately before and after every dispatch of the kernel classes
 * below (so the interval includes the dispatch's start latency), and
 * accumulates launch counts, elapsed time and the algorithmic bytes / min-plus
 * updates of each launch.  HTG_PROFILE_ATTACH=1 in the environment attaches
 * the events to the dispatch itself instead (hipExtLaunchKernelGGL), which
 * slows the timed kernel about twofold on gfx950.  Enabling resets the
 * counters. */
enum {
    HTG_K_LEAF_APSP = 0,
    HTG_K_CLOSURE_PIVOTS = 1,
    HTG_K_CLOSURE_UPDATE = 2,
    HTG_K_Q_PROJECT = 3,
    HTG_K_Q_MEET = 4,
    HTG_K_Q_LIFT = 5, /* fold of row-chunk partials + lift (only when a level splits its rows) */
    HTG_K_COVER_RELAX = 6, /* update: cover rows relaxed through the uncovered endpoints */
    HTG_K_COVER_PAIRS = 7, /* update: second hop on the cover x cover block */
    HTG_K_COUNT = 8
};
typedef struct htg_kernel_stats {
    uint64_t launches;
    double total_ms;
    double bytes;   /* algorithmic bytes moved (unique reads + writes) */
    double updates; /* min-plus updates (one saturating add + one min each) */
} htg_kernel_stats;
/* on: 0 off, 1 every operation, n > 1 one query batch / update / build in n
 * (timing events cost host time per launch; counters and durations come from
 * the same sampled operations). */
int htg_profile_enable(htg_map* map, int on);
int htg_profile_get(htg_map* map, int kernel_class, htg_kernel_stats* out);

/* ---- serving ------------------------------------------------------------
 * A server keeps every admitted map resident and serves requests from worker
 * threads.  Strict mode: requests of one session are served in submission
 * order (a query sees exactly the updates submitted before it); consecutive
 * queries are coalesced into one batched launch, consecutive updates into one
 * reveal batch.  Snapshot mode: each session is double-buffered like the
 * reference simulator's oracle pair (src/robot_simulator.cpp:897-1058):
 * queries never wait for updates and see the newest applied prefix of them.
 * Up to n_slots query batches are in flight at once on their own streams,
 * and requests that arrive meanwhile join the next batch as soon as a slot
 * frees.  Each stream needs a hardware queue of its own (the process's
 * GPU_MAX_HW_QUEUES >= n_slots + update_streams plus any other streams;
 * HIP's default is 4): streams that share a queue run in order behind each
 * other.  No counterpart in the reference's C++ API. */
enum {
    HTG_CONSISTENCY_STRICT = 0,  /* a query sees exactly the updates submitted before it */
    HTG_CONSISTENCY_SNAPSHOT = 1 /* double-buffered: a query sees the newest applied prefix of updates */
};
typedef struct htg_server_config {
    uint64_t hbm_budget;          /* admission limit for all sessions' maps; 0 = no limit */
    int32_t n_slots;              /* query batches in flight (1..64) */
    int32_t max_batch_queries;    /* queries coalesced per batch */
    int32_t max_batch_viewpoints; /* viewpoints per batch (a larger query still runs alone) */
    int32_t consistency;          /* HTG_CONSISTENCY_* */
    int32_t update_cu_percent;    /* builds/updates run on this share of the CUs (1..99); 0 or 100 = all */
    int32_t update_tick_us;       /* snapshot mode: a session's update batches start at least this far apart,
                                     so more reveals merge into one closure (0 = as soon as possible) */
    int32_t update_streams;       /* builds and updates run on this many streams (sessions are dealt to them
                                     round robin, and in snapshot mode as many update threads serve them), so
                                     updates of different sessions overlap on the device; 0 = 2, at most 8 */
} htg_server_config;
typedef struct htg_server_stats {
    uint64_t query_batches, queries, update_batches, updates, hbm_used;
    int32_t sessions;
    int32_t recovered_faults; /* update batches that failed and were recovered by rebuilding the copy */
    double plan_seconds;   /* worker-thread time spent planning and launching query batches */
    double update_seconds; /* time spent applying update batches (catch-up included) */
    double gpu_wait_seconds; /* slot time from the last launch until the batch's results were on the host */
    double finish_seconds;   /* slot time copying results out and completing requests */
} htg_server_stats;
int htg_server_create(htg_ctx* ctx, const htg_server_config* cfg, htg_server** out);
int htg_server_destroy(htg_server* s);
/* Admits and builds a session; HTG_E_NOMEM when it does not fit the budget. */
int htg_server_add_map(htg_server* s, const htg_config* cfg, int32_t rows, int32_t cols, const uint8_t* cells,
                       int32_t* session);
/* Waits for the session's requests, then frees its memory for new sessions. */
int htg_server_remove_map(htg_server* s, int32_t session);
/* out (k*k, host) must stay valid until the request completes. */
int htg_server_submit_query(htg_server* s, int32_t session, const int32_t* rc, int32_t k, uint16_t* out,
                            uint64_t* ticket);
int htg_server_submit_update(htg_server* s, int32_t session, const int32_t* rc, int32_t n, uint64_t* ticket);
/* Blocks until the request completed; latency_s = completion - submission. */
int htg_server_wait(htg_server* s, uint64_t ticket, double* latency_s);
/* Same; version = for a query, how many of its session's updates (in
 * submission order) its result includes; for an update, its index. */
int htg_server_wait_ex(htg_server* s, uint64_t ticket, double* latency_s, int64_t* version);
int htg_server_stats_get(htg_server* s, htg_server_stats* out);
/* Cancels a request that has not started: it completes with HTG_E_CANCELLED
 * (htg_server_wait returns that code) and has no effect on its session; a
 * cancelled update is never applied.  HTG_E_STATE when the request already
 * runs or has completed, HTG_E_RANGE for an unknown ticket.  A failed update
 * batch is recovered by rebuilding the session's copy from its cell state
 * (which includes the batch): its requests complete with HTG_OK and
 * recovered_faults counts it; other sessions are not touched. */
int htg_server_cancel(htg_server* s, uint64_t ticket);
/* Test hook: arms htg_inject_fault on the session's map copies (both copies
 * in snapshot mode: the next update of either). */
int htg_server_inject_fault(htg_server* s, int32_t session, int32_t site, int32_t countdown);
/* The session's map, for reading tables; do not use it while requests run. */
int htg_server_map(htg_server* s, int32_t session, htg_map** out);
/* Trace replay: submits request i at t_arrival[i] seconds after the call
 * (busy-wait pacing on the calling thread), waits for all, and writes each
 * completion time (seconds after the call) to t_done.  Request i is a query
 * (type 0, count = k) or an update (type 1, count = cells) on session[i]
 * with its (row, col) pairs at rc + 2*rc_off[i]; a query's matrix goes to
 * out + out_off[i] when out_off[i] >= 0.  version (nullable) receives each
 * request's version as in htg_server_wait_ex. */
int htg_server_replay(htg_server* s, int32_t n, const double* t_arrival, const int32_t* session, const int32_t* type,
                      const int32_t* count, const int64_t* rc_off, const int32_t* rc, const int64_t* out_off,
                      uint16_t* out, double* t_done, int64_t* version);

/* ---- build tuning (the project's autotuner picks these) --------------- */
typedef struct htg_tuning {
    int32_t pivot_block;   /* pivots per closure block: 16, 32 (default) or 64; blocks of up to 32 run each
                              block's update in the same launch as the next block's pivot rows */
    int32_t zq_tile;       /* viewpoints per query-projection work item: 1..16 */
    int32_t use_graph;     /* run the build as one captured hipGraph (replayed while its launch shape holds) */
    int32_t update_pivots; /* incremental reveal closes each node over: 0 (default) a vertex cover of the
                              decreased entries (new ports + endpoints of old-old decreases) after a two-hop
                              pre-relaxation through the uncovered endpoints; 1 every row with a decreased
                              entry (the reference's pivot set, src/internal_node.cpp:332-413).  Both are exact. */
    int32_t update_rebuild; /* with the cover (update_pivots 0): a level of an incremental update is instead
                               rebuilt from its children (the build's closure over the passable seam
                               endpoints) when 0 (default) an estimate says that is cheaper, 1 always, -1
                               never.  Every table after a reveal equals a fresh build, so all are exact. */
} htg_tuning;
int htg_get_tuning(const htg_map* map, htg_tuning* out);
int htg_set_tuning(htg_map* map, const htg_tuning* t);
/* Autotuner: rebuilds the map with every pivot block (16/32/64), with and
 * without graph capture, and runs the sample batch (nq queries, as in
 * htg_query_batch) with every projection tile (4/8/16), reps timed runs
 * each; keeps the fastest of each on the map and returns them in out.  The
 * tables and results do not depend on the tuning.  No counterpart in the
 * reference (whose kernels have no tunable shape). */
int htg_autotune(htg_map* map, const int32_t* rc, const int32_t* ks, int32_t nq, int32_t reps, htg_tuning* out);

#ifdef __cplusplus
}
#endif
#endif /* HTRIP_GPU_H */
