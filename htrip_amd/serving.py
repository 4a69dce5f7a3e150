"""Serving: ctypes binding of the htg_server C ABI and the request traces the
benchmarks replay.

Trace model (the serving workload): a seeded ground-truth grid, 16 simulated
robots that each start with an observed disk and reveal free cells in their
own frontier (BFS) order, and a Poisson request stream in which each request
is either an update batch (one robot's next 16..512 revealed cells) or a
queryAllPairs with k drawn log-uniform in [8, 256] over the cells observed
free at that point of the stream.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, grids
from ._lib import Config, check, ptr


STRICT, SNAPSHOT = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("n_slots", C.c_int32), ("max_batch_queries", C.c_int32),
                ("max_batch_viewpoints", C.c_int32), ("consistency", C.c_int32), ("update_cu_percent", C.c_int32),
                ("update_tick_us", C.c_int32), ("update_streams", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("query_batches", C.c_uint64), ("queries", C.c_uint64), ("update_batches", C.c_uint64),
                ("updates", C.c_uint64), ("hbm_used", C.c_uint64), ("sessions", C.c_int32),
                ("recovered_faults", C.c_int32),
                ("plan_seconds", C.c_double), ("update_seconds", C.c_double), ("gpu_wait_seconds", C.c_double),
                ("finish_seconds", C.c_double)]


def _bind(lib):
    if getattr(lib, "_server_bound", False):
        return
    vp, i32, u64, f64 = C.c_void_p, C.c_int32, C.c_uint64, C.c_double
    pi32, pi64, pu16, pu8, pf64 = (C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint16),
                                   C.POINTER(C.c_uint8), C.POINTER(C.c_double))
    sig = {
        "htg_server_create": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
        "htg_server_destroy": (C.c_int, [vp]),
        "htg_server_add_map": (C.c_int, [vp, C.POINTER(Config), i32, i32, pu8, pi32]),
        "htg_server_remove_map": (C.c_int, [vp, i32]),
        "htg_server_submit_query": (C.c_int, [vp, i32, pi32, i32, pu16, C.POINTER(u64)]),
        "htg_server_submit_update": (C.c_int, [vp, i32, pi32, i32, C.POINTER(u64)]),
        "htg_server_wait": (C.c_int, [vp, u64, pf64]),
        "htg_server_wait_ex": (C.c_int, [vp, u64, pf64, pi64]),
        "htg_server_stats_get": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "htg_server_map": (C.c_int, [vp, i32, C.POINTER(vp)]),
        "htg_server_replay": (C.c_int, [vp, i32, pf64, pi32, pi32, pi32, pi64, pi32, pi64, pu16, pf64, pi64]),
        "htg_server_cancel": (C.c_int, [vp, u64]),
        "htg_server_inject_fault": (C.c_int, [vp, i32, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._server_bound = True


SERVER_EXPORTS = ["htg_server_create", "htg_server_destroy", "htg_server_add_map", "htg_server_remove_map",
                  "htg_server_submit_query", "htg_server_submit_update", "htg_server_wait", "htg_server_wait_ex",
                  "htg_server_stats_get", "htg_server_map", "htg_server_replay", "htg_server_cancel",
                  "htg_server_inject_fault"]


class Server:
    """Resident sessions + continuous batching (htg_server)."""

    def __init__(self, hbm_budget: int = 0, n_slots: int = 4, max_batch_queries: int = 64,
                 max_batch_viewpoints: int = 8192, consistency: int = STRICT, update_cu_percent: int = 0,
                 update_tick_us: int = 0, update_streams: int = 0, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self.lib = self.ctx.lib
        _bind(self.lib)
        self.h = C.c_void_p()
        cfg = ServerConfig(hbm_budget, n_slots, max_batch_queries, max_batch_viewpoints, consistency,
                           update_cu_percent, update_tick_us, update_streams)
        check(self.lib.htg_server_create(self.ctx.h, C.byref(cfg), C.byref(self.h)))
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.htg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_map(self, cells: np.ndarray, b: int, g: int = 2) -> int:
        cells = np.ascontiguousarray(cells, dtype=np.uint8)
        sid = C.c_int32()
        cfg = Config(b, g, g, 0)
        check(self.lib.htg_server_add_map(self.h, C.byref(cfg), cells.shape[0], cells.shape[1],
                                          ptr(cells, C.c_uint8), C.byref(sid)))
        return sid.value

    def remove_map(self, sid: int) -> None:
        check(self.lib.htg_server_remove_map(self.h, sid))

    def submit_query(self, sid: int, vps) -> tuple:
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        out = np.empty((len(vps), len(vps)), dtype=np.uint16)
        t = C.c_uint64()
        check(self.lib.htg_server_submit_query(self.h, sid, ptr(vps, C.c_int32), len(vps), ptr(out, C.c_uint16),
                                               C.byref(t)))
        self._keep[t.value] = (vps, out)
        return t.value, out

    def submit_update(self, sid: int, cells) -> int:
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 2)
        t = C.c_uint64()
        check(self.lib.htg_server_submit_update(self.h, sid, ptr(cells, C.c_int32), len(cells), C.byref(t)))
        return t.value

    def wait(self, ticket: int, with_version: bool = False):
        lat = C.c_double()
        ver = C.c_int64()
        rc = self.lib.htg_server_wait_ex(self.h, ticket, C.byref(lat), C.byref(ver))
        self._keep.pop(ticket, None)
        check(rc)
        return (lat.value, ver.value) if with_version else lat.value

    def stats(self) -> dict:
        s = ServerStats()
        check(self.lib.htg_server_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def cancel(self, ticket: int) -> int:
        """htg_server_cancel: HTG_OK when the request was cancelled before it
        ran, else the (negative) reason code; never raises."""
        return int(self.lib.htg_server_cancel(self.h, ticket))

    def inject_fault(self, sid: int, site: int, countdown: int = 1) -> None:
        check(self.lib.htg_server_inject_fault(self.h, sid, site, countdown))

    def wait_status(self, ticket: int) -> int:
        """Waits for a request and returns its status code instead of raising."""
        lat = C.c_double()
        ver = C.c_int64()
        rc = self.lib.htg_server_wait_ex(self.h, ticket, C.byref(lat), C.byref(ver))
        self._keep.pop(ticket, None)
        return int(rc)

    def map_of(self, sid: int, cells: np.ndarray, b: int, g: int = 2) -> _lib.GpuMap:
        """A non-owning GpuMap view of a session (for reading tables between requests)."""
        h = C.c_void_p()
        check(self.lib.htg_server_map(self.h, sid, C.byref(h)))
        m = _lib.GpuMap(cells, b, g, self.ctx, _handle=h)
        m.close = lambda: None  # owned by the server
        return m

    def replay(self, trace: "Trace") -> np.ndarray:
        """Replay a trace with native arrival pacing; returns completion times (s)
        and stores each request's version in trace.version."""
        n = len(trace.t)
        t_done = np.zeros(n, dtype=np.float64)
        trace.version = np.zeros(n, dtype=np.int64)
        check(self.lib.htg_server_replay(
            self.h, n, ptr(trace.t, C.c_double), ptr(trace.session, C.c_int32), ptr(trace.type, C.c_int32),
            ptr(trace.count, C.c_int32), ptr(trace.rc_off, C.c_int64), ptr(trace.rc, C.c_int32),
            ptr(trace.out_off, C.c_int64), ptr(trace.out, C.c_uint16), ptr(t_done, C.c_double),
            ptr(trace.version, C.c_int64)))
        return t_done


@dataclass
class Trace:
    t: np.ndarray            # arrival times (s), float64
    session: np.ndarray      # int32
    type: np.ndarray         # int32: 0 query, 1 update
    count: np.ndarray        # int32: k or number of cells
    rc_off: np.ndarray       # int64 offsets (in pairs) into rc
    rc: np.ndarray           # int32 (n_pairs * 2)
    out_off: np.ndarray      # int64 offsets into out (-1: none)
    out: np.ndarray          # uint16 results
    meta: dict = field(default_factory=dict)
    version: np.ndarray | None = None  # filled by Server.replay

    def query_result(self, i: int) -> np.ndarray:
        k = int(self.count[i])
        o = int(self.out_off[i])
        return self.out[o:o + k * k].reshape(k, k)

    def cells_of(self, i: int) -> np.ndarray:
        o, n = int(self.rc_off[i]), int(self.count[i])
        return self.rc[2 * o:2 * (o + n)].reshape(n, 2)


class RobotReveal:
    """16 robots revealing the ground truth from their start disks, each in its own frontier order."""

    def __init__(self, truth: np.ndarray, n_robots: int, radius: int, seed: int, limit: int = 20000):
        self.truth = truth
        starts = grids.sample_free(truth, n_robots, seed)
        self.observed = grids.disk_observed(truth, [tuple(p) for p in starts], radius)
        self.orders = []
        for p in starts:
            one = grids.disk_observed(truth, [tuple(p)], radius)
            self.orders.append(grids.frontier_order(truth, one, limit))
        self.ptr = [0] * n_robots
        self.free_list = list(np.flatnonzero(self.observed.ravel() == grids.FREE))

    def next_batch(self, robot: int, m: int) -> np.ndarray:
        order, cols = self.orders[robot], self.truth.shape[1]
        out = []
        p = self.ptr[robot]
        while len(out) < m and p < len(order):
            r, c = order[p]
            p += 1
            if self.observed[r, c] != grids.FREE:
                self.observed[r, c] = grids.FREE
                out.append((r, c))
                self.free_list.append(int(r) * cols + int(c))
        self.ptr[robot] = p
        return np.array(out, dtype=np.int32).reshape(-1, 2)

    def sample_viewpoints(self, k: int, rng: grids.SplitMix64) -> np.ndarray:
        cols = self.truth.shape[1]
        n = len(self.free_list)
        idx = [self.free_list[rng.below(n)] for _ in range(k)]
        return np.array([(i // cols, i % cols) for i in idx], dtype=np.int32)


def loguniform_k(rng: grids.SplitMix64, lo: int = 8, hi: int = 256) -> int:
    u = rng.next() / 2.0 ** 64
    return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))


def make_trace(session_of, reveal: RobotReveal, n: int, rate: float, seed: int, update_frac: float = 0.1,
               k_lo: int = 8, k_hi: int = 256, upd_lo: int = 16, upd_hi: int = 512, record_states: int = 0):
    """Poisson request stream for one session (session_of: int session id)."""
    rng = grids.SplitMix64(seed)
    t, typ, cnt, rc_off, out_off, parts = [], [], [], [], [], []
    now, npairs, nout = 0.0, 0, 0
    states = []
    robots = len(reveal.orders)
    for i in range(n):
        u = (rng.next() + 1) / 2.0 ** 64
        now += -math.log(u) / rate
        is_upd = (rng.next() / 2.0 ** 64) < update_frac
        if is_upd:
            m = loguniform_k(rng, upd_lo, upd_hi)
            cells = reveal.next_batch(rng.below(robots), m)
            if len(cells) == 0:
                is_upd = False
        if is_upd:
            typ.append(1)
            cnt.append(len(cells))
            rc_off.append(npairs)
            parts.append(cells)
            npairs += len(cells)
            out_off.append(-1)
        else:
            k = loguniform_k(rng, k_lo, k_hi)
            vps = reveal.sample_viewpoints(k, rng)
            typ.append(0)
            cnt.append(k)
            rc_off.append(npairs)
            parts.append(vps)
            npairs += k
            out_off.append(nout)
            nout += k * k
            if len(states) < record_states and k <= 96:
                states.append((i, reveal.observed.copy()))
        t.append(now)
    rc = np.concatenate(parts).astype(np.int32).ravel() if parts else np.zeros(0, np.int32)
    tr = Trace(np.array(t), np.full(n, session_of, np.int32), np.array(typ, np.int32), np.array(cnt, np.int32),
               np.array(rc_off, np.int64), np.ascontiguousarray(rc), np.array(out_off, np.int64),
               np.zeros(max(nout, 1), np.uint16))
    tr.meta["states"] = states
    return tr


def latency_summary(trace: Trace, t_done: np.ndarray) -> dict:
    lat = t_done - trace.t
    q = trace.type == 0
    ql = lat[q]
    span = float(t_done.max() - trace.t.min()) if len(t_done) else 0.0
    return {
        "queries": int(q.sum()),
        "updates": int((~q).sum()),
        "span_s": span,
        "matrices_per_s": float(q.sum()) / span if span > 0 else 0.0,
        "p50_ms": float(np.percentile(ql, 50) * 1e3) if len(ql) else None,
        "p99_ms": float(np.percentile(ql, 99) * 1e3) if len(ql) else None,
        "max_ms": float(ql.max() * 1e3) if len(ql) else None,
        "update_p50_ms": float(np.percentile(lat[~q], 50) * 1e3) if (~q).any() else None,
        "update_p99_ms": float(np.percentile(lat[~q], 99) * 1e3) if (~q).any() else None,
    }


def merge_traces(traces) -> Trace:
    """Interleave several sessions' traces by arrival time (each session keeps
    its own request order)."""
    pair_base = out_base = 0
    parts = []
    for tr in traces:
        parts.append((tr, pair_base, out_base))
        pair_base += len(tr.rc) // 2
        out_base += len(tr.out)
    t = np.concatenate([tr.t for tr in traces])
    order = np.argsort(t, kind="stable")
    cat = lambda f: np.concatenate([getattr(tr, f) for tr in traces])  # noqa: E731
    rc_off = np.concatenate([tr.rc_off + pb for tr, pb, _ in parts])
    out_off = np.concatenate([np.where(tr.out_off >= 0, tr.out_off + ob, -1) for tr, _, ob in parts])
    merged = Trace(t[order], cat("session")[order].astype(np.int32), cat("type")[order].astype(np.int32),
                   cat("count")[order].astype(np.int32), rc_off[order].astype(np.int64),
                   np.ascontiguousarray(np.concatenate([tr.rc for tr in traces]), dtype=np.int32),
                   out_off[order].astype(np.int64), np.zeros(max(out_base, 1), np.uint16))
    return merged


class Exploration:
    """A fleet of robots exploring a ground truth from their start disks to
    the end: each robot reveals cells in its own BFS level order from its
    disk (grids.frontier_levels), skipping cells another robot revealed
    first.  Vectorised for long request streams; RobotReveal is the
    per-cell original used by the tests."""

    def __init__(self, truth: np.ndarray, n_robots: int, radius: int, seed: int, _orders=None):
        self.truth = truth
        self.cols = truth.shape[1]
        self.args = (n_robots, radius, seed)
        starts = grids.sample_free(truth, n_robots, seed)
        self.observed = grids.disk_observed(truth, [tuple(p) for p in starts], radius)
        if _orders is None:
            _orders = []
            for p in starts:
                o = grids.frontier_levels(truth, grids.disk_observed(truth, [tuple(p)], radius))
                _orders.append(o[:, 0].astype(np.int64) * self.cols + o[:, 1])
        self.orders = _orders
        self.ptr = [0] * n_robots
        self.obs_flat = (self.observed == grids.FREE).ravel().copy()
        n_free = int((truth == grids.FREE).sum())
        self.free_list = np.empty(n_free, dtype=np.int64)
        init = np.flatnonzero(self.obs_flat)
        self.free_list[:len(init)] = init
        self.n_free = len(init)
        self.total_free = n_free

    def fresh(self) -> "Exploration":
        """The same fleet back at its start disks (the robot orders are shared)."""
        return Exploration(self.truth, *self.args, _orders=self.orders)

    def advance_to(self, frac: float, batch: int = 143) -> None:
        """Reveal in robot round robin until `frac` of the free cells are observed."""
        r = 0
        while self.n_free < frac * self.total_free:
            if len(self.reveal(r % len(self.orders), batch)) == 0:
                break
            r += 1

    def observed_now(self) -> np.ndarray:
        """The observed grid at this point of the exploration."""
        obs = np.where(self.obs_flat.reshape(self.truth.shape), self.truth, grids.UNKNOWN)
        return obs.astype(np.uint8)

    def reveal(self, robot: int, m: int) -> np.ndarray:
        """Up to m newly revealed cells (flat indices) from robot `robot`, then
        from the next robots that still have cells; empty when all are done."""
        R = len(self.orders)
        got = []
        need = m
        for t in range(R):
            r = (robot + t) % R
            order, p = self.orders[r], self.ptr[r]
            while need > 0 and p < len(order):
                win = order[p:p + 4 * need + 16]
                new = win[~self.obs_flat[win]]
                if len(new) > need:
                    cut = int(np.searchsorted(np.cumsum(~self.obs_flat[win]), need))  # last window cell taken
                    win = win[:cut + 1]
                    new = win[~self.obs_flat[win]]
                p += len(win)
                self.obs_flat[new] = True
                got.append(new)
                need -= len(new)
            self.ptr[r] = p
            if need == 0:
                break
        cells = np.concatenate(got) if got else np.zeros(0, np.int64)
        self.free_list[self.n_free:self.n_free + len(cells)] = cells
        self.n_free += len(cells)
        return cells


def exploration_trace(sessions, explorations, n: int, rate: float, seed: int, update_frac: float = 0.1,
                      k_lo: int = 8, k_hi: int = 256, upd_lo: int = 16, upd_hi: int = 512, ring: int = 2048,
                      check_every: int = 1500):
    """Poisson stream of n requests spread uniformly over the sessions, each
    session exploring its own map (Exploration): with probability update_frac
    a request is an update batch of m ~ logU[upd_lo, upd_hi] cells from a
    random robot, else a queryAllPairs with k ~ logU[k_lo, k_hi] viewpoints
    drawn uniformly from the cells its session has revealed by then.  Once a
    session's map is fully revealed its update requests become queries.

    Results of most queries share a ring of `ring` result regions (the server
    still computes and delivers every matrix); every `check_every`-th query
    and the queries of the last ring window get regions of their own, listed
    in meta["checks"], so they can be compared afterwards."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    sess = rng.integers(0, len(sessions), n)
    upd = rng.random(n) < update_frac
    lk = np.log(k_lo) + rng.random(n) * (np.log(k_hi) - np.log(k_lo))
    ks = np.rint(np.exp(lk)).astype(np.int64)
    lu = np.log(upd_lo) + rng.random(n) * (np.log(upd_hi) - np.log(upd_lo))
    ms = np.rint(np.exp(lu)).astype(np.int64)
    robot = rng.integers(0, 1 << 30, n)
    typ = np.zeros(n, np.int32)
    cnt = np.zeros(n, np.int32)
    payload = [None] * n
    q_free = np.zeros(n, np.int64)  # revealed cells of the session when the query arrives
    upd_cells = [[] for _ in sessions]
    for i in range(n):
        ex = explorations[sess[i]]
        if upd[i]:
            cells = ex.reveal(int(robot[i] % len(ex.orders)), int(ms[i]))
            if len(cells):
                typ[i] = 1
                cnt[i] = len(cells)
                payload[i] = cells
                upd_cells[sess[i]].append(cells)
                continue
        cnt[i] = int(ks[i])
        q_free[i] = ex.n_free
    # viewpoints: uniform over the session's revealed cells at arrival time
    qi = np.flatnonzero(typ == 0)
    for s_idx, ex in enumerate(explorations):
        mine = qi[sess[qi] == s_idx]
        if len(mine) == 0:
            continue
        kk = cnt[mine].astype(np.int64)
        lim = np.repeat(q_free[mine], kk)
        pick = ex.free_list[(rng.random(int(kk.sum())) * lim).astype(np.int64)]
        parts = np.split(pick, np.cumsum(kk)[:-1])
        for j, i in enumerate(mine):
            payload[i] = parts[j]
    cols = explorations[0].cols
    rc_off = np.zeros(n, np.int64)
    rc_off[1:] = np.cumsum(cnt.astype(np.int64))[:-1]
    flat = np.concatenate([p for p in payload]) if n else np.zeros(0, np.int64)
    rc = np.stack([flat // cols, flat % cols], axis=1).astype(np.int32).ravel()
    # result regions: ring for most queries, own regions for the checked ones
    checks = [int(i) for j, i in enumerate(qi) if j % check_every == 0]
    last = [int(i) for i in qi[-ring:]]
    own = sorted(set(checks) | set(last))
    kmax2 = k_hi * k_hi
    out_off = np.full(n, -1, np.int64)
    ring_q = qi[~np.isin(qi, own)]
    out_off[ring_q] = (np.arange(len(ring_q)) % ring) * kmax2
    base = ring * kmax2
    for i in own:
        out_off[i] = base
        base += int(cnt[i]) * int(cnt[i])
    tr = Trace(t, np.asarray(sessions, np.int32)[sess], typ, cnt, rc_off, np.ascontiguousarray(rc), out_off,
               np.zeros(max(base, 1), np.uint16))
    tr.meta.update({"checks": own, "session_index": sess, "upd_cells": upd_cells,
                    "explored": [ex.n_free / ex.total_free for ex in explorations]})
    return tr


def state_at(init: np.ndarray, upd_cells, v: int) -> np.ndarray:
    """The observed grid after the first v update batches (flat cell lists)."""
    st = init.copy().ravel()
    for c in upd_cells[:v]:
        st[c] = grids.FREE
    return st.reshape(init.shape)
