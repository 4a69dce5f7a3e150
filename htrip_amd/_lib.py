"""ctypes binding of libhtrip_gpu.so (include/htrip_gpu.h).

The library is loaded from this package directory only; there is no CPU
fallback: if it is missing or the device is absent, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhtrip_gpu.so")
INF = 60000

HTG_OK, HTG_E_INVALID, HTG_E_NOMEM, HTG_E_HIP, HTG_E_STATE, HTG_E_RANGE, HTG_E_CANCELLED = 0, -1, -2, -3, -4, -5, -6
FAULT_UPDATE, FAULT_QUERY, FAULT_BUILD = 1, 2, 3

# Every symbol include/htrip_gpu.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "htg_ctx_create", "htg_ctx_destroy", "htg_last_error", "htg_ctx_mem_info",
    "htg_map_create", "htg_map_destroy", "htg_map_clone", "htg_map_footprint", "htg_map_info_get",
    "htg_build", "htg_open_cells", "htg_open_cell", "htg_block_cells", "htg_query_all_pairs", "htg_query_batch",
    "htg_lift_point", "htg_lift_root", "htg_inc_create", "htg_inc_destroy", "htg_query_all_pairs_incremental", "htg_last_batch_stats", "htg_sync",
    "htg_read_leaf_D", "htg_read_leaf_S", "htg_node_dims", "htg_read_node_DU", "htg_read_node_SU",
    "htg_profile_enable", "htg_profile_get",
    "htg_get_tuning", "htg_set_tuning", "htg_autotune",
    "htg_query_ws_create", "htg_query_ws_destroy", "htg_query_ws_bytes", "htg_query_all_pairs_ws",
    "htg_query_batch_ws", "htg_inject_fault", "htg_query_batch_noalloc",
]

K_LEAF_APSP, K_CLOSURE_PIVOTS, K_CLOSURE_UPDATE, K_Q_PROJECT, K_Q_MEET, K_Q_LIFT, K_COVER_RELAX, K_COVER_PAIRS = range(8)
KERNEL_NAMES = {
    K_LEAF_APSP: "k_leaf_apsp", K_CLOSURE_PIVOTS: "k_closure_pivots", K_CLOSURE_UPDATE: "k_closure_update",
    K_Q_PROJECT: "k_q_z", K_Q_MEET: "k_q_meet", K_Q_LIFT: "k_q_fold", K_COVER_RELAX: "k_cover_relax",
    K_COVER_PAIRS: "k_cover_pairs",
}


class HtgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"htg error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("tile_size_b", C.c_int32), ("group_w", C.c_int32), ("group_h", C.c_int32),
                ("hybrid_k_threshold", C.c_int32)]


class MapInfo(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("S", C.c_int32), ("H", C.c_int32),
                ("num_leaf_tiles_x", C.c_int32), ("num_leaves", C.c_int32), ("built", C.c_int32),
                ("reserved", C.c_int32), ("total_memory_bytes", C.c_uint64), ("device_bytes", C.c_uint64)]


class BatchStats(C.Structure):
    _fields_ = [("n_touched_leaves", C.c_int32), ("n_dirty_ancestors", C.c_int32), ("n_pivots", C.c_int32)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_double),
                ("updates", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("pivot_block", C.c_int32), ("zq_tile", C.c_int32), ("use_graph", C.c_int32),
                ("update_pivots", C.c_int32), ("update_rebuild", C.c_int32)]


_lib = None


def load(path: str = LIB_PATH):
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m htrip_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, i32, u64 = C.c_void_p, C.c_int32, C.c_uint64
    pi32, pu16, pu8 = C.POINTER(C.c_int32), C.POINTER(C.c_uint16), C.POINTER(C.c_uint8)
    sig = {
        "htg_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "htg_ctx_destroy": (C.c_int, [vp]),
        "htg_last_error": (C.c_char_p, []),
        "htg_ctx_mem_info": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "htg_map_create": (C.c_int, [vp, C.POINTER(Config), i32, i32, pu8, u64, C.POINTER(vp)]),
        "htg_map_destroy": (C.c_int, [vp]),
        "htg_map_clone": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "htg_map_footprint": (C.c_int, [C.POINTER(Config), i32, i32, C.POINTER(u64)]),
        "htg_map_info_get": (C.c_int, [vp, C.POINTER(MapInfo)]),
        "htg_build": (C.c_int, [vp]),
        "htg_open_cells": (C.c_int, [vp, pi32, i32]),
        "htg_open_cell": (C.c_int, [vp, i32, i32]),
        "htg_block_cells": (C.c_int, [vp, pi32, i32]),
        "htg_query_all_pairs": (C.c_int, [vp, pi32, i32, pu16]),
        "htg_query_batch": (C.c_int, [vp, pi32, pi32, i32, pu16]),
        "htg_lift_point": (C.c_int, [vp, i32, i32, pu16, C.c_int64, pi32, pi32, pi32, pi32]),
        "htg_lift_root": (C.c_int, [C.POINTER(vp), i32, i32, pu8]),
        "htg_inc_create": (C.c_int, [C.POINTER(vp)]),
        "htg_inc_destroy": (C.c_int, [vp]),
        "htg_query_all_pairs_incremental": (C.c_int, [vp, vp, pi32, i32, pu16, pi32]),
        "htg_last_batch_stats": (C.c_int, [vp, C.POINTER(BatchStats)]),
        "htg_sync": (C.c_int, [vp]),
        "htg_read_leaf_D": (C.c_int, [vp, i32, pu16]),
        "htg_read_leaf_S": (C.c_int, [vp, i32, pu16]),
        "htg_node_dims": (C.c_int, [vp, i32, pi32, pi32, pi32]),
        "htg_read_node_DU": (C.c_int, [vp, i32, i32, pu16]),
        "htg_read_node_SU": (C.c_int, [vp, i32, i32, pu16]),
        "htg_profile_enable": (C.c_int, [vp, C.c_int]),
        "htg_profile_get": (C.c_int, [vp, C.c_int, C.POINTER(KernelStats)]),
        "htg_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "htg_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "htg_autotune": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                   C.POINTER(Tuning)]),
        "htg_query_ws_create": (C.c_int, [vp, C.POINTER(vp)]),
        "htg_query_ws_destroy": (C.c_int, [vp]),
        "htg_query_ws_bytes": (C.c_int, [vp, C.POINTER(u64)]),
        "htg_query_all_pairs_ws": (C.c_int, [vp, vp, pi32, i32, pu16]),
        "htg_query_batch_ws": (C.c_int, [vp, vp, pi32, pi32, i32, pu16]),
        "htg_inject_fault": (C.c_int, [vp, i32, i32]),
        "htg_query_batch_noalloc": (C.c_int, [vp, pi32, pi32, i32, pu16]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != HTG_OK:
        msg = _lib.htg_last_error().decode() if _lib is not None else ""
        raise HtgError(rc, msg)


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """One device (htg_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.htg_ctx_create(device, C.byref(self.h)))

    def mem_info(self):
        f, t = C.c_uint64(), C.c_uint64()
        check(self.lib.htg_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def close(self):
        if self.h:
            self.lib.htg_ctx_destroy(self.h)
            self.h = C.c_void_p()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def footprint(b: int, g: int, rows: int, cols: int) -> int:
    lib = load()
    cfg = Config(b, g, g, 0)
    out = C.c_uint64()
    check(lib.htg_map_footprint(C.byref(cfg), rows, cols, C.byref(out)))
    return out.value


class IncState:
    """The caller-owned state of queryAllPairsIncremental (htg_inc)."""

    def __init__(self):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.htg_inc_create(C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.htg_inc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QueryWorkspace:
    """Caller-owned query workspace (htg_query_ws; the reference's
    HTripQueryScratch): its own stream and device buffers, so that queries
    through distinct workspaces run on one map from several threads."""

    def __init__(self, ctx: "Context | None" = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.h = C.c_void_p()
        check(self.lib.htg_query_ws_create(self.ctx.h, C.byref(self.h)))

    def device_bytes(self) -> int:
        n = C.c_uint64()
        check(self.lib.htg_query_ws_bytes(self.h, C.byref(n)))
        return n.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.htg_query_ws_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuMap:
    """One resident hierarchy (htg_map): the GPU MinPlusOracle over a cell array.

    Method names mirror oracle/oracle.py so parity tests read the same against
    either side.
    """

    def __init__(self, cells: np.ndarray, b: int, g: int = 2, ctx: Context | None = None, budget: int = 0,
                 _handle=None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.b, self.g = b, g
        self.cells = np.ascontiguousarray(cells, dtype=np.uint8).copy()
        if _handle is not None:
            self.h = _handle
            return
        self.h = C.c_void_p()
        cfg = Config(b, g, g, 0)
        check(self.lib.htg_map_create(self.ctx.h, C.byref(cfg), self.cells.shape[0], self.cells.shape[1],
                                      ptr(self.cells, C.c_uint8), budget, C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.htg_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clone(self, budget: int = 0) -> "GpuMap":
        h = C.c_void_p()
        check(self.lib.htg_map_clone(self.h, budget, C.byref(h)))
        return GpuMap(self.cells, self.b, self.g, self.ctx, _handle=h)

    def build(self) -> bool:
        rc = self.lib.htg_build(self.h)
        if rc == HTG_E_INVALID:
            return False
        check(rc)
        return True

    def open_cells(self, rc) -> None:
        """onBatchCellsOpened.  self.cells follows the map's cell state: it
        takes the batch unless the map refused the call (HTG_E_STATE), also
        when the update failed part way (the map's state holds the batch)."""
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        code = self.lib.htg_open_cells(self.h, ptr(rc, C.c_int32), len(rc))
        if code != HTG_E_STATE:
            ok = (rc[:, 0] >= 0) & (rc[:, 1] >= 0) & (rc[:, 0] < self.cells.shape[0]) & \
                (rc[:, 1] < self.cells.shape[1])
            self.cells[rc[ok, 0], rc[ok, 1]] = 1
        check(code)

    def open_cell(self, r: int, c: int) -> None:
        code = self.lib.htg_open_cell(self.h, r, c)
        if code != HTG_E_STATE and 0 <= r < self.cells.shape[0] and 0 <= c < self.cells.shape[1]:
            self.cells[r, c] = 1
        check(code)

    def block_cells(self, rc) -> None:
        """onCellBlocked for each listed cell: passable cells become Obstacle."""
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        ok = (rc[:, 0] >= 0) & (rc[:, 1] >= 0) & (rc[:, 0] < self.cells.shape[0]) & (rc[:, 1] < self.cells.shape[1])
        r, c = rc[ok, 0], rc[ok, 1]
        self.cells[r, c] = np.where(self.cells[r, c] == 1, 2, self.cells[r, c])
        check(self.lib.htg_block_cells(self.h, ptr(rc, C.c_int32), len(rc)))

    def query(self, vps) -> np.ndarray:
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        k = len(vps)
        out = np.empty((k, k), dtype=np.uint16)
        check(self.lib.htg_query_all_pairs(self.h, ptr(vps, C.c_int32), k, ptr(out, C.c_uint16)))
        return out

    def query_ws(self, ws: QueryWorkspace, vps) -> np.ndarray:
        """queryAllPairs through a caller-owned workspace (thread-safe across workspaces)."""
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        k = len(vps)
        out = np.empty((k, k), dtype=np.uint16)
        check(self.lib.htg_query_all_pairs_ws(self.h, ws.h, ptr(vps, C.c_int32), k, ptr(out, C.c_uint16)))
        return out

    def inject_fault(self, site: int, countdown: int = 1) -> None:
        check(self.lib.htg_inject_fault(self.h, site, countdown))

    def query_batch(self, vps_list, ws: QueryWorkspace | None = None) -> list:
        ks = np.array([len(v) for v in vps_list], dtype=np.int32)
        allv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32).reshape(-1, 2) for v in vps_list])
                                    if len(vps_list) else np.zeros((0, 2), np.int32), dtype=np.int32)
        out = np.empty(int((ks.astype(np.int64) ** 2).sum()), dtype=np.uint16)
        if ws is None:
            check(self.lib.htg_query_batch(self.h, ptr(allv, C.c_int32), ptr(ks, C.c_int32), len(ks),
                                           ptr(out, C.c_uint16)))
        else:
            check(self.lib.htg_query_batch_ws(self.h, ws.h, ptr(allv, C.c_int32), ptr(ks, C.c_int32), len(ks),
                                              ptr(out, C.c_uint16)))
        res, o = [], 0
        for k in ks:
            res.append(out[o:o + k * k].reshape(k, k))
            o += int(k) * int(k)
        return res

    def lift_root(self, cells: np.ndarray) -> None:
        """liftRoot over the grown grid ``cells``: the square grows by g and
        gains a level.  Inside the old square the map keeps its own state."""
        cells = np.ascontiguousarray(cells, dtype=np.uint8)
        S0 = self.S
        r0, c0 = self.cells.shape
        h = C.c_void_p(self.h.value)
        check(self.lib.htg_lift_root(C.byref(h), cells.shape[0], cells.shape[1], ptr(cells, C.c_uint8)))
        self.h = h
        merged = cells.copy()
        band = merged[:S0, :S0]
        band[band == 1] = 0  # the old square's cells beyond the old grid stay impassable
        merged[:r0, :c0] = self.cells
        self.cells = merged

    def query_incremental(self, state: "IncState", vps):
        """queryAllPairsIncremental: (k x k matrix, recomputed) where recomputed
        is k when the device ran the query and 0 when the cached matrix was
        still exact."""
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        k = len(vps)
        out = np.empty((k, k), dtype=np.uint16)
        n = C.c_int32()
        check(self.lib.htg_query_all_pairs_incremental(self.h, state.h, ptr(vps, C.c_int32), k,
                                                       ptr(out, C.c_uint16), C.byref(n)))
        return out, n.value

    def lift_point(self, r: int, c: int):
        """liftPoint: (X, parent_node_idx, child_quad, child_offset); X[l] is the
        distance vector of level l (leaf ports, then each ancestor's exterior)."""
        H = self.H
        cap = 4 * self.b - 4 + sum(self.node_dims(l)[2] for l in range(1, H + 1))
        x = np.empty(cap, dtype=np.uint16)
        lens, par, quad, off = (np.empty(H + 1, dtype=np.int32) for _ in range(4))
        check(self.lib.htg_lift_point(self.h, int(r), int(c), ptr(x, C.c_uint16), cap, ptr(lens, C.c_int32),
                                      ptr(par, C.c_int32), ptr(quad, C.c_int32), ptr(off, C.c_int32)))
        o = np.concatenate([[0], np.cumsum(lens)])
        return [x[o[l]:o[l + 1]] for l in range(H + 1)], par, quad, off

    def info(self) -> MapInfo:
        i = MapInfo()
        check(self.lib.htg_map_info_get(self.h, C.byref(i)))
        return i

    @property
    def S(self) -> int:
        return self.info().S

    @property
    def H(self) -> int:
        return self.info().H

    def num_leaves(self) -> int:
        return self.info().num_leaves

    def node_dims(self, lvl: int):
        n, v, e = C.c_int32(), C.c_int32(), C.c_int32()
        check(self.lib.htg_node_dims(self.h, lvl, C.byref(n), C.byref(v), C.byref(e)))
        return n.value, v.value, e.value

    def num_nodes(self, lvl: int) -> int:
        return self.node_dims(lvl)[0]

    def leaf_D(self, idx: int) -> np.ndarray:
        n = self.b * self.b
        out = np.empty((n, n), dtype=np.uint16)
        check(self.lib.htg_read_leaf_D(self.h, idx, ptr(out, C.c_uint16)))
        return out

    def leaf_S(self, idx: int) -> np.ndarray:
        p = 4 * self.b - 4
        out = np.empty((p, p), dtype=np.uint16)
        check(self.lib.htg_read_leaf_S(self.h, idx, ptr(out, C.c_uint16)))
        return out

    def node_DU(self, lvl: int, idx: int) -> np.ndarray:
        _, nv, _ = self.node_dims(lvl)
        out = np.empty((nv, nv), dtype=np.uint16)
        check(self.lib.htg_read_node_DU(self.h, lvl, idx, ptr(out, C.c_uint16)))
        return out

    def node_SU(self, lvl: int, idx: int) -> np.ndarray:
        _, _, ne = self.node_dims(lvl)
        out = np.empty((ne, ne), dtype=np.uint16)
        check(self.lib.htg_read_node_SU(self.h, lvl, idx, ptr(out, C.c_uint16)))
        return out

    def last_batch_stats(self):
        s = BatchStats()
        check(self.lib.htg_last_batch_stats(self.h, C.byref(s)))
        return (s.n_touched_leaves, s.n_dirty_ancestors, s.n_pivots)

    def total_memory_bytes(self) -> int:
        return int(self.info().total_memory_bytes)

    def tuning(self) -> Tuning:
        t = Tuning()
        check(self.lib.htg_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        check(self.lib.htg_set_tuning(self.h, C.byref(t)))

    def autotune(self, sample=(), reps: int = 3) -> Tuning:
        """htg_autotune: pick pivot block, graph capture and projection tile by timing."""
        ks = np.array([len(v) for v in sample], dtype=np.int32)
        allv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32).reshape(-1, 2) for v in sample])
                                    if len(sample) else np.zeros((0, 2), np.int32), dtype=np.int32)
        t = Tuning()
        check(self.lib.htg_autotune(self.h, ptr(allv, C.c_int32), ptr(ks, C.c_int32), len(ks), reps, C.byref(t)))
        return t

    def sync(self) -> None:
        check(self.lib.htg_sync(self.h))

    def profile(self, on: bool = True, every: int = 1) -> None:
        """Kernel timing; every > 1 samples one operation (query batch, update, build) in `every`."""
        check(self.lib.htg_profile_enable(self.h, max(1, every) if on else 0))

    def kernel_stats(self, cls: int) -> dict:
        s = KernelStats()
        check(self.lib.htg_profile_get(self.h, cls, C.byref(s)))
        return {"launches": int(s.launches), "total_ms": s.total_ms, "bytes": s.bytes, "updates": s.updates}
