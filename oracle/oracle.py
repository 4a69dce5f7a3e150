"""TEST INFRASTRUCTURE ONLY -- ctypes access to the CPU oracle libraries.

``Oracle("port")`` loads liboracle.so (this repository's restatement of the
reference algorithm, htrip_oracle.cpp); ``Oracle("reference")`` loads
_ref/libhtrip_ref.so (the reference's own sources compiled by the Makefile
here).  Both expose the same calls.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"port": os.path.join(HERE, "liboracle.so"), "reference": os.path.join(HERE, "_ref", "libhtrip_ref.so")}
PREFIX = {"port": "orc_", "reference": "ref_"}
INF = 60000

_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_libs: dict = {}


def build(reference_root: str = "/root/reference") -> None:
    """Compile liboracle.so (and _ref/ when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, f"REF={reference_root}", "-j8"], check=True)


def available(kind: str) -> bool:
    return os.path.exists(PATHS[kind])


def _cpu_has_avx2() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return " avx2" in f.read()
    except OSError:
        return False


def _lib(kind: str):
    if kind in _libs:
        return _libs[kind]
    if kind == "reference" and not _cpu_has_avx2():
        raise RuntimeError("reference build needs an AVX2 host")
    lib = C.CDLL(PATHS[kind])
    p = PREFIX[kind]

    def fn(name, res, *args):
        f = getattr(lib, p + name)
        f.restype = res
        f.argtypes = list(args)
        return f

    vp = C.c_void_p
    ci = C.c_int
    api = {
        "create": fn("create", vp, ci, ci, ci, ci, _u8p),
        "destroy": fn("destroy", None, vp),
        "build": fn("build", ci, vp),
        "open_cells": fn("open_cells", None, vp, _i32p, ci),
        "open_cell": fn("open_cell", None, vp, ci, ci),
        "query": fn("query", None, vp, _i32p, ci, _u16p),
        "bfs": fn("bfs", None, vp, _i32p, ci, _u16p),
        "S": fn("S", ci, vp),
        "H": fn("H", ci, vp),
        "num_leaves": fn("num_leaves", ci, vp),
        "num_nodes": fn("num_nodes", ci, vp, ci),
        "node_nV": fn("node_nV", ci, vp, ci, ci),
        "node_nExt": fn("node_nExt", ci, vp, ci, ci),
        "leaf_D": fn("leaf_D", None, vp, ci, _u16p),
        "leaf_S": fn("leaf_S", None, vp, ci, _u16p),
        "node_DU": fn("node_DU", None, vp, ci, ci, _u16p),
        "node_SU": fn("node_SU", None, vp, ci, ci, _u16p),
        "last_batch_stats": fn("last_batch_stats", None, vp, _i32p),
        "total_memory_bytes": fn("total_memory_bytes", C.c_uint64, vp),
        "lift_point": fn("lift_point", None, vp, ci, ci, _u16p, _i32p, _i32p, _i32p, _i32p, _i32p),
    }
    if kind == "port":
        api["region_bfs"] = fn("region_bfs", None, vp, ci, ci, ci, ci, ci, ci, _u16p)
        api["grid_bfs"] = fn("grid_bfs", None, ci, ci, _u8p, _i32p, ci, _u16p)
        api["grid_bfs_mt"] = fn("grid_bfs_mt", None, ci, ci, _u8p, _i32p, ci, _u16p, ci)
    else:
        api["import_begin"] = fn("import_begin", None, vp)
        api["import_leaf"] = fn("import_leaf", None, vp, ci, _u16p)
        api["import_node"] = fn("import_node", None, vp, ci, ci, _u16p)
        api["time_queries"] = fn("time_queries", C.c_double, vp, _i32p, _i32p, ci, ci, C.c_void_p)
        api["time_lifts"] = fn("time_lifts", C.c_double, vp, _i32p, ci)
        api["load_movingai"] = fn("load_movingai", ci, C.c_char_p, _i32p, _i32p, _u8p, C.c_int64)
        api["world"] = fn("world", ci, ci, ci, ci, C.c_uint64, ci, ci, C.c_double, C.c_double, C.c_uint64, _u8p,
                          _u8p)
        api["lift_root"] = fn("lift_root", None, vp, ci, ci, _u8p)
        api["frontiers"] = fn("frontiers", ci, ci, ci, _u8p, ci, ci, _i32p, ci, _i32p, ci, _u8p, _u8p)
        api["inc_create"] = fn("inc_create", vp)
        api["inc_destroy"] = fn("inc_destroy", None, vp)
        api["inc_query"] = fn("inc_query", C.c_double, vp, vp, _i32p, ci, _u16p)
        api["effective_memory_bytes"] = fn("effective_memory_bytes", C.c_uint64, vp, ci, ci, _u8p)
    _libs[kind] = api
    return api


def reference_load_movingai(path: str):
    """The compiled reference's GridMap::loadMovingAI on a file: its ground
    truth codes, or None where it returns false."""
    api = _lib("reference")
    rows, cols = np.zeros(1, np.int32), np.zeros(1, np.int32)
    if not api["load_movingai"](path.encode(), rows, cols, np.zeros(1, np.uint8), 0):
        return None
    out = np.empty(int(rows[0]) * int(cols[0]), dtype=np.uint8)
    api["load_movingai"](path.encode(), rows, cols, out, out.size)
    return out.reshape(int(rows[0]), int(cols[0]))


def reference_world(rows: int, cols: int, room: int, seed: int, start=(0, 0), coverage: float = 0.0,
                    island: float = 0.1, seed2: int = 12345):
    """The compiled reference's GridMap::generateProceduralRooms and, when
    coverage > 0, initExplorationSnapshot: (ok, truth codes, observed codes)."""
    api = _lib("reference")
    truth = np.empty(rows * cols, dtype=np.uint8)
    obs = np.empty(rows * cols, dtype=np.uint8)
    ok = api["world"](rows, cols, room, seed, start[0], start[1], coverage, island, seed2, truth, obs)
    return bool(ok), truth.reshape(rows, cols), obs.reshape(rows, cols)


def reference_frontiers(obs: np.ndarray, min_size: int = 2, max_size: int = 16):
    """The compiled reference's FrontierManager on an observed grid:
    (clearance, frontier flags, [(id, (r, c), cells)])."""
    api = _lib("reference")
    obs = np.ascontiguousarray(obs, dtype=np.uint8)
    R, Cn = obs.shape
    clr = np.empty_like(obs)
    front = np.empty_like(obs)
    cap = R * Cn
    meta = np.empty(4 * cap, dtype=np.int32)
    cells = np.empty(2 * cap, dtype=np.int32)
    n = api["frontiers"](R, Cn, obs, min_size, max_size, meta, cap, cells, cap, clr, front)
    out, w = [], 0
    for i in range(n):
        cid, r, c, sz = (int(v) for v in meta[4 * i:4 * i + 4])
        out.append((cid, (r, c), [tuple(int(v) for v in cells[2 * j:2 * j + 2]) for j in range(w, w + sz)]))
        w += sz
    return clr, front.astype(bool), out


def bfs_threads() -> int:
    """Host threads for the BFS checker: the CPU share (at most 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def grid_bfs(cells: np.ndarray, vps: np.ndarray, threads: int = 1) -> np.ndarray:
    """Global BFS all-pairs on a cell array: min(d, INF), INF for impassable
    endpoints; threads > 1 spreads the sources over host threads (same rows)."""
    api = _lib("port")
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
    k = len(vps)
    out = np.empty((k, k), dtype=np.uint16)
    if threads > 1:
        api["grid_bfs_mt"](cells.shape[0], cells.shape[1], cells, vps, k, out, threads)
    else:
        api["grid_bfs"](cells.shape[0], cells.shape[1], cells, vps, k, out)
    return out


class Oracle:
    """One CPU hierarchy (port or reference) over a cell array."""

    def __init__(self, kind: str, cells: np.ndarray, b: int, g: int = 2):
        self.kind = kind
        self.api = _lib(kind)
        self.cells = np.ascontiguousarray(cells, dtype=np.uint8).copy()
        self.b, self.g = b, g
        self.h = self.api["create"](b, g, self.cells.shape[0], self.cells.shape[1], self.cells)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.api["destroy"](h)
            self.h = None

    def build(self) -> bool:
        return bool(self.api["build"](self.h))

    def open_cells(self, rc) -> None:
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        ok = (rc[:, 0] >= 0) & (rc[:, 1] >= 0) & (rc[:, 0] < self.cells.shape[0]) & (rc[:, 1] < self.cells.shape[1])
        self.cells[rc[ok, 0], rc[ok, 1]] = 1
        self.api["open_cells"](self.h, rc, len(rc))

    def open_cell(self, r: int, c: int) -> None:
        if 0 <= r < self.cells.shape[0] and 0 <= c < self.cells.shape[1]:
            self.cells[r, c] = 1
        self.api["open_cell"](self.h, r, c)

    def query(self, vps) -> np.ndarray:
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        k = len(vps)
        out = np.empty((k, k), dtype=np.uint16)
        self.api["query"](self.h, vps, k, out)
        return out

    def bfs(self, vps) -> np.ndarray:
        vps = np.ascontiguousarray(vps, dtype=np.int32).reshape(-1, 2)
        k = len(vps)
        out = np.empty((k, k), dtype=np.uint16)
        self.api["bfs"](self.h, vps, k, out)
        return out

    @property
    def S(self) -> int:
        return self.api["S"](self.h)

    @property
    def H(self) -> int:
        return self.api["H"](self.h)

    def num_leaves(self) -> int:
        return self.api["num_leaves"](self.h)

    def num_nodes(self, lvl: int) -> int:
        return self.api["num_nodes"](self.h, lvl)

    def leaf_D(self, idx: int) -> np.ndarray:
        n = self.b * self.b
        out = np.empty((n, n), dtype=np.uint16)
        self.api["leaf_D"](self.h, idx, out)
        return out

    def leaf_S(self, idx: int) -> np.ndarray:
        p = 4 * self.b - 4
        out = np.empty((p, p), dtype=np.uint16)
        self.api["leaf_S"](self.h, idx, out)
        return out

    def node_DU(self, lvl: int, idx: int) -> np.ndarray:
        n = self.api["node_nV"](self.h, lvl, idx)
        out = np.empty((n, n), dtype=np.uint16)
        self.api["node_DU"](self.h, lvl, idx, out)
        return out

    def node_SU(self, lvl: int, idx: int) -> np.ndarray:
        n = self.api["node_nExt"](self.h, lvl, idx)
        out = np.empty((n, n), dtype=np.uint16)
        self.api["node_SU"](self.h, lvl, idx, out)
        return out

    def last_batch_stats(self):
        out = np.zeros(3, dtype=np.int32)
        self.api["last_batch_stats"](self.h, out)
        return tuple(int(x) for x in out)

    def total_memory_bytes(self) -> int:
        return int(self.api["total_memory_bytes"](self.h))

    def lift_point(self, r: int, c: int):
        """liftPoint: (X, parent_node_idx, child_quad, child_offset, active) with
        X[l] and active[l] per level (src/minplus_oracle.cpp:731-823)."""
        H = self.H
        cap = 4 * self.b - 4 + sum(self.api["node_nExt"](self.h, l, 0) for l in range(1, H + 1))
        x = np.empty(cap, dtype=np.uint16)
        act = np.empty(cap, dtype=np.int32)
        lens, par, quad, off = (np.empty(H + 1, dtype=np.int32) for _ in range(4))
        self.api["lift_point"](self.h, int(r), int(c), x, lens, par, quad, off, act)
        o = np.concatenate([[0], np.cumsum(lens)])
        X = [x[o[l]:o[l + 1]].copy() for l in range(H + 1)]
        A = [act[o[l]:o[l + 1]][act[o[l]:o[l + 1]] >= 0].copy() for l in range(H + 1)]
        return X, par, quad, off, A

    def lift_root(self, cells: np.ndarray) -> None:
        """liftRoot over the grown grid (compiled reference only)."""
        self.cells = np.ascontiguousarray(cells, dtype=np.uint8).copy()
        self.api["lift_root"](self.h, self.cells.shape[0], self.cells.shape[1], self.cells)

    def time_lifts(self, rc) -> float:
        """Seconds for liftPoint on each listed cell (compiled reference only)."""
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        return float(self.api["time_lifts"](self.h, rc, len(rc)))

    def region_bfs(self, R0: int, C0: int, h: int, w: int, sr: int, sc: int) -> np.ndarray:
        out = np.empty((h, w), dtype=np.uint16)
        _lib("port")["region_bfs"](self.h, R0, C0, h, w, sr, sc, out)
        return out
