"""htrip_amd -- MI355X-native HTRIP min-plus distance oracle.

Python mirror of the reference's hot-path interface
(include/htrip/minplus_oracle.hpp in the reference): ``HierarchyConfig``,
``MinPlusOracle.build / onCellOpened / onBatchCellsOpened / queryAllPairs``
over ``GridMap`` cell arrays, backed by libhtrip_gpu.so (hand-written gfx950
kernels behind the C ABI in include/htrip_gpu.h).  The C++20 drop-in with the
same names is htrip_amd/include/htrip/minplus_oracle.hpp.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import INF, Context, GpuMap, HtgError, QueryWorkspace, default_context, footprint  # noqa: F401

UNKNOWN, FREE, OBSTACLE = 0, 1, 2


@dataclass
class HierarchyConfig:
    """include/htrip/types.hpp:199-225."""

    tile_size_b: int = 16
    group_w: int = 2
    group_h: int = 2
    hybrid_k_threshold: int = 0

    def isValid(self) -> bool:  # noqa: N802 - reference name
        return (self.tile_size_b > 0 and self.group_w >= 2 and self.group_h >= 2
                and self.group_w == self.group_h and self.group_w * self.group_h <= 256)


class GridMap:
    """Observed occupancy layer (include/htrip/grid_map.hpp:31-421, the parts the oracle reads)."""

    def __init__(self, rows: int = 0, cols: int = 0, initial: int = UNKNOWN, cells: np.ndarray | None = None):
        if cells is not None:
            self.observed = np.ascontiguousarray(cells, dtype=np.uint8).copy()
        else:
            self.observed = np.full((rows, cols), initial, dtype=np.uint8)

    @property
    def rows(self) -> int:
        return self.observed.shape[0]

    @property
    def cols(self) -> int:
        return self.observed.shape[1]

    def inBounds(self, r: int, c: int) -> bool:  # noqa: N802
        return 0 <= r < self.rows and 0 <= c < self.cols

    def isPassable(self, r: int, c: int) -> bool:  # noqa: N802
        return self.inBounds(r, c) and self.observed[r, c] == FREE

    def openDoorway(self, r: int, c: int) -> bool:  # noqa: N802
        if not self.inBounds(r, c):
            return False
        self.observed[r, c] = FREE
        return True

    def blockCorridor(self, r: int, c: int) -> bool:  # noqa: N802
        if not self.inBounds(r, c):
            return False
        self.observed[r, c] = OBSTACLE
        return True


class HTripQueryScratch:
    """Caller-owned query workspace (include/htrip/minplus_oracle.hpp:189-238).

    It owns a device workspace (``QueryWorkspace``: stream, plan and lift
    buffers, result staging), created by the first query through it, so that
    queries with distinct scratch objects may run on one oracle from several
    threads at once, as the reference allows (ctypes releases the GIL during
    the call).
    """

    def __init__(self):
        self.ws: QueryWorkspace | None = None

    def workspace(self, ctx: Context | None) -> QueryWorkspace:
        if self.ws is None:
            self.ws = QueryWorkspace(ctx)
        return self.ws

    def close(self) -> None:
        if self.ws is not None:
            self.ws.close()
            self.ws = None


class MinPlusOracle:
    """GPU MinPlusOracle with the reference's method names
    (include/htrip/minplus_oracle.hpp:78-390); points are (r, c) pairs."""

    def __init__(self, config: HierarchyConfig | None = None, ctx: Context | None = None, budget: int = 0):
        self.config = config or HierarchyConfig()
        self.ctx = ctx
        self.budget = budget
        self.map: GpuMap | None = None
        self.S = 0
        self.H = 0
        self.num_leaf_tiles_x = self.num_leaf_tiles_y = 0
        self.query_scratch = HTripQueryScratch()
        self.update_counter = 0
        self.epoch = 0

    def build(self, grid: GridMap) -> bool:
        """src/minplus_oracle.cpp:48-137; False for an invalid config or empty grid."""
        if not self.config.isValid() or grid.rows <= 0 or grid.cols <= 0:
            return False
        if self.map is not None:
            self.map.close()
        self.map = GpuMap(grid.observed, self.config.tile_size_b, self.config.group_w, self.ctx, self.budget)
        ok = self.map.build()
        info = self.map.info()
        self.S, self.H = info.S, info.H
        self.num_leaf_tiles_x = self.num_leaf_tiles_y = info.num_leaf_tiles_x
        self.epoch += 1
        return ok

    def depth(self) -> int:
        return self.H

    def findLcaLevel(self, p1, p2) -> int:  # noqa: N802 - src/minplus_oracle.cpp:24-46
        b, g = self.config.tile_size_b, self.config.group_w
        ax, ay, bx, by = p1[1] // b, p1[0] // b, p2[1] // b, p2[0] // b
        if (ax, ay) == (bx, by):
            return 0
        for lvl in range(1, self.H + 1):
            ax, ay, bx, by = ax // g, ay // g, bx // g, by // g
            if (ax, ay) == (bx, by):
                return lvl
        return self.H + 1

    def onCellOpened(self, p, grid: GridMap) -> None:  # noqa: N802 - src/minplus_oracle.cpp:139-177
        if self.map is None or grid.observed[p[0], p[1]] == OBSTACLE:
            return
        self.map.open_cell(int(p[0]), int(p[1]))
        self.update_counter += 1

    def onBatchCellsOpened(self, cells, grid: GridMap) -> None:  # noqa: N802 - src/minplus_oracle.cpp:530-729
        if self.map is None:
            return
        cells = np.asarray(cells, dtype=np.int32).reshape(-1, 2)
        if len(cells) == 0:
            return
        inb = (cells[:, 0] >= 0) & (cells[:, 1] >= 0) & (cells[:, 0] < grid.rows) & (cells[:, 1] < grid.cols)
        keep = ~inb.copy()
        keep[inb] = grid.observed[cells[inb, 0], cells[inb, 1]] != OBSTACLE  # no edges into obstacles
        self.map.open_cells(cells[keep])
        self.update_counter += 1
        self.epoch += 1

    def onCellBlocked(self, p, grid: GridMap) -> None:  # noqa: N802 - src/minplus_oracle.cpp:247-302
        if self.map is None:
            return
        self.map.block_cells(np.array([p], dtype=np.int32))
        self.update_counter += 1
        self.epoch += 1

    def queryAllPairs(self, viewpoints, grid: GridMap, out=None,  # noqa: N802 - src/minplus_oracle.cpp:934-1426
                      scratch: HTripQueryScratch | None = None):
        """k x k matrix (row-major); with ``out`` (a 1-D uint16 array) the
        result is written there, resized to k*k, as the reference resizes its
        vector.  Runs on the scratch's device workspace (the oracle's own
        scratch when none is given)."""
        vps = np.asarray(viewpoints, dtype=np.int32).reshape(-1, 2)
        sc = scratch if scratch is not None else self.query_scratch
        if self.map is None:
            m = np.full((len(vps), len(vps)), INF, dtype=np.uint16)
        else:
            m = self.map.query_ws(sc.workspace(self.map.ctx), vps)
        if out is not None:
            out.resize(m.size, refcheck=False)
            out[:] = m.ravel()
            return out
        return m

    def queryDistance(self, s, t, grid: GridMap) -> int:  # noqa: N802 - src/minplus_oracle.cpp:829-928
        if not grid.isPassable(*s) or not grid.isPassable(*t):
            return INF
        if tuple(s) == tuple(t):
            return 0
        b = self.config.tile_size_b
        for p in (s, t):
            if p[0] < 0 or p[1] < 0 or p[1] // b >= self.num_leaf_tiles_x or p[0] // b >= self.num_leaf_tiles_y:
                return INF
        return int(self.queryAllPairs([s, t], grid)[0, 1])

    @property
    def last_batch_stats(self):
        return self.map.last_batch_stats()

    def totalMemoryBytes(self) -> int:  # noqa: N802 - src/minplus_oracle.cpp:2116-2134
        return self.map.total_memory_bytes() if self.map is not None else 0

    def effectiveMemoryBytes(self, grid: GridMap) -> int:  # noqa: N802 - src/minplus_oracle.cpp:2136-2182
        """Leaf D + S and node D_U + S_U bytes of the tiles inside ``grid``
        that hold a Free cell of it."""
        if self.map is None:
            return 0
        b = self.config.tile_size_b
        free = grid.observed == FREE
        total = 0

        def has_free(R0, C0, side):
            return R0 < grid.rows and C0 < grid.cols and bool(free[R0:R0 + side, C0:C0 + side].any())

        pl = 4 * b - 4
        for ty in range(self.num_leaf_tiles_y):
            for tx in range(self.num_leaf_tiles_x):
                if has_free(ty * b, tx * b, b):
                    total += (b ** 4 + pl * pl) * 2
        side = b
        for lvl in range(1, self.H + 1):
            side *= self.config.group_w
            nn, nv, ne = self.map.node_dims(lvl)
            nt = max(1, self.S // side)
            for ty in range(nt):
                for tx in range(nt):
                    if has_free(ty * side, tx * side, side):
                        total += (nv * nv + ne * ne) * 2
        return total
