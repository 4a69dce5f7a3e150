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

from ._lib import INF, Context, GpuMap, HtgError, default_context, footprint  # noqa: F401

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

    The device workspace lives with the map; this object keeps the host-side
    result buffer so repeated calls allocate nothing new.
    """

    def __init__(self):
        self.out = np.empty(0, dtype=np.uint16)


class MinPlusOracle:
    """GPU MinPlusOracle with the reference's method names."""

    def __init__(self, config: HierarchyConfig | None = None, ctx: Context | None = None, budget: int = 0):
        self.config = config or HierarchyConfig()
        self.ctx = ctx
        self.budget = budget
        self.map: GpuMap | None = None
        self.S = 0
        self.H = 0

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
        return ok

    def depth(self) -> int:
        return self.H

    def onCellOpened(self, p, grid: GridMap) -> None:  # noqa: N802
        self.map.open_cell(int(p[0]), int(p[1]))

    def onBatchCellsOpened(self, cells, grid: GridMap) -> None:  # noqa: N802
        self.map.open_cells(cells)

    def queryAllPairs(self, viewpoints, grid: GridMap, out=None, scratch: HTripQueryScratch | None = None):  # noqa: N802
        m = self.map.query(viewpoints)
        if out is not None:
            out.resize(m.size, refcheck=False)
            out[:] = m.ravel()
            return out
        return m

    @property
    def last_batch_stats(self):
        return self.map.last_batch_stats()

    def totalMemoryBytes(self) -> int:  # noqa: N802
        return self.map.total_memory_bytes()
