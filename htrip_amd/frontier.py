"""Frontier viewpoints: the request producer of a multi-robot explorer.

Restates the reference's FrontierManager (include/htrip/frontier.hpp,
src/frontier.cpp) and GridMap::getClearance (include/htrip/grid_map.hpp:174-203)
on an observed cell array (0 Unknown, 1 Free, 2 Obstacle).  The clusters'
representatives are the viewpoints an explorer asks the distance oracle
about; tests/test_frontier.py pins every step against the reference's own
code.
"""
from __future__ import annotations

import math
from collections import deque
from dataclasses import dataclass, field

import numpy as np

UNKNOWN, FREE, OBSTACLE = 0, 1, 2


def clearance(obs: np.ndarray) -> np.ndarray:
    """Per cell (include/htrip/grid_map.hpp:174-203): 0 impassable; 1 when an
    8-neighbour is an obstacle or off the grid; 2 when a cell at Chebyshev
    distance 2 is; 3 otherwise.  Unknown cells do not count as obstacles."""
    R, C = obs.shape
    blocked = np.ones((R + 4, C + 4), dtype=bool)  # off-grid counts as an obstacle
    blocked[2:-2, 2:-2] = obs == OBSTACLE
    near1 = np.zeros((R, C), dtype=bool)
    near2 = np.zeros((R, C), dtype=bool)
    for dr in range(-2, 3):
        for dc in range(-2, 3):
            if dr == 0 and dc == 0:
                continue
            sh = blocked[2 + dr:2 + dr + R, 2 + dc:2 + dc + C]
            if max(abs(dr), abs(dc)) == 1:
                near1 |= sh
            else:
                near2 |= sh
    out = np.where(near1, 1, np.where(near2, 2, 3)).astype(np.uint8)
    out[obs != FREE] = 0
    return out


def frontier_cells(obs: np.ndarray, clr: np.ndarray | None = None) -> np.ndarray:
    """FrontierManager::isFrontierCell (src/frontier.cpp:12-89) for every cell:
    free, a 4-neighbour Unknown, at least 4 Unknown cells in the 5x5 patch,
    not a wall-hugging cell with little unknown space, and an Unknown
    4-neighbour that has at least two Unknown 4-neighbours of its own."""
    R, C = obs.shape
    if clr is None:
        clr = clearance(obs)
    pad = np.full((R + 4, C + 4), -1, dtype=np.int16)  # -1 = off the grid
    pad[2:-2, 2:-2] = obs
    unk = pad == UNKNOWN
    obst = (pad == OBSTACLE) | (pad == -1)

    def at(a, dr, dc):
        return a[2 + dr:2 + dr + R, 2 + dc:2 + dc + C]

    four = [(-1, 0), (1, 0), (0, -1), (0, 1)]
    has_unk4 = np.zeros((R, C), dtype=bool)
    for dr, dc in four:
        has_unk4 |= at(unk, dr, dc)
    n_unk = np.zeros((R, C), dtype=np.int32)
    n_obs = np.zeros((R, C), dtype=np.int32)
    for dr in range(-2, 3):
        for dc in range(-2, 3):
            if dr == 0 and dc == 0:
                continue
            n_unk += at(unk, dr, dc)
            n_obs += at(obst, dr, dc)
    # Unknown cells with >= 2 Unknown 4-neighbours (stage 4)
    deep = np.zeros((R + 4, C + 4), dtype=bool)
    cnt = np.zeros((R + 2, C + 2), dtype=np.int32)
    for dr, dc in four:
        cnt += unk[1 + dr:1 + dr + R + 2, 1 + dc:1 + dc + C + 2]
    deep[1:-1, 1:-1] = unk[1:-1, 1:-1] & (cnt >= 2)
    deep_adj = np.zeros((R, C), dtype=bool)
    for dr, dc in four:
        deep_adj |= at(deep, dr, dc)
    ok = (obs == FREE) & has_unk4 & (n_unk >= 4)
    ok &= ~((clr <= 1) & ((n_obs >= 6) | (n_unk < 6)))
    return ok & deep_adj


def _find_safe(obs, clr, base):
    """src/frontier.cpp:95-117: the cell itself when its clearance is >= 2,
    else the passable cell of the 7x7 patch with the highest clearance, the
    nearest among equals (first in row order on exact ties)."""
    r0, c0 = base
    if clr[r0, c0] >= 2:
        return base
    R, C = obs.shape
    best, best_clr, min_d = base, int(clr[r0, c0]), 1e9
    for dr in range(-3, 4):
        for dc in range(-3, 4):
            r, c = r0 + dr, c0 + dc
            if not (0 <= r < R and 0 <= c < C) or obs[r, c] != FREE:
                continue
            v = int(clr[r, c])
            d = math.hypot(dr, dc)
            if v > best_clr or (v == best_clr and d < min_d):
                best_clr, min_d, best = v, d, (r, c)
    return best


def medoid(obs: np.ndarray, clr: np.ndarray, cells: list[tuple[int, int]]) -> tuple[int, int]:
    """FrontierManager::computeMedoid (src/frontier.cpp:91-168)."""
    if not cells:
        return (0, 0)
    if len(cells) == 1:
        return _find_safe(obs, clr, cells[0])
    best = cells[0]
    if len(cells) <= 128:  # L1 medoid with a wall-proximity penalty
        min_cost = None
        for (r, c) in cells:
            if obs[r, c] != FREE:
                continue
            s = sum(abs(r - r2) + abs(c - c2) for r2, c2 in cells)
            v = int(clr[r, c])
            cost = s + (1000 if v <= 1 else 100 if v == 2 else 0)
            if min_cost is None or cost < min_cost:
                min_cost, best = cost, (r, c)
        return _find_safe(obs, clr, best)
    ar = sum(r for r, _ in cells) / len(cells)  # large clusters: nearest to the centroid
    ac = sum(c for _, c in cells) / len(cells)
    min_cost = None
    for (r, c) in cells:
        if obs[r, c] != FREE:
            continue
        v = int(clr[r, c])
        cost = (r - ar) ** 2 + (c - ac) ** 2 + (200.0 if v <= 1 else 20.0 if v == 2 else 0.0)
        if min_cost is None or cost < min_cost:
            min_cost, best = cost, (r, c)
    return _find_safe(obs, clr, best)


@dataclass
class FrontierCluster:
    id: int
    representative: tuple[int, int] = (0, 0)
    cells: list = field(default_factory=list)


def detect_and_cluster(obs: np.ndarray, min_size: int = 2, max_size: int = 16) -> list[FrontierCluster]:
    """FrontierManager::detectAndCluster (src/frontier.cpp:174-261): frontier
    cells in 8-connected components (row-major seeds, breadth-first order),
    components of at least min_size cells with a cell of clearance >= 2,
    split into runs of max_size cells (runs shorter than min_size dropped),
    each with its medoid as the representative."""
    obs = np.ascontiguousarray(obs, dtype=np.uint8)
    R, C = obs.shape
    clr = clearance(obs)
    front = frontier_cells(obs, clr)
    seen = np.zeros_like(front)
    out: list[FrontierCluster] = []
    cid = 0
    n8 = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]
    for r, c in zip(*np.nonzero(front)):
        r, c = int(r), int(c)
        if seen[r, c]:
            continue
        cl = FrontierCluster(cid)
        cid += 1
        q = deque([(r, c)])
        seen[r, c] = True
        while q:
            cr, cc = q.popleft()
            cl.cells.append((cr, cc))
            for dr, dc in n8:
                nr, nc = cr + dr, cc + dc
                if 0 <= nr < R and 0 <= nc < C and front[nr, nc] and not seen[nr, nc]:
                    seen[nr, nc] = True
                    q.append((nr, nc))
        if len(cl.cells) < min_size:
            continue
        if max(int(clr[p]) for p in cl.cells) <= 1:
            continue  # wall-trapped
        if max_size > 0 and len(cl.cells) > max_size:
            for s in range(0, len(cl.cells), max_size):
                part = cl.cells[s:s + max_size]
                if len(part) >= min_size:
                    sub = FrontierCluster(cid, cells=part)
                    cid += 1
                    sub.representative = medoid(obs, clr, part)
                    out.append(sub)
        else:
            cl.representative = medoid(obs, clr, cl.cells)
            out.append(cl)
    return out


def frontier_viewpoints(obs: np.ndarray, robots, k: int | None = None) -> np.ndarray:
    """Query viewpoints of an exploration step: the robots' positions, then the
    clusters' representatives (in cluster order), at most k in all."""
    vps = [tuple(int(v) for v in p) for p in robots]
    vps += [c.representative for c in detect_and_cluster(obs)]
    if k is not None:
        vps = vps[:k]
    return np.array(vps, dtype=np.int32).reshape(-1, 2)
