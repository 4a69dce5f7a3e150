"""Seeded synthetic occupancy grids and request streams.

The reference ships no fixture grids for its hot path; its tests and benches
build grids in code (apps/test_exactness.cpp:370-415, src/grid_map.cpp:78-217)
or load MovingAI maps (src/grid_map.cpp:16-76).  These generators produce the
workloads the benchmarks are quoted on, with a portable PRNG (SplitMix64 with
an explicit integer mapping) so every grid replays bit for bit on any host:

* ``street_grid``  -- free background, axis-aligned obstacle blocks of side
  4..40 until about 30% is blocked, only the largest 4-connected free
  component kept free.
* ``maze_grid``    -- randomized depth-first (recursive backtracker) maze with
  corridor and wall width ``w``.
* ``reveal_order`` -- an initial observed disk plus the frontier order in
  which the remaining free cells are revealed (BFS over ground truth from the
  observed set).
* ``sample_free``  -- distinct viewpoints drawn uniformly from free cells.
* ``load_movingai`` -- a MovingAI .map file (or .map.gz) as ground truth,
  parsed as the reference's GridMap::loadMovingAI does.

Cell codes follow include/htrip/types.hpp:101-105: 0 Unknown, 1 Free,
2 Obstacle.  Grids are ``numpy.uint8`` arrays of shape (rows, cols).
"""
from __future__ import annotations

import numpy as np

UNKNOWN, FREE, OBSTACLE = 0, 1, 2
_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); ``below(n)`` maps by 128-bit multiply-high."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return (self.next() * n) >> 64


def _largest_component(free: np.ndarray) -> np.ndarray:
    from scipy import ndimage

    lab, n = ndimage.label(free, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    if n == 0:
        return free
    counts = np.bincount(lab.ravel())
    counts[0] = 0
    return lab == int(np.argmax(counts))


def street_grid(n: int, seed: int, blocked_frac: float = 0.30, side_lo: int = 4, side_hi: int = 40) -> np.ndarray:
    """Street-like grid: obstacle rectangles until ``blocked_frac`` is blocked."""
    rng = SplitMix64(seed)
    blocked = np.zeros((n, n), dtype=bool)
    target = int(blocked_frac * n * n)
    count = 0
    while count < target:
        h = side_lo + rng.below(side_hi - side_lo + 1)
        w = side_lo + rng.below(side_hi - side_lo + 1)
        r = rng.below(n)
        c = rng.below(n)
        sl = blocked[r:r + h, c:c + w]
        count += int(sl.size - np.count_nonzero(sl))
        sl[...] = True
    free = _largest_component(~blocked)
    return np.where(free, FREE, OBSTACLE).astype(np.uint8)


def maze_grid(n: int, seed: int, width: int = 1) -> np.ndarray:
    """Recursive-backtracker maze; lattice cells and carved walls are width x width blocks."""
    period = 2 * width
    L = n // period
    rng = SplitMix64(seed)
    free = np.zeros((n, n), dtype=bool)
    visited = np.zeros((L, L), dtype=bool)

    def carve(i, j):
        free[period * i:period * i + width, period * j:period * j + width] = True

    stack = [(0, 0)]
    visited[0, 0] = True
    carve(0, 0)
    while stack:
        i, j = stack[-1]
        nb = []
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            a, b = i + di, j + dj
            if 0 <= a < L and 0 <= b < L and not visited[a, b]:
                nb.append((a, b))
        if not nb:
            stack.pop()
            continue
        a, b = nb[rng.below(len(nb))]
        visited[a, b] = True
        carve(a, b)
        # the wall block between (i, j) and (a, b)
        r0 = period * min(i, a) + (width if a != i else 0)
        c0 = period * min(j, b) + (width if b != j else 0)
        free[r0:r0 + width, c0:c0 + width] = True
        stack.append((a, b))
    return np.where(free, FREE, OBSTACLE).astype(np.uint8)


def sample_free(cells: np.ndarray, k: int, seed: int) -> np.ndarray:
    """``k`` distinct Free cells, uniform, as an int32 array of (r, c) rows."""
    idx = np.flatnonzero(cells.ravel() == FREE)
    if len(idx) == 0:
        return np.zeros((0, 2), dtype=np.int32)
    k = min(k, len(idx))
    rng = SplitMix64(seed)
    picked, seen = [], set()
    while len(picked) < k:
        t = int(idx[rng.below(len(idx))])
        if t not in seen:
            seen.add(t)
            picked.append(t)
    cols = cells.shape[1]
    return np.array([(t // cols, t % cols) for t in picked], dtype=np.int32).reshape(-1, 2)


def disk_observed(truth: np.ndarray, centers, radius: int) -> np.ndarray:
    """Observed layer: ground truth inside the disks, Unknown outside."""
    rows, cols = truth.shape
    inside = np.zeros_like(truth, dtype=bool)
    for (r, c) in centers:  # each disk only touches its bounding box
        r0, r1 = max(0, int(r) - radius), min(rows, int(r) + radius + 1)
        c0, c1 = max(0, int(c) - radius), min(cols, int(c) + radius + 1)
        if r0 >= r1 or c0 >= c1:
            continue
        rr, cc = np.mgrid[r0:r1, c0:c1]
        inside[r0:r1, c0:c1] |= (rr - r) ** 2 + (cc - c) ** 2 <= radius * radius
    return np.where(inside, truth, UNKNOWN).astype(np.uint8)


def frontier_order(truth: np.ndarray, observed: np.ndarray, limit: int | None = None) -> np.ndarray:
    """Free-but-unobserved cells in BFS order from the observed free set (int32 (n, 2)),
    at most ``limit`` of them."""
    from collections import deque

    rows, cols = truth.shape
    tf = (truth == FREE).ravel()
    seen = (observed == FREE).ravel().copy()
    q = deque(int(i) for i in np.flatnonzero(seen))
    order = []
    while q:
        u = q.popleft()
        r, c = divmod(u, cols)
        for v, ok in ((u - cols, r > 0), (u + cols, r < rows - 1), (u - 1, c > 0), (u + 1, c < cols - 1)):
            if ok and tf[v] and not seen[v]:
                seen[v] = True
                order.append(v)
                q.append(v)
        if limit is not None and len(order) >= limit:
            break
    o = np.array(order, dtype=np.int64)
    return np.stack([o // cols, o % cols], axis=1).astype(np.int32) if len(o) else np.zeros((0, 2), np.int32)


def frontier_levels(truth: np.ndarray, observed: np.ndarray) -> np.ndarray:
    """Every free-but-unobserved cell in BFS level order from the observed free
    set (cells of one level in ascending row-major index), int32 (n, 2).
    The same level sets as ``frontier_order`` without the queue's order inside
    a level; vectorised per level, for whole-map reveals."""
    rows, cols = truth.shape
    free = (truth == FREE).ravel()
    seen = (observed == FREE).ravel().copy()
    front = np.flatnonzero(seen)
    parts = []
    while len(front):
        r, c = front // cols, front % cols
        nb = np.concatenate([front[r > 0] - cols, front[r < rows - 1] + cols, front[c > 0] - 1,
                             front[c < cols - 1] + 1])
        nb = np.unique(nb[free[nb] & ~seen[nb]])
        seen[nb] = True
        if len(nb):
            parts.append(nb)
        front = nb
    o = np.concatenate(parts) if parts else np.zeros(0, np.int64)
    return np.stack([o // cols, o % cols], axis=1).astype(np.int32)


def reveal_scenario(truth: np.ndarray, seed: int, radius: int = 64):
    """Initial observed disk around a seeded free start, plus the frontier reveal order."""
    start = sample_free(truth, 1, seed)[0]
    obs = disk_observed(truth, [tuple(start)], radius)
    return obs, frontier_order(truth, obs)


def _second_int(line: bytes) -> int:
    """The integer a stream extraction reads after the first token (0 when it fails)."""
    import re

    parts = line.split(None, 1)
    m = re.match(rb"[+-]?\d+", parts[1]) if len(parts) > 1 else None
    return int(m.group(0)) if m else 0


def load_movingai(src) -> np.ndarray:
    """Ground truth of a MovingAI .map (a path, ``.map.gz`` path, bytes or str)
    as cell codes, following GridMap::loadMovingAI (src/grid_map.cpp:16-76):
    header lines up to one that is exactly ``map`` (trailing CR, space and tab
    ignored), where lines starting with ``height`` / ``width`` give the size;
    then one line per row, where '.', 'G' and 'S' are Free (1) and every other
    character, and every cell past a short line or a missing row, is an
    Obstacle (2).  Raises ValueError where the reference returns false."""
    import gzip
    import os

    if isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        with open(src, "rb") as f:
            data = f.read()
        if str(src).endswith(".gz"):
            data = gzip.decompress(data)
    elif isinstance(src, bytes):
        data = src
    else:
        data = str(src).encode()
    lines = data.split(b"\n")
    h = w = 0
    i = 0
    while i < len(lines):
        line = lines[i].rstrip(b"\r \t")
        i += 1
        if line.startswith(b"height"):
            h = _second_int(line)
        elif line.startswith(b"width"):
            w = _second_int(line)
        elif line == b"map":
            break
    if h <= 0 or w <= 0:
        raise ValueError("not a MovingAI map: missing or non-positive height / width")
    out = np.full((h, w), 2, dtype=np.uint8)
    free = np.frombuffer(b".GS", dtype=np.uint8)
    # std::getline yields a last, empty piece only when the file does not end
    # with a newline; a trailing "\n" ends the stream
    rows = lines[i:]
    if rows and rows[-1] == b"" and data.endswith(b"\n"):
        rows = rows[:-1]
    for r, line in enumerate(rows[:h]):
        row = np.frombuffer(line[:w], dtype=np.uint8)
        out[r, :len(row)] = np.where(np.isin(row, free), 1, 2)
    return out

