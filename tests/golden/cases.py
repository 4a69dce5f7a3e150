"""Golden-fixture case definitions (shared by make_golden.py and the tests).

Seeded cases use this repository's generators.  The "suite*" cases rebuild
the hand-drawn grids of the reference's known-answer suites
(apps/test_exactness.cpp) from their published construction, with the
known answers recorded in KNOWN.
"""
from __future__ import annotations

import numpy as np

from htrip_amd import grids

FREE, OBST = 1, 2


def _blocked(n, m=None):
    return np.full((n, m or n), OBST, dtype=np.uint8)


def suite_cells(name: str) -> np.ndarray:
    if name == "suite1":  # test_exactness.cpp:97-146
        g = _blocked(32)
        g[4, 16] = FREE
        return g
    if name == "suite1b":  # :148-201
        g = _blocked(32)
        g[16, 8] = FREE
        return g
    if name == "suite1c":  # :203-260
        g = _blocked(32)
        g[0, 0:16] = FREE
        g[15, 0:32] = FREE
        g[1:15, 0] = FREE
        g[1:15, 15] = FREE
        g[8, 8:14] = FREE
        g[9:15, 13] = FREE
        g[14, 8:13] = FREE
        g[15:21, 20] = FREE
        g[20, 20] = FREE
        return g
    if name == "suite8":  # :594-667
        g = _blocked(32)
        g[2, 16:32] = FREE
        g[2:16, 24] = FREE
        g[16:21, 24] = FREE
        g[20, 16:32] = FREE
        g[16:32, 0:16] = FREE
        return g
    if name == "suite10":  # :760-795
        g = _blocked(16)
        g[8, 2:14] = FREE
        return g
    if name == "suite12":  # :842-893
        g = np.full((32, 32), FREE, dtype=np.uint8)
        g[:, 16] = OBST
        return g
    if name == "suite16_1":  # :1031-1069
        g = _blocked(16)
        g[:, 0] = FREE
        return g
    if name == "suite16_3":  # :1132-1163
        g = np.full((64, 64), FREE, dtype=np.uint8)
        g[:, 31] = OBST
        return g
    raise KeyError(name)


SUITE_VPS = {
    "suite1": [(4, 10), (4, 13), (4, 15), (4, 16), (4, 18), (4, 20)],
    "suite1b": [(10, 8), (12, 8), (15, 8), (16, 8), (18, 8), (22, 8)],
    "suite1c": [(8, 9), (14, 9), (9, 13), (20, 20), (15, 10)],
    "suite8": [(2, 24), (25, 10), (20, 20), (8, 24)],
    "suite10": [(8, 2), (8, 6), (8, 10), (8, 13)],
    "suite12": [(4, 4), (8, 12), (4, 20), (8, 28), (20, 16)],
    "suite16_1": [(0, 0), (5, 0), (10, 0), (15, 0), (5, 10), (10, 10)],
    "suite16_3": [(32, 10), (32, 45), (0, 0), (63, 63)],
}

SUITE_REVEAL = {
    "suite1": [{"single": True, "cells": [(4, 15)]}, {"single": True, "cells": [(4, c) for c in range(10, 21)]}],
    "suite1b": [{"single": True, "cells": [(15, 8)]}, {"single": True, "cells": [(r, 8) for r in range(10, 23)]}],
    "suite1c": [{"single": True, "cells": [(9, 8), (10, 8), (11, 8), (12, 8)]}],
    "suite8": [],
    "suite10": [{"single": True, "cells": [(7, 6), (6, 6), (5, 6), (4, 6), (9, 10), (10, 10), (11, 10), (12, 10)]}],
    "suite12": [{"single": True, "cells": [(8, 16)]}],
    "suite16_1": [{"single": False, "cells": [p for c in range(1, 11) for p in ((5, c), (10, c))]}],
    "suite16_3": [{"single": False, "cells": [(r, 31) for r in range(64)]}],
}

SUITE_CFG = {"suite1": 8, "suite1b": 8, "suite1c": 8, "suite8": 16, "suite10": 16, "suite12": 16,
             "suite16_1": 16, "suite16_3": 16}

# Known answers stated by the reference's suites: (case, step, vp_i, vp_j, distance).
# step 0 = right after build, step s = after the s-th reveal batch.
KNOWN = [
    ("suite1", 1, 2, 3, 1),        # d((4,15),(4,16)) = 1 after opening (4,15) (:124-131)
    ("suite1b", 1, 2, 3, 1),       # d((15,8),(16,8)) = 1 (:179-186)
    ("suite8", 0, 0, 1, 37),       # d((2,24),(25,10)) = 37 (:633-648)
    ("suite12", 0, 4, 4, 60000),   # blocked viewpoint: INF diagonal (:873-879)
    ("suite12", 0, 0, 2, 60000),   # separated halves before the door opens
    ("suite16_3", 1, 0, 1, 35),    # fully open 64x64 after the batch: Manhattan |45-10|
]

CASES = [
    {"name": "street64_b8", "kind": "street", "n": 64, "seed": 101, "b": 8, "g": 2, "k": 24,
     "reveal": {"radius": 14, "batch": 40, "steps": 6}},
    {"name": "street96_b16", "kind": "street", "n": 96, "seed": 102, "b": 16, "g": 2, "k": 24,
     "reveal": {"radius": 20, "batch": 64, "steps": 5}},
    {"name": "maze64_b8", "kind": "maze", "n": 64, "seed": 103, "b": 8, "g": 2, "k": 24,
     "reveal": {"radius": 10, "batch": 30, "steps": 5}},
    {"name": "street64_b8_g4", "kind": "street", "n": 64, "seed": 104, "b": 8, "g": 4, "k": 24,
     "reveal": {"radius": 12, "batch": 33, "steps": 4}},
    {"name": "street_nonsquare_b8", "kind": "street_crop", "n": 64, "seed": 105, "rows": 20, "cols": 44,
     "b": 8, "g": 2, "k": 12},
    {"name": "maze128w2_b16", "kind": "maze", "n": 128, "seed": 106, "width": 2, "b": 16, "g": 2, "k": 24},
] + [{"name": s, "kind": "suite", "b": SUITE_CFG[s], "g": 2} for s in SUITE_VPS]


def make_cells(c: dict) -> np.ndarray:
    k = c["kind"]
    if k == "street":
        truth = grids.street_grid(c["n"], c["seed"])
    elif k == "maze":
        truth = grids.maze_grid(c["n"], c["seed"], c.get("width", 1))
    elif k == "street_crop":
        truth = grids.street_grid(c["n"], c["seed"])[: c["rows"], : c["cols"]].copy()
    elif k == "suite":
        return suite_cells(c["name"])
    else:
        raise KeyError(k)
    if "reveal" in c:
        start = grids.sample_free(truth, 1, c["seed"])[0]
        return grids.disk_observed(truth, [tuple(start)], c["reveal"]["radius"])
    return truth


def truth_cells(c: dict) -> np.ndarray:
    k = c["kind"]
    if k == "street":
        return grids.street_grid(c["n"], c["seed"])
    if k == "maze":
        return grids.maze_grid(c["n"], c["seed"], c.get("width", 1))
    return make_cells(c)


def viewpoints(c: dict, cells: np.ndarray, salt: int = 0) -> np.ndarray:
    if c["kind"] == "suite":
        return np.array(SUITE_VPS[c["name"]], dtype=np.int32)
    v = grids.sample_free(cells, c["k"], c["seed"] * 7 + salt)
    return np.vstack([v, [[0, 0]]]).astype(np.int32)


def reveal_batches(c: dict, cells: np.ndarray) -> list:
    if c["kind"] == "suite":
        return SUITE_REVEAL[c["name"]]
    if "reveal" not in c:
        return []
    order = grids.frontier_order(truth_cells(c), cells)
    rv = c["reveal"]
    out = []
    for s in range(rv["steps"]):
        chunk = order[s * rv["batch"]:(s + 1) * rv["batch"]]
        if len(chunk):
            out.append({"single": s == rv["steps"] - 1, "cells": [tuple(map(int, p)) for p in chunk[:8]]
                        if s == rv["steps"] - 1 else [tuple(map(int, p)) for p in chunk]})
    return out
