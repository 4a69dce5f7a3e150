"""Writes tests/golden/golden.json from the compiled reference (oracle/_ref).

The reference ships no fixture files for its hot path (its tests build grids
in code and compare against BFS).  This script runs the reference's own C++
sources, compiled by oracle/Makefile, on this repository's seeded grids and on
the hand-drawn grids of the reference's known-answer suites, and stores:
  * sha256 digests of every leaf D / S and node D_U / S_U table,
  * the query matrices (hex), and last_batch_stats after reveal batches,
  * S, H and totalMemoryBytes.
Grids are regenerated from their seeds when the fixture is checked, so a
sha256 of each grid is stored too (detects generator drift).

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from htrip_amd import grids  # noqa: E402
from tests.golden import cases  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def table_digest(o) -> dict:
    h_leaf = hashlib.sha256()
    for i in range(o.num_leaves()):
        h_leaf.update(o.leaf_D(i).tobytes())
        h_leaf.update(o.leaf_S(i).tobytes())
    h_node = hashlib.sha256()
    for lvl in range(1, o.H + 1):
        for i in range(o.num_nodes(lvl)):
            h_node.update(o.node_DU(lvl, i).tobytes())
            h_node.update(o.node_SU(lvl, i).tobytes())
    return {"leaves": h_leaf.hexdigest(), "nodes": h_node.hexdigest()}


def run_case(c: dict) -> dict:
    cells = cases.make_cells(c)
    out = {"grid_sha256": sha(cells)}
    o = Oracle("reference", cells, c["b"], c["g"])
    o.build()
    out["S"], out["H"], out["total_memory_bytes"] = o.S, o.H, o.total_memory_bytes()
    out["tables"] = table_digest(o)
    vps = cases.viewpoints(c, cells)
    out["query_hex"] = o.query(vps).tobytes().hex()
    steps = []
    for batch in cases.reveal_batches(c, cells):
        if batch.get("single"):
            for r, cc in batch["cells"]:
                o.open_cell(int(r), int(cc))
        else:
            o.open_cells(np.array(batch["cells"], dtype=np.int32).reshape(-1, 2))
        vq = cases.viewpoints(c, o.cells, salt=len(steps) + 1)
        steps.append({"stats": list(o.last_batch_stats()), "query_hex": o.query(vq).tobytes().hex()})
    out["steps"] = steps
    if steps:
        out["tables_after"] = table_digest(o)
    return out


def main():
    res = {"generator": "compiled reference (oracle/_ref: src/*.cpp of the reference + oracle/ref_shim.cpp)",
           "cases": {}}
    for c in cases.CASES:
        res["cases"][c["name"]] = {"spec": c, **run_case(c)}
        print("golden", c["name"], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
