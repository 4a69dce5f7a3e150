"""GPU parity: every table and every query of libhtrip_gpu.so against the oracle.

Bar: bit-exact (integer tropical arithmetic).  The oracle is the CPU
restatement in oracle/ (pinned against the compiled reference and BFS in
tests/test_oracle.py); the compiled reference itself is used too when present.
"""
import numpy as np
import pytest

from htrip_amd import grids
from htrip_amd._lib import GpuMap
from oracle.oracle import Oracle, grid_bfs

pytestmark = pytest.mark.gpu

CASES = [
    # (kind, n, b, g, seed)
    ("street", 64, 8, 2, 1),
    ("street", 96, 16, 2, 2),
    ("street", 64, 8, 4, 3),
    ("maze", 64, 8, 2, 4),
    ("street", 128, 32, 2, 5),
    ("maze", 128, 16, 2, 6),
    ("street", 16, 16, 2, 7),   # H = 0
    ("street", 40, 4, 2, 8),    # small tiles
]


def make(kind, n, seed):
    return grids.street_grid(n, seed) if kind == "street" else grids.maze_grid(n, seed)


def assert_tables_equal(G, O):
    assert G.S == O.S and G.H == O.H
    for i in range(O.num_leaves()):
        np.testing.assert_array_equal(G.leaf_D(i), O.leaf_D(i), err_msg=f"leaf D {i}")
        np.testing.assert_array_equal(G.leaf_S(i), O.leaf_S(i), err_msg=f"leaf S {i}")
    for lvl in range(1, O.H + 1):
        for i in range(O.num_nodes(lvl)):
            np.testing.assert_array_equal(G.node_DU(lvl, i), O.node_DU(lvl, i), err_msg=f"D_U {lvl},{i}")
            np.testing.assert_array_equal(G.node_SU(lvl, i), O.node_SU(lvl, i), err_msg=f"S_U {lvl},{i}")


@pytest.mark.parametrize("kind,n,b,g,seed", CASES)
def test_build_tables_and_query(gpu_ctx, kind, n, b, g, seed):
    print(f"case {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    cells = make(kind, n, seed)
    O = Oracle("port", cells, b, g)
    assert O.build()
    G = GpuMap(cells, b, g)
    assert G.build()
    assert_tables_equal(G, O)
    vps = grids.sample_free(cells, 40, seed + 11)
    vps = np.vstack([vps, [[0, 0], [n - 1, n - 1], [-1, 3], [n + 5, 0]], vps[:2]]).astype(np.int32)
    expect = grid_bfs(cells, vps)
    np.testing.assert_array_equal(O.query(vps), expect)
    np.testing.assert_array_equal(G.query(vps), expect)
    assert G.total_memory_bytes() == O.total_memory_bytes()
