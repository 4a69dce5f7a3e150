"""GPU parity on the six MovingAI maps the reference ships (datasets/movingai,
stored under tests/golden/maps): fully revealed builds and queries against
BFS, the 256^2 map's tables against the oracle, liftPoint at the root against
BFS over the whole map, and frontier reveals with incremental refreshes."""
import os

import numpy as np
import pytest

from htrip_amd import frontier, grids
from htrip_amd._lib import GpuMap, IncState
from oracle.oracle import INF, Oracle, grid_bfs
from tests.test_gpu_parity import assert_tables_equal

pytestmark = pytest.mark.gpu

MAPDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "maps")
MAPS = [("Boston_0_256", 16), ("Paris_0_512", 16), ("maze512-1-0", 16), ("maze512-16-0", 16),
        ("Boston_0_1024", 32), ("Berlin_0_1024", 32)]


def load(name):
    return grids.load_movingai(os.path.join(MAPDIR, name + ".map.gz"))


def viewpoints(truth, k, seed):
    """k cells of the largest free component, plus free cells outside it and obstacles."""
    comp = grids._largest_component(truth == 1)
    main = grids.sample_free(np.where(comp, 1, 2).astype(np.uint8), k, seed)
    rest = np.argwhere((truth == 1) & ~comp)[:3]
    walls = np.argwhere(truth == 2)[:2]
    return np.vstack([main, rest, walls]).astype(np.int32)


@pytest.mark.parametrize("name,b", MAPS)
def test_fully_revealed_map(gpu_ctx, name, b):
    print(f"map {name} b={b}", flush=True)
    truth = load(name)
    G = GpuMap(truth, b, 2)
    assert G.build()
    vps = viewpoints(truth, 48, 11)
    m = G.query(vps)
    np.testing.assert_array_equal(m, grid_bfs(truth, vps))
    assert m[:48, :48].max() < INF
    if truth.shape[0] == 256:
        O = Oracle("port", truth, b, 2)
        assert O.build()
        assert_tables_equal(G, O)
    # root level of liftPoint = BFS over the whole map at its boundary ports
    n = truth.shape[0]
    if G.S == n:
        ring = [(0, c) for c in range(n)] + [(n - 1, c) for c in range(n)]
        ring += [(r, 0) for r in range(1, n - 1)] + [(r, n - 1) for r in range(1, n - 1)]
        r0, c0 = (int(v) for v in vps[0])
        X = G.lift_point(r0, c0)[0]
        reg = Oracle("port", truth, b, 2).region_bfs(0, 0, n, n, r0, c0)  # one BFS, no build
        np.testing.assert_array_equal(X[G.H], np.array([reg[r, c] for r, c in ring], np.uint16))


@pytest.mark.parametrize("name,b", [("Berlin_0_1024", 32), ("maze512-1-0", 16)])
def test_frontier_reveal_on_a_shipped_map(gpu_ctx, name, b):
    """Robots' disks revealed, then frontier batches: after every batch the
    matrix equals BFS on the observed map, the incremental refresh equals the
    full query, and so does a query over the robots and the frontier
    viewpoints (htrip_amd.frontier)."""
    print(f"reveal {name} b={b}", flush=True)
    truth = load(name)
    comp = grids._largest_component(truth == 1)
    starts = grids.sample_free(np.where(comp, 1, 2).astype(np.uint8), 3, 5)
    obs = grids.disk_observed(truth, [tuple(s) for s in starts], 48)
    G = GpuMap(obs, b, 2)
    assert G.build()
    order = grids.frontier_order(truth, obs)
    vps = grids.sample_free(np.where(comp, 1, 2).astype(np.uint8), 32, 6)
    st = IncState()
    for step in range(6):
        G.open_cells(order[step * 700:(step + 1) * 700])
        full = G.query(vps)
        np.testing.assert_array_equal(full, grid_bfs(G.cells, vps), err_msg=f"step {step}")
        inc, _ = G.query_incremental(st, vps)
        np.testing.assert_array_equal(inc, full)
        # an explorer's query: the robots and the frontier clusters' viewpoints
        fv = frontier.frontier_viewpoints(G.cells, starts, 64)
        assert len(fv) > len(starts)
        np.testing.assert_array_equal(G.query(fv), grid_bfs(G.cells, fv), err_msg=f"frontier step {step}")
    st.close()
