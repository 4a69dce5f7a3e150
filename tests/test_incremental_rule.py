"""The rule behind the incremental all-pairs refresh, checked on the oracle.

A query result for viewpoints (u, v) is a function of the leaf tables of u's
and v's leaves and the D_U tables of their ancestors (lift, meets at common
ancestors, same-leaf lookup; src/minplus_oracle.cpp:934-1419).  So when an
update leaves every one of those tables unchanged, the distance d(u, v) is
unchanged.  The library's refresh (htg_query_all_pairs_incremental) returns
its cached matrix exactly when no table on any viewpoint's chain was
rewritten.  This test replays reveals and blockings on the oracle, finds the
tables that changed by comparing them, and requires every pair with an
unchanged chain to keep its BFS distance.  It also shows that the ancestors
matter: a rule that looked at the leaves alone would miss changed distances.
"""
import numpy as np
import pytest

from htrip_amd import grids
from oracle.oracle import Oracle, grid_bfs


def tables(O):
    leaves = [O.leaf_D(i) for i in range(O.num_leaves())]
    nodes = {(lv, w): O.node_DU(lv, w) for lv in range(1, O.H + 1) for w in range(O.num_nodes(lv))}
    return leaves, nodes


def chain(O, r, c):
    """The leaf and every ancestor node of cell (r, c)."""
    b, g = O.b, O.g
    nl = O.S // b
    out = [("leaf", (r // b) * nl + c // b)]
    for lv in range(1, O.H + 1):
        s = b * g ** lv
        out.append((lv, (r // s) * max(1, O.S // s) + c // s))
    return out


REPLAYS = [("street", 64, 8, 3), ("maze", 64, 8, 4), ("street", 96, 8, 5)]


def replay(kind, n, b, seed):
    """Returns (pairs whose chains were unchanged, changed distances a leaf-only rule would miss)."""
    truth = grids.street_grid(n, seed) if kind == "street" else grids.maze_grid(n, seed)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, seed)[0])], n // 4)
    O = Oracle("port", obs, b, 2)
    O.build()
    order = grids.frontier_order(truth, obs)
    vps = grids.sample_free(truth, 40, seed + 1)
    chains = [chain(O, int(r), int(c)) for r, c in vps]
    rng = np.random.default_rng(seed)
    kept = leaf_only_misses = 0
    pos = 0
    for step in range(14):
        before = tables(O)
        d0 = grid_bfs(O.cells, vps)
        if step % 2 == 0:  # a reveal batch
            m = int(rng.integers(1, 25))
            O.open_cells(order[pos:pos + m])
            pos += m
        else:  # a single opening
            r, c = order[pos]
            pos += 1
            O.open_cell(int(r), int(c))
        after = tables(O)
        d1 = grid_bfs(O.cells, vps)
        changed = {("leaf", i) for i in range(len(before[0])) if not np.array_equal(before[0][i], after[0][i])}
        changed |= {key for key in before[1] if not np.array_equal(before[1][key], after[1][key])}
        for i in range(len(vps)):
            for j in range(len(vps)):
                ci, cj = set(chains[i]), set(chains[j])
                if not (ci | cj) & changed:
                    assert d1[i, j] == d0[i, j], f"step {step}: pair ({i},{j}) changed with unchanged chains"
                    kept += 1
                elif d1[i, j] != d0[i, j] and not {x for x in ci | cj if x[0] == "leaf"} & changed:
                    leaf_only_misses += 1
    return kept, leaf_only_misses


@pytest.mark.parametrize("kind,n,b,seed", REPLAYS)
def test_unchanged_chains_keep_their_distances(kind, n, b, seed):
    kept, _ = replay(kind, n, b, seed)
    assert kept > 0


def test_the_ancestors_matter():
    """Across the replays above, some pair's distance changes while both leaves
    are untouched: a leaf-only rule would return a stale matrix."""
    assert sum(replay(*c)[1] for c in REPLAYS) > 0
