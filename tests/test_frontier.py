"""Frontier viewpoints (htrip_amd.frontier) against the reference's
FrontierManager and GridMap::getClearance (src/frontier.cpp,
include/htrip/grid_map.hpp:174-203), plus hand-worked cases; the C++
drop-in's FrontierManager and BfsOracle against both."""
import os
import subprocess

import numpy as np
import pytest

from htrip_amd import frontier, grids
from oracle import oracle as orc

U, F, X = 0, 1, 2
HERE = os.path.dirname(os.path.abspath(__file__))
MAPDIR = os.path.join(HERE, "golden", "maps")


def test_clearance_by_hand():
    obs = np.full((7, 7), F, np.uint8)
    obs[3, 3] = X
    obs[0, 6] = U  # unknown is not an obstacle
    clr = frontier.clearance(obs)
    assert clr[3, 3] == 0 and clr[0, 6] == 0
    assert clr[2, 2] == 1 and clr[3, 4] == 1       # 8-neighbours of the obstacle
    assert clr[1, 3] == 2 and clr[5, 5] == 2        # Chebyshev 2 from the obstacle / from the off-grid row 7
    assert clr[6, 6] == 1
    assert clr[0, 0] == 1                            # off-grid neighbours count as obstacles
    big = np.full((9, 9), F, np.uint8)
    assert frontier.clearance(big)[4, 4] == 3 and frontier.clearance(big)[2, 2] == 3


def test_straight_front_by_hand():
    """A free band below an unknown band: the boundary row is one frontier run;
    with max_size 16 it splits into runs, each represented by its L1 medoid."""
    obs = np.full((12, 40), U, np.uint8)
    obs[6:, :] = F
    front = frontier.frontier_cells(obs)
    assert front[6, 5:35].all() and not front[7:, :].any() and not front[:6, :].any()
    cl = frontier.detect_and_cluster(obs, 2, 16)
    assert sum(len(c.cells) for c in cl) == int(front.sum())
    assert all(len(c.cells) <= 16 for c in cl)
    for c in cl:  # an L1 medoid of a run on one row (clearance 3 there, so it stays put)
        cols = [p[1] for p in c.cells]
        cost = lambda x: sum(abs(x - y) for y in cols)
        assert c.representative[0] == 6 and cost(c.representative[1]) == min(cost(x) for x in cols)


def observed_cases():
    t = grids.street_grid(96, 5)
    yield "street96", grids.disk_observed(t, [tuple(grids.sample_free(t, 1, 6)[0])], 30)
    m = grids.maze_grid(64, 7, width=3)
    yield "maze64w3", grids.disk_observed(m, [tuple(grids.sample_free(m, 1, 8)[0])], 25)
    b = grids.load_movingai(os.path.join(MAPDIR, "Boston_0_256.map.gz"))
    yield "boston256", grids.disk_observed(b, [tuple(p) for p in grids.sample_free(b, 3, 9)], 40)
    r = grids.street_grid(80, 10)[:50, :]
    yield "nonsquare", grids.disk_observed(r, [tuple(grids.sample_free(r, 1, 11)[0])], 20)


@pytest.mark.skipif(not orc.available("reference"), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name,obs", list(observed_cases()), ids=[n for n, _ in observed_cases()])
@pytest.mark.parametrize("min_size,max_size", [(2, 16), (1, 0), (3, 5)])
def test_matches_the_reference(name, obs, min_size, max_size):
    clr, front, ref = orc.reference_frontiers(obs, min_size, max_size)
    np.testing.assert_array_equal(frontier.clearance(obs), clr)
    np.testing.assert_array_equal(frontier.frontier_cells(obs), front)
    got = frontier.detect_and_cluster(obs, min_size, max_size)
    assert len(ref) > 0
    assert [(c.id, c.representative, c.cells) for c in got] == ref


def test_viewpoints_for_queries():
    t = grids.street_grid(96, 12)
    robots = grids.sample_free(t, 2, 13)
    obs = grids.disk_observed(t, [tuple(p) for p in robots], 25)
    vps = frontier.frontier_viewpoints(obs, robots, 20)
    assert len(vps) >= 3 and (vps[:2] == robots).all()
    assert (obs[vps[:, 0], vps[:, 1]] == F).all()


@pytest.fixture(scope="module")
def cpp_frontier_bfs(tmp_path_factory):
    """The C++ drop-in FrontierManager and BfsOracle (htrip_amd/include/htrip/
    frontier.hpp, bfs_oracle.hpp) through tests/cpp/frontier_bfs.cpp."""
    d = tmp_path_factory.mktemp("frontier_cpp")
    exe = str(d / "frontier_bfs")
    subprocess.run(["g++", "-std=c++20", "-O2", "-I", os.path.join(HERE, "..", "htrip_amd", "include"),
                    os.path.join(HERE, "cpp", "frontier_bfs.cpp"), "-o", exe], check=True)

    def run(obs, vps, min_size, max_size):
        R, Cn = obs.shape
        path = str(d / "in.bin")
        with open(path, "wb") as f:
            f.write(f"{R} {Cn} {len(vps)} {min_size} {max_size}\n".encode())
            f.write(np.ascontiguousarray(obs, dtype=np.uint8).tobytes())
            f.write(" ".join(f"{r} {c}" for r, c in vps).encode())
        out = subprocess.run([exe, path], capture_output=True, check=True).stdout
        front = np.frombuffer(out[:R * Cn], dtype=np.uint8).reshape(R, Cn).astype(bool)
        lines = out[R * Cn:].decode().splitlines()
        n = int(lines[0])
        clusters = []
        for ln in lines[1:1 + n]:
            v = [int(x) for x in ln.split()]
            clusters.append((v[0], (v[1], v[2]), [(v[4 + 2 * j], v[5 + 2 * j]) for j in range(v[3])]))
        m = np.array([int(x) for x in lines[1 + n].split()], dtype=np.int64).reshape(len(vps), len(vps))
        return front, clusters, m

    return run


@pytest.mark.parametrize("name,obs", list(observed_cases()), ids=[n for n, _ in observed_cases()])
@pytest.mark.parametrize("min_size,max_size", [(2, 16), (1, 0), (3, 5)])
def test_cpp_dropin_frontiers_and_bfs(name, obs, min_size, max_size, cpp_frontier_bfs):
    """The C++ drop-in gives the Python restatement's frontier cells and
    clusters (ids, representatives, cell order) and, when oracle/_ref is built,
    the reference's; its BfsOracle gives the oracle's BFS matrix, with
    impassable, off-grid and repeated viewpoints."""
    rng = np.random.default_rng(len(name) * 7 + min_size)
    R, Cn = obs.shape
    vps = [tuple(int(v) for v in p) for p in np.argwhere(obs == F)[rng.integers(0, int((obs == F).sum()), 10)]]
    vps += [(0, 0), (R - 1, Cn - 1), vps[0], (int(np.argwhere(obs == 2)[0][0]), int(np.argwhere(obs == 2)[0][1]))]
    front, clusters, m = cpp_frontier_bfs(obs, vps, min_size, max_size)
    np.testing.assert_array_equal(front, frontier.frontier_cells(obs))
    got = frontier.detect_and_cluster(obs, min_size, max_size)
    assert clusters == [(c.id, c.representative, c.cells) for c in got]
    np.testing.assert_array_equal(m, orc.grid_bfs(obs, np.array(vps, dtype=np.int32)))
    if orc.available("reference"):
        _, rfront, ref = orc.reference_frontiers(obs, min_size, max_size)
        np.testing.assert_array_equal(front, rfront)
        assert clusters == ref
