"""Synthetic worlds of the reference's test and benchmark programs:
GridMap::generateProceduralRooms and GridMap::initExplorationSnapshot
(src/grid_map.cpp:78-200), restated in the C++ drop-in
(htrip_amd/include/htrip/grid_map.hpp) so that the same seeds give the same
grids.  The drop-in's output is checked against the compiled reference
(oracle/_ref) cell for cell when that is built, and always against what the
definitions fix: room interiors free, every room reachable through its
doorways, a snapshot that reveals a connected set of truth-free cells around
its start, no more than its budget, and obstacles only next to revealed
cells.  GridMap::getClearance (grid_map.hpp:174-212) is checked on the same
grids against its definition and the reference."""
import os
import subprocess
from collections import deque

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
FREE, OBST, UNKNOWN = 1, 2, 0

CASES = [
    # rows, cols, room, seed, start, coverage, island, seed2
    (64, 64, 16, 42, (0, 0), 0.0, 0.1, 12345),      # rooms only
    (20, 44, 8, 411, (2, 2), 0.65, 0.10, 12),       # non-square, as the reference's suite 11
    (32, 32, 16, 814, (1, 1), 0.70, 0.10, 16),      # as suite 14
    (40, 40, 16, 915, (2, 2), 0.68, 0.10, 18),      # as suite 15
    (37, 53, 5, 7, (30, 50), 0.45, 0.3, 99),        # ragged room grid, islands
    (128, 96, 12, 123, (64, 48), 0.5, 0.0, 5),      # no island draws
    (16, 16, 3, 1, (0, 0), 0.9, 0.5, 2),            # room size below 4 is raised to 4
    (10, 10, 16, 1, (5, 5), 0.5, 0.1, 1),           # one room larger than the grid: nothing free
]


@pytest.fixture(scope="module")
def gen(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("world") / "gen_world")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(HERE, "..", "htrip_amd", "include"),
                    os.path.join(HERE, "cpp", "gen_world.cpp"), "-o", exe], check=True)

    def run(rows, cols, room, seed, start, coverage, island, seed2):
        out = subprocess.run([exe, str(rows), str(cols), str(room), str(seed), str(start[0]), str(start[1]),
                              repr(coverage), repr(island), str(seed2)], capture_output=True, check=True).stdout
        head, _, body = out.partition(b"\n")
        ok, r, c = (int(v) for v in head.split())
        cells = np.frombuffer(body, dtype=np.uint8)
        return bool(ok), cells[:r * c].reshape(r, c), cells[r * c:2 * r * c].reshape(r, c), \
            cells[2 * r * c:].reshape(r, c)

    return run


def components(mask):
    """4-connected component labels of a boolean grid (-1 outside)."""
    lab = np.full(mask.shape, -1, np.int64)
    n = 0
    for s in zip(*np.nonzero(mask)):
        if lab[s] >= 0:
            continue
        lab[s] = n
        q = deque([s])
        while q:
            r, c = q.popleft()
            for dr, dc in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                y, x = r + dr, c + dc
                if 0 <= y < mask.shape[0] and 0 <= x < mask.shape[1] and mask[y, x] and lab[y, x] < 0:
                    lab[y, x] = n
                    q.append((y, x))
        n += 1
    return lab, n


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}-room{c[2]}-cov{c[5]}" for c in CASES])
def test_world_matches_the_reference_and_its_definition(case, gen):
    rows, cols, room, seed, start, coverage, island, seed2 = case
    ok, truth, obs, clr = gen(*case)
    rs = max(room, 4)
    ny, nx = rows // rs, cols // rs
    # every room interior is free, and the rooms form one component through the doorways
    for ry in range(ny):
        for rx in range(nx):
            y1, x1 = min((ry + 1) * rs - 1, rows - 1), min((rx + 1) * rs - 1, cols - 1)
            assert (truth[ry * rs + 1:y1, rx * rs + 1:x1] == FREE).all()
    lab, n = components(truth == FREE)
    assert n == (1 if ny * nx > 0 and (truth == FREE).any() else 0)
    assert set(np.unique(truth)) <= {FREE, OBST}
    if coverage > 0:
        if not (truth == FREE).any():
            assert not ok
        else:
            assert ok
            seen = obs == FREE
            assert (truth[seen] == FREE).all() and (truth[obs == OBST] == OBST).all()
            budget = max(10, int(int((truth == FREE).sum()) * coverage))
            assert 1 <= seen.sum() <= budget
            lab_o, n_o = components(seen)
            assert n_o == 1
            grown = seen.copy()  # revealed obstacles touch a revealed free cell
            grown[1:] |= seen[:-1]
            grown[:-1] |= seen[1:]
            grown[:, 1:] |= seen[:, :-1]
            grown[:, :-1] |= seen[:, 1:]
            assert grown[obs == OBST].all()
            if island == 0.0 and seen.sum() < budget:
                # nothing skipped: the snapshot reached the start's whole component
                sr, sc = start if truth[start] == FREE else np.argwhere(truth == FREE)[0]
                assert seen.sum() == (lab == lab[sr, sc]).sum()
    else:
        assert ok and (obs == truth).all()
    # getClearance: 0 off the passable cells, else the Chebyshev ring (1 or 2)
    # of the nearest obstacle or grid edge, 3 beyond
    pad = np.pad(obs == OBST, 2, constant_values=True)
    for r, c in np.argwhere(np.ones_like(obs, bool)):
        want = 0
        if obs[r, c] == FREE:
            win = pad[r:r + 5, c:c + 5]
            want = 1 if win[1:4, 1:4].any() else 2 if win.any() else 3
        assert clr[r, c] == want
    if orc.available("reference"):
        rok, rtruth, robs = orc.reference_world(rows, cols, room, seed, start, coverage, island, seed2)
        assert rok == ok
        np.testing.assert_array_equal(truth, rtruth)
        np.testing.assert_array_equal(obs, robs)
        np.testing.assert_array_equal(clr, orc.reference_frontiers(obs)[0])
