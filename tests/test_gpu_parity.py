"""GPU parity: every table and every query of libhtrip_gpu.so against the oracle.

Bar: bit-exact (integer tropical arithmetic).  The oracle is the CPU
restatement in oracle/ (pinned against the compiled reference and BFS in
tests/test_oracle.py); test_tables_equal_the_compiled_reference also compares
the GPU directly with the compiled reference when oracle/_ref is built.
"""
import numpy as np
import pytest

from htrip_amd import grids, serving
from htrip_amd._lib import HTG_E_INVALID, HTG_E_RANGE, GpuMap, HtgError, IncState
from oracle import oracle as orc
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


REVEAL = [
    # (kind, n, b, g, seed, radius, batch, steps)
    ("street", 96, 8, 2, 21, 20, 37, 12),
    ("street", 128, 16, 2, 22, 24, 64, 10),
    ("maze", 64, 8, 2, 23, 12, 29, 10),
    ("street", 64, 8, 4, 24, 14, 33, 8),
    ("street", 128, 32, 2, 25, 30, 101, 6),
]


@pytest.mark.parametrize("kind,n,b,g,seed,radius,batch,steps", REVEAL)
def test_incremental_reveal(gpu_ctx, kind, n, b, g, seed, radius, batch, steps):
    """onBatchCellsOpened: tables and stats equal the oracle after every step;
    tables equal a fresh build (monotone exactness); queries equal BFS."""
    print(f"reveal {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    start = grids.sample_free(truth, 1, seed)[0]
    obs = grids.disk_observed(truth, [tuple(start)], radius)
    order = grids.frontier_order(truth, obs)
    O = Oracle("port", obs, b, g)
    O.build()
    G = GpuMap(obs, b, g)
    G.build()
    for st in range(steps):
        cells = order[st * batch:(st + 1) * batch]
        if len(cells) == 0:
            break
        O.open_cells(cells)
        G.open_cells(cells)
        assert G.last_batch_stats() == O.last_batch_stats(), st
        vps = grids.sample_free(O.cells, 24, seed + st)
        np.testing.assert_array_equal(G.query(vps), grid_bfs(O.cells, vps), err_msg=f"step {st}")
    assert_tables_equal(G, O)
    F = GpuMap(O.cells, b, g)
    F.build()
    assert_tables_equal(F, O)


@pytest.mark.parametrize("kind,n,b,seed,robots,batch,steps", [
    ("street", 256, 16, 41, 6, 400, 4),  # ~2400 cells a batch: cover > 64 rows, > 512 uncovered ends
    ("maze", 128, 8, 43, 4, 120, 6),
    ("street", 128, 32, 44, 3, 60, 8),
    ("street", 64, 8, 45, 2, 9, 10),     # small batches: mostly single-block covers
])
def test_update_pivot_sets_agree(gpu_ctx, kind, n, b, seed, robots, batch, steps):
    """Incremental reveal closed over the vertex cover (default: each level
    updated or rebuilt by the cost estimate), over the cover with every level
    always updated, with every level always rebuilt from its children, and
    over every decreased row (the reference's pivot set): identical tables and
    stats after every step, all equal to the oracle at the end."""
    print(f"pivot sets {kind} n={n} b={b} seed={seed} batch={robots}x{batch}", flush=True)
    truth = make(kind, n, seed)
    rv = serving.RobotReveal(truth, robots, 16, seed + 1, limit=100000)
    obs = rv.observed.copy()
    O = Oracle("port", obs, b, 2)
    O.build()
    A = GpuMap(obs, b, 2)
    A.build()
    R = GpuMap(obs, b, 2)
    R.set_tuning(update_pivots=1)
    R.build()
    U = GpuMap(obs, b, 2)
    U.set_tuning(update_rebuild=-1)
    U.build()
    W = GpuMap(obs, b, 2)
    W.set_tuning(update_rebuild=1)
    W.build()
    assert A.tuning().update_pivots == 0 and R.tuning().update_pivots == 1
    assert (A.tuning().update_rebuild, U.tuning().update_rebuild, W.tuning().update_rebuild) == (0, -1, 1)
    for st in range(steps):
        cells = np.vstack([rv.next_batch(r, batch) for r in range(robots)]).astype(np.int32)
        O.open_cells(cells)
        for M in (A, R, U, W):
            M.open_cells(cells)
        assert A.last_batch_stats() == R.last_batch_stats() == U.last_batch_stats() == W.last_batch_stats() == \
            O.last_batch_stats(), st
        for lvl in range(1, O.H + 1):
            for i in range(O.num_nodes(lvl)):
                ref = A.node_DU(lvl, i)
                for M in (R, U, W):
                    np.testing.assert_array_equal(M.node_DU(lvl, i), ref, err_msg=f"step {st} D_U {lvl},{i}")
    for M in (A, R, U, W):
        assert_tables_equal(M, O)
    vps = grids.sample_free(O.cells, 32, seed)
    np.testing.assert_array_equal(A.query(vps), grid_bfs(O.cells, vps))


@pytest.mark.parametrize("kind,n,b,seed", [("street", 96, 8, 51), ("maze", 64, 8, 52), ("street", 128, 16, 53)])
def test_blocked_cells_equal_a_fresh_build(gpu_ctx, kind, n, b, seed):
    """onCellBlocked (batched): after each blocking step every table equals a
    fresh build of the same grid (no other node depends on the blocked cells),
    queries equal BFS, last_batch_stats is untouched; reveals after blocking
    (including of blocked cells) stay exact."""
    print(f"block {kind} n={n} b={b} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    start = grids.sample_free(truth, 1, seed)[0]
    obs = grids.disk_observed(truth, [tuple(start)], n // 3)
    G = GpuMap(obs, b, 2)
    G.build()
    order = grids.frontier_order(truth, obs)
    G.open_cells(order[:60])
    stats = G.last_batch_stats()
    rng = np.random.default_rng(seed)
    for step, m in enumerate([1, 1, 5, 40, 3]):
        free = np.argwhere(G.cells == 1)
        pick = free[rng.choice(len(free), size=m, replace=False)].astype(np.int32)
        if step == 1:  # a port cell on a leaf boundary, when one is free
            edge = free[(free[:, 0] % b == 0) | (free[:, 1] % b == b - 1)]
            if len(edge):
                pick = edge[:1].astype(np.int32)
        G.block_cells(pick)
        assert G.last_batch_stats() == stats
        O = Oracle("port", G.cells, b, 2)
        assert O.build()
        assert_tables_equal(G, O)
        vps = grids.sample_free(G.cells, 24, seed + step)
        np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps), err_msg=f"step {step}")
    # reveal again, reopening some blocked cells too
    blocked = np.argwhere((G.cells == 2) & (truth == 1))[:10].astype(np.int32)
    G.open_cells(np.vstack([blocked, order[60:200]]))
    O = Oracle("port", G.cells, b, 2)
    assert O.build()
    assert_tables_equal(G, O)
    vps = grids.sample_free(G.cells, 32, seed + 99)
    np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps))


def test_single_cell_openings(gpu_ctx):
    """onCellOpened one cell at a time (Suite 5 style) including seam cells."""
    truth = grids.street_grid(64, 31)
    obs = grids.disk_observed(truth, [(32, 32)], 12)
    order = grids.frontier_order(truth, obs)[:60]
    O = Oracle("port", obs, 8, 2)
    O.build()
    G = GpuMap(obs, 8, 2)
    G.build()
    before = G.last_batch_stats()
    for r, c in order:
        O.open_cell(int(r), int(c))
        G.open_cell(int(r), int(c))
    assert G.last_batch_stats() == before
    assert_tables_equal(G, O)


def test_query_batch_and_edge_cases(gpu_ctx):
    """Coalesced queries of mixed k equal the same queries run alone."""
    cells = grids.street_grid(128, 41)
    G = GpuMap(cells, 16, 2)
    G.build()
    qs = [grids.sample_free(cells, k, 100 + k) for k in (1, 2, 7, 33, 64, 5)]
    qs.append(np.zeros((0, 2), np.int32))
    qs.append(np.array([[0, 0], [0, 0], [127, 127], [-3, -3]], np.int32))
    alone = [G.query(q) for q in qs]
    mixed = G.query_batch(qs)
    for a, m, q in zip(alone, mixed, qs):
        np.testing.assert_array_equal(a, m)
        np.testing.assert_array_equal(m, grid_bfs(cells, q))


def test_clone_is_independent(gpu_ctx):
    truth = grids.street_grid(64, 51)
    obs = grids.disk_observed(truth, [(20, 20)], 16)
    order = grids.frontier_order(truth, obs)
    G = GpuMap(obs, 8, 2)
    G.build()
    C = G.clone()
    G.open_cells(order[:80])
    vps = grids.sample_free(obs, 20, 7)
    np.testing.assert_array_equal(C.query(vps), grid_bfs(obs, vps))
    np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps))


def test_saturation_maze(gpu_ctx):
    """Long maze distances: sums of two near-diameter distances must clamp, never wrap."""
    cells = grids.maze_grid(256, 61)
    G = GpuMap(cells, 16, 2)
    G.build()
    vps = grids.sample_free(cells, 48, 62)
    expect = grid_bfs(cells, vps)
    assert expect[expect < 60000].max() > 4000
    np.testing.assert_array_equal(G.query(vps), expect)


@pytest.mark.parametrize("kind,n,b,g,seed", CASES)
def test_lift_point_equals_the_oracle(gpu_ctx, kind, n, b, g, seed):
    """liftPoint (src/minplus_oracle.cpp:731-823): every level's vector and the
    ancestor metadata equal the oracle's, for free, blocked, border and
    padding cells, after the build, after reveals and after a blocking step."""
    print(f"lift {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, seed)[0])], n // 3)
    G = GpuMap(obs, b, g)
    G.build()
    order = grids.frontier_order(truth, obs)
    S = G.S

    def check(tag):
        O = Oracle("port", G.cells, b, g)
        assert O.build()
        pts = [tuple(p) for p in grids.sample_free(G.cells, 8, seed)] + [(0, 0), (n - 1, n - 1), (S - 1, S - 1)]
        if b < n:
            pts.append((b - 1, b))
        for r, c in pts:
            X, par, quad, off = G.lift_point(r, c)
            Xo, po, qo, oo, _ = O.lift_point(r, c)
            assert len(X) == len(Xo) == G.H + 1
            for lvl in range(G.H + 1):
                np.testing.assert_array_equal(X[lvl], Xo[lvl], err_msg=f"{tag} ({r},{c}) level {lvl}")
            assert (par == po).all() and (quad == qo).all() and (off == oo).all(), (tag, r, c)

    check("built")
    G.open_cells(order[:150])
    check("revealed")
    free = np.argwhere(G.cells == 1)
    G.block_cells(free[:: max(1, len(free) // 7)][:7].astype(np.int32))
    check("blocked")
    with pytest.raises(HtgError) as ei:
        G.lift_point(S, 0)
    assert ei.value.code == HTG_E_RANGE


@pytest.mark.parametrize("kind,n,b,seed", [("street", 128, 16, 61), ("maze", 96, 8, 62), ("street", 64, 8, 63)])
def test_incremental_query_equals_the_full_query(gpu_ctx, kind, n, b, seed):
    """queryAllPairsIncremental (src/minplus_oracle.cpp:1444-2002): after reveal
    batches, single openings, blockings and a rebuild, the refreshed matrix
    equals the full query and BFS; with no update in between, or updates that
    touch no viewpoint's chain, it is returned from the cache."""
    print(f"incremental {kind} n={n} b={b} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, seed)[0])], n // 3)
    G = GpuMap(obs, b, 2)
    G.build()
    order = grids.frontier_order(truth, obs)
    vps = grids.sample_free(truth, 40, seed + 1)  # some not revealed yet
    st = IncState()
    m, nr = G.query_incremental(st, vps)
    assert nr == len(vps)
    np.testing.assert_array_equal(m, grid_bfs(G.cells, vps))
    m2, nr = G.query_incremental(st, vps)
    assert nr == 0 and np.array_equal(m, m2)
    rng = np.random.default_rng(seed)
    pos, cached = 0, 0
    for step in range(24):
        kind_of = step % 4
        if kind_of == 0:
            m_ = int(rng.integers(5, 40))
            G.open_cells(order[pos:pos + m_])
            pos += m_
        elif kind_of == 1:
            r, c = order[pos]
            pos += 1
            G.open_cell(int(r), int(c))
        elif kind_of == 2 and step % 8 == 2:
            free = np.argwhere(G.cells == 1)
            G.block_cells(free[rng.integers(len(free))][None].astype(np.int32))
        m, nr = G.query_incremental(st, vps)
        assert nr in (0, len(vps))
        cached += nr == 0
        np.testing.assert_array_equal(m, G.query(vps), err_msg=f"step {step}")
        np.testing.assert_array_equal(m, grid_bfs(G.cells, vps), err_msg=f"step {step}")
    assert cached >= 3  # at least the steps without an update
    # other viewpoints, a rebuild, another map: recomputed; k = 0 clears
    _, nr = G.query_incremental(st, vps[:30])
    assert nr == 30
    G.build()
    m, nr = G.query_incremental(st, vps[:30])
    assert nr == 30 and np.array_equal(m, grid_bfs(G.cells, vps[:30]))
    H = G.clone()
    m, nr = H.query_incremental(st, vps[:30])
    assert nr == 30 and np.array_equal(m, grid_bfs(G.cells, vps[:30]))
    m, nr = H.query_incremental(st, vps[:0])
    assert m.shape == (0, 0) and nr == 0
    _, nr = H.query_incremental(st, vps[:30])
    assert nr == 30
    st.close()


@pytest.mark.parametrize("kind,n,b,g,seed", [("street", 64, 8, 2, 71), ("street", 8, 8, 2, 72), ("maze", 64, 8, 2, 73),
                                              ("street", 16, 4, 4, 74), ("street", 64, 16, 2, 75)])
def test_lift_root_equals_a_fresh_build(gpu_ctx, kind, n, b, g, seed):
    """liftRoot (src/minplus_oracle.cpp:2008-2114): after reveals, the square
    grows by g with one more level; every table equals the oracle's fresh
    build of the grown grid, queries (across the old border too) equal BFS,
    incremental states recompute, and reveals, lifts and clones afterwards
    stay exact."""
    print(f"lift root {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, seed)[0])], max(3, n // 2))
    G = GpuMap(obs, b, g)
    assert G.build()
    order = grids.frontier_order(truth, obs)
    G.open_cells(order[:len(order) // 2])
    S0, H0 = G.S, G.H
    assert S0 == n
    st = IncState()
    vps0 = grids.sample_free(G.cells, 12, seed)
    G.query_incremental(st, vps0)
    S1 = S0 * g
    big_truth = make(kind, S1, seed + 1)
    big_truth[:n, :n] = truth
    big_obs = big_truth.copy()
    big_obs[:n, :n] = G.cells
    big_obs[n:n + max(2, n // 4), :] = np.where(big_obs[n:n + max(2, n // 4), :] == 1, 0, big_obs[n:n + max(2, n // 4), :])
    G.lift_root(big_obs)
    assert G.S == S1 and G.H == H0 + 1
    np.testing.assert_array_equal(G.cells, big_obs)
    O = Oracle("port", G.cells, b, g)
    assert O.build()
    assert_tables_equal(G, O)
    vps = np.vstack([grids.sample_free(G.cells[:n, :n], 8, seed + 2), grids.sample_free(G.cells, 24, seed + 3)])
    np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps))
    m, nr = G.query_incremental(st, vps0)
    assert nr == len(vps0) and np.array_equal(m, grid_bfs(G.cells, vps0))
    # reveals afterwards: the unknown band and the rest of the old frontier
    band = np.argwhere((G.cells == 0) & (big_truth == 1)).astype(np.int32)
    G.open_cells(np.vstack([band, order[len(order) // 2:]]).astype(np.int32))
    O = Oracle("port", G.cells, b, g)
    assert O.build()
    assert_tables_equal(G, O)
    np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps))
    r, c = (int(v) for v in grids.sample_free(G.cells, 1, seed + 4)[0])
    X, par, quad, off = G.lift_point(r, c)
    Xo, po, qo, oo, _ = O.lift_point(r, c)
    for lvl in range(G.H + 1):
        np.testing.assert_array_equal(X[lvl], Xo[lvl])
    C2 = G.clone()
    assert C2.S == S1
    np.testing.assert_array_equal(C2.query(vps), grid_bfs(G.cells, vps))
    # a grid that does not contain the current one is refused and the map stays usable
    with pytest.raises(HtgError) as ei:
        G.lift_root(G.cells[:S1 // 2, :S1 // 2])
    assert ei.value.code == HTG_E_INVALID and G.S == S1
    np.testing.assert_array_equal(G.query(vps), grid_bfs(G.cells, vps))
    st.close()


@pytest.mark.skipif(not orc.available("reference"), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("kind,n,b,g,seed", [("street", 64, 8, 2, 81), ("maze", 64, 16, 2, 82), ("street", 64, 4, 4, 83)])
def test_tables_equal_the_compiled_reference(gpu_ctx, kind, n, b, g, seed):
    """Against the reference's own CPU AVX2 code (oracle/_ref): every leaf and
    node table after the build and after each reveal batch, last_batch_stats,
    and the query matrices."""
    print(f"reference {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    truth = make(kind, n, seed)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, seed)[0])], n // 3)
    G = GpuMap(obs, b, g)
    assert G.build()
    R = Oracle("reference", obs, b, g)
    assert R.build()
    assert_tables_equal(G, R)
    order = grids.frontier_order(truth, obs)
    for st in range(3):
        batch = order[st * 90:(st + 1) * 90]
        G.open_cells(batch)
        R.open_cells(batch)
        assert G.last_batch_stats() == R.last_batch_stats(), st
        assert_tables_equal(G, R)
        vps = grids.sample_free(G.cells, 24, seed + st)
        np.testing.assert_array_equal(G.query(vps), R.query(vps))


def test_cpp_dropin_program(gpu_ctx):
    """The reference-API C++ program (tests/cpp/test_dropin.cpp) against the GPU library."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "cpp", "test_dropin")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr


def test_reference_exactness_program_on_the_dropin(gpu_ctx):
    """The reference's own exactness program (apps/test_exactness.cpp: suites 1
    to 16 over its procedural worlds, its frontier viewpoints and its BfsOracle),
    unchanged, compiled against the drop-in headers and linked with the GPU
    library (oracle/_ref/test_exactness_dropin, built by oracle/Makefile where
    the reference sources are present)."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "test_exactness_dropin")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/test_exactness_dropin not built (reference sources absent at build time)")
    print("reference exactness program on the drop-in", flush=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=280)
    print(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    assert "ALL EXACTNESS VERIFICATION SUITES PASSED" in res.stdout


@pytest.mark.parametrize("args", [[], ["--seeds", "4", "--size", "128", "--tile", "8", "--steps", "16"],
                                  ["--seeds", "3", "--size", "64", "--tile", "8", "--group", "4"]])
def test_reference_property_program_on_the_dropin(gpu_ctx, args):
    """The reference's property-based differential program
    (apps/test_property_based.cpp: incremental updates against rebuilds and
    BFS over seeded worlds), unchanged, against the drop-in and the GPU library."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "test_property_based_dropin")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/test_property_based_dropin not built (reference sources absent at build time)")
    print(f"reference property program on the drop-in {args}", flush=True)
    res = subprocess.run([exe] + args, capture_output=True, text=True, timeout=280)
    print(res.stdout[-2000:])
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    assert "ALL PROPERTY-BASED CHECKS PASSED" in res.stdout


@pytest.mark.parametrize("kind,n,b,g,seed", [("maze", 128, 16, 2, 31), ("street", 96, 8, 2, 32),
                                              ("street", 64, 8, 4, 33)])
def test_graph_build_and_tuning_keep_tables(gpu_ctx, kind, n, b, g, seed):
    """The build captured as a hipGraph (and replayed), every pivot block, and
    the autotuner's choices give the oracle's tables and BFS distances."""
    print(f"graph build {kind} n={n} b={b} g={g} seed={seed}", flush=True)
    cells = make(kind, n, seed)
    O = Oracle("port", cells, b, g)
    assert O.build()
    G = GpuMap(cells, b, g)
    for P in (16, 32, 64):
        G.set_tuning(pivot_block=P, use_graph=1)
        assert G.build()
        assert G.build()  # replays the captured graph
        assert_tables_equal(G, O)
    vps = grids.sample_free(cells, 48, seed + 3)
    t = G.autotune([vps, vps[:17]], reps=1)
    assert t.pivot_block in (16, 32, 64) and t.zq_tile in (4, 8, 16) and t.use_graph in (0, 1)
    assert_tables_equal(G, O)
    np.testing.assert_array_equal(G.query(vps), grid_bfs(cells, vps))
