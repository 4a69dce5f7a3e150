"""The boundary on the GPU: caller-owned query workspaces used from several
host threads at once, failure handling (injected HIP errors, the map state
they leave, recovery), server cancellation, and the Python mirror of the
reference API.  Bar: bit-exact against BFS and the oracle."""
import threading

import numpy as np
import pytest

import htrip_amd as ht
from htrip_amd import _lib, grids, serving
from oracle import oracle as orc
from oracle.oracle import Oracle, grid_bfs
from tests.test_gpu_parity import assert_tables_equal

pytestmark = pytest.mark.gpu


def test_threads_query_one_map_through_their_own_workspaces(gpu_ctx):
    """The reference's const queryAllPairs with distinct scratch from several
    threads (src/robot_simulator.cpp:1541-1546): 8 threads, one workspace
    each, 12 queries each of mixed k on one map, every result equal to BFS."""
    print("threads x workspaces", flush=True)
    cells = grids.street_grid(256, 11)
    G = _lib.GpuMap(cells, 16, 2)
    G.build()
    qs = [[grids.sample_free(cells, 8 + (7 * t + 5 * i) % 90, 100 * t + i) for i in range(12)] for t in range(8)]
    expect = [[grid_bfs(cells, v) for v in row] for row in qs]
    got = [[None] * 12 for _ in range(8)]
    errors = []
    start = threading.Barrier(8)

    def worker(t):
        try:
            ws = _lib.QueryWorkspace(G.ctx)
            start.wait()
            for i, v in enumerate(qs[t]):
                got[t][i] = G.query_ws(ws, v)
            ws.close()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    for t in range(8):
        for i in range(12):
            np.testing.assert_array_equal(got[t][i], expect[t][i], err_msg=f"thread {t} query {i}")
    # the map's own workspace and the plain entry points still agree
    np.testing.assert_array_equal(G.query(qs[0][0]), expect[0][0])


def test_workspace_sees_updates_and_serves_any_map(gpu_ctx):
    """A query through a workspace sees every update finished before it; one
    workspace serves maps of different shapes and grows as needed."""
    truth = grids.street_grid(128, 12)
    obs = grids.disk_observed(truth, [(64, 64)], 30)
    order = grids.frontier_order(truth, obs)
    A = _lib.GpuMap(obs, 8, 2)
    A.build()
    B = _lib.GpuMap(grids.maze_grid(96, 13), 16, 2)
    B.build()
    ws = _lib.QueryWorkspace(A.ctx)
    for step in range(5):
        A.open_cells(order[step * 150:(step + 1) * 150])
        v = grids.sample_free(A.cells, 20 + 30 * step, step)
        np.testing.assert_array_equal(A.query_ws(ws, v), grid_bfs(A.cells, v), err_msg=f"step {step}")
        w = grids.sample_free(B.cells, 10 + 11 * step, step + 50)
        np.testing.assert_array_equal(B.query_ws(ws, w), grid_bfs(B.cells, w))
    assert ws.device_bytes() > 0
    mixed = B.query_batch([w, w[:3], np.zeros((0, 2), np.int32)], ws=ws)
    np.testing.assert_array_equal(mixed[0], grid_bfs(B.cells, w))
    assert mixed[2].shape == (0, 0)
    ws.close()


def test_failed_update_leaves_a_map_that_refuses_until_rebuilt(gpu_ctx):
    """An update that fails part way (an injected HIP error after its first
    level's kernels) reports the error; the map then answers HTG_E_STATE to
    queries, updates, lifts and clones until htg_build, which rebuilds it from
    its cell state (the batch included): tables equal the oracle, queries
    equal BFS, and reveals continue exactly."""
    truth = grids.street_grid(128, 21)
    obs = grids.disk_observed(truth, [(60, 70)], 35)
    order = grids.frontier_order(truth, obs)
    G = _lib.GpuMap(obs, 16, 2)
    G.build()
    G.open_cells(order[:100])
    G.inject_fault(_lib.FAULT_UPDATE, 1)
    with pytest.raises(_lib.HtgError) as ei:
        G.open_cells(order[100:300])
    assert ei.value.code == _lib.HTG_E_HIP and "injected" in str(ei.value)
    v = grids.sample_free(G.cells, 16, 3)
    for call in (lambda: G.query(v), lambda: G.open_cells(order[300:310]), lambda: G.lift_point(60, 70),
                 lambda: G.clone()):
        with pytest.raises(_lib.HtgError) as ei:
            call()
        assert ei.value.code == _lib.HTG_E_STATE
    assert (G.cells[order[300:310, 0], order[300:310, 1]] != 1).all()  # refused: not applied
    assert (G.cells[order[100:300, 0], order[100:300, 1]] == 1).all()  # failed: the state holds the batch
    assert G.build()
    O = Oracle("port", G.cells, 16, 2)
    assert O.build()
    assert_tables_equal(G, O)
    np.testing.assert_array_equal(G.query(v), grid_bfs(G.cells, v))
    G.open_cells(order[300:600])
    O = Oracle("port", G.cells, 16, 2)
    assert O.build()
    assert_tables_equal(G, O)
    # a failed query launch fails that call only; a failed build leaves the old tables
    G.inject_fault(_lib.FAULT_QUERY, 1)
    with pytest.raises(_lib.HtgError) as ei:
        G.query(v)
    assert ei.value.code == _lib.HTG_E_HIP
    np.testing.assert_array_equal(G.query(v), grid_bfs(G.cells, v))
    G.inject_fault(_lib.FAULT_BUILD, 1)
    with pytest.raises(_lib.HtgError):
        G.build()
    np.testing.assert_array_equal(G.query(v), grid_bfs(G.cells, v))
    G.inject_fault(_lib.FAULT_UPDATE, 3)  # the third update fails
    G.open_cells(order[600:650])
    G.open_cells(order[650:700])
    G.inject_fault(_lib.FAULT_UPDATE, 0)  # disarmed before it fired
    G.open_cells(order[700:750])
    np.testing.assert_array_equal(G.query(v), grid_bfs(G.cells, v))


def states_by_version(trace, init, sid):
    obs = init.copy()
    out = [obs.copy()]
    for i in range(len(trace.t)):
        if trace.type[i] == 1 and trace.session[i] == sid:
            c = trace.cells_of(i)
            obs[c[:, 0], c[:, 1]] = 1
            out.append(obs.copy())
    return out


@pytest.mark.parametrize("consistency", [serving.STRICT, serving.SNAPSHOT])
def test_server_recovers_a_failed_update_and_other_sessions_stay_exact(gpu_ctx, consistency):
    """A HIP error injected into one session's update: the server rebuilds
    that copy from its cell state, the update completes, and every result of
    both sessions equals BFS on its reported state."""
    print(f"server fault consistency {consistency}", flush=True)
    srv = serving.Server(n_slots=3, max_batch_queries=16, consistency=consistency)
    sess = []
    for seed, n, b in ((31, 160, 16), (32, 128, 8)):
        truth = grids.street_grid(n, seed)
        rv = serving.RobotReveal(truth, 3, 20, seed + 1, limit=3000)
        sess.append((srv.add_map(rv.observed, b), rv, rv.observed.copy()))
    srv.inject_fault(sess[0][0], _lib.FAULT_UPDATE, 1)
    trs = [serving.make_trace(sid, rv, 90, 4000.0, 40 + i, update_frac=0.3, k_lo=2, k_hi=40, upd_lo=8, upd_hi=60)
           for i, (sid, rv, _) in enumerate(sess)]
    tr = serving.merge_traces(trs)
    srv.replay(tr)  # raises if any request failed
    assert srv.stats()["recovered_faults"] == 1
    for sid, _, init in sess:
        sv = states_by_version(tr, init, sid)
        for i in range(len(tr.t)):
            if tr.session[i] != sid or tr.type[i] != 0:
                continue
            np.testing.assert_array_equal(tr.query_result(i), grid_bfs(sv[int(tr.version[i])], tr.cells_of(i)),
                                          err_msg=f"session {sid} request {i}")
    # a failed query batch fails its own requests only
    srv.inject_fault(sess[1][0], _lib.FAULT_QUERY, 1)
    v = grids.sample_free(sess[1][2], 12, 5)
    t, _ = srv.submit_query(sess[1][0], v)
    assert srv.wait_status(t) == _lib.HTG_E_HIP
    final = states_by_version(tr, sess[1][2], sess[1][0])[-1]
    t, out = srv.submit_query(sess[1][0], v)
    srv.wait(t)
    np.testing.assert_array_equal(out, grid_bfs(final, v))
    srv.close()


def test_server_cancels_requests_that_have_not_started(gpu_ctx):
    """Cancelled requests complete with HTG_E_CANCELLED; a cancelled update is
    never applied; every other result equals BFS on the state of the updates
    that were not cancelled (strict order)."""
    srv = serving.Server(n_slots=1, max_batch_queries=1, consistency=serving.STRICT)
    truth = grids.street_grid(256, 41)
    rv = serving.RobotReveal(truth, 2, 30, 42, limit=5000)
    init = rv.observed.copy()
    sid = srv.add_map(init, 16)
    reqs = []  # (kind, ticket, out or cells)
    for i in range(60):
        if i % 6 == 3:
            cells = rv.next_batch(i % 2, 40)
            reqs.append(("u", srv.submit_update(sid, cells), cells))
        else:
            vps = rv.sample_viewpoints(24, grids.SplitMix64(i))
            t, out = srv.submit_query(sid, vps)
            reqs.append(("q", t, (vps, out)))
    cancelled = set()
    for j in range(len(reqs) - 1, 20, -4):  # from the back: these have not started yet
        rc = srv.cancel(reqs[j][1])
        assert rc in (_lib.HTG_OK, _lib.HTG_E_STATE)
        if rc == _lib.HTG_OK:
            cancelled.add(j)
    assert len(cancelled) >= 5
    assert srv.cancel(reqs[max(cancelled)][1]) == _lib.HTG_E_STATE  # already completed (cancelled)
    state = init.copy()
    for j, (kind, t, payload) in enumerate(reqs):
        st = srv.wait_status(t)
        if j in cancelled:
            assert st == _lib.HTG_E_CANCELLED
            continue
        assert st == _lib.HTG_OK
        if kind == "u":
            state[payload[:, 0], payload[:, 1]] = 1
        else:
            np.testing.assert_array_equal(payload[1], grid_bfs(state, payload[0]), err_msg=f"request {j}")
    assert srv.cancel(123456789) == _lib.HTG_E_RANGE
    srv.close()


def test_python_mirror_of_the_reference_api(gpu_ctx):
    """htrip_amd.MinPlusOracle: build / onBatchCellsOpened / onCellOpened /
    onCellBlocked / queryAllPairs (with out and scratch) / queryDistance /
    findLcaLevel / effectiveMemoryBytes, against BFS, the oracle and the
    compiled reference."""
    truth = grids.street_grid(200, 51)
    obs = grids.disk_observed(truth, [(100, 90)], 40)
    grid = ht.GridMap(cells=obs)
    o = ht.MinPlusOracle(ht.HierarchyConfig(16, 2, 2))
    assert not ht.MinPlusOracle(ht.HierarchyConfig(0, 2, 2)).build(grid)
    assert o.build(grid) and o.depth() == o.H == 4 and o.S == 256
    order = grids.frontier_order(truth, obs)
    grid.observed[order[:300, 0], order[:300, 1]] = ht.FREE
    o.onBatchCellsOpened(order[:300], grid)
    O = Oracle("port", obs, 16, 2)
    O.build()
    O.open_cells(order[:300])
    assert o.last_batch_stats == O.last_batch_stats()
    r, c = order[300]
    grid.observed[r, c] = ht.FREE
    o.onCellOpened((r, c), grid)
    vps = grids.sample_free(grid.observed, 30, 7)
    scratch = ht.HTripQueryScratch()
    out = np.empty(0, dtype=np.uint16)
    o.queryAllPairs(vps, grid, out, scratch)
    expect = grid_bfs(grid.observed, vps)
    np.testing.assert_array_equal(out.reshape(30, 30), expect)
    np.testing.assert_array_equal(o.queryAllPairs(vps, grid), expect)
    assert o.queryDistance(tuple(vps[0]), tuple(vps[1]), grid) == expect[0, 1]
    assert o.queryDistance(tuple(vps[2]), tuple(vps[2]), grid) == 0
    blocked = tuple(int(x) for x in np.argwhere(grid.observed != ht.FREE)[0])
    assert o.queryDistance(blocked, tuple(vps[1]), grid) == ht.INF
    assert o.queryDistance((-1, 5), tuple(vps[1]), grid) == ht.INF
    assert o.findLcaLevel((0, 0), (0, 1)) == 0 and o.findLcaLevel((0, 0), (0, 16)) == 1
    assert o.findLcaLevel((0, 0), (255, 255)) == 4
    blk = tuple(vps[3])
    grid.observed[blk] = ht.OBSTACLE
    o.onCellBlocked(blk, grid)
    np.testing.assert_array_equal(o.queryAllPairs(vps, grid, scratch=scratch), grid_bfs(grid.observed, vps))
    if orc.available("reference"):
        R = Oracle("reference", grid.observed, 16, 2)
        assert R.build()
        for g in (grid.observed, obs, obs[:150, :120]):
            g = np.ascontiguousarray(g)
            want = int(R.api["effective_memory_bytes"](R.h, g.shape[0], g.shape[1], g))
            assert o.effectiveMemoryBytes(ht.GridMap(cells=g)) == want
    assert o.totalMemoryBytes() == O.total_memory_bytes()
    scratch.close()
