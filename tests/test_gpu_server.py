"""Serving loop on the GPU: continuous batching of mixed-k queries and update
batches, admission by HBM budget, several sessions at once.

Exactness bar: every query result equals BFS on the map state that contains
exactly the updates submitted before it -- the same result the query gives
when run alone on that state (bit-exact).
"""
import numpy as np
import pytest

from htrip_amd import _lib, grids, serving
from oracle.oracle import grid_bfs

pytestmark = pytest.mark.gpu


def states_after(trace, init_obs):
    """Observed grid after each request prefix (applies update cells in order)."""
    obs = init_obs.copy()
    out = []
    for i in range(len(trace.t)):
        if trace.type[i] == 1:
            c = trace.cells_of(i)
            obs[c[:, 0], c[:, 1]] = 1
        out.append(obs.copy() if trace.type[i] == 0 else None)
    return out


def states_by_version(trace, init_obs):
    """Observed grid after the first v update requests, for every v."""
    obs = init_obs.copy()
    out = [obs.copy()]
    for i in range(len(trace.t)):
        if trace.type[i] == 1:
            c = trace.cells_of(i)
            obs[c[:, 0], c[:, 1]] = 1
            out.append(obs.copy())
    return out


@pytest.mark.parametrize("rate,slots,batch_q", [(5e4, 4, 16), (2e3, 2, 4), (1e9, 3, 64)])
def test_mixed_stream_alone_equals_in_the_mix(gpu_ctx, rate, slots, batch_q):
    print(f"server stream rate={rate} slots={slots} batch={batch_q}", flush=True)
    truth = grids.street_grid(192, 71)
    rv = serving.RobotReveal(truth, 4, 20, 72, limit=4000)
    init = rv.observed.copy()
    srv = serving.Server(n_slots=slots, max_batch_queries=batch_q, max_batch_viewpoints=2048)
    sid = srv.add_map(init, 16)
    tr = serving.make_trace(sid, rv, 160, rate, 73, update_frac=0.25, k_lo=2, k_hi=48, upd_lo=4, upd_hi=80)
    t_done = srv.replay(tr)
    assert (t_done >= tr.t - 1e-6).all()
    st = states_after(tr, init)
    checked = 0
    for i in range(len(tr.t)):
        if tr.type[i] != 0:
            continue
        vps = tr.cells_of(i)
        np.testing.assert_array_equal(tr.query_result(i), grid_bfs(st[i], vps), err_msg=f"request {i}")
        checked += 1
        if checked % 9 == 0:  # the same query alone on a fresh map of that state
            alone = _lib.GpuMap(st[i], 16)
            alone.build()
            np.testing.assert_array_equal(alone.query(vps), tr.query_result(i))
            alone.close()
    s = srv.stats()
    assert s["queries"] == int((tr.type == 0).sum()) and s["updates"] == int((tr.type == 1).sum())
    srv.close()


@pytest.mark.parametrize("rate", [3e4, 1e9])
def test_snapshot_stream_each_result_matches_its_snapshot(gpu_ctx, rate):
    """Snapshot (double-buffered) mode: queries never wait for updates; each
    result equals BFS on the state with exactly the update prefix it reports,
    which equals running it alone on that state."""
    print(f"snapshot stream rate={rate}", flush=True)
    truth = grids.street_grid(192, 75)
    rv = serving.RobotReveal(truth, 4, 20, 76, limit=4000)
    init = rv.observed.copy()
    srv = serving.Server(n_slots=3, max_batch_queries=16, max_batch_viewpoints=2048,
                         consistency=serving.SNAPSHOT)
    sid = srv.add_map(init, 16)
    tr = serving.make_trace(sid, rv, 200, rate, 77, update_frac=0.3, k_lo=2, k_hi=48, upd_lo=4, upd_hi=80)
    srv.replay(tr)
    sv = states_by_version(tr, init)
    n_upd = int((tr.type == 1).sum())
    seen = []
    prev_max = 0
    for i in range(len(tr.t)):
        if tr.type[i] != 0:
            continue
        v = int(tr.version[i])
        assert 0 <= v <= n_upd
        np.testing.assert_array_equal(tr.query_result(i), grid_bfs(sv[v], tr.cells_of(i)), err_msg=f"req {i} v {v}")
        seen.append(v)
    if rate < 1e6:  # paced arrivals: updates land while queries run (a burst may finish first)
        assert max(seen) > 0
    # after draining, a fresh query sees every update
    vps = grids.sample_free(sv[-1], 30, 5)
    t, out = srv.submit_query(sid, vps)
    _, v = srv.wait(t, with_version=True)
    assert v == n_upd
    np.testing.assert_array_equal(out, grid_bfs(sv[-1], vps))
    srv.close()


def test_admission_by_memory_budget(gpu_ctx):
    cells = grids.street_grid(128, 81)
    need = _lib.footprint(16, 2, 128, 128)
    srv = serving.Server(hbm_budget=int(2.5 * need), n_slots=2)
    a = srv.add_map(cells, 16)
    b = srv.add_map(cells, 16)
    with pytest.raises(_lib.HtgError) as ei:
        srv.add_map(cells, 16)
    assert ei.value.code == _lib.HTG_E_NOMEM
    vps = grids.sample_free(cells, 20, 3)
    t, out = srv.submit_query(b, vps)
    srv.wait(t)
    np.testing.assert_array_equal(out, grid_bfs(cells, vps))
    srv.remove_map(a)  # frees its share: a new session fits again
    c = srv.add_map(cells, 16)
    t, out = srv.submit_query(c, vps)
    srv.wait(t)
    np.testing.assert_array_equal(out, grid_bfs(cells, vps))
    assert srv.stats()["sessions"] == 2
    srv.close()


def test_sessions_overlap_and_keep_order(gpu_ctx):
    """Two sessions with interleaved traffic; a query submitted before an update
    does not see it, one submitted after does."""
    srv = serving.Server(n_slots=4, max_batch_queries=32)
    t1 = grids.street_grid(96, 91)
    t2 = grids.maze_grid(96, 92)
    o1 = grids.disk_observed(t1, [tuple(grids.sample_free(t1, 1, 1)[0])], 25)
    o2 = grids.disk_observed(t2, [tuple(grids.sample_free(t2, 1, 2)[0])], 25)
    s1, s2 = srv.add_map(o1, 8), srv.add_map(o2, 16)
    f1 = grids.frontier_order(t1, o1, 600)
    f2 = grids.frontier_order(t2, o2, 600)
    pending = []
    state = {s1: o1.copy(), s2: o2.copy()}
    for step in range(6):
        for sid, fr in ((s1, f1), (s2, f2)):
            vps = grids.sample_free(state[sid], 24, step * 7 + sid)
            t, out = srv.submit_query(sid, vps)
            pending.append((t, out, grid_bfs(state[sid], vps)))
            cells = fr[step * 90:(step + 1) * 90]
            srv.submit_update(sid, cells)
            state[sid][cells[:, 0], cells[:, 1]] = 1
    for t, out, expect in pending:
        srv.wait(t)
        np.testing.assert_array_equal(out, expect)
    srv.close()
