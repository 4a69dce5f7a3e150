#!/usr/bin/env python3
"""Benchmark of the MI355X min-plus distance oracle (one JSON line on stdout).

Workload (serving configuration): a 1024 x 1024 street-like seeded grid
(seed 3), leaf b = 32, group 2 x 2 (H = 5).  16 simulated robots start with
observed disks of radius 64 and reveal free cells in their frontier order.
The hierarchy is built once and stays resident in HBM.

Modes
  served (default) -- the continuous-batching server replays a Poisson request
      stream: queryAllPairs with k log-uniform in [8, 256] interleaved with
      update batches of 16..512 revealed cells (10% of requests).  Each
      session's 16 robots explore from their start disks to the fully revealed
      map; the stream runs over as many resident sessions as it needs so that
      the timed part holds at least 1 s of arrivals and 50 000 requests
      (--steps windows of at least --window requests, after --warmup untimed
      windows).  `value` = query matrices/s over the timed part at the highest
      tested arrival rate whose query p99 < 1 ms and update p99 < 25 ms.  The
      plain per-call GPU path is measured with the same k mix (and k = 64) at
      the start, half and fully revealed states, and short served trials at
      each of those states give the ratio per state.
  plain -- one step = one plain per-call queryAllPairs (k = 64) through the
      drop-in C ABI (host viewpoints in, host matrix out, synchronous).
  update -- onBatchCellsOpened latency on the serving map (robot batches).
  small -- 256^2 street grid, b=16: build + queryAllPairs k=64.
  build -- 512^2 maze (width 1), b=16: batched build (graph-captured,
      autotuned) + queryAllPairs k=128.
  reveal -- 512^2 street grid, b=8: 500 steps of 256 reveals, each followed by
      queryAllPairs k=32; every step's matrix is checked against BFS, and
      the reference replays the steps within its CPU budget.
  multi -- 8 resident sessions of mixed shapes, mixed traffic (k <= 512,
      reveals <= 1024 cells), BFS spot checks.

Timing: request latency = completion - arrival on the server's clock
(steady_clock); kernel durations from hipEvents recorded on the stream each
kernel runs on.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# One hardware queue per stream: the served mode runs 4 query slots, the update
# stream and the plain map's stream.  With HIP's default of 4, two query slots
# share a queue and run in order behind each other.  Read when the HIP runtime
# starts, so it is set before the library is loaded.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

PEAK_HBM_GBS = 8000.0
# 256 CUs x 64 lanes x 2 u16 per packed op x 2.4 GHz = 78.6 T u16 ops/s; one
# min-plus update is a packed add + a packed min: 39.3 T updates/s.
PEAK_UPDATES = 256 * 64 * 2 * 2.4e9 / 2

N, B, G, SEED, ROBOTS, RADIUS = 1024, 32, 2, 3, 16, 64
# queries per batch of the kernel phase (the workload of the rocprofv3 summaries in profiles/)
KERNEL_BATCH = 16


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_stats_of(gmap):
    from htrip_amd import _lib

    return {c: gmap.kernel_stats(c) for c in range(len(_lib.KERNEL_NAMES))}


TRAFFIC_FILE = "profiles/k_q_z_traffic_kernel_phase_1024_street.txt"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the kernel-phase workload, from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    TRAFFIC_FILE (None when that file is absent)."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_FILE)):
            if line.startswith(f"traffic_bytes_per_launch {kernel} = "):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def roofline_of(stats):
    """Roofline of the query projection (k_q_z) from accumulated kernel stats:
    bytes and min-plus updates are counted on the device, time from start/stop
    events recorded on the query stream around each dispatch."""
    from htrip_amd import _lib

    z = stats[_lib.K_Q_PROJECT]
    n = max(z["launches"], 1)
    avg_ms = z["total_ms"] / n
    gbs = (z["bytes"] / n) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    ups = (z["updates"] / n) / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "traffic": measured_traffic("k_q_z"), "kernel": "k_q_z", "avg_launch_ms": avg_ms,
            "launches": z["launches"], "traffic_source": TRAFFIC_FILE}
    valu = {"kernel": "k_q_z", "achieved_updates_per_s": ups, "peak_updates_per_s": PEAK_UPDATES,
            "frac": ups / PEAK_UPDATES}
    kernels = {_lib.KERNEL_NAMES[c]: stats[c] for c in range(len(_lib.KERNEL_NAMES))}
    return roof, valu, kernels


def kernel_roofline(gmap):
    return roofline_of(kernel_stats_of(gmap))


def add_stats(acc, st):
    for c, v in st.items():
        a = acc.setdefault(c, {"launches": 0, "total_ms": 0.0, "bytes": 0.0, "updates": 0.0})
        for f in a:
            a[f] += v[f]
    return acc


def plain_rate(gmap, cells, k, steps, warmup, seed=1000):
    from htrip_amd import grids

    qs = [grids.sample_free(cells, k, seed + i) for i in range(steps + warmup)]
    for q in qs[:warmup]:
        gmap.query(q)
    gmap.sync()
    t0 = time.perf_counter()
    res = [gmap.query(q) for q in qs[warmup:]]
    gmap.sync()
    dt = time.perf_counter() - t0
    return steps / dt, dt, qs[warmup:], res


def kernel_phase(gmap, src, batch_q, n_batches, seed=4242, events=True):
    """Dispatch timing of served-shaped query batches (batch_q queries of k ~
    logU[8,256], viewpoints from src's revealed cells), one batch at a time on
    the map's own stream with start/stop events on every dispatch.  The
    latency trials run without events: a dispatch that carries them can hold
    the other queues back for milliseconds, and under contention a kernel's
    duration measures its neighbours as much as itself."""
    from htrip_amd import grids, serving

    rng = grids.SplitMix64(seed)
    batches = [[src.sample_viewpoints(serving.loguniform_k(rng, 8, 256), rng) for _ in range(batch_q)]
               for _ in range(n_batches + 10)]
    for b in batches[:10]:
        gmap.query_batch(b)
    gmap.profile(events, every=1)
    t0 = time.perf_counter()
    for b in batches[10:]:
        gmap.query_batch(b)
    dt = time.perf_counter() - t0
    stats = kernel_stats_of(gmap)
    gmap.profile(False)
    return stats, n_batches * batch_q / dt


def import_reference(gmap, cells):
    """The compiled reference (oracle/_ref) laid out over tables copied from the
    GPU, so that its query path runs without its CPU build; None without it."""
    from oracle import oracle as orc

    if not orc.available("reference"):
        return None
    R = orc.Oracle("reference", cells, B, G)
    api = R.api
    t0 = time.time()
    api["import_begin"](R.h)
    n = B * B
    buf = np.empty((n, n), dtype=np.uint16)
    for i in range(gmap.num_leaves()):
        gmap.lib.htg_read_leaf_D(gmap.h, i, buf.ctypes.data_as(C.POINTER(C.c_uint16)))
        api["import_leaf"](R.h, i, buf)
    for lvl in range(1, gmap.H + 1):
        for i in range(gmap.num_nodes(lvl)):
            api["import_node"](R.h, lvl, i, gmap.node_DU(lvl, i))
    log(f"cpu baseline: GPU tables imported into the reference in {time.time() - t0:.1f}s")
    return R


def cpu_baseline(gmap, cells, queries, budget_s, expect, served_checks=()):
    """The reference's own CPU AVX2 queryAllPairs (oracle/_ref), one thread, on
    a bounded sample of the same queries, over tables copied from the GPU.
    Also the exactness gate: every timed CPU result must equal the GPU result,
    and the sampled served results must equal BFS on their map state."""
    from oracle import oracle as orc

    for state, vps, got in served_checks:
        if not np.array_equal(orc.grid_bfs(state, vps), got):
            raise SystemExit("served result differs from BFS on its map state")
    if served_checks:
        log(f"exactness: {len(served_checks)} served matrices equal BFS on their snapshot states")
    R = import_reference(gmap, cells)
    if R is None:
        return None
    api = R.api
    done, spent = 0, 0.0
    for _ in range(50):  # the query list, cyclically, until the CPU budget is spent
        for q, vps in enumerate(queries):
            vps = np.ascontiguousarray(vps, dtype=np.int32)
            ks = np.array([len(vps)], dtype=np.int32)
            out = np.empty(len(vps) * len(vps), dtype=np.uint16)
            spent += api["time_queries"](R.h, vps, ks, 1, 1, out.ctypes.data)
            if not np.array_equal(out.reshape(len(vps), -1), expect[q]):
                raise SystemExit("reference CPU result differs from the GPU result")
            done += 1
            if spent >= budget_s:
                break
        if spent >= budget_s:
            break
    log(f"cpu baseline: {done} queries in {spent:.2f}s, all equal to the GPU results")
    ks = sorted({len(v) for v in queries})
    kdesc = f"k={ks[0]}" if len(ks) == 1 else f"k in [{ks[0]}, {ks[-1]}] (the served mix)"
    one = {"value": done / spent, "unit": "matrices/s", "cores": 1, "kind": "reference",
           "sample": f"{done} queryAllPairs {kdesc} (1 thread, reference AVX2 sources) on the same map"}
    one["all_cores"] = cpu_all_cores(R, queries, expect)
    return one


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))  # the GPU box's CPU share


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_all_cores(R, queries, expect, reps=10):
    """The reference's queryAllPairs on every host core (independent queries
    over std::thread, each with its own scratch), same exactness gate; the
    query list `reps` times over."""
    api = R.api
    cores = host_cores()
    expect = list(expect) * reps
    queries = list(queries) * reps
    ks = np.array([len(v) for v in queries], dtype=np.int32)
    allv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32) for v in queries]), dtype=np.int32)
    out = np.empty(int((ks.astype(np.int64) ** 2).sum()), dtype=np.uint16)
    spent = api["time_queries"](R.h, allv, ks, len(ks), cores, out.ctypes.data)
    o = 0
    for q, k in enumerate(ks):
        if not np.array_equal(out[o:o + k * k].reshape(k, k), expect[q]):
            raise SystemExit("reference CPU result (all cores) differs from the GPU result")
        o += int(k) * int(k)
    log(f"cpu baseline, {cores} threads: {len(ks)} queries in {spent:.2f}s")
    return {"value": len(ks) / spent, "unit": "matrices/s", "cores": cores, "kind": "reference",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{len(ks)} queryAllPairs ({cores} threads over independent queries)"}


def run_plain(args):
    from htrip_amd import _lib, grids

    if args.map:  # a shipped MovingAI map, fully revealed (campaign_support.cpp:150-159)
        cells = grids.load_movingai(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                                 "maps", args.map + ".map.gz"))
    else:
        cells = grids.street_grid(N, SEED)
    gmap = _lib.GpuMap(cells, B, G)
    gmap.profile(True)
    t0 = time.time()
    gmap.build()
    log(f"build {1e3 * (time.time() - t0):.1f} ms (S={gmap.S} H={gmap.H})")
    gmap.profile(False)
    rate, dt, qs, res = plain_rate(gmap, cells, args.k, args.steps, args.warmup)
    log(f"plain: {args.steps} queries k={args.k}: {rate:.0f} matrices/s")
    # kernel figures from a second pass over the same queries with every
    # dispatch timed (the timed pass above carries no events)
    gmap.profile(True, every=1)
    for q in qs:
        gmap.query(q)
    roof, valu, kernels = kernel_roofline(gmap)
    cpu = None if args.no_cpu else cpu_baseline(gmap, cells, qs, args.cpu_seconds, res)
    return {
        "metric": "k x k query matrices/s (plain per-call GPU path, k=64, 1024^2 grid)",
        "value": rate, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u16", "data": "synthetic",
        "config": {"workload": (f"MovingAI {args.map} fully revealed" if args.map else
                                "1024x1024 street-like grid seed 3") + f", b=32, g=2x2, plain queryAllPairs k={args.k}",
                   "mode": "plain", "k": args.k},
        "roofline": roof, "valu": valu, "cpu_baseline": cpu, "kernels": kernels,
    }


def mix_queries(cells, n, seed, k_lo=8, k_hi=256):
    """n queries with k ~ logU[k_lo, k_hi] over the free cells of `cells`."""
    from htrip_amd import grids, serving

    rng = grids.SplitMix64(seed)
    return [grids.sample_free(cells, serving.loguniform_k(rng, k_lo, k_hi), seed * 7919 + i) for i in range(n)]


def plain_rate_of(gmap, qs, warmup):
    """The plain per-call path: one synchronous htg_query_all_pairs per query."""
    for q in qs[:warmup]:
        gmap.query(q)
    gmap.sync()
    t0 = time.perf_counter()
    res = [gmap.query(q) for q in qs[warmup:]]
    gmap.sync()
    dt = time.perf_counter() - t0
    return (len(qs) - warmup) / dt, res


def served_trial(srv, explorations, n_warm, n_timed, rate, seed, uf, fresh_sessions):
    """One Poisson trial over fresh sessions: n_warm untimed requests, then
    n_timed timed ones (statistics over the timed part only)."""
    from htrip_amd import serving

    sids, exs, inits = fresh_sessions(explorations)
    tr = serving.exploration_trace(sids, exs, n_warm + n_timed, rate, seed, update_frac=uf)
    st0 = srv.stats()
    c0, w0 = os.times(), time.perf_counter()
    td = srv.replay(tr)
    c1, w1 = os.times(), time.perf_counter()
    st1 = srv.stats()
    timed = serving.Trace(tr.t[n_warm:], tr.session[n_warm:], tr.type[n_warm:], tr.count[n_warm:],
                          tr.rc_off[n_warm:], tr.rc, tr.out_off[n_warm:], tr.out)
    s = serving.latency_summary(timed, td[n_warm:])
    s["rate"] = rate
    s["sessions"] = len(sids)
    s["requests"] = n_warm + n_timed
    s["warmup_requests"] = n_warm
    s["timed_arrival_span_s"] = float(tr.t[-1] - tr.t[n_warm])
    s["explored_at_end"] = [round(x, 3) for x in tr.meta["explored"]]
    s["cpu_cores"] = ((c1.user - c0.user) + (c1.system - c0.system)) / max(w1 - w0, 1e-9)
    nb = st1["query_batches"] - st0["query_batches"]
    s["mean_batch_queries"] = (st1["queries"] - st0["queries"]) / max(nb, 1)
    s["update_busy"] = (st1["update_seconds"] - st0["update_seconds"]) / max(w1 - w0, 1e-9)
    s["batch_ms"] = {k: 1e3 * (st1[f] - st0[f]) / max(nb, 1) for k, f in
                     (("plan_launch", "plan_seconds"), ("gpu_wait", "gpu_wait_seconds"), ("finish", "finish_seconds"))}
    return tr, td, s, sids, inits


def passes(s, lag_ms):
    """Query p99 under 1 ms and the updates keep up (their p99 lag stays within
    a few merged batches; a growing backlog would mean the stream is not
    sustained)."""
    return s["p99_ms"] < 1.0 and (s["update_p99_ms"] is None or s["update_p99_ms"] < lag_ms)


def run_served(args):
    """The served form on the 1024^2 serving map (see the module docstring)."""
    from htrip_amd import _lib, grids, serving
    from oracle import oracle as orc

    t0 = time.time()
    truth = grids.street_grid(N, SEED)
    n_free = int((truth == grids.FREE).sum())
    uf = args.update_frac
    mode = serving.SNAPSHOT if args.consistency == "snapshot" else serving.STRICT
    srv = None
    fleets = []  # one robot fleet per session, seeds 5, 12, 19, ... (orders computed once)

    def fleet(j):
        while len(fleets) <= j:
            fleets.append(serving.Exploration(truth, ROBOTS, RADIUS, seed=5 + 7 * len(fleets)))
        return fleets[j]

    def fresh_sessions(exs):
        sids, inits = [], []
        for ex in exs:
            obs = ex.observed_now()
            sid = srv.add_map(obs, B, G)
            # warm the session's slots with queries only (the map stays as built)
            warm = serving.exploration_trace([sid], [ex.fresh()], 64, 1e6, 11 + sid, update_frac=0.0)
            srv.replay(warm)
            sids.append(sid)
            inits.append(obs)
        return sids, exs, inits

    def close(sids):
        for sid in sids:
            srv.remove_map(sid)

    log(f"grid {N}x{N}: {n_free} free; robot fleets ({time.time() - t0:.1f}s)")
    # ---- exploration states: start disks, half revealed, fully revealed
    ex0 = fleet(0).fresh()
    half = fleet(0).fresh()
    half.advance_to(0.5)
    states = {"start": ex0.observed_now(), "half": half.observed_now(), "full": truth.copy()}
    plain = {}
    plain_maps = {}
    for name, cells in states.items():
        gm = _lib.GpuMap(cells, B, G)
        gm.build()
        q64 = [grids.sample_free(cells, 64, 1000 + i) for i in range(args.plain_steps + args.warmup)]
        qmix = mix_queries(cells, args.plain_steps + args.warmup, 77)
        r64, res64 = plain_rate_of(gm, q64, args.warmup)
        rmix, resmix = plain_rate_of(gm, qmix, args.warmup)
        plain[name] = {"explored": round(float((cells == 1).sum()) / n_free, 3), "k64_matrices_per_s": r64,
                       "kmix_matrices_per_s": rmix}
        plain_maps[name] = (cells, qmix[args.warmup:], resmix)
        gm.close()  # its stream would take one of the process's hardware queues from the server
        log(f"plain per-call at {name} ({100 * plain[name]['explored']:.0f}% explored): k=64 {r64:.0f}/s, "
            f"k~logU[8,256] {rmix:.0f}/s")
    srv = serving.Server(n_slots=args.slots, max_batch_queries=args.batch_queries,
                         max_batch_viewpoints=args.batch_viewpoints, consistency=mode,
                         update_cu_percent=args.update_cu, update_tick_us=args.update_tick_us,
                         update_streams=args.update_streams or 1)
    # plain figure over the exploration: harmonic mean of the three states
    plain_mix = 3.0 / sum(1.0 / plain[n]["kmix_matrices_per_s"] for n in plain)
    plain_64 = 3.0 / sum(1.0 / plain[n]["k64_matrices_per_s"] for n in plain)

    # ---- capacity estimate: a burst at the start state
    sids, exs, _ = fresh_sessions([fleet(0).fresh()])
    burst = serving.exploration_trace(sids, exs, args.window * 16, 1e7, 12, update_frac=uf)
    tb = srv.replay(burst)
    close(sids)
    cap = serving.latency_summary(burst, tb)["matrices_per_s"] / 0.9
    log(f"burst capacity ~{cap:.0f} requests/s")

    # ---- headline: whole explorations, at least 1 s of arrivals and 50K requests timed
    per_expl = (n_free - int((states["start"] == 1).sum())) / 143.1 / max(uf, 1e-9)  # requests per exploration
    trials, chosen = [], None
    for frac in args.rate_fracs:
        rate = cap * frac
        n_timed = int(max(args.steps * args.window, 50000, rate * 1.0))
        n_warm = args.warmup * int(math.ceil(n_timed / args.steps))
        n_sess = max(1, int(math.ceil((n_warm + n_timed) * 1.05 / per_expl))) if uf > 0 else 1
        tr, td, s, sids, inits = served_trial(srv, [fleet(j).fresh() for j in range(n_sess)], n_warm, n_timed, rate,
                                             100 + len(trials), uf, fresh_sessions)
        close(sids)
        trials.append((tr, td, s, sids, inits))
        log(f"rate {rate:.0f}/s x {s['timed_arrival_span_s']:.2f}s ({n_sess} sessions, {n_timed} timed + {n_warm} "
            f"warmup requests): {s['matrices_per_s']:.0f} matrices/s, p50 {s['p50_ms']:.3f} ms, p99 "
            f"{s['p99_ms']:.3f} ms, updates p99 {s['update_p99_ms']} ms, batch {s['mean_batch_queries']:.2f}, "
            f"explored at end {s['explored_at_end']}, cpu {s['cpu_cores']:.1f} cores")
        if passes(s, args.update_lag_ms):
            chosen = len(trials) - 1
            break
    if chosen is None:
        chosen = len(trials) - 1
    tr, td, s, sids, inits = trials[chosen]
    # exactness of the chosen trial: its checked results against BFS on the
    # state of the update prefix each reports
    n_checked = 0
    if not args.no_cpu:
        sidx = tr.meta["session_index"]
        for i in tr.meta["checks"][::max(1, len(tr.meta["checks"]) // 24)]:
            if tr.type[i] != 0 or tr.count[i] > 128:
                continue
            j = int(sidx[i])
            st = serving.state_at(inits[j], tr.meta["upd_cells"][j], int(tr.version[i]))
            if not np.array_equal(orc.grid_bfs(st, tr.cells_of(i), orc.bfs_threads()), tr.query_result(i)):
                raise SystemExit(f"served result {i} differs from BFS on its map state")
            n_checked += 1
        log(f"exactness: {n_checked} served matrices of the chosen trial equal BFS on their states")

    # ---- per exploration state: served (short trials, one session) against plain
    by_state = {}
    for name in (() if args.quick else ("start", "half", "full")):
        best = None
        for frac in (1.3, 1.1, 0.95, 0.8, 0.65, 0.5, 0.35, 0.25):
            ex = fleet(0).fresh()
            if name == "half":
                ex.advance_to(0.5)
            elif name == "full":
                ex.advance_to(1.0)
            _, _, st, ss, _ = served_trial(srv, [ex], args.window * 2, 20000, cap * frac, 500 + len(by_state), uf,
                                           fresh_sessions)
            close(ss)
            if passes(st, args.update_lag_ms):
                best = st
                break
        best = best or st
        by_state[name] = {"served_matrices_per_s": best["matrices_per_s"], "p50_ms": best["p50_ms"],
                          "p99_ms": best["p99_ms"], "update_p99_ms": best["update_p99_ms"],
                          "explored_at_end": best["explored_at_end"], **plain[name],
                          "served_over_plain_kmix": best["matrices_per_s"] / plain[name]["kmix_matrices_per_s"],
                          "served_over_plain_k64": best["matrices_per_s"] / plain[name]["k64_matrices_per_s"]}
        log(f"state {name}: served {best['matrices_per_s']:.0f}/s (p99 {best['p99_ms']:.3f} ms) vs plain k-mix "
            f"{plain[name]['kmix_matrices_per_s']:.0f}/s: {by_state[name]['served_over_plain_kmix']:.1f}x")

    server_stats = srv.stats()
    srv.close()
    if args.quick:
        return {"metric": "served (quick exploration: headline trials only)", "value": s["matrices_per_s"],
                "p99_ms": s["p99_ms"], "update_p99_ms": s["update_p99_ms"], "rate_trials": [t[2] for t in trials]}
    # ---- kernel figures: fixed served-shaped batches on the start-state map, every dispatch timed
    cells0, qmix_0, resmix_0 = plain_maps["start"]
    gm0 = _lib.GpuMap(cells0, B, G)
    gm0.build()
    kst, kphase_rate = kernel_phase(gm0, serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000),
                                    KERNEL_BATCH, args.kernel_batches)
    roof, valu, kernels = roofline_of(kst)
    log(f"kernel phase: {args.kernel_batches} batches of {KERNEL_BATCH} queries, k_q_z "
        f"{1e3 * roof['avg_launch_ms']:.1f} us per launch, {roof['achieved']:.0f} GB/s")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(gm0, cells0, qmix_0, args.cpu_seconds, resmix_0)

    value = s["matrices_per_s"]
    return {
        "metric": "k x k query matrices/s (served, query p99 < 1 ms, update p99 < "
                  f"{args.update_lag_ms:g} ms) + p99 latency, 1024^2 grid",
        "value": value, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": s["span_s"] * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": "1024x1024 street-like grid seed 3, b=32, g=2x2; per session 16 robots exploring "
                               "from radius-64 start disks to the fully revealed map; Poisson stream of "
                               "queryAllPairs k~logU[8,256] + 10% update batches of 16..512 cells; sessions "
                               "resident in HBM, as many as the timed stream needs (>= 1 s, >= 50K requests)",
                   "mode": "served", "consistency": args.consistency, "window": int(s["queries"] + s["updates"]) //
                   args.steps, "slots": args.slots, "max_batch_queries": args.batch_queries,
                   "update_cu_percent": args.update_cu, "update_tick_us": args.update_tick_us, "update_frac": uf,
                   "update_lag_ms": args.update_lag_ms, "update_streams": args.update_streams or 1},
        "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"], "max_ms": s["max_ms"],
        "update_p50_ms": s["update_p50_ms"], "update_p99_ms": s["update_p99_ms"],
        "arrival_rate_per_s": s["rate"], "timed_arrival_span_s": s["timed_arrival_span_s"],
        "timed_requests": int(s["queries"] + s["updates"]), "sessions": s["sessions"],
        "plain_matrices_per_s_kmix": plain_mix, "plain_matrices_per_s_k64": plain_64,
        "served_over_plain": value / plain_mix, "served_over_plain_k64": value / plain_64,
        "served_over_plain_note": "plain = the per-call path on the same k mix, harmonic mean over the start, half "
                                  "and fully revealed states; _k64 compares with plain k = 64 queries",
        "by_state": by_state, "served_checks": n_checked,
        "server_stats": server_stats,
        "rate_trials": [t[2] for t in trials],
        "roofline": roof, "valu": valu, "cpu_baseline": cpu, "kernels": kernels,
        "kernel_phase": {"batch_queries": KERNEL_BATCH, "batches": args.kernel_batches,
                         "matrices_per_s_one_stream": kphase_rate},
    }


def run_kernel(args):
    """The kernel phase of the served mode alone: served-shaped batches on the
    serving map at its start state, every dispatch timed (the command whose
    rocprofv3 summary backs the served line's roofline)."""
    from htrip_amd import _lib, grids, serving

    truth = grids.street_grid(N, SEED)
    rv = serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000)
    gmap = _lib.GpuMap(rv.observed, B, G)
    gmap.build()
    kst, rate = kernel_phase(gmap, rv, args.batch_size, args.kernel_batches, events=not args.no_events)
    roof, valu, kernels = roofline_of(kst)
    if args.calibrate_copy_mib > 0:
        # one device-to-device copy of known size by the runtime's copy kernel,
        # the reference for scaling hardware byte counters in a profiled run
        hip = C.CDLL("libamdhip64.so")
        nbytes = args.calibrate_copy_mib << 20
        bufs = [C.c_void_p(), C.c_void_p()]
        for bp in bufs:
            if hip.hipMalloc(C.byref(bp), C.c_size_t(nbytes)) != 0:
                raise SystemExit("hipMalloc for the calibration copy failed")
        if hip.hipMemcpy(bufs[1], bufs[0], C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise SystemExit("calibration copy failed")
        for bp in bufs:
            hip.hipFree(bp)
        log(f"calibration copy: {args.calibrate_copy_mib} MiB device to device")
    log(f"kernel phase: {args.kernel_batches} batches of {args.batch_size} queries: {rate:.0f} matrices/s, "
        f"k_q_z {1e3 * roof['avg_launch_ms']:.1f} us per launch, {roof['achieved']:.0f} GB/s")
    return {
        "metric": "k x k query matrices/s (served-shaped batches, one stream, every dispatch timed), 1024^2 grid",
        "value": rate, "unit": "matrices/s", "n_gpus": 1, "steps": args.kernel_batches, "warmup": 10,
        "ms_per_step": 1e3 * args.batch_size / rate, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"1024x1024 street-like grid seed 3, b=32, g=2x2, batches of {args.batch_size} "
                               "queryAllPairs k~logU[8,256] at the robots' start state", "mode": "kernel"},
        "roofline": roof, "valu": valu, "cpu_baseline": None, "kernels": kernels,
    }


PIVOT_SETS = {"cover": 0, "all": 1}


def run_update(args):
    """Incremental reveal latency on the serving map: robot batches of 16..512 cells."""
    from htrip_amd import _lib, grids, serving

    truth = grids.street_grid(N, SEED)

    def stream(timed):
        """The same reveal sequence on a fresh map: latencies without timing
        events, or kernel figures with events around every dispatch."""
        rv = serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000)
        gmap = _lib.GpuMap(rv.observed, B, G)
        gmap.set_tuning(update_pivots=PIVOT_SETS[args.update_pivots])
        gmap.build()
        rng = grids.SplitMix64(77)
        lat, sizes, stats = [], [], []
        for i in range(args.warmup + args.steps):
            parts = []
            for _ in range(args.update_requests):  # the server coalesces queued requests like this
                m = serving.loguniform_k(rng, 16, 512)
                parts.append(rv.next_batch(rng.below(ROBOTS), m))
            cells = np.vstack(parts)
            if timed and i == args.warmup:
                gmap.profile(True, every=1)
            t0 = time.perf_counter()
            gmap.open_cells(cells)
            dt = time.perf_counter() - t0
            if i >= args.warmup:
                lat.append(dt)
                sizes.append(len(cells))
                stats.append(gmap.last_batch_stats())
        kernels = kernel_roofline(gmap)[2] if timed else None
        gmap.close()
        return np.array(lat), sizes, stats, kernels

    lat, sizes, stats, _ = stream(False)
    _, _, _, kernels = stream(True)
    log(f"update: p50 {1e3 * np.median(lat):.3f} ms, p99 {1e3 * np.percentile(lat, 99):.3f} ms, "
        f"mean cells {np.mean(sizes):.0f}")
    for k, v in kernels.items():
        log(f"  {k:18s} launches {v['launches']:6d} total {v['total_ms']:9.2f} ms")
    per = f"{args.update_requests} robot batches" if args.update_requests > 1 else "a robot batch"
    return {"metric": f"onBatchCellsOpened latency (1024^2 street, b=32, {per} of 16..512 cells)",
            "value": float(np.median(lat) * 1e3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": float(lat.mean() * 1e3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
            "config": {"workload": "1024x1024 street seed 3, b=32, 16 robots, reveal batches logU[16,512]",
                       "mode": "update", "update_pivots": args.update_pivots,
                       "requests_per_update": args.update_requests},
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_cells": float(np.mean(sizes)),
            "mean_stats": [float(x) for x in np.mean(np.array(stats), axis=0)], "kernels": kernels}


def ref_oracle(cells, b, g):
    """The compiled reference (oracle/_ref) built on the CPU, one thread: (oracle, seconds)."""
    from oracle import oracle as orc

    if not orc.available("reference"):
        return None, None
    R = orc.Oracle("reference", cells, b, g)
    t0 = time.perf_counter()
    assert R.build()
    return R, time.perf_counter() - t0


def ref_query_rate(R, queries, expect, budget_s):
    """Reference queryAllPairs, one thread, over a bounded sample; exactness gate against expect."""
    api = R.api
    done, spent = 0, 0.0
    for q, vps in enumerate(queries):
        vps = np.ascontiguousarray(vps, dtype=np.int32)
        ks = np.array([len(vps)], dtype=np.int32)
        out = np.empty(len(vps) * len(vps), dtype=np.uint16)
        spent += api["time_queries"](R.h, vps, ks, 1, 1, out.ctypes.data)
        if not np.array_equal(out.reshape(len(vps), -1), expect[q]):
            raise SystemExit("reference CPU result differs from the GPU result")
        done += 1
        if spent >= budget_s:
            break
    return done, spent


def build_times(gmap, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        gmap.build()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def run_small(args):
    """256^2 street-like grid (seed 1), b=16: build, then queryAllPairs at k=64."""
    from htrip_amd import _lib, grids

    cells = grids.street_grid(256, 1)
    gmap = _lib.GpuMap(cells, 16, 2)
    gmap.build()
    build_ms = 1e3 * build_times(gmap, 5)
    rate, dt, qs, res = plain_rate(gmap, cells, 64, args.steps, args.warmup)
    log(f"256^2 street b=16: build {build_ms:.2f} ms, query k=64 {rate:.0f} matrices/s")
    cpu = None
    if not args.no_cpu:
        R, tb = ref_oracle(cells, 16, 2)
        if R is not None:
            done, spent = ref_query_rate(R, qs, res, args.cpu_seconds)
            cpu = {"value": done / spent, "unit": "matrices/s", "cores": 1, "kind": "reference",
                   "sample": f"{done} queryAllPairs k=64 (1 thread)", "build_ms": 1e3 * tb,
                   "all_cores": cpu_all_cores(R, qs[:max(done, 1) * 4], res)}
    return {"metric": "k x k query matrices/s (plain GPU path, k=64, 256^2 street grid, b=16)", "value": rate,
            "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u16", "data": "synthetic",
            "config": {"workload": "256x256 street-like grid seed 1, b=16, g=2x2; build + queryAllPairs k=64",
                       "mode": "small"},
            "build_ms": build_ms, "cpu_baseline": cpu}


def run_lift(args):
    """liftPoint on the 1024^2 street-like grid (b=32): one call at a time, each
    synchronous and including the copy of its H+1 vectors to the host, as the
    reference's API returns them.  The reference's liftPoint runs beside it over
    the same tables, and must return the same vectors."""
    from htrip_amd import _lib, grids

    cells = grids.street_grid(N, SEED)
    gmap = _lib.GpuMap(cells, B, G)
    gmap.build()
    H = gmap.H
    cap = 4 * B - 4 + sum(gmap.node_dims(l)[2] for l in range(1, H + 1))
    x = np.empty(cap, dtype=np.uint16)
    meta = [np.empty(H + 1, dtype=np.int32) for _ in range(4)]
    P = C.POINTER(C.c_int32)
    xp = x.ctypes.data_as(C.POINTER(C.c_uint16))
    mp = [m.ctypes.data_as(P) for m in meta]
    pts = grids.sample_free(cells, 20 * args.steps + args.warmup, 77).astype(np.int32)  # 800 timed by default
    lat = []
    for i, (r, c) in enumerate(pts):
        t0 = time.perf_counter()
        rc = gmap.lib.htg_lift_point(gmap.h, int(r), int(c), xp, cap, *mp)
        t1 = time.perf_counter()
        if rc != 0:
            raise SystemExit(f"htg_lift_point failed: {rc}")
        if i >= args.warmup:
            lat.append(t1 - t0)
    lat = np.array(lat)
    rate = len(lat) / lat.sum()
    log(f"lift: {len(lat)} calls, {rate:.0f} lifts/s, p50 {1e6 * np.median(lat):.1f} us, "
        f"p99 {1e6 * np.percentile(lat, 99):.1f} us")
    cpu = None
    if not args.no_cpu:
        R = import_reference(gmap, cells)
        if R is not None:
            for r, c in pts[:16]:  # exactness gate
                Xg = gmap.lift_point(int(r), int(c))[0]
                Xr = R.lift_point(int(r), int(c))[0]
                if any(not np.array_equal(a, b) for a, b in zip(Xg, Xr)):
                    raise SystemExit("reference liftPoint differs from the GPU")
            log("cpu baseline: 16 reference lifts equal the GPU lifts")
            done, spent = 0, 0.0
            while spent < args.cpu_seconds and done < len(pts):
                chunk = pts[done:done + 64]
                spent += R.time_lifts(chunk)
                done += len(chunk)
            cpu = {"value": done / spent, "unit": "lifts/s", "cores": 1, "kind": "reference",
                   "sample": f"{done} liftPoint calls (1 thread, reference sources) on the same map"}
    return {"metric": "liftPoint calls/s (one synchronous call at a time, 1024^2 grid)", "value": rate,
            "unit": "lifts/s", "n_gpus": 1, "steps": len(lat), "warmup": args.warmup,
            "ms_per_step": 1e3 / rate, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u16", "data": "synthetic",
            "config": {"workload": "1024x1024 street-like grid seed 3, b=32, g=2x2, liftPoint of free cells",
                       "mode": "lift"},
            "p50_us": 1e6 * float(np.median(lat)), "p99_us": 1e6 * float(np.percentile(lat, 99)),
            "cpu_baseline": cpu}


def run_incremental(args):
    """queryAllPairsIncremental after each single-cell opening: 256^2
    street-like grid (seed 1), b=16, k=64 fixed viewpoints, robots revealing
    the frontier one cell at a time.  GPU: onCellOpened, then the refresh (one
    batched query of all k^2 pairs).  Reference: the same openings, then its
    version-stamped incremental refresh.  The GPU refresh is
    htg_query_all_pairs_incremental: the cached matrix when no viewpoint's
    leaf or ancestor table was rewritten, one batched query otherwise.  Every
    refreshed matrix is compared with the reference's."""
    from htrip_amd import _lib, grids
    from oracle import oracle as orc

    truth = grids.street_grid(256, 1)
    obs = grids.disk_observed(truth, [tuple(grids.sample_free(truth, 1, 5)[0])], 60)
    order = grids.frontier_order(truth, obs)
    vps = grids.sample_free(truth, 64, 6).astype(np.int32)
    gmap = _lib.GpuMap(obs, 16, 2)
    gmap.build()
    steps = min(10 * args.steps, len(order) - args.warmup)  # 400 refreshes by default
    R = None
    if not args.no_cpu and orc.available("reference"):
        R = orc.Oracle("reference", obs, 16, 2)
        assert R.build()
        inc = R.api["inc_create"]()
    t_gpu, t_ref = [], []
    recomputed = 0
    out_r = np.empty(64 * 64, dtype=np.uint16)
    st = _lib.IncState()
    m = np.empty((64, 64), dtype=np.uint16)  # caller-owned output, as the reference's out_matrix
    nr = C.c_int32()
    P32 = C.POINTER(C.c_int32)
    vp_p, m_p = vps.ctypes.data_as(P32), m.ctypes.data_as(C.POINTER(C.c_uint16))
    refresh = gmap.lib.htg_query_all_pairs_incremental
    for i in range(args.warmup + steps):
        r, c = (int(v) for v in order[i])
        gmap.open_cell(r, c)
        gmap.sync()
        t0 = time.perf_counter()
        rc = refresh(gmap.h, st.h, vp_p, 64, m_p, C.byref(nr))
        t1 = time.perf_counter()
        if rc != 0:
            raise SystemExit(f"htg_query_all_pairs_incremental failed: {rc}")
        if i >= args.warmup:
            recomputed += nr.value > 0
        if R is not None:
            R.open_cell(r, c)
            dt = R.api["inc_query"](R.h, inc, vps, 64, out_r)
            if not np.array_equal(out_r.reshape(64, 64), m):
                raise SystemExit("reference incremental refresh differs from the GPU")
        if i >= args.warmup:
            t_gpu.append(t1 - t0)
            if R is not None:
                t_ref.append(dt)
    if R is not None:
        R.api["inc_destroy"](inc)
    t_gpu = np.array(t_gpu)
    rate = len(t_gpu) / t_gpu.sum()
    st.close()
    log(f"incremental: {len(t_gpu)} refreshes ({recomputed} recomputed), GPU p50 {1e6 * np.median(t_gpu):.1f} us"
        + (f", reference p50 {1e6 * np.median(t_ref):.0f} us" if t_ref else ""))
    cpu = None
    if t_ref:
        cpu = {"value": len(t_ref) / float(np.sum(t_ref)), "unit": "refreshes/s", "cores": 1, "kind": "reference",
               "sample": f"{len(t_ref)} queryAllPairsIncremental refreshes (1 thread), every result equal to the GPU's",
               "p50_us": 1e6 * float(np.median(t_ref)), "p99_us": 1e6 * float(np.percentile(t_ref, 99))}
    return {"metric": "incremental k=64 refreshes/s after single-cell openings (256^2 street grid, b=16)",
            "value": rate, "unit": "refreshes/s", "n_gpus": 1, "steps": len(t_gpu), "warmup": args.warmup,
            "ms_per_step": 1e3 / rate, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u16", "data": "synthetic",
            "config": {"workload": "256x256 street-like grid seed 1, b=16, g=2x2, k=64, one cell opened per refresh",
                       "mode": "incremental"},
            "p50_us": 1e6 * float(np.median(t_gpu)), "p99_us": 1e6 * float(np.percentile(t_gpu, 99)),
            "recomputed": recomputed, "cpu_baseline": cpu}


def run_liftroot(args):
    """liftRoot: a 256^2 corner of the 512^2 street-like grid (seed 3), fully
    revealed, b=16, grows to the whole 512^2 grid (S 256 -> 512, H 4 -> 5).
    GPU: htg_lift_root (new map, kept tables copied, 3/4 of the leaves and
    nodes and the new root built); reference: its liftRoot on its own build of
    the same corner.  Every lifted map must answer queries as BFS does."""
    from htrip_amd import _lib, grids
    from oracle import oracle as orc

    big = grids.street_grid(512, SEED)
    small = big[:256, :256].copy()
    vps = grids.sample_free(big, 32, 9)
    expect = orc.grid_bfs(big, vps)
    ts = []
    for rep in range(args.warmup + 5):
        gmap = _lib.GpuMap(small, 16, 2)
        gmap.build()
        gmap.sync()
        t0 = time.perf_counter()
        gmap.lift_root(big)
        t1 = time.perf_counter()
        if not np.array_equal(gmap.query(vps), expect):
            raise SystemExit("lifted GPU map differs from BFS")
        if rep >= args.warmup:
            ts.append(t1 - t0)
        gmap.close()
    tc = []  # a new map of the grown grid, created and built (the lift allocates its map too)
    for _ in range(5):
        t0 = time.perf_counter()
        fresh = _lib.GpuMap(big, 16, 2)
        fresh.build()
        tc.append(time.perf_counter() - t0)
        fresh.close()
    fresh_ms = 1e3 * float(np.median(tc))
    gpu_ms = 1e3 * float(np.median(ts))
    log(f"liftroot: GPU {gpu_ms:.2f} ms (creating and building the grown map instead: {fresh_ms:.2f} ms)")
    cpu = None
    if not args.no_cpu and orc.available("reference"):
        R = orc.Oracle("reference", small, 16, 2)
        assert R.build()
        t0 = time.perf_counter()
        R.lift_root(big)
        t_ref = time.perf_counter() - t0
        if not np.array_equal(R.query(vps), expect):
            raise SystemExit("reference lifted map differs from BFS")
        log(f"liftroot: reference {1e3 * t_ref:.1f} ms")
        cpu = {"value": 1e3 * t_ref, "unit": "ms", "cores": 1, "kind": "reference",
               "sample": "one liftRoot 256^2 -> 512^2 (1 thread), queries equal BFS"}
    return {"metric": "liftRoot latency, 256^2 -> 512^2 street grid, b=16", "value": gpu_ms, "unit": "ms",
            "n_gpus": 1, "steps": len(ts), "warmup": args.warmup, "ms_per_step": gpu_ms, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
            "config": {"workload": "512x512 street-like grid seed 3 (its 256x256 corner built first), b=16, g=2x2",
                       "mode": "liftroot"},
            "fresh_create_build_ms": fresh_ms, "cpu_baseline": cpu}


def run_build(args):
    """512^2 DFS maze (seed 2, corridor and wall width 1), b=16: the batched
    build (every node of a level per launch, the whole build as one captured
    graph, autotuned), then queryAllPairs at k=128."""
    from htrip_amd import _lib, grids

    cells = grids.maze_grid(512, 2, 1)
    gmap = _lib.GpuMap(cells, 16, 2)
    gmap.build()
    default_ms = 1e3 * build_times(gmap, 5)
    sample = [grids.sample_free(cells, 128, 900 + i) for i in range(4)]
    t = gmap.autotune(sample, reps=3)
    tuned_ms = 1e3 * build_times(gmap, 5)
    log(f"512^2 maze b=16: build {default_ms:.2f} ms (default), {tuned_ms:.2f} ms autotuned "
        f"(pivot block {t.pivot_block}, graph {t.use_graph}, projection tile {t.zq_tile})")
    rate, dt, qs, res = plain_rate(gmap, cells, 128, args.steps, args.warmup)
    log(f"query k=128: {rate:.0f} matrices/s")
    gmap.profile(True)
    gmap.build()
    _, _, kernels = kernel_roofline(gmap)
    cpu = None
    if not args.no_cpu:
        R, tb = ref_oracle(cells, 16, 2)
        if R is not None:
            done, spent = ref_query_rate(R, qs, res, args.cpu_seconds)
            cpu = {"value": 1.0 / tb, "unit": "builds/s", "cores": 1, "kind": "reference",
                   "sample": "one full build (1 thread, reference AVX2 sources)", "build_ms": 1e3 * tb,
                   "query_k128_matrices_per_s": done / spent,
                   "query_all_cores": cpu_all_cores(R, qs[:max(done, 1) * 4], res)}
            log(f"reference build {1e3 * tb:.0f} ms, query k=128 {done / spent:.1f} matrices/s (1 thread)")
    return {"metric": "full build time, 512^2 maze b=16 (batched, graph-captured, autotuned)",
            "value": tuned_ms, "unit": "ms", "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": tuned_ms,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
            "config": {"workload": "512x512 DFS maze seed 2 width 1, b=16, g=2x2; build + queryAllPairs k=128",
                       "mode": "build"},
            "default_build_ms": default_ms,
            "tuning": {"pivot_block": t.pivot_block, "use_graph": t.use_graph, "zq_tile": t.zq_tile},
            "query_k128_matrices_per_s": rate, "build_kernels": kernels, "cpu_baseline": cpu}


def run_reveal(args):
    """512^2 street-like grid (seed 3), b=8: observed area starts as a disk of
    radius 64; 500 steps of 256 frontier-order reveals, each followed by
    queryAllPairs at k=32.  The reference replays the same steps for a bounded
    sample and every compared matrix must be equal."""
    from htrip_amd import _lib, grids

    truth = grids.street_grid(512, 3)
    start = grids.sample_free(truth, 1, 31)[0]
    obs = grids.disk_observed(truth, [tuple(start)], 64)
    order = grids.frontier_order(truth, obs)
    steps = min(500, len(order) // 256)
    gmap = _lib.GpuMap(obs, 8, 2)
    gmap.set_tuning(update_pivots=PIVOT_SETS[args.update_pivots])
    gmap.build()
    state = obs.copy()
    batches, views, lat = [], [], []
    t_all = time.perf_counter()
    for st in range(steps):
        cells = order[st * 256:(st + 1) * 256]
        state[cells[:, 0], cells[:, 1]] = 1
        vps = grids.sample_free(state, 32, 5000 + st)
        t0 = time.perf_counter()
        gmap.open_cells(cells)
        m = gmap.query(vps)
        lat.append(time.perf_counter() - t0)
        batches.append(cells)
        views.append((vps, m))
    total = time.perf_counter() - t_all
    lat = np.array(lat) * 1e3
    log(f"512^2 street b=8 reveal: {steps} steps, step p50 {np.percentile(lat, 50):.3f} ms, "
        f"p99 {np.percentile(lat, 99):.3f} ms")
    # kernel figures: the same steps on a second map with every dispatch timed
    # (the timed steps above carry no events)
    gk = _lib.GpuMap(obs, 8, 2)
    gk.set_tuning(update_pivots=PIVOT_SETS[args.update_pivots])
    gk.build()
    gk.profile(True, every=1)
    for st in range(steps):
        gk.open_cells(batches[st])
        gk.query(views[st][0])
    _, _, kernels = kernel_roofline(gk)
    gk.close()
    cpu = None
    checked = 0
    if not args.no_cpu:
        from oracle import oracle as orc

        # every step's matrix against BFS on that step's grid
        state = obs.copy()
        for st in range(steps):
            c = batches[st]
            state[c[:, 0], c[:, 1]] = 1
            if not np.array_equal(orc.grid_bfs(state, views[st][0], orc.bfs_threads()), views[st][1]):
                raise SystemExit(f"reveal step {st}: GPU matrix differs from BFS")
            checked += 1
        log(f"exactness: all {checked} step matrices equal BFS")
        if orc.available("reference"):
            R = orc.Oracle("reference", obs, 8, 2)
            assert R.build()
            spent, n = 0.0, 0
            for st in range(steps):
                t0 = time.perf_counter()
                R.open_cells(batches[st])
                got = R.query(views[st][0])
                spent += time.perf_counter() - t0
                if not np.array_equal(got, views[st][1]):
                    raise SystemExit(f"reveal step {st}: GPU matrix differs from the reference")
                n += 1
                if spent >= args.cpu_seconds:
                    break
            cpu = {"value": n / spent, "unit": "steps/s", "cores": 1, "kind": "reference",
                   "sample": f"first {n} steps (onBatchCellsOpened 256 cells + queryAllPairs k=32), all equal"}
            log(f"reference: {n} steps in {spent:.2f}s, every matrix equal")
    return {"metric": "reveal steps/s (onBatchCellsOpened 256 cells + queryAllPairs k=32), 512^2 street b=8",
            "value": steps / total, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": total * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u16", "data": "synthetic",
            "config": {"workload": "512x512 street-like grid seed 3, b=8, g=2x2, disk radius 64, 256 frontier "
                                   "reveals per step + queryAllPairs k=32", "mode": "reveal",
                       "update_pivots": args.update_pivots},
            "step_p50_ms": float(np.percentile(lat, 50)), "step_p99_ms": float(np.percentile(lat, 99)),
            "steps_checked_against_bfs": checked, "kernels": kernels, "cpu_baseline": cpu}


def run_multi(args):
    """Eight resident sessions admitted by HBM budget: 2x 256^2 street (b=16),
    2x 512^2 DFS maze width 1 and 2x width 16 (b=16), 2x 1024^2 street (b=8,
    H=7).  Mixed traffic: queryAllPairs k log-uniform in [8, 512] and reveal
    batches of 16..1024 cells (10%), 4 robots per map, Poisson arrivals."""
    from htrip_amd import _lib, grids, serving

    maps = [("street", 256, 16, 101), ("street", 256, 16, 102), ("maze1", 512, 16, 103), ("maze1", 512, 16, 104),
            ("maze16", 512, 16, 105), ("maze16", 512, 16, 106), ("street", 1024, 8, 107), ("street", 1024, 8, 108)]
    truths = []
    for kind, n, b, seed in maps:
        if kind == "street":
            truths.append(grids.street_grid(n, seed))
        else:
            truths.append(grids.maze_grid(n, seed, 1 if kind == "maze1" else 16))
    need = sum(_lib.footprint(b, 2, n, n) for _, n, b, _ in maps) * 2
    srv = serving.Server(hbm_budget=int(need * 1.05), n_slots=args.slots, max_batch_queries=args.batch_queries,
                         max_batch_viewpoints=args.batch_viewpoints, consistency=serving.SNAPSHOT,
                         update_streams=args.update_streams)
    log(f"8 sessions, HBM budget {need * 1.05 / 2**30:.1f} GiB (two copies each)")

    def fresh():
        sess = []
        t0 = time.time()
        for (kind, n, b, seed), truth in zip(maps, truths):
            rv = serving.RobotReveal(truth, 4, n // 16, seed=seed, limit=20000)
            sid = srv.add_map(rv.observed, b, 2)
            sess.append((sid, rv, rv.observed.copy()))
        log(f"sessions admitted and built in {time.time() - t0:.2f}s")
        return sess

    def traffic(sess, n_each, rate, seed):
        trs = [serving.make_trace(sid, rv, n_each, rate / len(sess), seed + i, update_frac=0.1, k_lo=8, k_hi=512,
                                  upd_lo=16, upd_hi=1024) for i, (sid, rv, _) in enumerate(sess)]
        return serving.merge_traces(trs)

    sess = fresh()
    burst = traffic(sess, 64, 1e8, 700)
    tb = srv.replay(burst)
    cap = serving.latency_summary(burst, tb)["matrices_per_s"] / 0.9
    for sid, _, _ in sess:
        srv.remove_map(sid)
    log(f"burst capacity ~{cap:.0f} requests/s")
    trials = []
    chosen = None
    for frac in args.rate_fracs:
        sess = fresh()
        tr = traffic(sess, args.steps * 8, cap * frac, 800 + len(trials))
        td = srv.replay(tr)
        s = serving.latency_summary(tr, td)
        s["rate"] = cap * frac
        checks = multi_checks(tr, sess) if not args.no_cpu else 0
        for sid, _, _ in sess:
            srv.remove_map(sid)
        trials.append(s)
        log(f"rate {s['rate']:.0f}/s: {s['matrices_per_s']:.0f} matrices/s, p50 {s['p50_ms']:.3f} ms, "
            f"p99 {s['p99_ms']:.3f} ms, updates p50 {s['update_p50_ms']} ms; {checks} spot checks equal BFS")
        if s["p99_ms"] < 1.0:  # the highest tested rate with query p99 < 1 ms, as in the served mode
            chosen = s
            break
    best = chosen if chosen is not None else trials[-1]
    return {"metric": "k x k query matrices/s over 8 resident sessions (mixed k <= 512 and reveals, p99 < 1 ms)",
            "value": best["matrices_per_s"], "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 0, "ms_per_step": best["span_s"] * 1e3 / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
            "config": {"workload": "8 sessions: 2x256^2 street b16, 2x512^2 maze w1 b16, 2x512^2 maze w16 b16, "
                                   "2x1024^2 street b8; k~logU[8,512], 10% reveals of 16..1024 cells",
                       "mode": "multi"},
            "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "rate_trials": trials}


def multi_checks(tr, sess, per_session=1, kmax=64):
    """Spot checks: small queries of each session against BFS on the state with
    exactly the update prefix the result reports."""
    from oracle import oracle as orc

    n = 0
    for sid, _, state0 in sess:
        mine = [i for i in range(len(tr.t)) if tr.session[i] == sid]
        upd = [i for i in mine if tr.type[i] == 1]
        got = 0
        for i in mine:
            if tr.type[i] != 0 or tr.count[i] > kmax:
                continue
            st = state0.copy()
            for j in upd[:int(tr.version[i])]:
                c = tr.cells_of(j)
                st[c[:, 0], c[:, 1]] = 1
            if not np.array_equal(orc.grid_bfs(st, tr.cells_of(i)), tr.query_result(i)):
                raise SystemExit(f"session {sid}: served result differs from BFS on its map state")
            n += 1
            got += 1
            if got == per_session:
                break
    return n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (served: windows of requests per rate trial, default 200; others 40)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", default="served",
                    choices=["plain", "served", "update", "small", "build", "reveal", "multi", "kernel", "lift",
                             "incremental", "liftroot"])
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--map", default="", help="plain mode: a shipped MovingAI map (tests/golden/maps) instead of "
                                              "the synthetic grid, e.g. Boston_0_1024")
    ap.add_argument("--window", type=int, default=64, help="requests per served step")
    ap.add_argument("--slots", type=int, default=4, help="query streams (the update stream makes one more)")
    ap.add_argument("--update-cu", type=int, default=0, help="percent of CUs for update kernels (0 = all)")
    ap.add_argument("--update-frac", type=float, default=0.1, help="share of update requests in the stream")
    ap.add_argument("--update-tick-us", type=int, default=0, help="snapshot mode: minimum spacing of update batches")
    ap.add_argument("--update-streams", type=int, default=0,
                    help="update streams (and snapshot update threads) the sessions are dealt to; 0 = the mode's "
                         "default: 1 for served (one update at a time merges more reveals per closure and keeps "
                         "heavy update kernels of different sessions from coinciding), 2 for multi")
    ap.add_argument("--update-pivots", default="cover", choices=sorted(PIVOT_SETS),
                    help="update and reveal modes: close over a vertex cover of the decreased entries, or over "
                         "every decreased row (the reference's pivot set)")
    ap.add_argument("--update-lag-ms", type=float, default=25.0,
                    help="served: a rate passes only if update p99 latency stays below this (updates keep up)")
    ap.add_argument("--update-requests", type=int, default=1,
                    help="update mode: robot requests coalesced into each onBatchCellsOpened call")
    ap.add_argument("--kernel-batches", type=int, default=400,
                    help="served and kernel modes: dispatch-timed served-shaped batches for the kernel figures")
    ap.add_argument("--batch-size", type=int, default=KERNEL_BATCH, help="kernel mode: queries per batch")
    ap.add_argument("--no-events", action="store_true", help="kernel mode: no timing events (for an outside profiler)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="kernel mode: end with one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--dump-trials", default="", help="served: save each trial's arrival/completion times here")
    ap.add_argument("--batch-queries", type=int, default=64)
    ap.add_argument("--batch-viewpoints", type=int, default=8192)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--consistency", default="snapshot", choices=["snapshot", "strict"])
    ap.add_argument("--rate-fracs", type=float, nargs="+",
                    default=[2.2, 2.0, 1.82, 1.66, 1.52, 1.4, 1.29, 1.19, 1.1, 1.02, 0.94, 0.86, 0.78, 0.72, 0.66,
                             0.61, 0.56, 0.52, 0.48, 0.44, 0.41, 0.38, 0.35, 0.32, 0.29, 0.26, 0.23])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="served: headline trials only (no per-state table, no "
                                                         "kernel phase, no CPU leg)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 200 if args.mode == "served" else 40
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU (one MI355X per box); --gpus must be 1"}))
        sys.exit(2)
    line = {"plain": run_plain, "served": run_served, "update": run_update, "small": run_small, "build": run_build,
            "reveal": run_reveal, "multi": run_multi, "kernel": run_kernel, "lift": run_lift,
            "incremental": run_incremental, "liftroot": run_liftroot}[args.mode](args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
