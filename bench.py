#!/usr/bin/env python3
"""Benchmark of the MI355X min-plus distance oracle (one JSON line on stdout).

Workload (serving configuration): a 1024 x 1024 street-like seeded grid
(seed 3), leaf b = 32, group 2 x 2 (H = 5).  16 simulated robots start with
observed disks of radius 64 and reveal free cells in their frontier order.
The hierarchy is built once and stays resident in HBM.

Modes
  served (default) -- the continuous-batching server replays a Poisson request
      stream: queryAllPairs with k log-uniform in [8, 256] interleaved with
      update batches of 16..512 revealed cells (10% of requests).  One step =
      one window of --window requests; `value` = query matrices/s over the
      timed steps at the highest tested arrival rate whose p99 query latency
      stays under 1 ms.  The plain per-call GPU path (k = 64) is measured on
      the same map state for the served/plain ratio.
  plain -- one step = one plain per-call queryAllPairs (k = 64) through the
      drop-in C ABI (host viewpoints in, host matrix out, synchronous).

Timing: request latency = completion - arrival on the server's clock
(steady_clock); kernel durations from hipEvents recorded on the stream each
kernel runs on.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# 256 CUs x 64 lanes x 2 u16 per packed op x 2.4 GHz = 78.6 T u16 ops/s; one
# min-plus update is a packed add + a packed min: 39.3 T updates/s.
PEAK_UPDATES = 256 * 64 * 2 * 2.4e9 / 2

N, B, G, SEED, ROBOTS, RADIUS = 1024, 32, 2, 3, 16, 64


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_stats_of(gmap):
    from htrip_amd import _lib

    return {c: gmap.kernel_stats(c) for c in range(6)}


def roofline_of(stats):
    """Roofline of the query projection (k_q_z) from accumulated kernel stats:
    bytes and min-plus updates are counted on the device, time from start/stop
    events attached to each dispatch."""
    from htrip_amd import _lib

    z = stats[_lib.K_Q_PROJECT]
    n = max(z["launches"], 1)
    avg_ms = z["total_ms"] / n
    gbs = (z["bytes"] / n) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    ups = (z["updates"] / n) / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "traffic": None, "kernel": "k_q_z", "avg_launch_ms": avg_ms, "launches": z["launches"]}
    valu = {"kernel": "k_q_z", "achieved_updates_per_s": ups, "peak_updates_per_s": PEAK_UPDATES,
            "frac": ups / PEAK_UPDATES}
    kernels = {_lib.KERNEL_NAMES[c]: stats[c] for c in range(6)}
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
    log(f"cpu baseline: {done} queries in {spent:.2f}s, all equal to the GPU results")
    return {"value": done / spent, "unit": "matrices/s", "cores": 1, "kind": "reference",
            "sample": f"{done} queryAllPairs k={len(queries[0])} (1 thread, reference AVX2 sources) on the same map"}


def served_checks(tr, state0, base, n=4, kmax=96):
    """(state, viewpoints, result) for the first n small queries of a trace,
    with the state rebuilt from the update prefix each result reports (the
    session had applied `base` updates when the trace started)."""
    if state0 is None:
        return []
    upd = [i for i in range(len(tr.t)) if tr.type[i] == 1]
    out = []
    for i in range(len(tr.t)):
        if tr.type[i] != 0 or tr.count[i] > kmax:
            continue
        st = state0.copy()
        for j in upd[:int(tr.version[i]) - base]:
            c = tr.cells_of(j)
            st[c[:, 0], c[:, 1]] = 1
        out.append((st, tr.cells_of(i), tr.query_result(i)))
        if len(out) == n:
            break
    return out


def run_plain(args):
    from htrip_amd import _lib, grids

    cells = grids.street_grid(N, SEED)
    gmap = _lib.GpuMap(cells, B, G)
    gmap.profile(True)
    t0 = time.time()
    gmap.build()
    log(f"build {1e3 * (time.time() - t0):.1f} ms (S={gmap.S} H={gmap.H})")
    gmap.profile(True)
    rate, dt, qs, res = plain_rate(gmap, cells, args.k, args.steps, args.warmup)
    log(f"plain: {args.steps} queries k={args.k}: {rate:.0f} matrices/s")
    roof, valu, kernels = kernel_roofline(gmap)
    cpu = None if args.no_cpu else cpu_baseline(gmap, cells, qs, args.cpu_seconds, res)
    return {
        "metric": "k x k query matrices/s (plain per-call GPU path, k=64, 1024^2 grid)",
        "value": rate, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u16", "data": "synthetic",
        "config": {"workload": "1024x1024 street-like grid seed 3, b=32, g=2x2, plain queryAllPairs k=64",
                   "mode": "plain", "k": args.k},
        "roofline": roof, "valu": valu, "cpu_baseline": cpu, "kernels": kernels,
    }


def run_served(args):
    from htrip_amd import _lib, grids, serving

    t0 = time.time()
    truth = grids.street_grid(N, SEED)
    init = serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000).observed.copy()
    log(f"grid {N}x{N}: {int((truth == 1).sum())} free, {int((init == 1).sum())} observed at start "
        f"({time.time() - t0:.1f}s)")
    mode = serving.SNAPSHOT if args.consistency == "snapshot" else serving.STRICT
    srv = serving.Server(n_slots=args.slots, max_batch_queries=args.batch_queries,
                         max_batch_viewpoints=args.batch_viewpoints, consistency=mode,
                         update_cu_percent=args.update_cu, update_tick_us=args.update_tick_us)
    uf = args.update_frac
    kstats = {}  # kernel figures over every session of the served phase

    def fresh_session():
        """Every trial starts from the same map state: a new session built from
        the robots' start disks (the reveal replays the same frontier order),
        warmed with queries only."""
        rv = serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000)
        t1 = time.time()
        sid = srv.add_map(rv.observed, B, G)
        build_ms = 1e3 * (time.time() - t1)
        smap = srv.map_of(sid, rv.observed, B, G)
        smap.profile(True)
        srv.replay(serving.make_trace(sid, rv, 100, 1e6, 11, update_frac=0.0))
        return rv, sid, smap, build_ms

    def close_session(sid, smap):
        add_stats(kstats, kernel_stats_of(smap))
        srv.remove_map(sid)

    # plain per-call path on the same map state (a separate resident copy)
    plain = _lib.GpuMap(init, B, G)
    plain.build()
    prate, pdt, pqs, pres = plain_rate(plain, init, 64, args.plain_steps, args.warmup)
    log(f"plain per-call k=64: {prate:.0f} matrices/s ({1e3 * pdt / args.plain_steps:.3f} ms/query)")

    # capacity estimate: a burst (all arrivals within ~1 ms)
    rv, sid, smap, build_ms = fresh_session()
    log(f"session admitted and built in {build_ms:.1f} ms")
    burst = serving.make_trace(sid, rv, args.window * 8, 1e7, 12, update_frac=uf)
    tb = srv.replay(burst)
    close_session(sid, smap)
    cap = serving.latency_summary(burst, tb)["matrices_per_s"] / 0.9
    log(f"burst capacity ~{cap:.0f} requests/s")

    # Poisson rate search: the highest tested rate with p99 < 1 ms
    trials = []
    chosen = None
    for frac in args.rate_fracs:
        rate = cap * frac
        rv, sid, smap, _ = fresh_session()
        state0 = rv.observed.copy() if not trials and not args.no_cpu else None
        tr = serving.make_trace(sid, rv, args.window * args.steps, rate, 100 + len(trials), update_frac=uf)
        st0 = srv.stats()
        td = srv.replay(tr)
        st1 = srv.stats()
        close_session(sid, smap)
        s = serving.latency_summary(tr, td)
        s["rate"] = rate
        nb = st1["query_batches"] - st0["query_batches"]
        s["mean_batch_queries"] = (st1["queries"] - st0["queries"]) / max(nb, 1)
        s["worker_plan_busy"] = (st1["plan_seconds"] - st0["plan_seconds"]) / max(s["span_s"], 1e-9)
        s["update_busy"] = (st1["update_seconds"] - st0["update_seconds"]) / max(s["span_s"], 1e-9)
        s["batch_ms"] = {k: 1e3 * (st1[f] - st0[f]) / max(nb, 1) for k, f in
                         (("plan_launch", "plan_seconds"), ("gpu_wait", "gpu_wait_seconds"),
                          ("finish", "finish_seconds"))}
        trials.append((tr, td, s, state0))
        log(f"rate {rate:.0f}/s: {s['matrices_per_s']:.0f} matrices/s, p50 {s['p50_ms']:.3f} ms, "
            f"p99 {s['p99_ms']:.3f} ms, updates p50 {s['update_p50_ms']} ms, batch {s['mean_batch_queries']:.2f}, "
            f"worker planning {100 * s['worker_plan_busy']:.0f}%, updating {100 * s['update_busy']:.0f}%, "
            f"batch ms " + " ".join(f"{k} {v:.3f}" for k, v in s["batch_ms"].items()))
        if s["p99_ms"] < 1.0:
            chosen = len(trials) - 1
            break
    if chosen is None:
        chosen = len(trials) - 1
    tr, td, s = trials[chosen][:3]
    roof, valu, kernels = roofline_of(kstats)

    cpu = None
    if not args.no_cpu:
        checks = served_checks(trials[0][0], trials[0][3], 0)
        cpu = cpu_baseline(plain, init, pqs, args.cpu_seconds, pres, checks)

    value = s["matrices_per_s"]
    span = float(td.max() - tr.t.min())
    return {
        "metric": "k x k query matrices/s (served, p99 < 1 ms target) + p99 latency, 1024^2 grid",
        "value": value, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": span * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u16", "data": "synthetic",
        "config": {"workload": "1024x1024 street-like grid seed 3, b=32, g=2x2, 16 robots; Poisson stream of "
                               "queryAllPairs k~logU[8,256] + 10% update batches of 16..512 cells; every rate "
                               "starts from the robots' start disks",
                   "mode": "served", "consistency": args.consistency, "window": args.window, "slots": args.slots,
                   "max_batch_queries": args.batch_queries, "update_cu_percent": args.update_cu,
                   "update_tick_us": args.update_tick_us, "update_frac": uf},
        "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"], "max_ms": s["max_ms"],
        "update_p50_ms": s["update_p50_ms"], "update_p99_ms": s["update_p99_ms"],
        "arrival_rate_per_s": s["rate"],
        "plain_matrices_per_s_k64": prate,
        "served_over_plain": value / prate,
        "server_stats": srv.stats(),
        "rate_trials": [t[2] for t in trials],
        "roofline": roof, "valu": valu, "cpu_baseline": cpu, "kernels": kernels,
    }


def run_update(args):
    """Incremental reveal latency on the serving map: robot batches of 16..512 cells."""
    from htrip_amd import _lib, grids, serving

    truth = grids.street_grid(N, SEED)
    rv = serving.RobotReveal(truth, ROBOTS, RADIUS, seed=5, limit=40000)
    gmap = _lib.GpuMap(rv.observed, B, G)
    gmap.build()
    rng = grids.SplitMix64(77)
    lat, sizes, stats = [], [], []
    for i in range(args.warmup + args.steps):
        m = serving.loguniform_k(rng, 16, 512)
        cells = rv.next_batch(rng.below(ROBOTS), m)
        if i == args.warmup:
            gmap.profile(True)
        t0 = time.perf_counter()
        gmap.open_cells(cells)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            lat.append(dt)
            sizes.append(len(cells))
            stats.append(gmap.last_batch_stats())
    lat = np.array(lat)
    _, _, kernels = kernel_roofline(gmap)
    log(f"update: p50 {1e3 * np.median(lat):.3f} ms, p99 {1e3 * np.percentile(lat, 99):.3f} ms, "
        f"mean cells {np.mean(sizes):.0f}")
    for k, v in kernels.items():
        log(f"  {k:18s} launches {v['launches']:6d} total {v['total_ms']:9.2f} ms")
    return {"metric": "onBatchCellsOpened latency (1024^2 street, b=32, robot batches 16..512 cells)",
            "value": float(np.median(lat) * 1e3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": float(lat.mean() * 1e3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
            "config": {"workload": "1024x1024 street seed 3, b=32, 16 robots, reveal batches logU[16,512]",
                       "mode": "update"},
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_cells": float(np.mean(sizes)),
            "mean_stats": [float(x) for x in np.mean(np.array(stats), axis=0)], "kernels": kernels}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", default="served", choices=["plain", "served", "update"])
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--window", type=int, default=64, help="requests per served step")
    ap.add_argument("--slots", type=int, default=3, help="query streams (the update stream makes one more)")
    ap.add_argument("--update-cu", type=int, default=0, help="percent of CUs for update kernels (0 = all)")
    ap.add_argument("--update-frac", type=float, default=0.1, help="share of update requests in the stream")
    ap.add_argument("--update-tick-us", type=int, default=0, help="snapshot mode: minimum spacing of update batches")
    ap.add_argument("--batch-queries", type=int, default=64)
    ap.add_argument("--batch-viewpoints", type=int, default=8192)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--consistency", default="snapshot", choices=["snapshot", "strict"])
    ap.add_argument("--rate-fracs", type=float, nargs="+", default=[1.2, 1.05, 0.9, 0.75, 0.6, 0.45, 0.3, 0.2])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU (one MI355X per box); --gpus must be 1"}))
        sys.exit(2)
    line = {"plain": run_plain, "served": run_served, "update": run_update}[args.mode](args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
