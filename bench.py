#!/usr/bin/env python3
"""Benchmark of the MI355X min-plus distance oracle (one JSON line on stdout).

Workload (the serving configuration): a 1024 x 1024 street-like seeded grid,
leaf b = 32, group 2 x 2 (H = 5), hierarchy built and resident in HBM.

Modes
  plain  -- one step = one plain per-call queryAllPairs (k = 64) through the
            drop-in C ABI (host viewpoints in, host matrix out, synchronous).
  served -- one step = one window of the request trace served by the
            continuous-batching loop (see htrip_amd/serving.py).

`value` is k x k query matrices per second.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# 256 CUs x 64 lanes x 2 u16 per packed op x 2.4 GHz = 78.6 T u16 ops/s; one
# min-plus update is a packed add + a packed min, so 39.3 T updates/s.
PEAK_UPDATES = 256 * 64 * 2 * 2.4e9 / 2

N, B, G, SEED = 1024, 32, 2, 3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(gmap, cells, queries, budget_s: float, expect):
    """The reference's own CPU AVX2 queryAllPairs (oracle/_ref), one thread,
    over a bounded sample of the same queries, on tables copied from the GPU."""
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
        gmap.lib.htg_read_leaf_D(gmap.h, i, buf.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_uint16)))
        api["import_leaf"](R.h, i, buf)
    for lvl in range(1, gmap.H + 1):
        for i in range(gmap.num_nodes(lvl)):
            api["import_node"](R.h, lvl, i, gmap.node_DU(lvl, i))
    log(f"cpu baseline: imported GPU tables into the reference in {time.time() - t0:.1f}s")
    done, spent = 0, 0.0
    for q, vps in enumerate(queries):
        vps = np.ascontiguousarray(vps, dtype=np.int32)
        ks = np.array([len(vps)], dtype=np.int32)
        out = np.empty(len(vps) * len(vps), dtype=np.uint16)
        dt = api["time_queries"](R.h, vps, ks, 1, 1, out.ctypes.data)
        if q < len(expect) and expect[q] is not None and not np.array_equal(out.reshape(len(vps), -1), expect[q]):
            raise SystemExit("reference CPU result differs from the GPU result")
        spent += dt
        done += 1
        if spent >= budget_s:
            break
    log(f"cpu baseline: {done} queries in {spent:.2f}s")
    return {"value": done / spent, "unit": "matrices/s", "cores": 1, "kind": "reference",
            "sample": f"{done} queryAllPairs k={len(queries[0])} on the same map, 1 thread, reference AVX2 build"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", default="plain", choices=["plain", "served"])
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from htrip_amd import _lib, grids

    t0 = time.time()
    cells = grids.street_grid(N, SEED)
    log(f"grid {N}x{N} street seed={SEED}: {int((cells == 1).sum())} free cells ({time.time() - t0:.1f}s)")
    gmap = _lib.GpuMap(cells, B, G)
    gmap.profile(True)
    t0 = time.time()
    gmap.build()
    build_s = time.time() - t0
    build_stats = {_lib.KERNEL_NAMES[c]: gmap.kernel_stats(c) for c in range(3)}
    log(f"build {build_s * 1e3:.1f} ms (S={gmap.S} H={gmap.H})")

    nq = args.warmup + args.steps
    queries = [grids.sample_free(cells, args.k, 1000 + i) for i in range(nq)]
    for i in range(args.warmup):
        gmap.query(queries[i])
    gmap.profile(True)
    gmap.sync()
    t0 = time.perf_counter()
    results = []
    for i in range(args.warmup, nq):
        results.append(gmap.query(queries[i]))
    gmap.sync()
    dt = time.perf_counter() - t0
    log(f"plain: {args.steps} queries k={args.k} in {dt * 1e3:.1f} ms")
    ks = {_lib.KERNEL_NAMES[c]: gmap.kernel_stats(c) for c in (3, 4, 5)}
    z = ks["k_q_z"]
    avg_ms = z["total_ms"] / max(z["launches"], 1)
    achieved_gbs = (z["bytes"] / max(z["launches"], 1)) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    upd_rate = (z["updates"] / max(z["launches"], 1)) / (avg_ms * 1e-3) if avg_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        sample = queries[args.warmup:]
        cpu = cpu_baseline(gmap, cells, sample, args.cpu_seconds, results)

    value = args.steps / dt
    line = {
        "metric": "k x k query matrices/s (plain per-call GPU path, k=64)",
        "value": value,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": "1024x1024 street-like grid seed 3, b=32, g=2x2, queryAllPairs k=64",
                   "mode": args.mode, "k": args.k},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / PEAK_HBM_GBS, "traffic": None, "kernel": "k_q_z",
                     "avg_launch_ms": avg_ms},
        "valu": {"kernel": "k_q_z", "achieved_updates_per_s": upd_rate, "peak_updates_per_s": PEAK_UPDATES,
                 "frac": upd_rate / PEAK_UPDATES},
        "cpu_baseline": cpu,
        "build_ms": build_s * 1e3,
        "kernels": ks,
        "build_kernels": build_stats,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
