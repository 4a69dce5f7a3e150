"""htg_query_batch_noalloc (the query entry point that allocates nothing)
against htg_query_batch, the allocating entry point it mirrors.

Bar: bit-exact (integer tropical arithmetic).  Two calls in a row, the
second with a smaller input, on two maps that share the per-GPU scratch
array; a batch beyond the array is refused with HTG_E_RANGE and leaves the
entry point working."""
import ctypes as C

import numpy as np
import pytest

from htrip_amd import _lib, grids, serving
from oracle.oracle import grid_bfs

pytestmark = pytest.mark.gpu


def noalloc(G, vps_list):
    ks = np.array([len(v) for v in vps_list], dtype=np.int32)
    allv = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32).reshape(-1, 2) for v in vps_list]),
                                dtype=np.int32)
    out = np.full(int((ks.astype(np.int64) ** 2).sum()), 12345, dtype=np.uint16)
    rc = G.lib.htg_query_batch_noalloc(G.h, _lib.ptr(allv, C.c_int32), _lib.ptr(ks, C.c_int32), len(ks),
                                       _lib.ptr(out, C.c_uint16))
    res, o = [], 0
    for k in ks:
        res.append(out[o:o + k * k].reshape(k, k))
        o += int(k) * int(k)
    return rc, res


def test_no_alloc_batches_equal_the_allocating_entry_point(gpu_ctx):
    print("no-alloc: served 1024^2 map and a 256^2 map", flush=True)
    truth = grids.street_grid(1024, 3)
    rv = serving.RobotReveal(truth, 16, 64, seed=5, limit=40000)
    A = _lib.GpuMap(rv.observed, 32, 2)
    assert A.build()
    small = grids.street_grid(256, 1)
    Bm = _lib.GpuMap(small, 16, 2)
    assert Bm.build()
    rng = grids.SplitMix64(77)
    big_batch = [rv.sample_viewpoints(serving.loguniform_k(rng, 8, 256), rng) for _ in range(8)]
    big_batch.append(rv.sample_viewpoints(256, rng))
    small_batch = big_batch[:2]
    for G, first, second in ((A, big_batch, small_batch),
                             (Bm, [grids.sample_free(small, 200, 3), grids.sample_free(small, 64, 4)],
                              [grids.sample_free(small, 17, 5)]),
                             (A, small_batch, [big_batch[0][:5]])):
        for batch in (first, second):  # two calls in a row, the second smaller
            rc, got = noalloc(G, batch)
            assert rc == _lib.HTG_OK, G.lib.htg_last_error().decode()
            want = G.query_batch(batch)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g, w)
    np.testing.assert_array_equal(noalloc(Bm, [small_batch[0][:0], grids.sample_free(small, 30, 9)])[1][1],
                                  grid_bfs(small, grids.sample_free(small, 30, 9)))
    assert noalloc(Bm, [small_batch[0][:0]])[0] == _lib.HTG_OK  # k = 0 only: nothing to do


def test_no_alloc_refuses_batches_beyond_its_array(gpu_ctx):
    small = grids.street_grid(256, 1)
    G = _lib.GpuMap(small, 16, 2)
    assert G.build()
    huge = grids.sample_free(small, 6000, 11)  # the 6000 x 6000 result alone is 72 MB
    rc, _ = noalloc(G, [huge])
    assert rc == _lib.HTG_E_RANGE
    msg = G.lib.htg_last_error().decode()
    assert "bytes of device scratch" in msg and str(64 << 20) in msg
    v = grids.sample_free(small, 40, 12)
    rc, got = noalloc(G, [v])
    assert rc == _lib.HTG_OK
    np.testing.assert_array_equal(got[0], grid_bfs(small, v))
    # the largest single query that fits on the 1024^2 serving map (printed for the record)
    truth = grids.street_grid(1024, 3)
    rv = serving.RobotReveal(truth, 16, 64, seed=5, limit=40000)
    A = _lib.GpuMap(rv.observed, 32, 2)
    assert A.build()
    pool = grids.sample_free(rv.observed, 4096, 13)
    lo, hi = 256, 4096
    assert noalloc(A, [pool[:lo]])[0] == _lib.HTG_OK
    while hi - lo > 16:
        mid = (lo + hi) // 2
        if noalloc(A, [pool[:mid]])[0] == _lib.HTG_OK:
            lo = mid
        else:
            hi = mid
    print(f"largest single query on the 1024^2 b=32 serving map: k ~ {lo}", flush=True)
