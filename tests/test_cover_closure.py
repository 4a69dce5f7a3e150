"""The update closure over a vertex cover, restated on the CPU and checked
against the oracle's tables.

An incremental reveal lowers entries of a node's closed table D (child
summary blocks and new seam edges).  The reference closes D over every row
with a decreased entry (src/internal_node.cpp:332-413).  The library closes it
over a cover C of the decreased entries instead: ports opened by the batch,
plus both endpoints of decreases between old ports.  First the C rows are
relaxed through the uncovered endpoints B, then the C x C block is relaxed
through B a second time.  This test replays that procedure in numpy on the
oracle's before/after tables and requires the oracle's closed table exactly.
It also checks that each hop is needed.
"""
import numpy as np
import pytest

from htrip_amd import grids, serving
from oracle.oracle import Oracle

INF = 60000


def port_cells(n, b, g, lvl, w):
    """V_U port cells of node w at level lvl in canonical order (children in
    quad order; per child: top row, bottom row, left column, right column)."""
    s = b * g ** lvl
    cs = s // g
    nt = n // s
    r0, c0 = (w // nt) * s, (w % nt) * s
    per = [(0, c) for c in range(cs)] + [(cs - 1, c) for c in range(cs)]
    per += [(r, 0) for r in range(1, cs - 1)] + [(r, cs - 1) for r in range(1, cs - 1)]
    out = []
    for q in range(g * g):
        qr, qc = q // g, q % g
        out += [(r0 + qr * cs + r, c0 + qc * cs + c) for r, c in per]
    return out


def child_summary(O, n, b, g, lvl, w, q):
    s = b * g ** lvl
    nt = n // s
    cx, cy = (w % nt) * g + q % g, (w // nt) * g + q // g
    if lvl == 1:
        return O.leaf_S(cy * (n // b) + cx).astype(np.int64)
    return O.node_SU(lvl - 1, cy * (n // (s // g)) + cx).astype(np.int64)


def minplus(a, b):
    return np.minimum(a[:, :, None] + b[None, :, :], INF).min(1)


def cover_close(D, cmark, bmark, hop1=True, hop2=True):
    """Hop 1, hop 2, then a closure over C alone in the form the GPU kernels use
    (R = C* (x) D[C][:], D = min(D, R^T (x) R): only rows of C are read)."""
    C = np.nonzero(cmark)[0]
    B = np.nonzero(bmark & ~cmark)[0]
    if len(C) == 0:
        return D
    if len(B) and hop1:
        D[C, :] = np.minimum(D[C, :], minplus(D[np.ix_(C, B)], D[B, :]))
    if len(B) and hop2:
        T = D[np.ix_(C, B)]
        D[np.ix_(C, C)] = np.minimum(D[np.ix_(C, C)], minplus(T, T.T))
    cstar = D[np.ix_(C, C)].copy()
    for k in range(len(C)):
        cstar = np.minimum(cstar, np.minimum(cstar[:, k:k + 1] + cstar[k:k + 1, :], INF))
    R = minplus(cstar, D[C, :])
    return np.minimum(D, minplus(R.T, R))


def replay(kind, n, b, seed, steps, lo, hi, hop1=True, hop2=True):
    """Reveal batches; every changed node table closed over the cover must equal
    the oracle's.  Returns (nodes checked, cover rows, reference pivot rows)."""
    g = 2
    truth = grids.street_grid(n, seed) if kind == "street" else grids.maze_grid(n, seed)
    rv = serving.RobotReveal(truth, 6, max(6, n // 10), seed + 1, limit=100000)
    O = Oracle("port", rv.observed.copy(), b, g)
    O.build()
    rng = grids.SplitMix64(seed + 2)
    checked = n_cover = n_ref = 0
    for _ in range(steps):
        cells = rv.next_batch(rng.below(6), serving.loguniform_k(rng, lo, hi))
        if len(cells) == 0:
            continue
        before = O.cells.copy()
        old = {(lv, w): O.node_DU(lv, w).astype(np.int64) for lv in range(1, O.H + 1) for w in range(O.num_nodes(lv))}
        O.open_cells(cells)
        new = {tuple(p) for p in cells.tolist() if before[tuple(p)] != 1}
        for lv in range(1, O.H + 1):
            for w in range(O.num_nodes(lv)):
                D = old[(lv, w)].copy()
                pcs = port_cells(n, b, g, lv, w)
                nv = len(pcs)
                pc = nv // (g * g)
                isnew = np.array([p in new for p in pcs])
                cmark = np.zeros(nv, bool)
                bmark = np.zeros(nv, bool)
                dec_rows = np.zeros(nv, bool)

                def mark(v1, v2):  # the marking of k_child_relax_live / k_edge_relax
                    to_c = isnew[v1] | ~isnew[v2]
                    cmark[v1[to_c]] = True
                    bmark[v1[~to_c]] = True
                    dec_rows[v1] = True

                for q in range(g * g):
                    sl = slice(q * pc, (q + 1) * pc)
                    sq = child_summary(O, n, b, g, lv, w, q)
                    dec = sq < D[sl, sl]
                    D[sl, sl] = np.minimum(D[sl, sl], sq)
                    v1, v2 = np.nonzero(dec)
                    mark(v1 + q * pc, v2 + q * pc)
                idx = {p: i for i, p in enumerate(pcs)}
                for i, p in enumerate(pcs):
                    if not isnew[i] or O.cells[p] != 1:
                        continue
                    for dr, dc in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                        j = idx.get((p[0] + dr, p[1] + dc))
                        if j is None or j // pc == i // pc or O.cells[pcs[j]] != 1 or D[i, j] <= 1:
                            continue
                        D[i, j] = D[j, i] = 1
                        mark(np.array([i, j]), np.array([j, i]))
                D = cover_close(D, cmark, bmark, hop1, hop2)
                if not (D == O.node_DU(lv, w)).all():
                    return None
                checked += 1
                n_cover += int(cmark.sum())
                n_ref += int(dec_rows.sum())
    return checked, n_cover, n_ref


@pytest.mark.parametrize("kind,n,b,seed,steps,lo,hi", [
    ("street", 64, 8, 3, 6, 8, 160),
    ("maze", 64, 8, 4, 6, 8, 160),
    ("street", 128, 16, 5, 4, 16, 400),
])
def test_cover_closure_equals_the_oracle(kind, n, b, seed, steps, lo, hi):
    got = replay(kind, n, b, seed, steps, lo, hi)
    assert got is not None, "a node table closed over the cover differs from the oracle's"
    checked, n_cover, n_ref = got
    assert checked > 0
    assert n_cover < n_ref  # the cover is smaller than the reference's pivot set


def test_each_hop_is_needed():
    """Without hop 1, or without hop 2, some node table comes out wrong."""
    assert replay("maze", 128, 8, 7, 8, 8, 300, hop1=False) is None
    assert replay("maze", 128, 8, 7, 8, 8, 300, hop2=False) is None
    assert replay("maze", 128, 8, 7, 8, 8, 300) is not None
