// Launch wrappers for the gfx950 kernels (kernels_build.hip, kernels_query.hip).
// The runtime checks shapes before calling; a wrapper whose kernel has a hard
// limit of its own returns false for arguments beyond it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace htg {

// Start/stop events attached to one kernel dispatch (hipExtLaunchKernelGGL),
// so the measured interval is the kernel's own execution.  Null events: no
// timing.  Launchers that take one return whether they launched.
struct KTimer {
    hipEvent_t a = nullptr, b = nullptr;
};

// Launch with the timer's start/stop events attached to the dispatch when it
// carries them, as a plain launch otherwise.
#define HTG_LAUNCH_T(K, GRID, BLOCK, LDS, ST, KT, ...)                                          \
    do {                                                                                         \
        if ((KT).a) hipExtLaunchKernelGGL(K, GRID, BLOCK, LDS, ST, (KT).a, (KT).b, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(K, GRID, BLOCK, LDS, ST, __VA_ARGS__);                           \
    } while (0)


// ---- build / update ------------------------------------------------------
// Tile-restricted BFS from every cell of each listed leaf; writes the leaf's
// D (b^2 x b^2) and port summary S ((4b-4)^2).  When changed != nullptr,
// changed[i] is set to 1 if listed leaf i's S decreased anywhere.
// Workgroups an update-path launch may occupy at once (of ~2048 resident on
// the chip): the grid strides over its work so that every workgroup is
// dispatched at once and query kernels on other streams find free slots.
constexpr int kBackgroundWG = 512;

// max_wg > 0 caps the grid (work items are strided over it); chg_idx
// (nullable) maps list position i to its changed[] index.
bool launch_leaf_apsp(const uint8_t* pass, int S, int b, int nl, const int* leaf_list, int n_list, uint16_t* leafD,
                      uint16_t* leafS, int* changed, hipStream_t st, KTimer kt = {}, int max_wg = 0,
                      const int* chg_idx = nullptr);
// Incremental reveal of a leaf with few new cells (b >= 8, b*b % 8 == 0):
// BFS rows of the new cells (items: leaf, first pivot, count <=
// leaf_sources_per_wg(b); pcell: pivot cell ids within the leaf) into the
// leaf's D and into P[pivot][*], then the rank-k relaxation of the leaf's D
// and S through them (leaves: leaf, first pivot, count, changed index).
// Leaves with more new cells than kLeafRelaxMax are recomputed whole.
constexpr int kLeafRelaxMax = 64;
int leaf_sources_per_wg(int b);
bool launch_leaf_pivot_rows(const uint8_t* pass, int S, int b, int nl, const int4* items, int n_items,
                            const int* pcell, uint16_t* leafD, uint16_t* P, uint16_t* leafS, int* changed,
                            hipStream_t st, KTimer kt = {}, int max_wg = 0);
bool launch_leaf_relax(int b, const int4* leaves, int n_leaves, const uint16_t* P, uint16_t* leafD, uint16_t* leafS,
                       int* changed, hipStream_t st, KTimer kt = {});

// D_U of each listed node := block-diagonal child summaries, INF elsewhere.
// Child summaries come from leafS (lvl 1, stride (4b-4)^2) or the previous
// level's S_U.
void launch_node_assemble(int lvl, int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum,
                          size_t child_stride, const int* node_list, int n_list, uint16_t* DU, size_t node_stride,
                          hipStream_t st, const int* keep_cnt = nullptr,
                          const int* keep_list = nullptr, int ng8 = 0);

// Seam edges: for each (u, v) pair of the level's seam template (relative
// cell coordinates in vu_rc) whose cells are both passable, D_U[u][v] and
// D_U[v][u] := min(., 1).  flags (per node, nV bytes) marks rows that
// decreased (nullable).
void launch_seam_relax(int nV, int ld, int s, int nt, const int2* seams, int n_seams, const short2* vu_rc,
                       const uint8_t* pass, int S, const int* node_list, int n_list, uint16_t* DU, size_t node_stride,
                       uint8_t* flags, hipStream_t st);

// Explicit seam edges (absolute V_U pairs per node, from the update path).
// flags: with plane > 0, three planes of n_list * nV bytes (C marks, B marks,
// new-port mask) for the cover of an incremental update; plane == 0 marks
// every decreased row as a pivot.
void launch_edge_relax(int nV, int ld, const int4* edges /* node slot, u, v, w */, int n_edges,
                       const int* node_list, uint16_t* DU, size_t node_stride, uint8_t* flags, size_t plane,
                       hipStream_t st);
void launch_mark_u8(const int2* items /* slot, port */, int n, int nV, uint8_t* plane, hipStream_t st);

// D_U child blocks := min(D_U block, child summary) for children whose mask bit
// is set; flags marks rows that decreased.
void launch_child_relax(int lvl, int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum,
                        size_t child_stride, const int* node_list, const uint32_t* child_mask, int n_list,
                        uint16_t* DU, size_t node_stride, uint8_t* flags, hipStream_t st);

// Flags -> pivot list per node (ascending V_U order); counts[i] = |K_i|;
// flags are cleared.
// Update path, over each listed node's live octets (max_live8 = the largest
// live octet count in the list): child relax for the children whose flag in
// the level below is set (child_slot[i*g^2 + q] = the child's index there,
// -1 when not listed), pivot compaction, and S_U refresh of the nodes with
// pivots (vu_ext[v] = exterior index of V_U port v or -1).
void launch_child_relax_live(int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum, size_t child_stride,
                             const int* node_list, const int* child_slot, const int* child_changed, int n_list,
                             const int* live8_cnt, const int* live8_list, int ng8, int max_live8, uint16_t* DU,
                             size_t node_stride, uint8_t* flags, size_t plane, hipStream_t st);
// C marks -> pivots, counts[slot]; with plane > 0 also B-only marks -> bpiv,
// counts[n_list + slot].
void launch_compact_flags_live(int nV, const int* node_list, int n_list, const int* live8_cnt, const int* live8_list,
                               int ng8, uint8_t* flags, size_t plane, int* pivots, int* bpiv, int piv_stride,
                               int* counts, hipStream_t st);
// Pre-relaxation of the cover rows C through the uncovered endpoints B
// (counts as left by launch_compact_flags_live), then the C x C second hop,
// over work units planned on the host: plan_cover_units appends, for listed
// nodes with cover counts c[i], uncovered-endpoint counts b[i] and live octet
// counts l8[i], the units (slot, row tile | column tile << 16, B begin, B end)
// of one hop and returns the rows per thread (1, 2 or 4) to launch them with.
int plan_cover_units(bool hop2, int n_list, const int* c, const int* b, const int* l8, std::vector<int4>& units);
bool launch_cover_relax(int nV, int ld, const int* node_list, int n_list, const int* counts, const int* cpiv,
                        const int* bpiv, int piv_stride, const int4* units, int n_units, int rpt,
                        const int* live8_cnt, const int* live8_list, int ng8, uint16_t* DU, size_t node_stride,
                        hipStream_t st, KTimer kt = {});
bool launch_cover_pairs(int nV, int ld, const int* node_list, int n_list, const int* counts, const int* cpiv,
                        const int* bpiv, int piv_stride, const int4* units, int n_units, int rpt, uint16_t* DU,
                        size_t node_stride, hipStream_t st, KTimer kt = {});
void launch_project_su_live(int nV, int ld, int ne, const int* vu_ext, const int* node_list, const int* counts,
                            int n_list, const int* live8_cnt, const int* live8_list, int ng8, int max_live8,
                            const uint16_t* DU, size_t node_stride, uint16_t* SU, int* changed, hipStream_t st);
void launch_compact_flags(int nV, const int* node_list, int n_list, uint8_t* flags, int* pivots, int piv_stride,
                          int* counts, hipStream_t st);

// One pivot block of the closure: for every listed node with more than t*P
// pivots, R := the node's block pivot rows after Floyd-Warshall over the
// block's pivots (P x nV, into Rbuf slot i), on the node's live octets
// (live8_list, live8_cnt; max_live8 = the largest count in the list).
bool launch_closure_pivots(int P, int nV, int ld, const int* node_list, int n_list, const int* pivots, int piv_stride,
                           const int* counts, int t, const uint16_t* DU, size_t node_stride, const int* live8_cnt,
                           const int* live8_list, int ng8, int max_live8, uint16_t* Rbuf, hipStream_t st, KTimer kt = {});
// Phase 2 of the same block, D := min(D, R^T (x) R), over live octets.
bool launch_closure_update(int P, int nV, int ld, const int* node_list, int n_list, const int* counts, int t,
                           uint16_t* DU, size_t node_stride, const int* live8_cnt, const int* live8_list, int ng8,
                           int max_live8, const uint16_t* Rbuf, hipStream_t st, KTimer kt = {});
// Phase 2 of block t (from Rcur) and phase 1 of block t + 1 (into Rnext, a
// different buffer) in one launch.
bool launch_closure_step(int P, int nV, int ld, const int* node_list, int n_list, const int* pivots, int piv_stride,
                         const int* counts, int t, uint16_t* DU, size_t node_stride, const int* live8_cnt,
                         const int* live8_list, int ng8, int max_live8, const uint16_t* Rcur, uint16_t* Rnext,
                         hipStream_t st, KTimer kt = {});

// S_U := D_U[ext][ext]; changed[i] = 1 if listed node i's S_U decreased.
void launch_project_su(int nV, int ld, int ne, const int* ext_idx, const int* node_list, int n_list,
                       const uint16_t* DU, size_t node_stride, uint16_t* SU, int* changed, hipStream_t st);

// Snapshot catch-up copies: whole items (item t = elems u16 at t*elems) and
// the live rows of node tables (D_U rows of live octets, S_U rows of the
// exterior ports among them).
void launch_copy_items(const int* items, int n_items, size_t elems, const uint16_t* src, uint16_t* dst,
                       hipStream_t st);
void launch_copy_live_rows(const int* nodes, int n_nodes, int nV, int ld, size_t node_stride, const int* live8_cnt,
                           const int* live8_list, int ng8, int max_live8, int ne, const int* vu_ext,
                           const uint16_t* srcDU, uint16_t* dstDU, const uint16_t* srcSU, uint16_t* dstSU,
                           hipStream_t st);

// p[0..n) = v (graph-capturable).
// src and dst 16-byte aligned; either may be pinned host memory.
void launch_copy_bytes(const void* src, void* dst, size_t bytes, hipStream_t st);
void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t st);

// pass[r*S + c] = 1 for the listed cells.
void launch_set_pass(uint8_t* pass, int S, const int2* cells, int n, hipStream_t st, uint8_t v = 1);

// ---- query -----------------------------------------------------------------
// Work counters of the query projection: kWorkSlots (bytes, updates) pairs,
// summed on the host.
constexpr int kWorkSlots = 64;

// Argument blocks of the query's work units.  The per-phase kernels take
// them by value; the persistent launch reads them from the batch's plan in
// device memory.
//
// Front: out matrices (INF, 0 on the diagonal of valid viewpoints, leaf D
// for same-leaf pairs; q_* per query, vp_* per viewpoint, all queries
// packed), X0[slot][p] = leaf D row of the slot's cell at leaf port p, and
// INF fills of n_fill ranges (start lo, start hi, n8, first unit) of n8
// units of 8 u16 from fill_base (n_fill <= 64: one ballot finds a unit's
// range); unit 0 clears wstat (nullable).  Units: nq * ib output blocks,
// leaf_blocks, fill_blocks.
struct QFrontArgs {
    const int* q_k;
    const long long* q_out;
    const int* q_vp;
    const int2* vp_rc;
    const uint8_t* vp_valid;
    const int2* slot_rc;
    const uint16_t* leafD;
    uint16_t* out;
    uint16_t* X0;
    const int4* fills;
    uint16_t* fill_base;
    unsigned long long* wstat;
    int nq, ib, leaf_blocks, n_slots, b, nl, n_fill, fill_blocks;
};
// One level's projection: Zp[chunk][slot][col] = min over the chunk's seam
// rows a of X[slot][a] (+) D_W[q*pc + a][cols(col)] for work items (two
// int4 each: (W, q | first-of-run flag << 16, first slot, count) and (live
// column groups of W, seam range lo, hi, first seam port of q)); ksplit row
// chunks per item, col_tiles tiles of 64 live groups (the largest item's
// count).  With ksplit == 1 the results go straight to Z (position
// colpos[col]) and the lifted Xn (exterior port colext[col]; ne = 0 at the
// root).  Units: n_items x col_tiles x ksplit, items fastest.
// Its fold (ksplit > 1): Z[slot][pos] = min_c Zp[c][slot][zsrc[pos]] (INF
// where zsrc[pos] < 0 or the column's group is not live in the slot's node)
// and Xn[slot][e] = min(folded column ext_col[e], X[slot][own port of e]).
// Units: ceil(n_slots * (nz + ne) / 256).
struct QZArgs {
    const uint16_t* DU;
    const int* seam_list;
    const int* cols;
    const int* live_list;
    const int4* items;
    const uint16_t* X;
    uint16_t* Zp;
    const int* colpos;
    const int* colext;
    const int* ext_child;
    const int* ext_cport;
    uint16_t* Z;
    uint16_t* Xn;
    unsigned long long* wstat;
    const int* zsrc;
    const int2* slot_rc;
    const uint8_t* live_map;
    const int* ext_col;
    size_t node_stride, part_stride;
    int ld, pc, ngrp, ldp, n_items, col_tiles, ksplit, ldz, ne, M;
    int n_slots, nz, side, nt, cs, g;
};
// Meet tiles (8 ints each: group size, q2, i0, j0, j_end, level | (16 x 16
// tile) << 16, the group's offset in glist, its query): out[i][j] = min(out,
// min over the q2 seam block of Z_level[i] (+) Z_level[j]); zgeo per level:
// Z offset lo, hi, ldz, 0, zoff[g2], zlen[g2].  One unit per tile; a tile
// covers 8 x 8 pairs, or 16 x 16 when its flag is set.
struct QMeetArgs {
    const int* tiles;
    const int* glist;
    const int* zgeo;
    const int* slot_i;
    const int* q_k;
    const long long* q_out;
    const uint16_t* Z;
    uint16_t* out;
    int n_tiles, g2;
};

bool launch_q_front(const QFrontArgs& a, hipStream_t st);
bool launch_q_z(const QZArgs& a, hipStream_t st, KTimer kt = {});
bool launch_q_fold(const QZArgs& a, hipStream_t st, KTimer kt = {});
bool launch_q_meet(const QMeetArgs& a, hipStream_t st, KTimer kt = {});
// liftPoint: X0[i] = row[cell of leaf port i] (row: the point's row of its leaf D).
void launch_lift_leaf(const uint16_t* row, int b, uint16_t* X0, hipStream_t st);
// liftPoint, one level: Xo[e] = min_p X[p] (+) D[row0 + p][column of exterior
// port e] over the pc rows of the point's child, in nchunks chunks of
// rows_per_chunk (<= kLiftRowsPerChunk) rows (part: nchunks x ld scratch).
constexpr int kLiftRowsPerChunk = 64;
void launch_lift_level(const uint16_t* D, int ld, int row0, int pc, int rows_per_chunk, int nchunks,
                       const uint16_t* X, uint16_t* part, const int* ext_child, const int* ext_cport, int ne,
                       uint16_t* Xo, hipStream_t st);

}  // namespace htg
