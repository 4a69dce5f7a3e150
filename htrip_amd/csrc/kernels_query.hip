// gfx950 kernels for the batched all-pairs query.
//
// The reference lifts each viewpoint level by level and meets viewpoint pairs
// through every common ancestor W (src/minplus_oracle.cpp:934-1419).  Here a
// viewpoint i in child q of W is first projected once per level onto W's
// port set:
//     Z_i[x] = min_{a in seam(q)} X_i[a] (+) D_W[a][x]
// where seam(q) are q's ports with a 4-neighbour in a sibling child (a path
// that leaves q inside W must leave through one of them).  Z then serves both
// the meet of every pair that shares W,
//     m_ij = min_{b in seam(q_j)} Z_i[b] (+) X_j[b],
// and the lift to W's exterior ports, X_i'[e] = min(Z_i[e], X_i[e] when e is
// one of q's own ports).  The projection is a min-plus GEMM over the
// viewpoints that share (W, q), so every D_W row read feeds up to M
// viewpoints from registers.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <hip/hip_ext.h>

#include "geometry.hpp"
#include "kernels.hpp"

namespace htg {

typedef unsigned short us2q __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t qrelax2(uint32_t acc, uint32_t a, uint32_t b) {
    us2q x = __builtin_bit_cast(us2q, a), y = __builtin_bit_cast(us2q, b), z = __builtin_bit_cast(us2q, acc);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(z, __builtin_elementwise_add_sat(x, y)));
}
__device__ __forceinline__ uint32_t qdup(uint32_t v) { return (v & 0xffffu) | (v << 16); }
__device__ __forceinline__ uint32_t qadd2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(us2q, a),
                                                                      __builtin_bit_cast(us2q, b)));
}
__device__ __forceinline__ uint32_t qmin2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2q, a),
                                                                  __builtin_bit_cast(us2q, b)));
}

// XCD-aware work order.  The hardware deals a launch's workgroups round-robin
// over the 8 XCDs, each with its own L2.  This bijection hands each XCD runs
// of G consecutive logical items (its k-th workgroup takes item
// (k / G) * 8G + x G + k % G), so that neighbouring items, which share rows,
// share one L2, while the runs still rotate over all XCDs and a range of
// expensive items does not land on one of them.  The last n mod 8G items keep
// the hardware order.  Runs of 8 fetch about 12 % fewer HBM bytes in the
// projection and the meet than runs of 4, at the same kernel time.  Runs of 2
// fetch more; runs of 16 fetch a little less again but move closer to whole
// ranges per XCD, which slowed batches under load.
constexpr unsigned kXcdRun = 8;
template <unsigned G>
__device__ __forceinline__ unsigned xcd_logical(unsigned hw, unsigned n) {
    constexpr unsigned X = 8;
    const unsigned full = n / (X * G) * (X * G);
    if (hw >= full) return hw;
    const unsigned x = hw % X, k = hw / X;
    return (k / G) * (X * G) + x * G + k % G;
}

// Front of the query, one 1-D launch with three block ranges:
//  [0, nq * ib): output init (src/minplus_oracle.cpp:940-965, same-leaf pairs
//     :1027-1048) of query blk / ib: INF, 0 on the diagonal of valid
//     viewpoints, leaf D for same-leaf pairs;
//  next leaf_blocks: level-0 lift, X0[slot][p] = leaf D row of the slot's cell
//     at leaf port p (src/minplus_oracle.cpp:1000-1019);
//  then INF fills of the buffers whose live-only writers leave gaps: fill
//     range r = (start lo, start hi, n8, first block), n8 units of 8 u16.
// Block 0 also clears the k_q_z work counters.
__global__ __launch_bounds__(256) void k_q_front(int nq, int ib, int leaf_blocks, const int* __restrict__ q_k,
                                                 const long long* __restrict__ q_out, const int* __restrict__ q_vp,
                                                 const int2* __restrict__ vp_rc, const uint8_t* __restrict__ vp_valid,
                                                 int n_slots, const int2* __restrict__ slot_rc, int b, int nl,
                                                 const uint16_t* __restrict__ leafD, uint16_t* __restrict__ out,
                                                 uint16_t* __restrict__ X0, int n_fill, const int4* __restrict__ fills,
                                                 uint16_t* __restrict__ fill_base,
                                                 unsigned long long* __restrict__ wstat) {
    const int n = b * b;
    const int blk = int(blockIdx.x);
    if (wstat && blk == 0 && threadIdx.x < 2 * kWorkSlots) wstat[threadIdx.x] = 0;
    if (blk < nq * ib) {
        const int Q = blk / ib;
        const int k = q_k[Q];
        const long long idx = (long long)(blk % ib) * 256 + threadIdx.x;
        if (idx >= (long long)k * k) return;
        const int i = int(idx / k), j = int(idx % k);
        const int vi = q_vp[Q] + i, vj = q_vp[Q] + j;
        uint16_t val = kInf;
        if (vp_valid[vi] && vp_valid[vj]) {
            if (i == j) {
                val = 0;
            } else {
                const int2 a = vp_rc[vi], c = vp_rc[vj];
                if (a.x / b == c.x / b && a.y / b == c.y / b) {
                    const size_t leaf = size_t(a.x / b) * nl + a.y / b;
                    val = leafD[leaf * n * n + size_t((a.x % b) * b + a.y % b) * n + (c.x % b) * b + c.y % b];
                }
            }
        }
        out[q_out[Q] + idx] = val;
        return;
    }
    const int lb = blk - nq * ib;
    if (lb < leaf_blocks) {
        const int pl = 4 * b - 4;
        const long long gid = (long long)lb * 256 + threadIdx.x;
        if (gid >= (long long)n_slots * pl) return;
        const int s = int(gid / pl), p = int(gid % pl);
        const int2 rc = slot_rc[s];
        int pr, pcc;
        port_rc(p, b, pr, pcc);
        const size_t leaf = size_t(rc.x / b) * nl + rc.y / b;
        X0[gid] = leafD[leaf * n * n + size_t((rc.x % b) * b + rc.y % b) * n + pr * b + pcc];
        return;
    }
    // the fill range of this block: ranges are in order of their first block,
    // so r counts the ranges after the first that start at or before fb (one
    // load per lane and a ballot instead of a walk of dependent loads)
    const int fb = lb - leaf_blocks;
    const int lane = int(threadIdx.x & 63u);
    const bool starts_before = lane >= 1 && lane < n_fill && fills[lane].w <= fb;
    const int r = __popcll(__ballot(starts_before));
    const int4 f = fills[r];
    const long long u = (long long)(fb - f.w) * 256 + threadIdx.x;
    if (u >= f.z) return;
    const size_t start = (size_t(unsigned(f.y)) << 32) | unsigned(f.x);
    reinterpret_cast<uint4*>(fill_base + start)[u] = make_uint4(kInf2, kInf2, kInf2, kInf2);
}

bool launch_q_front(int nq, const int* q_k, const long long* q_out, const int* q_vp, const int2* vp_rc,
                    const uint8_t* vp_valid, int kmax, int n_slots, const int2* slot_rc, int b, int nl,
                    const uint16_t* leafD, uint16_t* out, uint16_t* X0, int n_fill, const int4* fills,
                    int fill_blocks, uint16_t* fill_base, unsigned long long* wstat, hipStream_t st) {
    // the block's fill range is found with one 64-lane ballot over the ranges
    if (n_fill > 64 || n_fill < 0) return false;
    if (nq == 0) return true;
    const int ib = int(((long long)kmax * kmax + 255) / 256);
    const int leaf_blocks = int(((long long)n_slots * (4 * b - 4) + 255) / 256);
    const long long blocks = (long long)nq * ib + leaf_blocks + fill_blocks;
    if (blocks == 0) return true;
    hipLaunchKernelGGL(k_q_front, dim3(unsigned(blocks)), dim3(256), 0, st, nq, ib, leaf_blocks, q_k, q_out, q_vp,
                       vp_rc, vp_valid, n_slots, slot_rc, b, nl, leafD, out, X0, n_fill, fills, fill_base, wstat);
    return true;
}

// Projection Z (see file comment), split over the seam rows: work item
// (W, q, first slot, count), column tile and row chunk (the grid's x, y, z in
// an XCD-aware order);
// four wavefronts share the columns and take interleaved rounds of rows.
// Only live column groups are computed: live_list[W] holds the groups of 4
// projection columns with at least one passable port (the others are INF and
// k_q_zred writes them so), 64 groups = 256 columns per workgroup.  Each
// wavefront keeps M x 4 packed accumulators and reads one 8-byte slice of a
// D_W row per seam row.  Seam rows whose X is INF for all M viewpoints add
// nothing and are dropped: each round stages 64 rows, compacts the finite
// ones (ballot + prefix count) into LDS with their X values duplicated into
// both u16 halves, and pads the count to PF rows with INF.  Partial minima go
// to Zp[chunk] and k_q_fold folds the chunks; with one chunk (ksplit == 1)
// the kernel folds in registers instead: column c goes to Z position
// colpos[c] and, below the root, lifts to exterior port colext[c] of W,
// Xn[slot][e] = min(Z, X[slot][own port of e] when e is a port of q)
// (src/minplus_oracle.cpp:1113-1154).  wstat (nullable) accumulates the
// work done: [0] bytes (D slices of distinct (W, q) items, X reads, Zp
// writes, each counted once), [1] min-plus updates on finite rows.
template <int M, int PF>
__global__ __launch_bounds__(256) void k_q_z(
    int ld, int pc, const uint16_t* __restrict__ DU, size_t node_stride, const int* __restrict__ seam_list,
    const int* __restrict__ cols, int ngrp, const int* __restrict__ live_list, int ldp,
    const int4* __restrict__ items, int ksplit,
    const uint16_t* __restrict__ X, uint16_t* __restrict__ Zp, size_t part_stride,
    const int* __restrict__ colpos, const int* __restrict__ colext, const int* __restrict__ ext_child,
    const int* __restrict__ ext_cport, uint16_t* __restrict__ Z, int ldz, uint16_t* __restrict__ Xn, int ne,
    unsigned long long* __restrict__ wstat) {
    constexpr int ACH = 64;  // seam rows staged per wavefront and round (PF: D rows in flight per lane)
    constexpr int NW = 4;    // wavefronts per workgroup: same columns, interleaved row rounds
    // Padding rows carry INF in xs: min(acc, INF (+) d) leaves acc unchanged,
    // so the loop runs branch-free over whole PF blocks.
    __shared__ __align__(16) uint32_t xs[NW][ACH][M];  // per wavefront, row-major: one row's M values contiguous
    __shared__ unsigned roff[NW][ACH];
    __shared__ uint2 red[NW - 1][M][64];  // accumulators of wavefronts 1..NW-1 for the final minimum
    // logical (item, column tile, row chunk): items of one (W, q) read the same
    // D rows, so runs of consecutive items go to one XCD
    const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned L = xcd_logical<kXcdRun>(lin, gridDim.x * gridDim.y * gridDim.z);
    const int bx = int(L % gridDim.x), by = int((L / gridDim.x) % gridDim.y), bz = int(L / (gridDim.x * gridDim.y));
    const int4 it = items[2 * bx], it2 = items[2 * bx + 1];
    const int W = it.x, q = it.y & 0xffff, s0 = it.z, m = it.w;
    const bool first_item = (it.y >> 16) & 1;  // first item of its (W, q) run: counts the D bytes once
    const int tid = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nlive = it2.x;  // live column groups of W
    const int gi = by * 64 + tid;
    if (by * 64 >= nlive) return;  // whole workgroup past W's live groups
    const bool active = gi < nlive;
    const int c0 = active ? live_list[size_t(W) * ngrp + gi] * 4 : 0;
    // Projection columns come in aligned groups of 4 consecutive V_U ports,
    // so one 8-byte load per seam row feeds the lane's 4 columns.
    const int base = cols[c0];
    const uint16_t* D = DU + size_t(W) * node_stride + base;
    // The epilogue's metadata (Z positions, exterior ports, the slot's own
    // X at those ports), loaded now by the wavefront that writes the result:
    // its dependent loads overlap the row loads instead of following them.
    int pos[4] = {-1, -1, -1, -1}, ext[4] = {-1, -1, -1, -1};
    uint16_t ownx[M][4];
#pragma unroll
    for (int v = 0; v < M; ++v)
#pragma unroll
        for (int k = 0; k < 4; ++k) ownx[v][k] = kInf;
    if (ksplit == 1 && w == 0 && active) {
        const int4 cp = *reinterpret_cast<const int4*>(colpos + c0);
        const int4 ce = *reinterpret_cast<const int4*>(colext + c0);
        pos[0] = cp.x, pos[1] = cp.y, pos[2] = cp.z, pos[3] = cp.w;
        ext[0] = ce.x, ext[1] = ce.y, ext[2] = ce.z, ext[3] = ce.w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int own = (ext[k] >= 0 && ext_child[ext[k]] == q) ? ext_cport[ext[k]] : -1;
            if (own < 0) continue;
#pragma unroll
            for (int v = 0; v < M; ++v)
                if (v < m) ownx[v][k] = X[size_t(s0 + v) * pc + own];
        }
    }
    const int* seam = seam_list + it2.y;  // child q's seam ports
    const int ns = it2.z - it2.y;
    const int a_lo = int((long long)ns * bz / ksplit), a_hi = int((long long)ns * (bz + 1) / ksplit);
    const unsigned pad_off = unsigned(q * pc + it2.w) * unsigned(ld);
    // rows per wavefront and round: the item's rows spread over all NW
    // wavefronts (a multiple of PF, at most ACH), so that a short seam list
    // costs each wavefront few dependent D-row round trips
    const int tot = a_hi - a_lo;
    const int chunk = min(ACH, max(PF, ((tot + NW - 1) / NW + PF - 1) / PF * PF));
    const int rounds = (tot + NW * chunk - 1) / (NW * chunk);  // the same for every wavefront
    int rows_done = 0;
    uint32_t acc[M][2];
#pragma unroll
    for (int v = 0; v < M; ++v) acc[v][0] = acc[v][1] = kInf2;
    for (int r = 0; r < rounds; ++r) {
        const int a0 = a_lo + (r * NW + w) * chunk;
        const int na = max(0, min(chunk, a_hi - a0));
        const int sr = tid < na ? seam[a0 + tid] : seam[0];
        uint32_t xv[M];
        bool fin = false;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            xv[v] = (v < m && tid < na) ? uint32_t(X[size_t(s0 + v) * pc + sr]) : uint32_t(kInf);
            fin |= xv[v] < kInf;
        }
        const unsigned long long bal = __ballot(fin);
        const int nk = __popcll(bal);
        const int nkp = (nk + PF - 1) / PF * PF;
        const int pos = __popcll(bal & ((1ull << tid) - 1ull));
        __syncthreads();  // previous round's readers are done with xs / roff
        if (fin) {
            roff[w][pos] = unsigned(q * pc + sr) * unsigned(ld);
#pragma unroll
            for (int v = 0; v < M; ++v) xs[w][pos][v] = qdup(xv[v]);
        }
        if (tid >= nk && tid < nkp) {
            roff[w][tid] = pad_off;
#pragma unroll
            for (int v = 0; v < M; ++v) xs[w][tid][v] = kInf2;
        }
        __syncthreads();
        rows_done += nk;
        if (!active) continue;
        // D rows double-buffered: the next PF rows are in flight while this
        // block's PF rows are folded in
        uint2 dv[PF], dn[PF];
        if (nkp > 0) {
#pragma unroll
            for (int u = 0; u < PF; ++u) dn[u] = *reinterpret_cast<const uint2*>(D + roff[w][u]);
        }
        for (int a = 0; a < nkp; a += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) dv[u] = dn[u];
            if (a + PF < nkp) {
#pragma unroll
                for (int u = 0; u < PF; ++u) dn[u] = *reinterpret_cast<const uint2*>(D + roff[w][a + PF + u]);
            }
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const uint32_t* xr = xs[w][a + u];
                // all saturating adds first, then the mins: independent
                // instructions back to back instead of add->min chains
                uint32_t t0[M], t1[M];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    t0[v] = qadd2(xr[v], dv[u].x);
                    t1[v] = qadd2(xr[v], dv[u].y);
                }
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    acc[v][0] = qmin2(acc[v][0], t0[v]);
                    acc[v][1] = qmin2(acc[v][1], t1[v]);
                }
            }
        }
    }
    // work counters: the wavefronts' figures meet in LDS and wavefront 0 adds
    // them to one of kWorkSlots counter pairs (one same-address atomic per
    // wavefront would serialise in the L2 and double the kernel's time)
    __shared__ unsigned long long wsum[NW][2];
    if (wstat && tid == 0) {
        const int nact = min(64, nlive - by * 64);
        const unsigned long long cols4 = 4ull * nact;
        unsigned long long bytes = first_item ? (unsigned long long)rows_done * cols4 * 2 : 0ull;
        if (w == 0) {
            if (by == 0) bytes += (unsigned long long)m * (a_hi - a_lo) * 2;
            bytes += (unsigned long long)m * cols4 * 2;
        }
        wsum[w][0] = bytes;
        wsum[w][1] = (unsigned long long)rows_done * cols4 * m;
    }
    // minimum over the wavefronts, in wavefront 0
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < M; ++v) red[w - 1][v][tid] = make_uint2(acc[v][0], acc[v][1]);
    }
    __syncthreads();
    if (w > 0) return;
    if (wstat && tid == 0) {
        const unsigned slot = (unsigned(bx) * 131u + unsigned(by) * 17u + unsigned(bz) * 7u) % unsigned(kWorkSlots);
        atomicAdd(&wstat[2 * slot], wsum[0][0] + wsum[1][0] + wsum[2][0] + wsum[3][0]);
        atomicAdd(&wstat[2 * slot + 1], wsum[0][1] + wsum[1][1] + wsum[2][1] + wsum[3][1]);
    }
#pragma unroll
    for (int j = 0; j < NW - 1; ++j)
#pragma unroll
        for (int v = 0; v < M; ++v) {
            const uint2 o = red[j][v][tid];
            acc[v][0] = qmin2(acc[v][0], o.x);
            acc[v][1] = qmin2(acc[v][1], o.y);
        }
    if (!active) return;
    if (ksplit > 1) {
        uint16_t* P = Zp + size_t(bz) * part_stride;
#pragma unroll
        for (int v = 0; v < M; ++v)
            if (v < m) *reinterpret_cast<uint2*>(P + size_t(s0 + v) * ldp + c0) = make_uint2(acc[v][0], acc[v][1]);
        return;
    }
#pragma unroll
    for (int v = 0; v < M; ++v) {
        if (v >= m) continue;
        const size_t sl = size_t(s0 + v);
        const uint16_t val[4] = {uint16_t(acc[v][0] & 0xffffu), uint16_t(acc[v][0] >> 16),
                                 uint16_t(acc[v][1] & 0xffffu), uint16_t(acc[v][1] >> 16)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (pos[k] >= 0) Z[sl * ldz + pos[k]] = val[k];
            if (ext[k] >= 0) Xn[sl * ne + ext[k]] = ownx[v][k] < val[k] ? ownx[v][k] : val[k];
        }
    }
}

bool launch_q_z(int M, int ld, int pc, const uint16_t* DU, size_t node_stride, const int* seam_list,
                const int* cols, int ngrp, const int* live_list, int col_tiles, int ldp, const int4* items,
                int n_items, int ksplit, const uint16_t* X, uint16_t* Zp,
                size_t part_stride, const int* colpos, const int* colext, const int* ext_child,
                const int* ext_cport, uint16_t* Z, int ldz, uint16_t* Xn, int ne, unsigned long long* wstat,
                hipStream_t st, KTimer kt) {
    if (n_items == 0 || col_tiles == 0) return false;
    const dim3 grid{unsigned(n_items), unsigned(col_tiles), unsigned(ksplit)};
#define HTG_QZ(MM)                                                                                                  \
    HTG_LAUNCH_T((k_q_z<MM, 8>), grid, dim3(256), 0, st, kt, ld, pc, DU, node_stride, seam_list, cols, ngrp,        \
                 live_list, ldp, items, ksplit, X, Zp, part_stride, colpos, colext, ext_child, ext_cport, Z, ldz,   \
                 Xn, ne, wstat)
    if (M <= 1) HTG_QZ(1);
    else if (M <= 2) HTG_QZ(2);
    else if (M <= 4) HTG_QZ(4);
    else if (M <= 8) HTG_QZ(8);
    else HTG_QZ(16);
#undef HTG_QZ
    return true;
}

// Fold of the row chunks, one launch with two index ranges per slot:
//  pos < nz: Z[slot][pos] = min_c Zp[c][slot][src[pos]], the layout of
//     contiguous per-child seam blocks then the other exterior ports; INF
//     where src[pos] < 0 (block padding) or where the column's group is not
//     live in the slot's node (no passable port: k_q_z skipped it);
//  pos = nz + e (below the root): lift to W's exterior port e
//     (src/minplus_oracle.cpp:1113-1154), Xnext[slot][e] = min(folded column
//     ext_col[e], X[slot][own port of e] when e is a port of the slot's child).
__device__ __forceinline__ uint16_t fold_col(int col, int s, int W, int ngrp, const uint8_t* __restrict__ live_map,
                                             const uint16_t* __restrict__ Zp, int ldp, int ksplit,
                                             size_t part_stride) {
    uint16_t v = kInf;
    if (col >= 0 && live_map[size_t(W) * ngrp + col / 4]) {
        const uint16_t* p = Zp + size_t(s) * ldp + col;
        for (int c = 0; c < ksplit; ++c) {
            const uint16_t x = p[size_t(c) * part_stride];
            v = x < v ? x : v;
        }
    }
    return v;
}

__global__ __launch_bounds__(256) void k_q_fold(int n, int nz, int ldz, const int* __restrict__ src, int ldp,
                                                int ksplit, const uint16_t* __restrict__ Zp, size_t part_stride,
                                                const int2* __restrict__ slot_rc, int side, int nt, int ngrp,
                                                const uint8_t* __restrict__ live_map, uint16_t* __restrict__ Z, int ne,
                                                const int* __restrict__ ext_col, const int* __restrict__ ext_child,
                                                const int* __restrict__ ext_cport, int pc, int cs, int g,
                                                const uint16_t* __restrict__ Xprev, uint16_t* __restrict__ Xnext) {
    const int per = nz + ne;
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)n * per) return;
    const int s = int(gid / per), pos = int(gid % per);
    const int2 rc = slot_rc[s];
    const int W = (rc.x / side) * nt + rc.y / side;
    if (pos < nz) {
        Z[size_t(s) * ldz + pos] = fold_col(src[pos], s, W, ngrp, live_map, Zp, ldp, ksplit, part_stride);
        return;
    }
    const int e = pos - nz;
    uint16_t v = fold_col(ext_col[e], s, W, ngrp, live_map, Zp, ldp, ksplit, part_stride);
    const int q = ((rc.x / cs) % g) * g + (rc.y / cs) % g;
    if (ext_child[e] == q) {
        const uint16_t d = Xprev[size_t(s) * pc + ext_cport[e]];
        v = d < v ? d : v;
    }
    Xnext[size_t(s) * ne + e] = v;
}

bool launch_q_fold(int n, int nz, int ldz, const int* src, int ldp, int ksplit, const uint16_t* Zp,
                   size_t part_stride, const int2* slot_rc, int side, int nt, int ngrp, const uint8_t* live_map,
                   uint16_t* Z, int ne, const int* ext_col, const int* ext_child, const int* ext_cport, int pc, int cs,
                   int g, const uint16_t* Xprev, uint16_t* Xnext, hipStream_t st, KTimer kt) {
    if (n == 0) return false;
    const long long total = (long long)n * (nz + ne);
    HTG_LAUNCH_T(k_q_fold, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, kt, n, nz, ldz,
                          src, ldp, ksplit, Zp, part_stride, slot_rc, side, nt, ngrp, live_map, Z, ne, ext_col,
                          ext_child, ext_cport, pc, cs, g, Xprev, Xnext);
    return true;
}

// Meet: m_ij = min_{b in seam(q_j)} Z_i[b] (+) Z_j[b].  Every candidate is the
// length of a real walk (i -> b -> j inside W), and Z_j[b] <= X_j[b] on q_j's
// seam ports (D_W[b][b] = 0), so the candidate is never above the
// reference's X_i-projection (+) X_j meet: the minimum over levels is exact.
// One launch serves every level: Z of each level stays in its own region.
// One 256-thread workgroup computes an 8 x 8 tile of (i, j) pairs of one
// (query, W, q2) group, 2 x 2 pairs per thread with 16 lanes per 2 x 2
// block on interleaved columns, streaming the seam block through LDS as
// packed pairs.
// Tiles are 8 ints (group, q2, i0, j0, j_end, level); zgeo holds per level
// (Z offset lo, hi, ldz, 0, zoff[g2], zlen[g2]).  A pair meets at several
// levels, so the result is merged with a 16-bit atomic minimum.
__device__ __forceinline__ void atomic_min_u16(uint16_t* o, uint16_t v) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(o);
    unsigned* w = reinterpret_cast<unsigned*>(a & ~uintptr_t(3));
    const int sh = (a & 2) ? 16 : 0;
    unsigned old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (((old >> sh) & 0xffffu) <= v) return;
        const unsigned nw = (old & ~(0xffffu << sh)) | (unsigned(v) << sh);
        const unsigned prev = atomicCAS(w, old, nw);
        if (prev == old) return;
        old = prev;
    }
}

__global__ __launch_bounds__(256) void k_q_meet(const int* __restrict__ tiles, const int* __restrict__ glist,
                                                const int* __restrict__ zgeo, int g2, const int* __restrict__ slot_i,
                                                const int* __restrict__ q_k, const long long* __restrict__ q_out,
                                                const uint16_t* __restrict__ Zall, uint16_t* __restrict__ out) {
    constexpr int CH = 256;      // packed pairs per round (512 columns)
    constexpr int LDR = CH + 8;  // row stride: rows two apart fall 16 banks apart
    __shared__ uint32_t A[8][LDR];
    __shared__ uint32_t Bm[8][LDR];
    // tile: group size, target child q2, first rows i0 / j0, j end, level,
    // the group's offset in glist, its query
    // logical tile: tiles of one group and target child share Z rows, so runs
    // of consecutive tiles go to one XCD
    const unsigned bt = xcd_logical<kXcdRun>(blockIdx.x, gridDim.x);
    const int4 t0 = *reinterpret_cast<const int4*>(tiles + size_t(bt) * 8);
    const int4 t1 = *reinterpret_cast<const int4*>(tiles + size_t(bt) * 8 + 4);
    const int n = t0.x, q2 = t0.y, i0 = t0.z, j0 = t0.w, jend = t1.x, lvl = t1.y, Q = t1.w;
    const int* lst = glist + t1.z;
    const int* zg = zgeo + size_t(lvl) * (4 + 2 * g2);
    const int tid = threadIdx.x;
    // 2 x 2 pairs per thread: block (bi, bj) = rows 2bi, 2bi+1 of the tile's
    // i side against rows 2bj, 2bj+1 of its j side, 16 lanes per block on
    // interleaved columns (four LDS reads feed four packed relaxations)
    const int l16 = tid & 15, blk = tid >> 4, bi = blk >> 2, bj = blk & 3;
    const int ni = min(8, n - i0), nj = min(8, jend - j0);
    // the result's address, resolved now (its dependent loads overlap the row
    // loads): lane l < 4 of a block writes pair (2bi + l / 2, 2bj + l % 2)
    const int pi = 2 * bi + (l16 >> 1), pj = 2 * bj + (l16 & 1);
    const bool writer = l16 < 4 && pi < ni && pj < nj;
    int si = 0, sj = 0;
    size_t oaddr = 0;
    if (writer) {
        si = lst[i0 + pi];
        sj = lst[j0 + pj];
        oaddr = size_t(q_out[Q]) + size_t(slot_i[si]) * q_k[Q] + slot_i[sj];
    }
    const uint16_t* Z = Zall + ((size_t(unsigned(zg[1])) << 32) | unsigned(zg[0]));
    const int ldz = zg[2];
    const int off = zg[4 + q2], len2 = zg[4 + g2 + q2] / 2;  // block length is even
    // staging: thread (row, l16) loads columns l16, l16 + 16, ... of one of
    // the 16 rows (8 of A, then 8 of B)
    const int row = tid >> 4;
    const bool is_a = row < 8;
    const int rr = row & 7;
    const int nrow = is_a ? ni : nj;
    const uint32_t* zr =
        rr < nrow ? reinterpret_cast<const uint32_t*>(Z + size_t(lst[(is_a ? i0 : j0) + rr]) * ldz + off) : nullptr;
    uint32_t* dst = is_a ? A[rr] : Bm[rr];
    uint32_t a00 = kInf2, a01 = kInf2, a10 = kInf2, a11 = kInf2;
    for (int c0 = 0; c0 < len2; c0 += CH) {
        const int cn = min(CH, len2 - c0);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < CH / 16; ++u) {
            const int cc = l16 + 16 * u;
            dst[cc] = (zr && cc < cn) ? zr[c0 + cc] : kInf2;
        }
        __syncthreads();
        const uint32_t* A0 = A[2 * bi];
        const uint32_t* A1 = A[2 * bi + 1];
        const uint32_t* B0 = Bm[2 * bj];
        const uint32_t* B1 = Bm[2 * bj + 1];
#pragma unroll 2
        for (int cc = l16; cc < cn; cc += 16) {
            const uint32_t x0 = A0[cc], x1 = A1[cc], y0 = B0[cc], y1 = B1[cc];
            a00 = qrelax2(a00, x0, y0);
            a01 = qrelax2(a01, x0, y1);
            a10 = qrelax2(a10, x1, y0);
            a11 = qrelax2(a11, x1, y1);
        }
    }
    // minimum over the block's 16 lanes (xor 1..8 stays inside the group)
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
        a00 = qmin2(a00, uint32_t(__shfl_xor(int(a00), m)));
        a01 = qmin2(a01, uint32_t(__shfl_xor(int(a01), m)));
        a10 = qmin2(a10, uint32_t(__shfl_xor(int(a10), m)));
        a11 = qmin2(a11, uint32_t(__shfl_xor(int(a11), m)));
    }
    if (!writer || si == sj) return;
    const uint32_t acc = l16 == 0 ? a00 : l16 == 1 ? a01 : l16 == 2 ? a10 : a11;
    const uint16_t best = min(uint16_t(acc & 0xffffu), uint16_t(acc >> 16));
    if (best >= kInf) return;
    atomic_min_u16(out + oaddr, best);
}

bool launch_q_meet(const int* tiles, int n_tiles, const int* glist, const int* zgeo, int g2, const int* slot_i,
                   const int* q_k, const long long* q_out, const uint16_t* Z, uint16_t* out, hipStream_t st,
                   KTimer kt) {
    if (n_tiles == 0) return false;
    HTG_LAUNCH_T(k_q_meet, dim3(unsigned(n_tiles)), dim3(256), 0, st, kt, tiles, glist, zgeo, g2, slot_i, q_k,
                 q_out, Z, out);
    return true;
}

// ---------------------------------------------------------------------------
// liftPoint (src/minplus_oracle.cpp:731-823): distances from one cell to the
// ports of its leaf and of every ancestor's exterior.
//
// Level 0: X0[i] = D_leaf[u][cell of port i] (one row of the leaf table).
__global__ __launch_bounds__(256) void k_lift_leaf(const uint16_t* __restrict__ row, int b,
                                                   uint16_t* __restrict__ X0) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 4 * b - 4) return;
    int r, c;
    port_rc(i, b, r, c);
    X0[i] = row[r * b + c];
}

// Level l: the rows of child q in the node's D_U, against the child's lift.
// Workgroup (x, y) covers 512 columns (two per lane, u32 loads) and the rows
// [y * rows_per_chunk, ...) of the child (at most kLiftRows):
//     part[y][j] = min_p X[p] (+) D[row0 + p][j].
// The chunk's X values sit in LDS; rows are read 16 at a time so that 16
// loads per lane are in flight, and rows with X[p] = INF are not read (the
// reference skips inactive ports the same way).
constexpr int kLiftRows = kLiftRowsPerChunk;
__global__ __launch_bounds__(256) void k_lift_rows(const uint16_t* __restrict__ D, int ld, int row0, int pc,
                                                   int rows_per_chunk, const uint16_t* __restrict__ X,
                                                   uint16_t* __restrict__ part) {
    __shared__ uint32_t xs[kLiftRows];
    const int p0 = blockIdx.y * rows_per_chunk;
    const int n = min(pc - p0, rows_per_chunk);
    if (int(threadIdx.x) < n) xs[threadIdx.x] = qdup(X[p0 + threadIdx.x]);
    __syncthreads();
    const int j2 = blockIdx.x * 256 + threadIdx.x;  // column pair
    if (2 * j2 >= ld) return;
    const size_t ldw = size_t(ld) / 2;
    const uint32_t* Dp = reinterpret_cast<const uint32_t*>(D + size_t(row0 + p0) * ld) + j2;
    uint32_t acc = kInf2;
    for (int p = 0; p < n; p += 16) {
        uint32_t x[16], d[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = p + k < n ? xs[p + k] : kInf2;
#pragma unroll
        for (int k = 0; k < 16; ++k) d[k] = x[k] != kInf2 ? Dp[size_t(p + k) * ldw] : kInf2;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc = qrelax2(acc, x[k], d[k]);
    }
    reinterpret_cast<uint32_t*>(part + size_t(blockIdx.y) * ld)[j2] = acc;
}

// Xo[e] = min over row chunks of part[chunk][column of exterior port e]
// (eight independent loads in flight per lane).
__global__ __launch_bounds__(256) void k_lift_ext(const uint16_t* __restrict__ part, int ld, int nchunks,
                                                  const int* __restrict__ ext_child,
                                                  const int* __restrict__ ext_cport, int pc, int ne,
                                                  uint16_t* __restrict__ Xo) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    const uint16_t* col = part + ext_child[e] * pc + ext_cport[e];
    uint16_t v = kInf;
    int ch = 0;
    for (; ch + 8 <= nchunks; ch += 8) {
        uint16_t t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = col[size_t(ch + k) * ld];
#pragma unroll
        for (int k = 0; k < 8; ++k) v = min(v, t[k]);
    }
    for (; ch < nchunks; ++ch) v = min(v, col[size_t(ch) * ld]);
    Xo[e] = v;
}

void launch_lift_leaf(const uint16_t* row, int b, uint16_t* X0, hipStream_t st) {
    hipLaunchKernelGGL(k_lift_leaf, dim3(unsigned((4 * b - 4 + 255) / 256)), dim3(256), 0, st, row, b, X0);
}

void launch_lift_level(const uint16_t* D, int ld, int row0, int pc, int rows_per_chunk, int nchunks,
                       const uint16_t* X, uint16_t* part, const int* ext_child, const int* ext_cport, int ne,
                       uint16_t* Xo, hipStream_t st) {
    hipLaunchKernelGGL(k_lift_rows, dim3(unsigned((ld / 2 + 255) / 256), unsigned(nchunks)), dim3(256), 0, st, D,
                       ld, row0, pc, rows_per_chunk, X, part);
    hipLaunchKernelGGL(k_lift_ext, dim3(unsigned((ne + 255) / 256)), dim3(256), 0, st, part, ld, nchunks, ext_child,
                       ext_cport, pc, ne, Xo);
}

}  // namespace htg
