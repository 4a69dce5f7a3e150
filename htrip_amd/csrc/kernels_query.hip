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

// Global-memory pointers.  The units read their pointers from argument
// blocks, which in the persistent launch live in device memory; typed as
// global, every access is a global_* instruction with a scalar base instead
// of a flat one with a 64-bit vector address.  Vector loads and stores use
// the compiler's own vector types (HIP's vector classes do not bind to a
// global-qualified reference).
#define HTG_G(T) __attribute__((address_space(1))) T*
template <typename T>
__device__ __forceinline__ HTG_G(T) gp(T* p) {
    return (HTG_G(T))p;
}
// The same memory viewed as elements of type V.
template <typename V, typename T>
__device__ __forceinline__ HTG_G(V) gpv(T* p) {
    return (HTG_G(V))p;
}
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Front (src/minplus_oracle.cpp:940-965, same-leaf pairs :1027-1048, level-0
// lift :1000-1019), unit blk of three ranges:
//  [0, nq * ib): output init of query blk / ib: INF, 0 on the diagonal of
//     valid viewpoints, leaf D for same-leaf pairs;
//  next leaf_blocks: level-0 lift, X0[slot][p] = leaf D row of the slot's
//     cell at leaf port p;
//  then INF fills of the buffers whose live-only writers leave gaps: fill
//     range r = (start lo, start hi, n8, first unit), n8 units of 8 u16.
// Unit 0 also clears the projection's work counters.
__device__ __forceinline__ void front_unit(const QFrontArgs& a, int blk) {
    const int b = a.b, n = b * b;
    const auto leafD = gp(a.leafD);
    if (a.wstat && blk == 0 && threadIdx.x < 2 * kWorkSlots) gp(a.wstat)[threadIdx.x] = 0;
    if (blk < a.nq * a.ib) {
        const int Q = blk / a.ib;
        const int k = gp(a.q_k)[Q];
        const long long idx = (long long)(blk % a.ib) * 256 + threadIdx.x;
        if (idx >= (long long)k * k) return;
        const int i = int(idx / k), j = int(idx % k);
        const int vp0 = gp(a.q_vp)[Q];
        const int vi = vp0 + i, vj = vp0 + j;
        const auto valid = gp(a.vp_valid);
        const auto vp_rc = gpv<const i32x2>(a.vp_rc);
        uint16_t val = kInf;
        if (valid[vi] && valid[vj]) {
            if (i == j) {
                val = 0;
            } else {
                const i32x2 p = vp_rc[vi], c = vp_rc[vj];
                if (p.x / b == c.x / b && p.y / b == c.y / b) {
                    const size_t leaf = size_t(p.x / b) * a.nl + p.y / b;
                    val = leafD[leaf * n * n + size_t((p.x % b) * b + p.y % b) * n + (c.x % b) * b + c.y % b];
                }
            }
        }
        gp(a.out)[gp(a.q_out)[Q] + idx] = val;
        return;
    }
    const int lb = blk - a.nq * a.ib;
    if (lb < a.leaf_blocks) {
        const int pl = 4 * b - 4;
        const long long gid = (long long)lb * 256 + threadIdx.x;
        if (gid >= (long long)a.n_slots * pl) return;
        const int s = int(gid / pl), p = int(gid % pl);
        const i32x2 rc = gpv<const i32x2>(a.slot_rc)[s];
        int pr, pcc;
        port_rc(p, b, pr, pcc);
        const size_t leaf = size_t(rc.x / b) * a.nl + rc.y / b;
        gp(a.X0)[gid] = leafD[leaf * n * n + size_t((rc.x % b) * b + rc.y % b) * n + pr * b + pcc];
        return;
    }
    // the fill range of this unit: ranges are in order of their first unit,
    // so r counts the ranges after the first that start at or before fb (one
    // load per lane and a ballot instead of a walk of dependent loads)
    const auto fills = gpv<const i32x4>(a.fills);
    const int fb = lb - a.leaf_blocks;
    const int lane = int(threadIdx.x & 63u);
    const bool starts_before = lane >= 1 && lane < a.n_fill && fills[lane].w <= fb;
    const int r = __popcll(__ballot(starts_before));
    const i32x4 f = fills[r];
    const long long u = (long long)(fb - f.w) * 256 + threadIdx.x;
    if (u >= f.z) return;
    const size_t start = (size_t(unsigned(f.y)) << 32) | unsigned(f.x);
    gpv<u32x4>(gp(a.fill_base) + start)[u] = u32x4{kInf2, kInf2, kInf2, kInf2};
}

// Projection Z (see file comment), split over the seam rows: unit (work item
// bx, column tile by, row chunk bz).  Four wavefronts share the columns and
// take interleaved rounds of rows.  Only live column groups are computed:
// live_list[W] holds the groups of 4 projection columns with at least one
// passable port (the others are INF, written so by the front's fills or the
// fold), 64 groups = 256 columns per unit.  Each wavefront keeps M x 4
// packed accumulators and reads one 8-byte slice of a D_W row per seam row.
// Seam rows whose X is INF for all M viewpoints add nothing and are dropped:
// each round stages 64 rows, compacts the finite ones (ballot + prefix
// count) into LDS with their X values duplicated into both u16 halves, and
// pads the count to PF rows with INF.  Partial minima go to Zp[chunk] and
// the fold combines the chunks; with one chunk (ksplit == 1) the unit folds
// in registers instead: column c goes to Z position colpos[c] and, below the
// root, lifts to exterior port colext[c] of W,
// Xn[slot][e] = min(Z, X[slot][own port of e] when e is a port of q)
// (src/minplus_oracle.cpp:1113-1154).  wstat (nullable) accumulates the
// work done: [0] bytes (D slices of distinct (W, q) items, X reads, Zp
// writes, each counted once), [1] min-plus updates on finite rows.
constexpr int kZRows = 64;  // seam rows staged per wavefront and round
constexpr int kZWaves = 4;  // wavefronts per workgroup: same columns, interleaved row rounds
template <int M>
struct ZLds {
    uint32_t xs[kZWaves][kZRows][M];  // per wavefront, row-major: one row's M values contiguous
    unsigned roff[kZWaves][kZRows];
    u32x2 red[kZWaves - 1][M][64];  // accumulators of wavefronts 1..3 for the final minimum
    unsigned long long wsum[kZWaves][2];
};

template <int M, int PF>
__device__ __forceinline__ void z_unit(const QZArgs& a, int bx, int by, int bz, ZLds<M>& sh) {
    // Padding rows carry INF in xs: min(acc, INF (+) d) leaves acc unchanged,
    // so the loop runs branch-free over whole PF blocks.
    constexpr int ACH = kZRows, NW = kZWaves;
    const int ld = a.ld, pc = a.pc, ksplit = a.ksplit;
    const auto items = gpv<const i32x4>(a.items);
    const i32x4 it = items[2 * bx], it2 = items[2 * bx + 1];
    const int W = it.x, q = it.y & 0xffff, s0 = it.z, m = it.w;
    const bool first_item = (it.y >> 16) & 1;  // first item of its (W, q) run: counts the D bytes once
    const int tid = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nlive = it2.x;  // live column groups of W
    const int gi = by * 64 + tid;
    if (by * 64 >= nlive) return;  // whole unit past W's live groups
    const bool active = gi < nlive;
    const int c0 = active ? gp(a.live_list)[size_t(W) * a.ngrp + gi] * 4 : 0;
    // Projection columns come in aligned groups of 4 consecutive V_U ports,
    // so one 8-byte load per seam row feeds the lane's 4 columns.
    const int base = gp(a.cols)[c0];
    const auto D = gp(a.DU) + size_t(W) * a.node_stride + base;
    const auto X = gp(a.X);
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
        const i32x4 cp = *gpv<const i32x4>(gp(a.colpos) + c0);
        const i32x4 ce = *gpv<const i32x4>(gp(a.colext) + c0);
        pos[0] = cp.x, pos[1] = cp.y, pos[2] = cp.z, pos[3] = cp.w;
        ext[0] = ce.x, ext[1] = ce.y, ext[2] = ce.z, ext[3] = ce.w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int own = (ext[k] >= 0 && gp(a.ext_child)[ext[k]] == q) ? gp(a.ext_cport)[ext[k]] : -1;
            if (own < 0) continue;
#pragma unroll
            for (int v = 0; v < M; ++v)
                if (v < m) ownx[v][k] = X[size_t(s0 + v) * pc + own];
        }
    }
    const auto seam = gp(a.seam_list) + it2.y;  // child q's seam ports
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
        const int pos_r = __popcll(bal & ((1ull << tid) - 1ull));
        __syncthreads();  // previous round's readers are done with xs / roff
        if (fin) {
            sh.roff[w][pos_r] = unsigned(q * pc + sr) * unsigned(ld);
#pragma unroll
            for (int v = 0; v < M; ++v) sh.xs[w][pos_r][v] = qdup(xv[v]);
        }
        if (tid >= nk && tid < nkp) {
            sh.roff[w][tid] = pad_off;
#pragma unroll
            for (int v = 0; v < M; ++v) sh.xs[w][tid][v] = kInf2;
        }
        __syncthreads();
        rows_done += nk;
        if (!active) continue;
        // D rows double-buffered: the next PF rows are in flight while this
        // block's PF rows are folded in
        u32x2 dv[PF], dn[PF];
        if (nkp > 0) {
#pragma unroll
            for (int u = 0; u < PF; ++u) dn[u] = *gpv<const u32x2>(D + sh.roff[w][u]);
        }
        for (int k0 = 0; k0 < nkp; k0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) dv[u] = dn[u];
            if (k0 + PF < nkp) {
#pragma unroll
                for (int u = 0; u < PF; ++u)
                    dn[u] = *gpv<const u32x2>(D + sh.roff[w][k0 + PF + u]);
            }
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const auto xr = sh.xs[w][k0 + u];
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
    if (a.wstat && tid == 0) {
        const int nact = min(64, nlive - by * 64);
        const unsigned long long cols4 = 4ull * nact;
        unsigned long long bytes = first_item ? (unsigned long long)rows_done * cols4 * 2 : 0ull;
        if (w == 0) {
            if (by == 0) bytes += (unsigned long long)m * (a_hi - a_lo) * 2;
            bytes += (unsigned long long)m * cols4 * 2;
        }
        sh.wsum[w][0] = bytes;
        sh.wsum[w][1] = (unsigned long long)rows_done * cols4 * m;
    }
    // minimum over the wavefronts, in wavefront 0
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < M; ++v) sh.red[w - 1][v][tid] = u32x2{acc[v][0], acc[v][1]};
    }
    __syncthreads();
    if (w > 0) return;
    if (a.wstat && tid == 0) {
        const unsigned slot = (unsigned(bx) * 131u + unsigned(by) * 17u + unsigned(bz) * 7u) % unsigned(kWorkSlots);
        const auto ws = gp(a.wstat);
        unsigned long long by_sum = 0, up_sum = 0;
#pragma unroll
        for (int j = 0; j < NW; ++j) by_sum += sh.wsum[j][0], up_sum += sh.wsum[j][1];
        __hip_atomic_fetch_add(ws + 2 * slot, by_sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(ws + 2 * slot + 1, up_sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int j = 0; j < NW - 1; ++j)
#pragma unroll
        for (int v = 0; v < M; ++v) {
            const u32x2 o = sh.red[j][v][tid];
            acc[v][0] = qmin2(acc[v][0], o.x);
            acc[v][1] = qmin2(acc[v][1], o.y);
        }
    if (!active) return;
    if (ksplit > 1) {
        const auto P = gp(a.Zp) + size_t(bz) * a.part_stride;
#pragma unroll
        for (int v = 0; v < M; ++v)
            if (v < m)
                *gpv<u32x2>(P + size_t(s0 + v) * a.ldp + c0) = u32x2{acc[v][0], acc[v][1]};
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
            if (pos[k] >= 0) gp(a.Z)[sl * a.ldz + pos[k]] = val[k];
            if (ext[k] >= 0) gp(a.Xn)[sl * a.ne + ext[k]] = ownx[v][k] < val[k] ? ownx[v][k] : val[k];
        }
    }
}

// Fold of the row chunks, unit blk of two index ranges per slot:
//  pos < nz: Z[slot][pos] = min_c Zp[c][slot][zsrc[pos]], the layout of
//     contiguous per-child seam blocks; INF
//     where zsrc[pos] < 0 (block padding) or where the column's group is not
//     live in the slot's node (no passable port: the projection skipped it);
//  pos = nz + e (below the root): lift to W's exterior port e
//     (src/minplus_oracle.cpp:1113-1154), Xn[slot][e] = min(folded column
//     ext_col[e], X[slot][own port of e] when e is a port of the slot's child).
__device__ __forceinline__ uint16_t fold_col(const QZArgs& a, int col, int s, int W) {
    uint16_t v = kInf;
    if (col >= 0 && gp(a.live_map)[size_t(W) * a.ngrp + col / 4]) {
        const auto p = gp(a.Zp) + size_t(s) * a.ldp + col;
        for (int c = 0; c < a.ksplit; ++c) {
            const uint16_t x = p[size_t(c) * a.part_stride];
            v = x < v ? x : v;
        }
    }
    return v;
}

__device__ __forceinline__ void fold_unit(const QZArgs& a, int blk) {
    const int per = a.nz + a.ne;
    const long long gid = (long long)blk * 256 + threadIdx.x;
    if (gid >= (long long)a.n_slots * per) return;
    const int s = int(gid / per), pos = int(gid % per);
    const i32x2 rc = gpv<const i32x2>(a.slot_rc)[s];
    const int W = (rc.x / a.side) * a.nt + rc.y / a.side;
    if (pos < a.nz) {
        gp(a.Z)[size_t(s) * a.ldz + pos] = fold_col(a, gp(a.zsrc)[pos], s, W);
        return;
    }
    const int e = pos - a.nz;
    uint16_t v = fold_col(a, gp(a.ext_col)[e], s, W);
    const int q = ((rc.x / a.cs) % a.g) * a.g + (rc.y / a.cs) % a.g;
    if (gp(a.ext_child)[e] == q) {
        const uint16_t d = gp(a.X)[size_t(s) * a.pc + gp(a.ext_cport)[e]];
        v = d < v ? d : v;
    }
    gp(a.Xn)[size_t(s) * a.ne + e] = v;
}

// Meet: m_ij = min_{b in seam(q_j)} Z_i[b] (+) Z_j[b].  Every candidate is the
// length of a real walk (i -> b -> j inside W), and Z_j[b] <= X_j[b] on q_j's
// seam ports (D_W[b][b] = 0), so the candidate is never above the
// reference's X_i-projection (+) X_j meet: the minimum over levels is exact.
// Z of each level stays in its own region.  One 256-thread unit computes a
// tile of (i, j) pairs of one (query, W, q2) group, streaming the seam block
// through LDS as packed pairs: 16 blocks of PB x PB pairs, 16 lanes per block
// on interleaved columns, so that 2 PB LDS reads feed PB^2 packed
// relaxations.  Groups of at least 16 viewpoints on both sides take 16 x 16
// tiles (PB = 4), smaller ones 8 x 8 (PB = 2).  A pair meets at several
// levels, so the result is merged with a 16-bit atomic minimum.
__device__ __forceinline__ void atomic_min_u16(HTG_G(uint16_t) o, uint16_t v) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(o);
    HTG_G(unsigned) const w = (HTG_G(unsigned))(a & ~uintptr_t(3));
    const int sh = (a & 2) ? 16 : 0;
    unsigned old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (((old >> sh) & 0xffffu) <= v) return;
        const unsigned nw = (old & ~(0xffffu << sh)) | (unsigned(v) << sh);
        if (__hip_atomic_compare_exchange_strong(w, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// Rows per tile side 4 PB; packed pairs per staging round CH (the LDS holds
// 8 PB rows of CH + 16 / PB words: rows PB apart fall 16 banks apart).
template <int PB>
struct MeetGeo {
    static constexpr int R = 4 * PB;
    static constexpr int CH = 512 / PB;
    static constexpr int LD = CH + 16 / PB;
};
struct MeetLds {
    uint32_t w[2 * MeetGeo<2>::R * MeetGeo<2>::LD > 2 * MeetGeo<4>::R * MeetGeo<4>::LD
                   ? 2 * MeetGeo<2>::R * MeetGeo<2>::LD
                   : 2 * MeetGeo<4>::R * MeetGeo<4>::LD];
};

template <int PB>
__device__ __forceinline__ void meet_tile(const QMeetArgs& a, int n, int q2, int i0, int j0, int jend, int lvl,
                                          int goff, int Q, MeetLds& sh) {
    constexpr int R = MeetGeo<PB>::R, CH = MeetGeo<PB>::CH, LD = MeetGeo<PB>::LD;
    uint32_t* const A = sh.w;
    uint32_t* const Bm = sh.w + R * LD;
    const auto lst = gp(a.glist) + goff;
    const auto zg = gp(a.zgeo) + size_t(lvl) * (4 + 2 * a.g2);
    const int tid = threadIdx.x;
    // block (bi, bj) = rows PB bi .. of the tile's i side against rows PB bj
    // .. of its j side; lane l16 of the block takes columns l16, l16 + 16, ...
    const int l16 = tid & 15, blk = tid >> 4, bi = blk >> 2, bj = blk & 3;
    const int ni = min(R, n - i0), nj = min(R, jend - j0);
    // the result's address, resolved now (its dependent loads overlap the row
    // loads): lane l < PB^2 of a block writes pair (PB bi + l / PB, PB bj + l % PB)
    const int pi = PB * bi + l16 / PB, pj = PB * bj + l16 % PB;
    const bool writer = l16 < PB * PB && pi < ni && pj < nj;
    int si = 0, sj = 0;
    size_t oaddr = 0;
    if (writer) {
        si = lst[i0 + pi];
        sj = lst[j0 + pj];
        const auto slot_i = gp(a.slot_i);
        oaddr = size_t(gp(a.q_out)[Q]) + size_t(slot_i[si]) * gp(a.q_k)[Q] + slot_i[sj];
    }
    const auto Z = gp(a.Z) + ((size_t(unsigned(zg[1])) << 32) | unsigned(zg[0]));
    const int ldz = zg[2];
    const int off = zg[4 + q2], len2 = zg[4 + a.g2 + q2] / 2;  // block length is even
    // staging: in pass p, thread (row, l16) loads columns l16, l16 + 16, ...
    // of staging row 16 p + row (rows 0 .. R-1 are the i side, then the j side)
    constexpr int PASSES = 2 * R / 16;
    HTG_G(const uint32_t) zr[PASSES];
    uint32_t* dst[PASSES];
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
        const int row = 16 * p + (tid >> 4);
        const bool is_a = row < R;
        const int rr = is_a ? row : row - R;
        zr[p] = rr < (is_a ? ni : nj)
                    ? gpv<const uint32_t>(Z + size_t(lst[(is_a ? i0 : j0) + rr]) * ldz + off)
                    : nullptr;
        dst[p] = (is_a ? A : Bm) + rr * LD;
    }
    uint32_t acc[PB][PB];
#pragma unroll
    for (int x = 0; x < PB; ++x)
#pragma unroll
        for (int y = 0; y < PB; ++y) acc[x][y] = kInf2;
    // the staging values of the next round are loaded into registers while
    // this round's relaxations run
    uint32_t nxt[PASSES][CH / 16];
    auto load_round = [&](int c0) {
        const int cn = min(CH, len2 - c0);
#pragma unroll
        for (int p = 0; p < PASSES; ++p)
#pragma unroll
            for (int u = 0; u < CH / 16; ++u) {
                const int cc = l16 + 16 * u;
                nxt[p][u] = (zr[p] && cc < cn) ? zr[p][c0 + cc] : kInf2;
            }
    };
    if (len2 > 0) load_round(0);
    for (int c0 = 0; c0 < len2; c0 += CH) {
        const int cn = min(CH, len2 - c0);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PASSES; ++p)
#pragma unroll
            for (int u = 0; u < CH / 16; ++u) dst[p][l16 + 16 * u] = nxt[p][u];
        __syncthreads();
        if (c0 + CH < len2) load_round(c0 + CH);
#pragma unroll 2
        for (int cc = l16; cc < cn; cc += 16) {
            uint32_t x[PB], y[PB];
#pragma unroll
            for (int k = 0; k < PB; ++k) {
                x[k] = A[(PB * bi + k) * LD + cc];
                y[k] = Bm[(PB * bj + k) * LD + cc];
            }
#pragma unroll
            for (int u = 0; u < PB; ++u)
#pragma unroll
                for (int v = 0; v < PB; ++v) acc[u][v] = qrelax2(acc[u][v], x[u], y[v]);
        }
    }
    // minimum over the block's 16 lanes (xor 1..8 stays inside the group)
#pragma unroll
    for (int m = 1; m < 16; m <<= 1)
#pragma unroll
        for (int u = 0; u < PB; ++u)
#pragma unroll
            for (int v = 0; v < PB; ++v) acc[u][v] = qmin2(acc[u][v], uint32_t(__shfl_xor(int(acc[u][v]), m)));
    if (!writer || si == sj) return;
    uint32_t mine = kInf2;
#pragma unroll
    for (int u = 0; u < PB; ++u)
#pragma unroll
        for (int v = 0; v < PB; ++v)
            if (l16 == u * PB + v) mine = acc[u][v];
    const uint16_t best = min(uint16_t(mine & 0xffffu), uint16_t(mine >> 16));
    if (best >= kInf) return;
    atomic_min_u16(gp(a.out) + oaddr, best);
}

// Tile bt: group size, target child q2, first rows i0 / j0, j end, level (bit
// 16: a 16 x 16 tile), the group's offset in glist, its query.
__device__ __forceinline__ void meet_unit(const QMeetArgs& a, unsigned bt, MeetLds& sh) {
    const auto tl = gpv<const i32x4>(gp(a.tiles) + size_t(bt) * 8);
    const i32x4 t0 = tl[0], t1 = tl[1];
    if ((t1.y >> 16) & 1)
        meet_tile<4>(a, t0.x, t0.y, t0.z, t0.w, t1.x, t1.y & 0xffff, t1.z, t1.w, sh);
    else
        meet_tile<2>(a, t0.x, t0.y, t0.z, t0.w, t1.x, t1.y & 0xffff, t1.z, t1.w, sh);
}

// ---- one launch per phase ---------------------------------------------------
__global__ __launch_bounds__(256) void k_q_front(QFrontArgs a) { front_unit(a, int(blockIdx.x)); }

// Units in an XCD-aware order: items of one (W, q) read the same D rows, so
// runs of consecutive items go to one XCD.
template <int M, int PF>
__global__ __launch_bounds__(256) void k_q_z(QZArgs a) {
    __shared__ ZLds<M> sh;
    const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned L = xcd_logical<kXcdRun>(lin, gridDim.x * gridDim.y * gridDim.z);
    z_unit<M, PF>(a, int(L % gridDim.x), int((L / gridDim.x) % gridDim.y), int(L / (gridDim.x * gridDim.y)), sh);
}

__global__ __launch_bounds__(256) void k_q_fold(QZArgs a) { fold_unit(a, int(blockIdx.x)); }

// Tiles of one group and target child share Z rows: runs of consecutive
// tiles go to one XCD.
__global__ __launch_bounds__(256) void k_q_meet(QMeetArgs a) {
    __shared__ MeetLds sh;
    meet_unit(a, xcd_logical<kXcdRun>(blockIdx.x, gridDim.x), sh);
}

bool launch_q_front(const QFrontArgs& a, hipStream_t st) {
    // the unit's fill range is found with one 64-lane ballot over the ranges
    if (a.n_fill > 64 || a.n_fill < 0) return false;
    const long long blocks = (long long)a.nq * a.ib + a.leaf_blocks + a.fill_blocks;
    if (blocks == 0) return true;
    hipLaunchKernelGGL(k_q_front, dim3(unsigned(blocks)), dim3(256), 0, st, a);
    return true;
}

bool launch_q_z(const QZArgs& a, hipStream_t st, KTimer kt) {
    if (a.n_items == 0 || a.col_tiles == 0) return false;
    const dim3 grid{unsigned(a.n_items), unsigned(a.col_tiles), unsigned(a.ksplit)};
    if (a.M <= 1) HTG_LAUNCH_T((k_q_z<1, 8>), grid, dim3(256), 0, st, kt, a);
    else if (a.M <= 2) HTG_LAUNCH_T((k_q_z<2, 8>), grid, dim3(256), 0, st, kt, a);
    else if (a.M <= 4) HTG_LAUNCH_T((k_q_z<4, 8>), grid, dim3(256), 0, st, kt, a);
    else if (a.M <= 8) HTG_LAUNCH_T((k_q_z<8, 8>), grid, dim3(256), 0, st, kt, a);
    else HTG_LAUNCH_T((k_q_z<16, 8>), grid, dim3(256), 0, st, kt, a);
    return true;
}

bool launch_q_fold(const QZArgs& a, hipStream_t st, KTimer kt) {
    const long long total = (long long)a.n_slots * (a.nz + a.ne);
    if (total == 0) return false;
    HTG_LAUNCH_T(k_q_fold, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, kt, a);
    return true;
}

bool launch_q_meet(const QMeetArgs& a, hipStream_t st, KTimer kt) {
    if (a.n_tiles == 0) return false;
    HTG_LAUNCH_T(k_q_meet, dim3(unsigned(a.n_tiles)), dim3(256), 0, st, kt, a);
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
