// gfx950 kernels for index construction and incremental update.
//
// Arithmetic is the reference's saturating u16 tropical semiring
// (include/htrip/types.hpp:66-70), evaluated two lanes at a time with
// v_pk_add_u16 (clamp) + v_pk_min_u16.  Every stored value is <= INF = 60000,
// so min(x, sat16(a + b)) == min(x, add_dist(a, b)) exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hip/hip_ext.h>

#include "geometry.hpp"
#include "kernels.hpp"

namespace htg {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t relax2(uint32_t acc, uint32_t a, uint32_t b) {
    us2 x = __builtin_bit_cast(us2, a), y = __builtin_bit_cast(us2, b), z = __builtin_bit_cast(us2, acc);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(z, __builtin_elementwise_add_sat(x, y)));
}
__device__ __forceinline__ uint32_t add2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b)));
}
__device__ __forceinline__ uint32_t min2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b)));
}
__device__ __forceinline__ uint32_t dup16(uint32_t v) { return (v & 0xffffu) | (v << 16); }
__device__ __forceinline__ uint16_t add1(uint16_t a, uint16_t b) {
    uint32_t s = uint32_t(a) + b;
    return s >= kInf ? kInf : uint16_t(s);
}

// Value of the previous / next lane of the wavefront through a DPP wavefront
// shift (wave_shr:1 / wave_shl:1, bound_ctrl: 0 past the ends): one VALU move
// instead of an LDS permute round trip.  Lanes at source-group boundaries are
// masked by the caller.
__device__ __forceinline__ uint32_t lane_prev(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ uint32_t lane_next(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ uint64_t lane_prev(uint64_t v) {
    return uint64_t(lane_prev(uint32_t(v))) | (uint64_t(lane_prev(uint32_t(v >> 32))) << 32);
}
__device__ __forceinline__ uint64_t lane_next(uint64_t v) {
    return uint64_t(lane_next(uint32_t(v))) | (uint64_t(lane_next(uint32_t(v >> 32))) << 32);
}

// ---------------------------------------------------------------------------
// Leaf BFS core (src/leaf_tile.cpp:56-153): one wavefront runs 64/LANES BFS
// sources at once; lane r of a source group owns tile row r as a bitmask and
// a BFS step is a shift/or of neighbour rows (bit-parallel frontier
// expansion).  Distances from the workgroup's SWG sources land in LDS
// (dist[slot][cell], INF where unreached); the caller streams them out.
template <int LANES, typename RowT>
__device__ __forceinline__ void leaf_bfs(const uint8_t* __restrict__ pass, int S, int b, int R0, int C0, bool valid,
                                         int src, uint16_t* dist) {
    constexpr int SPW = 64 / LANES;
    constexpr int SWG = 4 * SPW;
    const int n = b * b;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = lane / LANES, r = lane % LANES;
    const int slot = wave * SPW + grp;
    if ((SWG * n) % 8 == 0) {
        uint4* d128 = reinterpret_cast<uint4*>(dist);
        for (int i = tid; i < SWG * n / 8; i += 256) d128[i] = make_uint4(kInf2, kInf2, kInf2, kInf2);
    } else {
        uint32_t* d32 = reinterpret_cast<uint32_t*>(dist);
        for (int i = tid; i < SWG * n / 2; i += 256) d32[i] = kInf2;
    }
    // The tile's passability rows as bit masks, each row read once per
    // workgroup with word loads (every source group of the workgroup shares
    // the tile): row loads one byte per lane touched a cache line per lane
    // and per byte, and bounded the whole kernel.
    __shared__ RowT rowmask[64];
    if (tid < b) {
        const uint8_t* p = pass + size_t(R0 + tid) * S + C0;
        RowT m = 0;
        if ((b & 3) == 0) {  // 4-aligned: C0 and S are multiples of b
            const uint32_t* p4 = reinterpret_cast<const uint32_t*>(p);
            for (int w = 0; w < b / 4; ++w) {
                const uint32_t x = p4[w];
                // 0x80 in each nonzero byte, then the four flags gathered
                // into bits 28..31 by one multiply
                const uint32_t nz = (((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u;
                m |= RowT(((nz >> 7) * 0x10204080u) >> 28) << (4 * w);
            }
        } else {
            for (int c = 0; c < b; ++c) m |= RowT(p[c] != 0) << c;
        }
        rowmask[tid] = m;
    }
    __syncthreads();
    const RowT prow = r < b ? rowmask[r] : RowT(0);
    const int sr = valid ? src / b : 0, sc = valid ? src % b : 0;
    const RowT srow = __shfl(prow, sr, LANES);
    RowT front = (valid && r == sr && ((srow >> sc) & 1)) ? (RowT(1) << sc) : RowT(0);
    RowT vis = front;
    if (valid && r == 0) dist[slot * n + src] = 0;  // D[i][i] = 0 for every cell
    uint16_t d = 0;
    while (__ballot(front != 0)) {
        ++d;
        RowT up = lane_prev(front);  // row r - 1 of the same source (r > 0)
        RowT dn = lane_next(front);  // row r + 1 (r < LANES - 1)
        if (r == 0) up = 0;
        if (r == LANES - 1) dn = 0;
        RowT nb = ((front << 1) | (front >> 1) | up | dn) & prow & ~vis;
        vis |= nb;
        front = nb;
        uint16_t* drow = dist + slot * n + r * b;
        while (nb) {
            int c = sizeof(RowT) == 8 ? __builtin_ctzll(uint64_t(nb)) : __builtin_ctz(uint32_t(nb));
            drow[c] = d;
            nb &= nb - 1;
        }
    }
    __syncthreads();
}

// Leaf all-pairs BFS (src/leaf_tile.cpp:56-153 + projection :198-207): work
// item (leaf, chunk of SWG consecutive sources); whole D rows are streamed
// out (coalesced) and port sources also write their S row.  changed[li]
// (li: position in leaf_list, or chg_idx[li]) is set when S decreased.
template <int LANES, typename RowT>
__global__ __launch_bounds__(256) void k_leaf_apsp(const uint8_t* __restrict__ pass, int S, int b, int nl,
                                                   const int* __restrict__ leaf_list, int chunks, int n_items,
                                                   uint16_t* __restrict__ leafD, uint16_t* __restrict__ leafS,
                                                   int* __restrict__ changed, const int* __restrict__ chg_idx) {
    constexpr int SPW = 64 / LANES;
    constexpr int SWG = 4 * SPW;
    extern __shared__ __align__(16) uint16_t dist[];  // [SWG][b*b]
    // work item = (leaf, chunk of SWG sources); a capped grid strides over them
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
    if (item != int(blockIdx.x)) __syncthreads();  // the previous item's readers of dist are done
    const int li = item / chunks, chunk = item % chunks;
    const int leaf = leaf_list ? leaf_list[li] : li;
    const int n = b * b;
    const int R0 = (leaf / nl) * b, C0 = (leaf % nl) * b;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slot = wave * SPW + lane / LANES;
    const int src = chunk * SWG + slot;
    leaf_bfs<LANES, RowT>(pass, S, b, R0, C0, src < n, src, dist);

    const int first = chunk * SWG;
    const int nvalid = min(SWG, n - first);
    uint16_t* gD = leafD + size_t(leaf) * n * n + size_t(first) * n;
    if ((n & 7) == 0) {
        const uint4* s4 = reinterpret_cast<const uint4*>(dist);
        uint4* g4 = reinterpret_cast<uint4*>(gD);
        for (int i = tid; i < nvalid * n / 8; i += 256) g4[i] = s4[i];
    } else {
        for (int i = tid; i < nvalid * n; i += 256) gD[i] = dist[i];
    }
    const int pl = 4 * b - 4;
    uint16_t* gS = leafS + size_t(leaf) * pl * pl;
    int* chg = changed ? changed + (chg_idx ? chg_idx[li] : li) : nullptr;
    for (int s = 0; s < nvalid; ++s) {
        const int cell = first + s;
        const int pi = port_index(cell / b, cell % b, b);
        if (pi < 0) continue;
        for (int pj = tid; pj < pl; pj += 256) {
            int pr, pc;
            port_rc(pj, b, pr, pc);
            const uint16_t v = dist[s * n + pr * b + pc];
            uint16_t* dst = gS + size_t(pi) * pl + pj;
            if (chg && v < *dst) *chg = 1;
            *dst = v;
        }
    }
    }
}

// Rows of a leaf's newly opened cells (incremental reveal of a leaf): work
// item (leaf, first pivot, count <= SWG, changed index); BFS from each pivot
// cell p over the leaf's current passability gives D'[p][*], written to the
// leaf's D row p and to P[pivot] for the rank-k relaxation; a pivot that is
// a port also writes its S row (the relaxation leaves row p as it is).
template <int LANES, typename RowT>
__global__ __launch_bounds__(256) void k_leaf_pivot_rows(const uint8_t* __restrict__ pass, int S, int b, int nl,
                                                         const int4* __restrict__ items, int n_items,
                                                         const int* __restrict__ pcell, uint16_t* __restrict__ leafD,
                                                         uint16_t* __restrict__ P, uint16_t* __restrict__ leafS,
                                                         int* __restrict__ changed) {
    constexpr int SPW = 64 / LANES;
    extern __shared__ __align__(16) uint16_t dist[];  // [SWG][b*b]
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        if (item != int(blockIdx.x)) __syncthreads();
        const int4 it = items[item];
        const int leaf = it.x, off = it.y, cnt = it.z;
        const int n = b * b;
        const int R0 = (leaf / nl) * b, C0 = (leaf % nl) * b;
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const int slot = wave * SPW + lane / LANES;
        const bool valid = slot < cnt;
        const int src = valid ? pcell[off + slot] : 0;
        leaf_bfs<LANES, RowT>(pass, S, b, R0, C0, valid, src, dist);
        const int n8 = n / 8;  // n % 8 == 0
        const uint4* s4 = reinterpret_cast<const uint4*>(dist);
        for (int i = tid; i < cnt * n8; i += 256) {
            const int sl = i / n8, q = i % n8;
            const uint4 v = s4[i];
            reinterpret_cast<uint4*>(leafD + size_t(leaf) * n * n + size_t(pcell[off + sl]) * n)[q] = v;
            reinterpret_cast<uint4*>(P + size_t(off + sl) * n)[q] = v;
        }
        const int pl = 4 * b - 4;
        uint16_t* gS = leafS + size_t(leaf) * pl * pl;
        for (int sl = 0; sl < cnt; ++sl) {
            const int cell = pcell[off + sl];
            const int pi = port_index(cell / b, cell % b, b);
            if (pi < 0) continue;
            for (int pj = tid; pj < pl; pj += 256) {
                int pr, pc;
                port_rc(pj, b, pr, pc);
                const uint16_t v = dist[sl * n + pr * b + pc];
                uint16_t* dst = gS + size_t(pi) * pl + pj;
                if (changed && v < *dst) changed[it.w] = 1;
                *dst = v;
            }
        }
    }
}

// Rank-k relaxation of a leaf through its new cells:
//   D'[i][j] = min(D[i][j], min_p P[p][i] + P[p][j]),  P[p] = D'[p][*]
// exact, since a new shortest path passes through at least one new cell and
// D'[i][p] = D'[p][i].  Work item: leaf (grid y: leaf, first pivot, count,
// changed index) x tile of 64 rows x 128 columns (grid x); thread = 4 rows x
// 8 columns in packed u16 pairs; pivots stream through LDS in chunks of 32
// (A = dup16(P[p][row]), X = P[p][column octet]).  Port x port entries also
// go to S, setting changed[] where S decreased.
constexpr int kLeafRelaxK = 32;
__global__ __launch_bounds__(256) void k_leaf_relax(int b, const int4* __restrict__ leaves, int col_tiles,
                                                    const uint16_t* __restrict__ P, uint16_t* __restrict__ leafD,
                                                    uint16_t* __restrict__ leafS, int* __restrict__ changed) {
    __shared__ uint4 A4[kLeafRelaxK][16];  // [pivot][row group]: 4 rows' dup16 values
    __shared__ uint4 X4[kLeafRelaxK][16];  // [pivot][column octet]
    const int4 lf = leaves[blockIdx.y];
    const int leaf = lf.x, off = lf.y, k = lf.z;
    const int n = b * b;
    const int I = blockIdx.x / col_tiles, J = blockIdx.x % col_tiles;
    const int i0 = I * 64, j0 = J * 128;
    if (i0 >= n || j0 >= n) return;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    uint16_t* D = leafD + size_t(leaf) * n * n;
    const int col = j0 + 8 * tx;
    const bool cok = col < n;  // n % 8 == 0: whole octets
    uint32_t acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = i0 + ty * 4 + r;
        uint4 v = make_uint4(kInf2, kInf2, kInf2, kInf2);
        if (cok && row < n) v = *reinterpret_cast<const uint4*>(D + size_t(row) * n + col);
        acc[r][0] = v.x, acc[r][1] = v.y, acc[r][2] = v.z, acc[r][3] = v.w;
    }
    uint32_t old[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int h = 0; h < 4; ++h) old[r][h] = acc[r][h];
    const uint16_t* Pl = P + size_t(off) * n;
    for (int k0 = 0; k0 < k; k0 += kLeafRelaxK) {
        const int m = min(kLeafRelaxK, k - k0);
        __syncthreads();  // the previous chunk's readers are done
        // A: 32 pivots x 64 rows (u16 -> dup16), 8 per thread; X: 32 x 16 octets, 2 per thread
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int idx = tid + 256 * e;  // (pivot, row group of 4)
            const int kk = idx >> 4, rg = idx & 15;
            uint4 a = make_uint4(kInf2, kInf2, kInf2, kInf2);
            const int row = i0 + rg * 4;
            if (kk < m && row < n) {
                const uint2 pv = *reinterpret_cast<const uint2*>(Pl + size_t(k0 + kk) * n + row);
                a = make_uint4(dup16(pv.x), dup16(pv.x >> 16), dup16(pv.y), dup16(pv.y >> 16));
            }
            A4[kk][rg] = a;
            const int oc = idx & 15;
            uint4 x = make_uint4(kInf2, kInf2, kInf2, kInf2);
            if (kk < m && j0 + 8 * oc < n) x = *reinterpret_cast<const uint4*>(Pl + size_t(k0 + kk) * n + j0 + 8 * oc);
            X4[kk][oc] = x;
        }
        __syncthreads();
        for (int kk = 0; kk < m; ++kk) {
            const uint4 a = A4[kk][ty];
            const uint4 x = X4[kk][tx];
            const uint32_t av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[r][0] = relax2(acc[r][0], av[r], x.x);
                acc[r][1] = relax2(acc[r][1], av[r], x.y);
                acc[r][2] = relax2(acc[r][2], av[r], x.z);
                acc[r][3] = relax2(acc[r][3], av[r], x.w);
            }
        }
    }
    if (!cok) return;
    const int pl = 4 * b - 4;
    uint16_t* Sl = leafS + size_t(leaf) * pl * pl;
    bool dec = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = i0 + ty * 4 + r;
        if (row >= n) continue;
        bool same = true;
#pragma unroll
        for (int h = 0; h < 4; ++h) same &= acc[r][h] == old[r][h];
        if (same) continue;
        *reinterpret_cast<uint4*>(D + size_t(row) * n + col) = make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
        const int pi = port_index(row / b, row % b, b);
        if (pi < 0) continue;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t v = (acc[r][q >> 1] >> (16 * (q & 1))) & 0xffffu;
            const uint32_t o = (old[r][q >> 1] >> (16 * (q & 1))) & 0xffffu;
            if (v == o) continue;
            const int pj = port_index((col + q) / b, (col + q) % b, b);
            if (pj < 0) continue;
            Sl[size_t(pi) * pl + pj] = uint16_t(v);
            dec = true;
        }
    }
    if (dec && changed) changed[lf.w] = 1;
}

bool launch_leaf_apsp(const uint8_t* pass, int S, int b, int nl, const int* leaf_list, int n_list, uint16_t* leafD,
                      uint16_t* leafS, int* changed, hipStream_t st, KTimer kt, int max_wg, const int* chg_idx) {
    if (n_list == 0) return false;
    const int n = b * b;
    auto go = [&](auto kern, int lanes) {
        const int swg = 4 * (64 / lanes);
        const int chunks = (n + swg - 1) / swg;
        const int items = n_list * chunks;
        const int grid = max_wg > 0 ? std::min(items, max_wg) : items;
        const uint32_t lds = uint32_t(size_t(swg) * n * sizeof(uint16_t));
        HTG_LAUNCH_T(kern, dim3(unsigned(grid)), dim3(256), lds, st, kt, pass, S, b, nl, leaf_list, chunks, items,
                     leafD, leafS, changed, chg_idx);
    };
    if (b <= 2) go(k_leaf_apsp<2, uint32_t>, 2);
    else if (b <= 4) go(k_leaf_apsp<4, uint32_t>, 4);
    else if (b <= 8) go(k_leaf_apsp<8, uint32_t>, 8);
    else if (b <= 16) go(k_leaf_apsp<16, uint32_t>, 16);
    else if (b <= 32) go(k_leaf_apsp<32, uint32_t>, 32);
    else go(k_leaf_apsp<64, uint64_t>, 64);
    return true;
}

int leaf_sources_per_wg(int b) {
    int lanes = 2;
    while (lanes < b) lanes *= 2;
    return 4 * (64 / lanes);
}

bool launch_leaf_pivot_rows(const uint8_t* pass, int S, int b, int nl, const int4* items, int n_items,
                            const int* pcell, uint16_t* leafD, uint16_t* P, uint16_t* leafS, int* changed,
                            hipStream_t st, KTimer kt, int max_wg) {
    if (n_items == 0) return false;
    const int n = b * b;
    const int grid = max_wg > 0 ? std::min(n_items, max_wg) : n_items;
    auto go = [&](auto kern) {
        const uint32_t lds = uint32_t(size_t(leaf_sources_per_wg(b)) * n * sizeof(uint16_t));
        HTG_LAUNCH_T(kern, dim3(unsigned(grid)), dim3(256), lds, st, kt, pass, S, b, nl, items, n_items, pcell, leafD,
                     P, leafS, changed);
    };
    if (b <= 8) go(k_leaf_pivot_rows<8, uint32_t>);
    else if (b <= 16) go(k_leaf_pivot_rows<16, uint32_t>);
    else if (b <= 32) go(k_leaf_pivot_rows<32, uint32_t>);
    else go(k_leaf_pivot_rows<64, uint64_t>);
    return true;
}

bool launch_leaf_relax(int b, const int4* leaves, int n_leaves, const uint16_t* P, uint16_t* leafD, uint16_t* leafS,
                       int* changed, hipStream_t st, KTimer kt) {
    if (n_leaves == 0) return false;
    const int n = b * b;
    const int rt = (n + 63) / 64, ct = (n + 127) / 128;
    HTG_LAUNCH_T(k_leaf_relax, dim3(unsigned(rt * ct), unsigned(n_leaves)), dim3(256), 0, st, kt, b, leaves, ct, P,
                 leafD, leafS, changed);
    return true;
}

// ---------------------------------------------------------------------------
// D_U assembly (src/internal_node.cpp:98-145): identity + child summaries on
// the diagonal blocks.  A child summary with no finite off-diagonal entry is
// the identity, so copying every child unconditionally gives the same D_U.
// An octet is in a node's sorted keep list (binary search).
__device__ __forceinline__ bool in_octets(const int* __restrict__ list, int n, int o) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < o) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && list[lo] == o;
}

__global__ __launch_bounds__(256) void k_node_assemble(int g, int pc, int nV, int ld, int nt, int child_nt,
                                                       const uint16_t* __restrict__ child_sum, size_t child_stride,
                                                       const int* __restrict__ node_list,
                                                       uint16_t* __restrict__ DU, size_t node_stride,
                                                       const int* __restrict__ keep_cnt,
                                                       const int* __restrict__ keep_list, int ng8) {
    const int groups_per_row = ld / 8;
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)nV * groups_per_row) return;
    const int v = int(gid / groups_per_row), u0 = int(gid % groups_per_row) * 8;
    const int node = node_list[blockIdx.y];
    if (keep_cnt) {  // entries whose row and column octets are both kept stay as they are
        const int* kl = keep_list + size_t(node) * ng8;
        const int kn = keep_cnt[node];
        if (in_octets(kl, kn, v / 8) && in_octets(kl, kn, u0 / 8)) return;
    }
    const int tx = node % nt, ty = node / nt;
    const int qv = v / pc, pv = v % pc;
    const int cidx = (ty * g + qv / g) * child_nt + (tx * g + qv % g);
    const uint16_t* src = child_sum + size_t(cidx) * child_stride + size_t(pv) * pc;
    uint16_t o[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int u = u0 + e;
        o[e] = (u < nV && u / pc == qv) ? src[u % pc] : kInf;
    }
    uint4 w;
    w.x = o[0] | (uint32_t(o[1]) << 16);
    w.y = o[2] | (uint32_t(o[3]) << 16);
    w.z = o[4] | (uint32_t(o[5]) << 16);
    w.w = o[6] | (uint32_t(o[7]) << 16);
    *reinterpret_cast<uint4*>(DU + size_t(node) * node_stride + size_t(v) * ld + u0) = w;
}

void launch_node_assemble(int lvl, int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum,
                          size_t child_stride, const int* node_list, int n_list, uint16_t* DU, size_t node_stride,
                          hipStream_t st, const int* keep_cnt, const int* keep_list, int ng8) {
    (void)lvl;
    const long long total = (long long)nV * (ld / 8);
    dim3 grid(unsigned((total + 255) / 256), unsigned(n_list));
    hipLaunchKernelGGL(k_node_assemble, grid, dim3(256), 0, st, g, pc, nV, ld, nt, nt * g, child_sum, child_stride,
                       node_list, DU, node_stride, keep_cnt, keep_list, ng8);
}

// ---------------------------------------------------------------------------
// Seam edges between sibling children (src/internal_node.cpp:147-173).
__global__ __launch_bounds__(256) void k_seam_relax(int nV, int ld, int s, int nt, const int2* __restrict__ seams,
                                                    int n_seams, const short2* __restrict__ vu_rc,
                                                    const uint8_t* __restrict__ pass, int S,
                                                    const int* __restrict__ node_list, uint16_t* __restrict__ DU,
                                                    size_t node_stride, uint8_t* __restrict__ flags) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_seams) return;
    const int node = node_list[blockIdx.y];
    const int R0 = (node / nt) * s, C0 = (node % nt) * s;
    const int2 e = seams[i];
    const short2 a = vu_rc[e.x], c = vu_rc[e.y];
    if (!pass[size_t(R0 + a.x) * S + C0 + a.y] || !pass[size_t(R0 + c.x) * S + C0 + c.y]) return;
    uint16_t* D = DU + size_t(node) * node_stride;
    uint16_t& uv = D[size_t(e.x) * ld + e.y];
    uint16_t& vu = D[size_t(e.y) * ld + e.x];
    if (uv > 1) uv = 1;
    if (vu > 1) vu = 1;
    if (flags) {  // every passable seam edge's endpoints are pivots
        flags[size_t(blockIdx.y) * nV + e.x] = 1;
        flags[size_t(blockIdx.y) * nV + e.y] = 1;
    }
}

void launch_seam_relax(int nV, int ld, int s, int nt, const int2* seams, int n_seams, const short2* vu_rc,
                       const uint8_t* pass, int S, const int* node_list, int n_list, uint16_t* DU, size_t node_stride,
                       uint8_t* flags, hipStream_t st) {
    if (n_seams == 0 || n_list == 0) return;
    dim3 grid(unsigned((n_seams + 255) / 256), unsigned(n_list));
    hipLaunchKernelGGL(k_seam_relax, grid, dim3(256), 0, st, nV, ld, s, nt, seams, n_seams, vu_rc, pass, S, node_list,
                       DU, node_stride, flags);
}

// Explicit edges (node slot, u, v, w): D_U[u][v] = D_U[v][u] = min(., w)
// (src/internal_node.cpp:356-365).  The update path lists each pair once.
// Cover marks (incremental reveal).  With `plane` > 0 the flag buffer holds
// three planes of n_list * nV bytes: C marks, B marks, and the new-port mask.
// A decreased entry (v1, v2) marks row v1 into C when v1 is a new port or
// neither endpoint is; an old row whose partner is new goes to B, the
// uncovered endpoints.  Every decreased pair then has an endpoint in C.  With
// `plane` == 0 every decreased row is a C mark (the reference's pivot set).
__device__ __forceinline__ void cover_mark(uint8_t* flags, size_t plane, size_t row_off, bool self_new,
                                           bool other_new) {
    if (plane == 0 || self_new || !other_new) flags[row_off] = 1;
    else flags[plane + row_off] = 1;
}

__global__ __launch_bounds__(256) void k_edge_relax(int nV, int ld, const int4* __restrict__ edges, int n_edges,
                                                    const int* __restrict__ node_list, uint16_t* __restrict__ DU,
                                                    size_t node_stride, uint8_t* __restrict__ flags, size_t plane) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_edges) return;
    const int4 e = edges[i];
    uint16_t* D = DU + size_t(node_list[e.x]) * node_stride;
    const uint16_t w = uint16_t(e.w);
    uint16_t& uv = D[size_t(e.y) * ld + e.z];
    if (w < uv) {
        uv = w;
        D[size_t(e.z) * ld + e.y] = w;
        const uint8_t* nw = flags + 2 * plane + size_t(e.x) * nV;
        const bool ny = plane && nw[e.y], nz = plane && nw[e.z];
        cover_mark(flags, plane, size_t(e.x) * nV + e.y, ny, nz);
        cover_mark(flags, plane, size_t(e.x) * nV + e.z, nz, ny);
    }
}

void launch_edge_relax(int nV, int ld, const int4* edges, int n_edges, const int* node_list, uint16_t* DU,
                       size_t node_stride, uint8_t* flags, size_t plane, hipStream_t st) {
    if (n_edges == 0) return;
    hipLaunchKernelGGL(k_edge_relax, dim3(unsigned((n_edges + 255) / 256)), dim3(256), 0, st, nV, ld, edges, n_edges,
                       node_list, DU, node_stride, flags, plane);
}

// items = (slot, port): plane[slot * nV + port] = 1.
__global__ __launch_bounds__(256) void k_mark_u8(const int2* __restrict__ items, int n, int nV,
                                                 uint8_t* __restrict__ plane) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int2 it = items[i];
    plane[size_t(it.x) * nV + it.y] = 1;
}

void launch_mark_u8(const int2* items, int n, int nV, uint8_t* plane, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_mark_u8, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, items, n, nV, plane);
}

// Child decreases (src/internal_node.cpp:343-354) expressed as an elementwise
// min of the child's new summary into its diagonal block: entries that did not
// decrease are already <= the summary, so only decreased entries move.
__global__ __launch_bounds__(256) void k_child_relax(int g, int pc, int nV, int ld, int nt, int child_nt,
                                                     const uint16_t* __restrict__ child_sum, size_t child_stride,
                                                     const int* __restrict__ node_list,
                                                     const uint32_t* __restrict__ child_mask,
                                                     uint16_t* __restrict__ DU, size_t node_stride,
                                                     uint8_t* __restrict__ flags) {
    const long long per_child = (long long)pc * pc;
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= per_child * g * g) return;
    const int q = int(gid / per_child);
    if (!((child_mask[blockIdx.y] >> q) & 1u)) return;
    const int a = int((gid % per_child) / pc), c = int(gid % pc);
    const int node = node_list[blockIdx.y];
    const int tx = node % nt, ty = node / nt;
    const int cidx = (ty * g + q / g) * child_nt + (tx * g + q % g);
    const uint16_t w = child_sum[size_t(cidx) * child_stride + size_t(a) * pc + c];
    uint16_t& dst = DU[size_t(node) * node_stride + size_t(q * pc + a) * ld + q * pc + c];
    if (w < dst) {
        dst = w;
        flags[size_t(blockIdx.y) * nV + q * pc + a] = 1;
    }
}

void launch_child_relax(int lvl, int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum,
                        size_t child_stride, const int* node_list, const uint32_t* child_mask, int n_list,
                        uint16_t* DU, size_t node_stride, uint8_t* flags, hipStream_t st) {
    (void)lvl;
    if (n_list == 0) return;
    const long long total = (long long)pc * pc * g * g;
    dim3 grid(unsigned((total + 255) / 256), unsigned(n_list));
    hipLaunchKernelGGL(k_child_relax, grid, dim3(256), 0, st, g, pc, nV, ld, nt, nt * g, child_sum, child_stride,
                       node_list, child_mask, DU, node_stride, flags);
}

// Child relax of an update over live octets: for V_U ports v1, v2 of the
// same child q whose summary changed (its flag in the level below is set),
// D_U[v1][v2] = min(D_U[v1][v2], S_q[v1 - q*pc][v2 - q*pc]); rows that
// decreased are flagged as pivots (src/internal_node.cpp:343-370).  Dead
// ports have INF rows and columns in S_q and D_U alike, so only live octet
// pairs are visited: thread = (row octet, column octet, row in the octet).
__global__ __launch_bounds__(256) void k_child_relax_live(int g, int pc, int nV, int ld, int nt, int child_nt,
                                                          const uint16_t* __restrict__ child_sum,
                                                          size_t child_stride, const int* __restrict__ node_list,
                                                          const int* __restrict__ child_slot,
                                                          const int* __restrict__ child_changed,
                                                          const int* __restrict__ live8_cnt,
                                                          const int* __restrict__ live8_list, int ng8,
                                                          uint16_t* __restrict__ DU, size_t node_stride,
                                                          uint8_t* __restrict__ flags, size_t plane) {
    const int i = blockIdx.y;
    const int node = node_list[i];
    const int n8 = live8_cnt[node];
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)n8 * n8 * 8) return;
    const int* l8 = live8_list + size_t(node) * ng8;
    const long long pr = t >> 3;
    const int v1 = 8 * l8[pr / n8] + int(t & 7);
    if (v1 >= nV) return;
    const int q = v1 / pc;
    const int cs = child_slot[i * g * g + q];
    if (cs < 0 || !child_changed[cs]) return;
    const int c0 = 8 * l8[pr % n8];
    const int lo = q * pc, hi = lo + pc;
    if (c0 + 8 <= lo || c0 >= hi) return;  // column octet outside child q
    const int tx = node % nt, ty = node / nt;
    const int cidx = (ty * g + q / g) * child_nt + (tx * g + q % g);
    const uint16_t* Srow = child_sum + size_t(cidx) * child_stride + size_t(v1 - lo) * pc;
    uint16_t* Drow = DU + size_t(node) * node_stride + size_t(v1) * ld;
    const uint8_t* nw = flags + 2 * plane + size_t(i) * nV;
    const bool n1 = plane && nw[v1];
    bool dec_c = false, dec_b = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int v2 = c0 + k;
        if (v2 < lo || v2 >= hi) continue;
        const uint16_t w = Srow[v2 - lo];
        if (w < Drow[v2]) {
            Drow[v2] = w;
            if (!plane || n1 || !nw[v2]) dec_c = true;
            else dec_b = true;
        }
    }
    if (dec_c) flags[size_t(i) * nV + v1] = 1;
    if (dec_b) flags[plane + size_t(i) * nV + v1] = 1;
}

void launch_child_relax_live(int g, int pc, int nV, int ld, int nt, const uint16_t* child_sum, size_t child_stride,
                             const int* node_list, const int* child_slot, const int* child_changed, int n_list,
                             const int* live8_cnt, const int* live8_list, int ng8, int max_live8, uint16_t* DU,
                             size_t node_stride, uint8_t* flags, size_t plane, hipStream_t st) {
    if (n_list == 0 || max_live8 == 0) return;
    const long long total = (long long)max_live8 * max_live8 * 8;
    const dim3 grid{unsigned((total + 255) / 256), unsigned(n_list), 1u};
    hipLaunchKernelGGL(k_child_relax_live, grid, dim3(256), 0, st, g, pc, nV, ld, nt, nt * g, child_sum, child_stride,
                       node_list, child_slot, child_changed, live8_cnt, live8_list, ng8, DU, node_stride, flags,
                       plane);
}

// Row flags -> ascending pivot list over the live octets (flags are only
// ever set on live rows); clears the flags it reads.  One workgroup per node.
// With `plane` > 0 the B marks of rows outside C form a second list (bpiv,
// counts[n_list + slot]).
__global__ __launch_bounds__(256) void k_compact_flags_live(int nV, const int* __restrict__ node_list, int n_list,
                                                            const int* __restrict__ live8_cnt,
                                                            const int* __restrict__ live8_list, int ng8,
                                                            uint8_t* __restrict__ flags, size_t plane,
                                                            int* __restrict__ pivots, int* __restrict__ bpiv,
                                                            int piv_stride, int* __restrict__ counts) {
    __shared__ int wsum[2][4];
    const int node = node_list[blockIdx.x];
    const int n = 8 * live8_cnt[node];
    const int* l8 = live8_list + size_t(node) * ng8;
    uint8_t* f = flags + size_t(blockIdx.x) * nV;
    uint8_t* fb = plane ? f + plane : nullptr;
    int* out = pivots + size_t(blockIdx.x) * piv_stride;
    int* outb = bpiv ? bpiv + size_t(blockIdx.x) * piv_stride : nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int base = 0, base_b = 0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int e = c0 + tid;
        const int v = e < n ? 8 * l8[e >> 3] + (e & 7) : nV;
        const bool on = v < nV && f[v] != 0;
        const bool on_b = fb && v < nV && !on && fb[v] != 0;
        const unsigned long long m = __ballot(on), mb = __ballot(on_b);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (lane == 0) {
            wsum[0][wave] = __popcll(m);
            wsum[1][wave] = __popcll(mb);
        }
        __syncthreads();
        int off = base, off_b = base_b;
        for (int w = 0; w < wave; ++w) off += wsum[0][w], off_b += wsum[1][w];
        if (on) {
            out[off + __popcll(m & below)] = v;
            f[v] = 0;
        }
        if (on_b) outb[off_b + __popcll(mb & below)] = v;
        if (fb && v < nV) fb[v] = 0;
        base += wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
        base_b += wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
        __syncthreads();
    }
    if (tid == 0) {
        counts[blockIdx.x] = base;
        if (fb) counts[n_list + blockIdx.x] = base_b;
    }
}

void launch_compact_flags_live(int nV, const int* node_list, int n_list, const int* live8_cnt, const int* live8_list,
                               int ng8, uint8_t* flags, size_t plane, int* pivots, int* bpiv, int piv_stride,
                               int* counts, hipStream_t st) {
    if (n_list == 0) return;
    hipLaunchKernelGGL(k_compact_flags_live, dim3(unsigned(n_list)), dim3(256), 0, st, nV, node_list, n_list,
                       live8_cnt, live8_list, ng8, flags, plane, pivots, plane ? bpiv : nullptr, piv_stride, counts);
}

// Pre-relaxation of the cover rows through the uncovered endpoints, before
// the closure over C (incremental reveal):
//   hop 1:  D[c][j]   = min(D[c][j],   min_{b in B} D[c][b] + D[b][j])   c in C, live j
//   hop 2:  D[c1][c2] = min(D[c1][c2], min_{b in B} D[c1][b] + D[c2][b]) c1, c2 in C
// Every shortest path of the updated node graph alternates closed old
// distances and decreased entries; each decreased entry has an endpoint in
// C, so between two consecutive C vertices (or a path end and its nearest C
// vertex) lie at most two vertices of B.  Hop 1 folds one of them into the C
// rows, hop 2 (on rows already relaxed, so D[c2][b] <= D[b][c2]) the second;
// after both, the closure over C alone is exact for every pair.  Hop 1 only
// reads B rows; a C row may be read by one workgroup while another lowers
// it, which only ever substitutes a smaller valid path length.  Hop 2 reads
// B columns and writes C columns.
//
// Both hops are one rectangular min-plus product, tiled like the closure
// update: a work unit owns 16*RPT C rows x 16 column octets of one node
// (thread = RPT rows x 8 columns, two u16 lanes per v_pk op) and a range of
// the inner set B, which it streams through LDS in chunks of 32: dup'd
// A[b][row] = D[row][b] and X[b][octet] (hop 1: the B row's live octet;
// hop 2: D[c2][b] for 8 consecutive cover entries c2).  Staging is
// software-pipelined: the gathers of chunk i+1 are issued before chunk i is
// computed, with their B indices loaded one chunk earlier still.  Waves whose
// rows are all past the cover size skip the arithmetic.  The host lists the
// units from the level's C and B counts, so no workgroup is spent on a tile
// past a node's cover, and it splits a node's B into parts of a few chunks:
// a unit's duration is a chain of dependent chunk stages, so one node with a
// long B would otherwise hold the launch for all of them.  Units over a part
// of B start from INF and fold their minima into D with compare-and-swap
// minima, which lose no smaller value whatever the order.
constexpr int kCoverOct = 16, kCoverK = 32;

// D[w] = min(D[w], v) for N packed words at once (null address: no word),
// atomically against other units doing the same: every word is read once,
// all compare-and-swaps of a round are issued before any result is examined,
// and only words that another unit lowered in between go round again.  Values
// only decrease, so the rounds end.
template <int N>
__device__ __forceinline__ void fold_min_words(uint32_t* (&wa)[N], const uint32_t (&wv)[N]) {
    uint32_t old[N], nv[N], prev[N];
#pragma unroll
    for (int i = 0; i < N; ++i) old[i] = wa[i] ? __hip_atomic_load(wa[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    for (;;) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            nv[i] = min2(old[i], wv[i]);
            if (nv[i] == old[i]) wa[i] = nullptr;  // nothing to lower
        }
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (wa[i]) prev[i] = atomicCAS(wa[i], old[i], nv[i]);
        bool again = false;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (wa[i]) {
                if (prev[i] == old[i]) wa[i] = nullptr;
                else old[i] = prev[i], again = true;
            }
        if (!again) return;
    }
}

template <int RPT> struct LdsRow;
template <> struct LdsRow<1> {
    typedef uint32_t T;
    static __device__ __forceinline__ void get(const T& v, uint32_t* o) { o[0] = v; }
};
template <> struct LdsRow<2> {
    typedef uint2 T;
    static __device__ __forceinline__ void get(const T& v, uint32_t* o) { o[0] = v.x, o[1] = v.y; }
};
template <> struct LdsRow<4> {
    typedef uint4 T;
    static __device__ __forceinline__ void get(const T& v, uint32_t* o) { o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w; }
};

template <bool HOP2, int RPT>
__global__ __launch_bounds__(256) void k_cover_product(int nV, int ld, const int* __restrict__ node_list, int n_list,
                                                       const int* __restrict__ counts, const int* __restrict__ cpiv,
                                                       const int* __restrict__ bpiv, int piv_stride,
                                                       const int4* __restrict__ units, const int* __restrict__ live8_cnt,
                                                       const int* __restrict__ live8_list, int ng8,
                                                       uint16_t* __restrict__ DU, size_t node_stride) {
    constexpr int R = 16 * RPT;               // rows per tile
    constexpr int SA = kCoverK * R / 256;     // A entries staged per thread
    constexpr int SX = kCoverK * kCoverOct / 256;  // X octets staged per thread
    typedef typename LdsRow<RPT>::T RowT;
    __shared__ RowT A4[kCoverK][16];  // dup16(D[row][b]): the RPT rows of a thread
    __shared__ uint4 X4[kCoverK][kCoverOct];
    __shared__ int rowv[R], colv[kCoverOct * 8];
    const int4 un = units[blockIdx.x];  // slot, row tile | column tile << 16, B range
    const int slot = un.x, I = un.y & 0xffff, J = un.y >> 16, kb = un.z, ke = un.w;
    const int nc = counts[slot], nb = counts[n_list + slot];
    if (kb >= ke || ke > nb || I * R >= nc) return;
    const bool part = kb > 0 || ke < nb;  // other units fold the rest of B into the same entries
    const int node = node_list[slot];
    const int n_cols = HOP2 ? nc : 8 * live8_cnt[node];  // columns this node has
    if (J * kCoverOct * 8 >= n_cols) return;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int* cl = cpiv + size_t(slot) * piv_stride;
    const int* bl = bpiv + size_t(slot) * piv_stride;
    uint16_t* D = DU + size_t(node) * node_stride;
    if (tid < R) rowv[tid] = I * R + tid < nc ? cl[I * R + tid] : -1;
    if (tid < kCoverOct * 8) {
        const int e = J * kCoverOct * 8 + tid;  // column entry of this tile
        int v = -1;
        if (HOP2) v = e < nc ? cl[e] : -1;
        else if (e < n_cols) v = 8 * live8_list[size_t(node) * ng8 + (e >> 3)] + (e & 7);
        colv[tid] = v < nV ? v : -1;
    }
    __syncthreads();
    const bool busy = (tid >> 6) * 4 * RPT < nc - I * R;  // this wave has rows
    // staging roles: A entry i is (k = tid/R + (256/R) i, row = tid%R); X
    // octet i is (k = tid/16 + 16 i, octet = tid%16)
    const int a_r = tid % R;
    const int a_row = rowv[a_r];
    const int x_o = tid & 15;
    int x_c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x_c[q] = colv[x_o * 8 + q];
    uint32_t acc[RPT][4];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int row = rowv[ty * RPT + r];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int c0 = colv[tx * 8 + 2 * h], c1 = colv[tx * 8 + 2 * h + 1];
            const uint32_t lo = !part && row >= 0 && c0 >= 0 ? D[size_t(row) * ld + c0] : kInf;
            const uint32_t hi = !part && row >= 0 && c1 >= 0 ? D[size_t(row) * ld + c1] : kInf;
            acc[r][h] = lo | (hi << 16);
        }
    }
    auto load_idx = [&](int k0, int* ia, int* ix) {
#pragma unroll
        for (int i = 0; i < SA; ++i) {
            const int k = k0 + tid / R + (256 / R) * i;
            ia[i] = k < ke ? bl[k] : -1;
        }
#pragma unroll
        for (int i = 0; i < SX; ++i) {
            const int k = k0 + (tid >> 4) + 16 * i;
            ix[i] = k < ke ? bl[k] : -1;
        }
    };
    auto gather = [&](const int* ia, const int* ix, uint32_t* va, uint4* vx) {
#pragma unroll
        for (int i = 0; i < SA; ++i)
            va[i] = ia[i] >= 0 && a_row >= 0 ? dup16(D[size_t(a_row) * ld + ia[i]]) : kInf2;
#pragma unroll
        for (int i = 0; i < SX; ++i) {
            uint4 v = make_uint4(kInf2, kInf2, kInf2, kInf2);
            if (ix[i] >= 0) {
                uint16_t* ve = reinterpret_cast<uint16_t*>(&v);
                if (HOP2) {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (x_c[q] >= 0) ve[q] = D[size_t(x_c[q]) * ld + ix[i]];
                } else if (x_c[0] >= 0) {
                    v = *reinterpret_cast<const uint4*>(D + size_t(ix[i]) * ld + x_c[0]);
#pragma unroll
                    for (int q = 0; q < 8; ++q)  // past nV: padding
                        if (x_c[q] < 0) ve[q] = kInf;
                }
            }
            vx[i] = v;
        }
    };
    int ia[SA], ix[SX];
    uint32_t va[SA];
    uint4 vx[SX];
    load_idx(kb, ia, ix);
    gather(ia, ix, va, vx);
    load_idx(kb + kCoverK, ia, ix);
    for (int k0 = kb; k0 < ke; k0 += kCoverK) {
        __syncthreads();  // the previous chunk's arithmetic is done
#pragma unroll
        for (int i = 0; i < SA; ++i)  // row a_r sits in thread group a_r / RPT, lane a_r % RPT
            reinterpret_cast<uint32_t*>(&A4[tid / R + (256 / R) * i][a_r / RPT])[a_r % RPT] = va[i];
#pragma unroll
        for (int i = 0; i < SX; ++i) X4[(tid >> 4) + 16 * i][x_o] = vx[i];
        __syncthreads();
        if (k0 + kCoverK < ke) {  // next chunk in flight during this one's arithmetic
            gather(ia, ix, va, vx);
            load_idx(k0 + 2 * kCoverK, ia, ix);
        }
        if (!busy) continue;
        const int m = min(kCoverK, ke - k0);
#pragma unroll 4
        for (int k = 0; k < m; ++k) {
            uint32_t av[RPT];
            LdsRow<RPT>::get(A4[k][ty], av);
            const uint4 x = X4[k][tx];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                acc[r][0] = relax2(acc[r][0], av[r], x.x);
                acc[r][1] = relax2(acc[r][1], av[r], x.y);
                acc[r][2] = relax2(acc[r][2], av[r], x.z);
                acc[r][3] = relax2(acc[r][3], av[r], x.w);
            }
        }
    }
    if (part) {  // one batch of up to 8 words per row
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int row = rowv[ty * RPT + r];
            if (row < 0) continue;
            uint16_t* Dr = D + size_t(row) * ld;
            uint32_t* wa[8];
            uint32_t wv[8];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int c0 = colv[tx * 8 + 2 * h], c1 = colv[tx * 8 + 2 * h + 1];
                wa[2 * h] = wa[2 * h + 1] = nullptr;
                wv[2 * h] = wv[2 * h + 1] = 0xffffffffu;
                if (c0 >= 0 && c1 == c0 + 1 && (c0 & 1) == 0) {  // an aligned pair: one word
                    wa[2 * h] = reinterpret_cast<uint32_t*>(Dr + c0);
                    wv[2 * h] = acc[r][h];
                    continue;
                }
                // single entries: the other half of the word is min'd with 0xffff
                const uint32_t v2[2] = {acc[r][h] & 0xffffu, acc[r][h] >> 16};
                const int c2[2] = {c0, c1};
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    if (c2[e] < 0) continue;
                    const uintptr_t u = reinterpret_cast<uintptr_t>(Dr + c2[e]);
                    wa[2 * h + e] = reinterpret_cast<uint32_t*>(u & ~uintptr_t(3));
                    wv[2 * h + e] = (u & 2) ? (v2[e] << 16) | 0xffffu : v2[e] | 0xffff0000u;
                }
            }
            fold_min_words(wa, wv);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int row = rowv[ty * RPT + r];
        if (row < 0) continue;
        uint16_t* Dr = D + size_t(row) * ld;
        if (!HOP2 && colv[tx * 8] >= 0 && colv[tx * 8 + 7] >= 0) {
            *reinterpret_cast<uint4*>(Dr + colv[tx * 8]) = make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
            continue;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = colv[tx * 8 + q];
            if (c >= 0) Dr[c] = uint16_t(acc[r][q >> 1] >> (16 * (q & 1)));
        }
    }
}

// Shortest B part, and the units a level's work is spread over: a launch
// lasts as long as its longest unit, while every unit beyond the few hundred
// that fill the chip holds compute units that query batches on other streams
// are waiting for.
constexpr int kCoverPart = 2 * kCoverK;
constexpr long long kCoverUnits = 512;

int plan_cover_units(bool hop2, int n_list, const int* c, const int* b, const int* l8, std::vector<int4>& units) {
    auto row_tiles = [&](int i, int rpt) { return (c[i] + 16 * rpt - 1) / (16 * rpt); };
    auto col_tiles = [&](int i) {
        return hop2 ? (c[i] + kCoverOct * 8 - 1) / (kCoverOct * 8) : (l8[i] + kCoverOct - 1) / kCoverOct;
    };
    auto tiles = [&](int i, int rpt) -> long long {
        return c[i] > 0 && b[i] > 0 ? (long long)row_tiles(i, rpt) * col_tiles(i) : 0;
    };
    // rows per thread: the tallest tile that still gives 512 or more tiles
    int rpt = 1;
    for (int r = 4; r > 1; r /= 2) {
        long long t = 0;
        for (int i = 0; i < n_list; ++i) t += tiles(i, r);
        if (t >= 512) {
            rpt = r;
            break;
        }
    }
    // B parts: a node's B is split into parts of about the level's work per
    // unit at kCoverUnits units (B entries x tiles), at least kCoverPart long,
    // so a long B no longer sets the launch's duration while short ones stay
    // whole; that makes at most kCoverUnits units more than there are tiles
    long long work = 0;
    for (int i = 0; i < n_list; ++i) work += tiles(i, rpt) * b[i];
    const long long len = std::max<long long>(kCoverPart, (work + kCoverUnits - 1) / kCoverUnits);
    auto parts = [&](int i) { return int(std::min<long long>(64, (b[i] + len - 1) / len)); };
    for (int i = 0; i < n_list; ++i) {
        if (tiles(i, rpt) == 0) continue;
        const int ks = parts(i), rt = row_tiles(i, rpt), ct = col_tiles(i);
        for (int z = 0; z < ks; ++z) {
            const int kb = int((long long)b[i] * z / ks), ke = int((long long)b[i] * (z + 1) / ks);
            for (int I = 0; I < rt; ++I)
                for (int J = 0; J < ct; ++J) units.push_back(make_int4(i, I | (J << 16), kb, ke));
        }
    }
    return rpt;
}

#define HTG_COVER_LAUNCH(HOP2, RPT, L8C, L8L, NG8)                                                                 \
    HTG_LAUNCH_T((k_cover_product<HOP2, RPT>), dim3(unsigned(n_units)), dim3(256), 0, st, kt, nV, ld, node_list,    \
                 n_list, counts, cpiv, bpiv, piv_stride, units, L8C, L8L, NG8, DU, node_stride)

bool launch_cover_relax(int nV, int ld, const int* node_list, int n_list, const int* counts, const int* cpiv,
                        const int* bpiv, int piv_stride, const int4* units, int n_units, int rpt,
                        const int* live8_cnt, const int* live8_list, int ng8, uint16_t* DU, size_t node_stride,
                        hipStream_t st, KTimer kt) {
    if (n_units <= 0) return false;
    if (rpt == 4) HTG_COVER_LAUNCH(false, 4, live8_cnt, live8_list, ng8);
    else if (rpt == 2) HTG_COVER_LAUNCH(false, 2, live8_cnt, live8_list, ng8);
    else HTG_COVER_LAUNCH(false, 1, live8_cnt, live8_list, ng8);
    return true;
}

bool launch_cover_pairs(int nV, int ld, const int* node_list, int n_list, const int* counts, const int* cpiv,
                        const int* bpiv, int piv_stride, const int4* units, int n_units, int rpt, uint16_t* DU,
                        size_t node_stride, hipStream_t st, KTimer kt) {
    if (n_units <= 0) return false;
    const int* nul = nullptr;
    if (rpt == 4) HTG_COVER_LAUNCH(true, 4, nul, nul, 0);
    else if (rpt == 2) HTG_COVER_LAUNCH(true, 2, nul, nul, 0);
    else HTG_COVER_LAUNCH(true, 1, nul, nul, 0);
    return true;
}
#undef HTG_COVER_LAUNCH

// ---------------------------------------------------------------------------
// Row flags -> ascending pivot list (one workgroup per node).
__global__ __launch_bounds__(256) void k_compact_flags(int nV, uint8_t* __restrict__ flags, int* __restrict__ pivots,
                                                       int piv_stride, int* __restrict__ counts) {
    __shared__ int wsum[4];
    uint8_t* f = flags + size_t(blockIdx.x) * nV;
    int* out = pivots + size_t(blockIdx.x) * piv_stride;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int base = 0;
    for (int c0 = 0; c0 < nV; c0 += 256) {
        const int i = c0 + tid;
        const bool on = i < nV && f[i] != 0;
        const unsigned long long m = __ballot(on);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (on) {
            out[off + before] = i;
            f[i] = 0;
        }
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) counts[blockIdx.x] = base;
}

void launch_compact_flags(int nV, const int* node_list, int n_list, uint8_t* flags, int* pivots, int piv_stride,
                          int* counts, hipStream_t st) {
    (void)node_list;
    if (n_list == 0) return;
    hipLaunchKernelGGL(k_compact_flags, dim3(unsigned(n_list)), dim3(256), 0, st, nV, flags, pivots, piv_stride,
                       counts);
}

// ---------------------------------------------------------------------------
// Blocked Floyd-Warshall over a pivot subset, phase 1: the block's pivot rows.
// With C* the closure of the block's P x P pivot submatrix,
//   R[k][j] = min_{k'} C*[k][k'] (+) D[piv_k'][j]
// are the pivot rows after FW over the block (src/internal_node.cpp:199-231
// runs the same pivots one at a time).
//
// The phase can run in the same launch as phase 2 of the previous block
// (Rprev set): it then reads the pivot rows it needs as they are after that
// block, D'[k][j] = min(D[k][j], min_{k'} Rprev[k'][k] (+) Rprev[k'][j]),
// whether or not the concurrent phase-2 workgroups have stored those rows yet
// (both give the same minimum).  Pivot ports are passable, so their columns
// of Rprev lie in live octets, which the previous block wrote.
template <int P>
struct PivotLds {
    int kr[P];
    alignas(16) uint16_t C[P][P];
    alignas(16) uint32_t CdT[P][P];  // CdT[k2][k] = dup(C*[k][k2])
    uint32_t Dt[P][64];              // pivot rows at this block's 128 columns (2 per lane)
    uint16_t Rk[P][P];               // Rprev[k'][pivot a] of the previous block
};

template <int P>
__device__ __forceinline__ void pivots_unit(PivotLds<P>& sh, int nV, int ld, const int* __restrict__ node_list,
                                            const int* __restrict__ pivots, int piv_stride,
                                            const int* __restrict__ counts, int t, const uint16_t* __restrict__ DU,
                                            size_t node_stride, const int* __restrict__ live8_cnt,
                                            const int* __restrict__ live8_list, int ng8, uint16_t* __restrict__ Rbuf,
                                            const uint16_t* __restrict__ Rprev, int bx, int slot) {
    // Workgroup = 16 live octets (128 columns) of one node; wave w computes
    // output pivot rows [w*P/4, (w+1)*P/4) for the lane's 2 columns.  Columns
    // of dead octets (no passable port) stay INF in every pivot row and are
    // skipped.
    constexpr int KW = P / 4;
    const int p = min(P, counts[slot] - t * P);
    if (p <= 0) return;
    const int node = node_list[slot];
    const int nl8 = live8_cnt[node];
    if (bx * 16 >= nl8) return;
    const int tid = threadIdx.x;
    const uint16_t* D = DU + size_t(node) * node_stride;
    const int* K = pivots + size_t(slot) * piv_stride + t * P;
    if (tid < P) sh.kr[tid] = tid < p ? K[tid] : 0;
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int oct = bx * 16 + (lane >> 2);
    const bool col_ok = oct < nl8;
    const int j0 = col_ok ? 8 * live8_list[size_t(node) * ng8 + oct] + (lane & 3) * 2 : 0;
    // Issue the pivot-row loads first (wave w takes rows w, w + 4, ...); they
    // land while the pivot block is closed below.
    constexpr int RPW = P / 4;
    uint32_t pre[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i)
        pre[i] = *reinterpret_cast<const uint32_t*>(D + size_t(sh.kr[wave + 4 * i]) * ld + j0);
    const uint16_t* Rp = Rprev ? Rprev + size_t(slot) * P * ld : nullptr;
    const int pp = Rprev ? min(P, counts[slot] - (t - 1) * P) : 0;  // rows of the previous block
    constexpr int GPT = P * P / 256;
    {  // the pivot block: every gather in flight before the first LDS store
        uint16_t gv[GPT];
#pragma unroll
        for (int e = 0; e < GPT; ++e) {
            const int i = tid + 256 * e, a = i / P, c = i % P;
            gv[e] = (a < p && c < p) ? D[size_t(sh.kr[a]) * ld + sh.kr[c]] : uint16_t(a == c ? 0 : kInf);
        }
#pragma unroll
        for (int e = 0; e < GPT; ++e) {
            const int i = tid + 256 * e;
            sh.C[i / P][i % P] = gv[e];
        }
    }
    if (Rp) {
        {  // the previous block's pivot rows at this block's pivot columns
            uint16_t rv[GPT];
#pragma unroll
            for (int e = 0; e < GPT; ++e) {
                const int i = tid + 256 * e, a = i / P, c = i % P;
                rv[e] = (a < pp && c < p) ? Rp[size_t(a) * ld + sh.kr[c]] : uint16_t(kInf);
            }
#pragma unroll
            for (int e = 0; e < GPT; ++e) {
                const int i = tid + 256 * e;
                sh.Rk[i / P][i % P] = rv[e];
            }
        }
        __syncthreads();
        // the lane's pivot-row pairs: pre[i] = min(pre[i], Rk[k'][row] (+) Rprev[k'][j0 pair])
#pragma unroll 2
        for (int k = 0; k < pp; ++k) {
            const uint32_t r2 = *reinterpret_cast<const uint32_t*>(Rp + size_t(k) * ld + j0);
#pragma unroll
            for (int i = 0; i < RPW; ++i) pre[i] = relax2(pre[i], dup16(sh.Rk[k][wave + 4 * i]), r2);
        }
        // C[a][c] = min(C[a][c], min_k' Rk[k'][a] (+) Rk[k'][c]) for the
        // thread's entries (fixed c, a = a0 + (256 / P) e)
        const int c = tid % P, a0 = tid / P;
        uint16_t cv[GPT];
#pragma unroll
        for (int e = 0; e < GPT; ++e) cv[e] = sh.C[a0 + (256 / P) * e][c];
#pragma unroll 1
        for (int k = 0; k < pp; ++k) {
            const uint16_t rc = sh.Rk[k][c];
#pragma unroll
            for (int e = 0; e < GPT; ++e) {
                const uint16_t v = add1(sh.Rk[k][a0 + (256 / P) * e], rc);
                cv[e] = v < cv[e] ? v : cv[e];
            }
        }
        __syncthreads();  // every thread has read C
#pragma unroll
        for (int e = 0; e < GPT; ++e) sh.C[a0 + (256 / P) * e][c] = cv[e];
    }
    __syncthreads();
    // Floyd-Warshall over the block's pivots.  Thread t keeps EPT entries of
    // row fa in registers (packed pairs); each step reads pivot row k and its
    // own column-k entry from LDS, relaxes, and publishes its entries.  Step
    // k leaves row k and column k unchanged (C[k][k] = 0), so reads of step k
    // never race with its writes and one barrier per step suffices.
    // Rows/columns >= p hold the identity and never change.
    constexpr int EPT = P * P / 256, TPR = P / EPT;
    const int fa = tid / TPR, fc0 = (tid % TPR) * EPT;
    if constexpr (EPT >= 2) {
        constexpr int EP2 = EPT / 2;
        uint32_t cr[EP2];
        const uint32_t* Crow = reinterpret_cast<const uint32_t*>(&sh.C[fa][fc0]);
#pragma unroll
        for (int e = 0; e < EP2; ++e) cr[e] = Crow[e];
        for (int k = 0; k < p; ++k) {
            const uint32_t aik = dup16(sh.C[fa][k]);
            const uint32_t* rk = reinterpret_cast<const uint32_t*>(&sh.C[k][fc0]);
#pragma unroll
            for (int e = 0; e < EP2; ++e) cr[e] = relax2(cr[e], aik, rk[e]);
            uint32_t* Cw = reinterpret_cast<uint32_t*>(&sh.C[fa][fc0]);
#pragma unroll
            for (int e = 0; e < EP2; ++e) Cw[e] = cr[e];
            __syncthreads();
        }
    } else {
        uint16_t cr = sh.C[fa][fc0];
        for (int k = 0; k < p; ++k) {
            const uint16_t v = add1(sh.C[fa][k], sh.C[k][fc0]);
            cr = v < cr ? v : cr;
            sh.C[fa][fc0] = cr;
            __syncthreads();
        }
    }
    for (int i = tid; i < P * P; i += 256) {
        const int k2 = i / P, k = i % P;
        sh.CdT[k2][k] = dup16(sh.C[k][k2]);
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) sh.Dt[wave + 4 * i][lane] = pre[i];
    __syncthreads();
    if (!col_ok) return;
    const int kbase = wave * KW;
    uint32_t acc[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) acc[k] = kInf2;
    for (int k2 = 0; k2 < p; ++k2) {
        const uint32_t dv = sh.Dt[k2][lane];
        // CdT rows k2 >= p hold INF off the diagonal; acc[k >= p] is never stored
        const uint4* c4 = reinterpret_cast<const uint4*>(&sh.CdT[k2][kbase]);
#pragma unroll
        for (int k = 0; k < KW; k += 4) {
            const uint4 cv = c4[k / 4];
            acc[k + 0] = relax2(acc[k + 0], cv.x, dv);
            acc[k + 1] = relax2(acc[k + 1], cv.y, dv);
            acc[k + 2] = relax2(acc[k + 2], cv.z, dv);
            acc[k + 3] = relax2(acc[k + 3], cv.w, dv);
        }
    }
    uint16_t* R = Rbuf + size_t(slot) * P * ld;
#pragma unroll
    for (int k = 0; k < KW; ++k)
        if (kbase + k < p) *reinterpret_cast<uint32_t*>(R + size_t(kbase + k) * ld + j0) = acc[k];
}

// Phase 2: D := min(D, R^T (x) R) over the live octets only (rows and
// columns of dead ports are INF off the diagonal and no pivot changes them).
// A 256-thread workgroup owns a tile of 16 x 16 live octets (128 x 128
// entries), 8 x 8 elements per thread held as 32 packed registers; R's row
// slices for the tile sit in LDS (rows pre-duplicated into both u16 halves).
template <int P>
struct UpdateLds {
    uint4 Rr4[P][32];  // dup'd R[k][rows of the tile's 16 row octets]
    uint4 Rc4[P][16];  // R[k][columns of the tile's 16 column octets]
    int ro[16], co[16];  // first port of each row / column octet, -1 past the live ones
    uint32_t rmin[128], cmin[64];
};

template <int P>
__device__ __forceinline__ void update_unit(UpdateLds<P>& sh, int nV, int ld, const int* __restrict__ node_list,
                                            const int* __restrict__ counts, int t, int tiles,
                                            uint16_t* __restrict__ DU, size_t node_stride,
                                            const int* __restrict__ live8_cnt, const int* __restrict__ live8_list,
                                            int ng8, const uint16_t* __restrict__ Rbuf, int bx, int slot) {
    const int p = min(P, counts[slot] - t * P);
    if (p <= 0) return;
    const int node = node_list[slot];
    const int nl8 = live8_cnt[node];
    const int I = bx / tiles, J = bx % tiles;
    if (I * 16 >= nl8 || J * 16 >= nl8) return;  // tile past this node's live octets
    const int tid = threadIdx.x;
    const int* l8 = live8_list + size_t(node) * ng8;
    if (tid < 16) sh.ro[tid] = I * 16 + tid < nl8 ? 8 * l8[I * 16 + tid] : -1;
    else if (tid < 32) sh.co[tid - 16] = J * 16 + tid - 16 < nl8 ? 8 * l8[J * 16 + tid - 16] : -1;
    __syncthreads();
    const uint16_t* R = Rbuf + size_t(slot) * P * ld;
    // stage with 16-byte loads: item i = (pivot k, octet slot o, side); every
    // load in flight before the first LDS store
    constexpr int SPT = P * 32 / 256;
    uint4 sv[SPT];
#pragma unroll
    for (int e = 0; e < SPT; ++e) {
        const int i = tid + 256 * e;
        const int k = i >> 5, o = (i >> 1) & 15, col_side = i & 1;
        const int base = col_side ? sh.co[o] : sh.ro[o];
        sv[e] = make_uint4(kInf2, kInf2, kInf2, kInf2);
        if (k < p && base >= 0) sv[e] = *reinterpret_cast<const uint4*>(R + size_t(k) * ld + base);
    }
#pragma unroll
    for (int e = 0; e < SPT; ++e) {
        const int i = tid + 256 * e;
        const int k = i >> 5, o = (i >> 1) & 15, col_side = i & 1;
        const int base = col_side ? sh.co[o] : sh.ro[o];
        uint4 v = sv[e];
        if (base >= 0 && base + 8 > nV) {  // octet straddles nV: entries past it are padding
            uint16_t* ev = reinterpret_cast<uint16_t*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (base + j >= nV) ev[j] = kInf;
        }
        if (col_side) {
            sh.Rc4[k][o] = v;
        } else {
            sh.Rr4[k][2 * o] = make_uint4(dup16(v.x), dup16(v.x >> 16), dup16(v.y), dup16(v.y >> 16));
            sh.Rr4[k][2 * o + 1] = make_uint4(dup16(v.z), dup16(v.z >> 16), dup16(v.w), dup16(v.w >> 16));
        }
    }
    __syncthreads();
    // Minima of the block's pivot rows at the tile's rows (dup'd) and column
    // pairs: every candidate R[k][i] (+) R[k][j] is >= rmin[i] (+) cmin[j].
    {
        uint32_t m = kInf2;
        if (tid < 128) {
            for (int k = 0; k < p; ++k) m = min2(m, reinterpret_cast<const uint32_t*>(&sh.Rr4[k][0])[tid]);
            sh.rmin[tid] = m;
        } else if (tid < 192) {
            for (int k = 0; k < p; ++k) m = min2(m, reinterpret_cast<const uint32_t*>(&sh.Rc4[k][0])[tid - 128]);
            sh.cmin[tid - 128] = m;
        }
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    const int col = sh.co[tx], row0 = sh.ro[ty];
    const bool ok = col >= 0 && row0 >= 0;
    uint16_t* D = DU + size_t(node) * node_stride;
    uint32_t acc[8][4];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = row0 + r;
        uint4 v = make_uint4(kInf2, kInf2, kInf2, kInf2);
        if (ok && i < nV) v = *reinterpret_cast<const uint4*>(D + size_t(i) * ld + col);
        acc[r][0] = v.x;
        acc[r][1] = v.y;
        acc[r][2] = v.z;
        acc[r][3] = v.w;
    }
    // Exact pruning: an entry with rmin[i] (+) cmin[j] >= D[i][j] cannot
    // decrease in this block.  A wavefront none of whose entries can skips
    // the pivot loop; a thread none of whose entries can skips its stores.
    bool need = false;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint32_t rm = sh.rmin[ty * 8 + r];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const uint32_t lb = add2(rm, sh.cmin[tx * 4 + h]);
            need |= min2(lb, acc[r][h]) != acc[r][h];
        }
    }
    if (__ballot(need) != 0) {
#pragma unroll 2
        for (int k = 0; k < p; ++k) {
            const uint4 a0 = sh.Rr4[k][ty * 2], a1 = sh.Rr4[k][ty * 2 + 1];
            const uint4 c = sh.Rc4[k][tx];
            const uint32_t av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                acc[r][0] = relax2(acc[r][0], av[r], c.x);
                acc[r][1] = relax2(acc[r][1], av[r], c.y);
                acc[r][2] = relax2(acc[r][2], av[r], c.z);
                acc[r][3] = relax2(acc[r][3], av[r], c.w);
            }
        }
    }
    if (!need) return;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = row0 + r;
        if (ok && i < nV)
            *reinterpret_cast<uint4*>(D + size_t(i) * ld + col) = make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_closure_pivots(int nV, int ld, const int* __restrict__ node_list,
                                                        const int* __restrict__ pivots, int piv_stride,
                                                        const int* __restrict__ counts, int t,
                                                        const uint16_t* __restrict__ DU, size_t node_stride,
                                                        const int* __restrict__ live8_cnt,
                                                        const int* __restrict__ live8_list, int ng8,
                                                        uint16_t* __restrict__ Rbuf) {
    __shared__ __align__(16) unsigned char lds[sizeof(PivotLds<P>)];
    pivots_unit<P>(*reinterpret_cast<PivotLds<P>*>(lds), nV, ld, node_list, pivots, piv_stride, counts, t, DU,
                   node_stride, live8_cnt, live8_list, ng8, Rbuf, nullptr, int(blockIdx.x), int(blockIdx.y));
}

template <int P>
__global__ __launch_bounds__(256) void k_closure_update(int nV, int ld, const int* __restrict__ node_list,
                                                        const int* __restrict__ counts, int t, int tiles,
                                                        uint16_t* __restrict__ DU, size_t node_stride,
                                                        const int* __restrict__ live8_cnt,
                                                        const int* __restrict__ live8_list, int ng8,
                                                        const uint16_t* __restrict__ Rbuf) {
    __shared__ __align__(16) unsigned char lds[sizeof(UpdateLds<P>)];
    update_unit<P>(*reinterpret_cast<UpdateLds<P>*>(lds), nV, ld, node_list, counts, t, tiles, DU, node_stride,
                   live8_cnt, live8_list, ng8, Rbuf, int(blockIdx.x), int(blockIdx.y));
}

// Phase 2 of block t and phase 1 of block t + 1 in one launch (a 1-D grid):
// workgroups [0, n_list * tiles) are the pivot column blocks of block t + 1
// of every listed node (reading Rcur for their pivot rows, writing Rnext),
// the rest the update tiles of block t (reading Rcur).  The pivot blocks of
// all nodes come first: their Floyd-Warshall chains are the longest paths of
// the launch, so they start with the first workgroups.
template <int P>
__global__ __launch_bounds__(256) void k_closure_step(int nV, int ld, const int* __restrict__ node_list, int n_list,
                                                      const int* __restrict__ pivots, int piv_stride,
                                                      const int* __restrict__ counts, int t, int tiles,
                                                      uint16_t* __restrict__ DU, size_t node_stride,
                                                      const int* __restrict__ live8_cnt,
                                                      const int* __restrict__ live8_list, int ng8,
                                                      const uint16_t* __restrict__ Rcur,
                                                      uint16_t* __restrict__ Rnext) {
    constexpr size_t kLds = sizeof(PivotLds<P>) > sizeof(UpdateLds<P>) ? sizeof(PivotLds<P>) : sizeof(UpdateLds<P>);
    __shared__ __align__(16) unsigned char lds[kLds];
    const int lin = int(blockIdx.x), npiv = n_list * tiles;
    if (lin < npiv) {
        pivots_unit<P>(*reinterpret_cast<PivotLds<P>*>(lds), nV, ld, node_list, pivots, piv_stride, counts, t + 1,
                       DU, node_stride, live8_cnt, live8_list, ng8, Rnext, Rcur, lin % tiles, lin / tiles);
    } else {
        const int u = lin - npiv, nt2 = tiles * tiles;
        update_unit<P>(*reinterpret_cast<UpdateLds<P>*>(lds), nV, ld, node_list, counts, t, tiles, DU, node_stride,
                       live8_cnt, live8_list, ng8, Rcur, u % nt2, u / nt2);
    }
}

bool launch_closure_pivots(int P, int nV, int ld, const int* node_list, int n_list, const int* pivots, int piv_stride,
                           const int* counts, int t, const uint16_t* DU, size_t node_stride, const int* live8_cnt,
                           const int* live8_list, int ng8, int max_live8, uint16_t* Rbuf, hipStream_t st, KTimer kt) {
    if (n_list == 0 || max_live8 == 0) return false;
    const dim3 grid{unsigned((max_live8 + 15) / 16), unsigned(n_list), 1u};
#define HTG_CP(PP)                                                                                                 \
    HTG_LAUNCH_T(k_closure_pivots<PP>, grid, dim3(256), 0, st, kt, nV, ld, node_list, pivots,  \
                          piv_stride, counts, t, DU, node_stride, live8_cnt, live8_list, ng8, Rbuf)
    if (P == 16) HTG_CP(16);
    else if (P == 32) HTG_CP(32);
    else HTG_CP(64);
#undef HTG_CP
    return true;
}

bool launch_closure_update(int P, int nV, int ld, const int* node_list, int n_list, const int* counts, int t,
                           uint16_t* DU, size_t node_stride, const int* live8_cnt, const int* live8_list, int ng8,
                           int max_live8, const uint16_t* Rbuf, hipStream_t st, KTimer kt) {
    if (n_list == 0 || max_live8 == 0) return false;
    const int tiles = (max_live8 + 15) / 16;
    const dim3 grid{unsigned(tiles * tiles), unsigned(n_list), 1u};
#define HTG_CU(PP)                                                                                             \
    HTG_LAUNCH_T(k_closure_update<PP>, grid, dim3(256), 0, st, kt, nV, ld, node_list, counts, t, \
                          tiles, DU, node_stride, live8_cnt, live8_list, ng8, Rbuf)
    if (P == 16) HTG_CU(16);
    else if (P == 32) HTG_CU(32);
    else HTG_CU(64);
#undef HTG_CU
    return true;
}

bool launch_closure_step(int P, int nV, int ld, const int* node_list, int n_list, const int* pivots, int piv_stride,
                         const int* counts, int t, uint16_t* DU, size_t node_stride, const int* live8_cnt,
                         const int* live8_list, int ng8, int max_live8, const uint16_t* Rcur, uint16_t* Rnext,
                         hipStream_t st, KTimer kt) {
    if (n_list == 0 || max_live8 == 0 || Rcur == Rnext) return false;
    const int tiles = (max_live8 + 15) / 16;
    const long long blocks = (long long)n_list * (tiles * tiles + tiles);
    if (blocks > 0x7fffffffll) return false;
    const dim3 grid{unsigned(blocks), 1u, 1u};
#define HTG_CS(PP)                                                                                              \
    HTG_LAUNCH_T(k_closure_step<PP>, grid, dim3(256), 0, st, kt, nV, ld, node_list, n_list, pivots, piv_stride, \
                 counts, t, tiles, DU, node_stride, live8_cnt, live8_list, ng8, Rcur, Rnext)
    if (P == 16) HTG_CS(16);
    else if (P == 32) HTG_CS(32);
    else HTG_CS(64);
#undef HTG_CS
    return true;
}

// ---------------------------------------------------------------------------
// S_U := D_U[ext][ext] (src/internal_node.cpp:233-242, :415-433).
__global__ __launch_bounds__(256) void k_project_su(int ld, int ne, const int* __restrict__ ext_idx,
                                                    const int* __restrict__ node_list,
                                                    const uint16_t* __restrict__ DU, size_t node_stride,
                                                    uint16_t* __restrict__ SU, int* __restrict__ changed) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)ne * ne) return;
    const int e1 = int(gid / ne), e2 = int(gid % ne);
    const int node = node_list[blockIdx.y];
    const uint16_t v = DU[size_t(node) * node_stride + size_t(ext_idx[e1]) * ld + ext_idx[e2]];
    uint16_t* dst = SU + size_t(node) * ne * ne + gid;
    if (changed && v < *dst) changed[blockIdx.y] = 1;
    *dst = v;
}

void launch_project_su(int nV, int ld, int ne, const int* ext_idx, const int* node_list, int n_list,
                       const uint16_t* DU, size_t node_stride, uint16_t* SU, int* changed, hipStream_t st) {
    (void)nV;
    if (n_list == 0) return;
    const long long total = (long long)ne * ne;
    dim3 grid(unsigned((total + 255) / 256), unsigned(n_list));
    hipLaunchKernelGGL(k_project_su, grid, dim3(256), 0, st, ld, ne, ext_idx, node_list, DU, node_stride, SU,
                       changed);
}

// S_U refresh of an update over live octet pairs: S_U[e1][e2] = D_U[v1][v2]
// for exterior ports e = vu_ext[v] of live octets (dead exterior entries
// never change); skips nodes without pivots (their D_U did not change);
// changed[i] = 1 when an entry decreased.
__global__ __launch_bounds__(256) void k_project_su_live(int nV, int ld, int ne, const int* __restrict__ vu_ext,
                                                         const int* __restrict__ node_list,
                                                         const int* __restrict__ counts,
                                                         const int* __restrict__ live8_cnt,
                                                         const int* __restrict__ live8_list, int ng8,
                                                         const uint16_t* __restrict__ DU, size_t node_stride,
                                                         uint16_t* __restrict__ SU, int* __restrict__ changed) {
    const int i = blockIdx.y;
    if (counts[i] == 0) return;
    const int node = node_list[i];
    const int n8 = live8_cnt[node];
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)n8 * n8 * 8) return;
    const int* l8 = live8_list + size_t(node) * ng8;
    const long long pr = t >> 3;
    const int v1 = 8 * l8[pr / n8] + int(t & 7);
    if (v1 >= nV) return;
    const int e1 = vu_ext[v1];
    if (e1 < 0) return;
    const int c0 = 8 * l8[pr % n8];
    const uint16_t* Drow = DU + size_t(node) * node_stride + size_t(v1) * ld;
    uint16_t* Srow = SU + size_t(node) * ne * ne + size_t(e1) * ne;
    bool dec = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int v2 = c0 + k;
        if (v2 >= nV) break;
        const int e2 = vu_ext[v2];
        if (e2 < 0) continue;
        const uint16_t v = Drow[v2];
        dec |= v < Srow[e2];
        Srow[e2] = v;
    }
    if (dec) changed[i] = 1;
}

void launch_project_su_live(int nV, int ld, int ne, const int* vu_ext, const int* node_list, const int* counts,
                            int n_list, const int* live8_cnt, const int* live8_list, int ng8, int max_live8,
                            const uint16_t* DU, size_t node_stride, uint16_t* SU, int* changed, hipStream_t st) {
    if (n_list == 0 || max_live8 == 0) return;
    const long long total = (long long)max_live8 * max_live8 * 8;
    const dim3 grid{unsigned((total + 255) / 256), unsigned(n_list), 1u};
    hipLaunchKernelGGL(k_project_su_live, grid, dim3(256), 0, st, nV, ld, ne, vu_ext, node_list, counts, live8_cnt,
                       live8_list, ng8, DU, node_stride, SU, changed);
}

// ---------------------------------------------------------------------------
// Table copies for bringing a snapshot copy level with its twin.
// Whole items (leaf tables): item t of the list is the contiguous block of
// elems u16 at src + t*elems; vec: elems is a multiple of 8 (16-byte copies).
template <bool VEC>
__global__ __launch_bounds__(256) void k_copy_items(const int* __restrict__ items, int n_items, size_t elems,
                                                    const uint16_t* __restrict__ src, uint16_t* __restrict__ dst) {
    const size_t units = VEC ? elems / 8 : elems;
    const size_t total = units * size_t(n_items);
    for (size_t g = size_t(blockIdx.x) * 256 + threadIdx.x; g < total; g += size_t(gridDim.x) * 256) {
        const size_t it = g / units, u = g - it * units;
        const size_t base = size_t(items[it]) * elems;
        if (VEC)
            reinterpret_cast<uint4*>(dst + base)[u] = reinterpret_cast<const uint4*>(src + base)[u];
        else
            dst[base + u] = src[base + u];
    }
}

void launch_copy_items(const int* items, int n_items, size_t elems, const uint16_t* src, uint16_t* dst,
                       hipStream_t st) {
    if (n_items == 0 || elems == 0) return;
    const bool vec = elems % 8 == 0;
    const size_t units = vec ? elems / 8 : elems;
    const unsigned blocks = unsigned(std::min<size_t>((units * size_t(n_items) + 255) / 256, kBackgroundWG));
    if (vec)
        hipLaunchKernelGGL(k_copy_items<true>, dim3(blocks), dim3(256), 0, st, items, n_items, elems, src, dst);
    else
        hipLaunchKernelGGL(k_copy_items<false>, dim3(blocks), dim3(256), 0, st, items, n_items, elems, src, dst);
}

// Live rows of node tables: for each listed node, the D_U rows of its live
// octets (ld u16 each) and the S_U rows of the exterior ports among them (ne
// u16 each, ne a multiple of 4).  Rows of dead ports never change.  Block =
// one row (blockIdx.x = live row slot, blockIdx.y = node).
__global__ __launch_bounds__(256) void k_copy_live_rows(const int* __restrict__ nodes, int nV, int ld,
                                                        size_t node_stride, const int* __restrict__ live8_cnt,
                                                        const int* __restrict__ live8_list, int ng8, int ne,
                                                        const int* __restrict__ vu_ext,
                                                        const uint16_t* __restrict__ srcDU, uint16_t* __restrict__ dstDU,
                                                        const uint16_t* __restrict__ srcSU,
                                                        uint16_t* __restrict__ dstSU, int rows_per_node,
                                                        int n_nodes) {
    // work item = (node slot, live row slot); a capped grid strides over them
    for (int w = blockIdx.x; w < rows_per_node * n_nodes; w += gridDim.x) {
        const int node = nodes[w / rows_per_node];
        const int t = w % rows_per_node;
        if (t >= 8 * live8_cnt[node]) continue;
        const int v = 8 * live8_list[size_t(node) * ng8 + t / 8] + t % 8;
        if (v >= nV) continue;
        const size_t roff = size_t(node) * node_stride + size_t(v) * ld;
        const uint4* s4 = reinterpret_cast<const uint4*>(srcDU + roff);
        uint4* d4 = reinterpret_cast<uint4*>(dstDU + roff);
        for (int u = threadIdx.x; u < ld / 8; u += 256) d4[u] = s4[u];
        const int e = vu_ext[v];
        if (e < 0) continue;
        const size_t soff = (size_t(node) * ne + e) * ne;
        const uint2* s2 = reinterpret_cast<const uint2*>(srcSU + soff);
        uint2* d2 = reinterpret_cast<uint2*>(dstSU + soff);
        for (int u = threadIdx.x; u < ne / 4; u += 256) d2[u] = s2[u];
    }
}

void launch_copy_live_rows(const int* nodes, int n_nodes, int nV, int ld, size_t node_stride, const int* live8_cnt,
                           const int* live8_list, int ng8, int max_live8, int ne, const int* vu_ext,
                           const uint16_t* srcDU, uint16_t* dstDU, const uint16_t* srcSU, uint16_t* dstSU,
                           hipStream_t st) {
    if (n_nodes == 0 || max_live8 == 0) return;
    const int rows = 8 * max_live8;
    const unsigned blocks = unsigned(std::min<long long>((long long)rows * n_nodes, kBackgroundWG));
    hipLaunchKernelGGL(k_copy_live_rows, dim3(blocks), dim3(256), 0, st, nodes, nV, ld, node_stride, live8_cnt,
                       live8_list, ng8, ne, vu_ext, srcDU, dstDU, srcSU, dstSU, rows, n_nodes);
}

// Copy by a kernel between device memory and pinned host memory (the device
// reads or writes the host pages directly).  Small per-batch transfers of the
// query path go this way: a runtime copy can wait behind other streams' work.
__global__ __launch_bounds__(256) void k_copy_bytes(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    size_t n16, const uint8_t* __restrict__ src_tail,
                                                    uint8_t* __restrict__ dst_tail, int tail) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n16; i += size_t(gridDim.x) * 256) dst[i] = src[i];
    if (blockIdx.x == 0 && int(threadIdx.x) < tail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

void launch_copy_bytes(const void* src, void* dst, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    const size_t n16 = bytes / 16;
    const int tail = int(bytes - n16 * 16);
    const unsigned blocks = unsigned(std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 64)));
    hipLaunchKernelGGL(k_copy_bytes, dim3(blocks), dim3(256), 0, st, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16, static_cast<const uint8_t*>(src) + n16 * 16,
                       static_cast<uint8_t*>(dst) + n16 * 16, tail);
}

// Byte fill (a kernel rather than hipMemsetAsync, so that it can sit in a
// captured graph).
__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* __restrict__ p, size_t n, uint8_t v) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) p[i] = v;
}

void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t st) {
    if (n == 0) return;
    const unsigned blocks = unsigned(std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks), dim3(256), 0, st, p, n, v);
}

__global__ void k_set_pass(uint8_t* pass, int S, const int2* cells, int n, uint8_t v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) pass[size_t(cells[i].x) * S + cells[i].y] = v;
}

void launch_set_pass(uint8_t* pass, int S, const int2* cells, int n, hipStream_t st, uint8_t v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_set_pass, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, pass, S, cells, n, v);
}

}  // namespace htg
