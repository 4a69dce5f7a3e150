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

#include "geometry.hpp"
#include "kernels.hpp"

namespace htg {

typedef unsigned short us2q __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t qrelax2(uint32_t acc, uint32_t a, uint32_t b) {
    us2q x = __builtin_bit_cast(us2q, a), y = __builtin_bit_cast(us2q, b), z = __builtin_bit_cast(us2q, acc);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(z, __builtin_elementwise_add_sat(x, y)));
}
__device__ __forceinline__ uint32_t qdup(uint32_t v) { return (v & 0xffffu) | (v << 16); }

// Output init (src/minplus_oracle.cpp:940-965, same-leaf pairs :1027-1048).
__global__ __launch_bounds__(256) void k_q_init(const int* __restrict__ q_k, const long long* __restrict__ q_out,
                                                const int* __restrict__ q_vp, const int2* __restrict__ vp_rc,
                                                const uint8_t* __restrict__ vp_valid, int b, int nl,
                                                const uint16_t* __restrict__ leafD, uint16_t* __restrict__ out) {
    const int Q = blockIdx.y;
    const int k = q_k[Q];
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
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
                const int n = b * b;
                const size_t leaf = size_t(a.x / b) * nl + a.y / b;
                val = leafD[leaf * n * n + size_t((a.x % b) * b + a.y % b) * n + (c.x % b) * b + c.y % b];
            }
        }
    }
    out[q_out[Q] + idx] = val;
}

void launch_q_init(int nq, const int* q_k, const long long* q_out, const int* q_vp, const int2* vp_rc,
                   const uint8_t* vp_valid, int b, int nl, const uint16_t* leafD, uint16_t* out, int kmax,
                   hipStream_t st) {
    if (nq == 0 || kmax == 0) return;
    const long long total = (long long)kmax * kmax;
    dim3 grid(unsigned((total + 255) / 256), unsigned(nq));
    hipLaunchKernelGGL(k_q_init, grid, dim3(256), 0, st, q_k, q_out, q_vp, vp_rc, vp_valid, b, nl, leafD, out);
}

// Level-0 lift: leaf D row of the viewpoint at the leaf's ports
// (src/minplus_oracle.cpp:1000-1019).
__global__ __launch_bounds__(256) void k_q_leaf(int n_slots, const int2* __restrict__ slot_rc, int b, int nl,
                                                const uint16_t* __restrict__ leafD, uint16_t* __restrict__ X0) {
    const int pl = 4 * b - 4;
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)n_slots * pl) return;
    const int s = int(gid / pl), p = int(gid % pl);
    const int2 rc = slot_rc[s];
    const int n = b * b;
    int pr, pcc;
    port_rc(p, b, pr, pcc);
    const size_t leaf = size_t(rc.x / b) * nl + rc.y / b;
    X0[gid] = leafD[leaf * n * n + size_t((rc.x % b) * b + rc.y % b) * n + pr * b + pcc];
}

void launch_q_leaf(int n, const int2* slot_rc, int b, int nl, const uint16_t* leafD, uint16_t* X0, hipStream_t st) {
    if (n == 0) return;
    const long long total = (long long)n * (4 * b - 4);
    hipLaunchKernelGGL(k_q_leaf, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, n, slot_rc, b, nl, leafD, X0);
}

// Projection Z (see file comment).  One wavefront owns 512 consecutive Z
// columns (8 per lane, one 16-byte D load per seam row) for up to M
// viewpoints of one (W, q) work item; the M viewpoints' X values for a chunk
// of seam rows are staged in LDS, duplicated into both u16 halves.
template <int M, bool GATHER>
__global__ __launch_bounds__(64) void k_q_z(int ld, int pc, const uint16_t* __restrict__ DU, size_t node_stride,
                                            const int* __restrict__ seam_list, const int* __restrict__ seam_off,
                                            const int* __restrict__ cols, int ncols, int ldz,
                                            const int4* __restrict__ items, const uint16_t* __restrict__ X,
                                            uint16_t* __restrict__ Z) {
    constexpr int ACH = 64;
    __shared__ uint32_t xs[M][ACH];
    __shared__ int rows[ACH];
    const int4 it = items[blockIdx.x];
    const int W = it.x, q = it.y, s0 = it.z, m = it.w;
    const int tid = threadIdx.x;
    const int c0 = (blockIdx.y * 64 + tid) * 8;
    const bool active = c0 < ncols;
    const uint16_t* D = DU + size_t(W) * node_stride;
    const int* seam = seam_list + seam_off[q];
    const int ns = seam_off[q + 1] - seam_off[q];
    int cidx[8];
    if (GATHER) {
#pragma unroll
        for (int e = 0; e < 8; ++e) cidx[e] = cols[(active && c0 + e < ncols) ? c0 + e : 0];
    }
    uint32_t acc[M][4];
#pragma unroll
    for (int v = 0; v < M; ++v) acc[v][0] = acc[v][1] = acc[v][2] = acc[v][3] = kInf2;
    for (int a0 = 0; a0 < ns; a0 += ACH) {
        const int na = min(ACH, ns - a0);
        __syncthreads();
        if (tid < na) rows[tid] = q * pc + seam[a0 + tid];
#pragma unroll
        for (int v = 0; v < M; ++v)
            xs[v][tid] = (v < m && tid < na) ? qdup(X[size_t(s0 + v) * pc + seam[a0 + tid]]) : kInf2;
        __syncthreads();
        if (!active) continue;
        for (int a = 0; a < na; ++a) {
            const uint16_t* drow = D + size_t(rows[a]) * ld;
            uint32_t d[4];
            if (!GATHER) {
                const uint4 v4 = *reinterpret_cast<const uint4*>(drow + c0);
                d[0] = v4.x;
                d[1] = v4.y;
                d[2] = v4.z;
                d[3] = v4.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) d[e] = uint32_t(drow[cidx[2 * e]]) | (uint32_t(drow[cidx[2 * e + 1]]) << 16);
            }
#pragma unroll
            for (int v = 0; v < M; ++v) {
                if (v < m) {
                    const uint32_t x = xs[v][a];
                    acc[v][0] = qrelax2(acc[v][0], x, d[0]);
                    acc[v][1] = qrelax2(acc[v][1], x, d[1]);
                    acc[v][2] = qrelax2(acc[v][2], x, d[2]);
                    acc[v][3] = qrelax2(acc[v][3], x, d[3]);
                }
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int v = 0; v < M; ++v)
        if (v < m)
            *reinterpret_cast<uint4*>(Z + size_t(s0 + v) * ldz + c0) = make_uint4(acc[v][0], acc[v][1], acc[v][2], acc[v][3]);
}

void launch_q_z(int M, int nV, int ld, int pc, const uint16_t* DU, size_t node_stride, const int* seam_list,
                const int* seam_off, const int* cols, int ncols, int ldz, bool gather, const int4* items, int n_items,
                const uint16_t* X, uint16_t* Z, hipStream_t st) {
    (void)nV;
    if (n_items == 0) return;
    dim3 grid(unsigned(n_items), unsigned((ncols + 511) / 512));
#define HTG_QZ(MM)                                                                                                   \
    do {                                                                                                             \
        if (gather)                                                                                                  \
            hipLaunchKernelGGL((k_q_z<MM, true>), grid, dim3(64), 0, st, ld, pc, DU, node_stride, seam_list, seam_off, \
                               cols, ncols, ldz, items, X, Z);                                                      \
        else                                                                                                         \
            hipLaunchKernelGGL((k_q_z<MM, false>), grid, dim3(64), 0, st, ld, pc, DU, node_stride, seam_list,        \
                               seam_off, cols, ncols, ldz, items, X, Z);                                             \
    } while (0)
    if (M <= 1) HTG_QZ(1);
    else if (M <= 2) HTG_QZ(2);
    else if (M <= 4) HTG_QZ(4);
    else if (M <= 8) HTG_QZ(8);
    else HTG_QZ(16);
#undef HTG_QZ
}

// Meet of one row viewpoint against every other viewpoint of its (query, W)
// group.  Pairs with i == j keep the diagonal from the init kernel.
__global__ __launch_bounds__(64) void k_q_meet(int pc, int cs, int g, const int* __restrict__ seam_list,
                                               const int* __restrict__ seam_off, const int* __restrict__ zpos,
                                               int ldz, const int2* __restrict__ row_items,
                                               const int4* __restrict__ groups, const int* __restrict__ glist,
                                               const int2* __restrict__ slot_rc, const int* __restrict__ slot_q,
                                               const int* __restrict__ slot_i, const int* __restrict__ q_k,
                                               const long long* __restrict__ q_out, const uint16_t* __restrict__ X,
                                               const uint16_t* __restrict__ Z, uint16_t* __restrict__ out) {
    const int2 ri = row_items[blockIdx.x];
    const int4 grp = groups[ri.x];  // (query, unused, list offset, count)
    const int* lst = glist + grp.z;
    const int n = grp.w;
    const int sa = lst[ri.y];
    const int Q = slot_q[sa];
    const int k = q_k[Q];
    uint16_t* orow = out + q_out[Q] + size_t(slot_i[sa]) * k;
    const uint16_t* zrow = Z + size_t(sa) * ldz;
    for (int bb = threadIdx.x; bb < n; bb += 64) {
        if (bb == ri.y) continue;
        const int sb = lst[bb];
        const int2 rc = slot_rc[sb];
        const int qb = ((rc.x / cs) % g) * g + (rc.y / cs) % g;
        const int off = seam_off[qb], ns = seam_off[qb + 1] - off;
        const uint16_t* xrow = X + size_t(sb) * pc;
        uint32_t best = kInf;
        for (int s = 0; s < ns; ++s) {
            const uint32_t v = uint32_t(zrow[zpos[off + s]]) + xrow[seam_list[off + s]];
            best = v < best ? v : best;
        }
        uint16_t* o = orow + slot_i[sb];
        if (best < *o) *o = uint16_t(best);
    }
}

void launch_q_meet(int pc, int cs, int g, const int* seam_list, const int* seam_off, const int* zpos, int ldz,
                   const int2* row_items, int n_rows, const int4* groups, const int* glist, const int2* slot_rc,
                   const int* slot_q, const int* slot_i, const int* q_k, const long long* q_out, const uint16_t* X,
                   const uint16_t* Z, uint16_t* out, hipStream_t st) {
    if (n_rows == 0) return;
    hipLaunchKernelGGL(k_q_meet, dim3(unsigned(n_rows)), dim3(64), 0, st, pc, cs, g, seam_list, seam_off, zpos, ldz,
                       row_items, groups, glist, slot_rc, slot_q, slot_i, q_k, q_out, X, Z, out);
}

// Lift to W's exterior ports (src/minplus_oracle.cpp:1113-1154).
__global__ __launch_bounds__(256) void k_q_lift(int n, int ne, const int* __restrict__ ext_zcol,
                                                const int* __restrict__ ext_child, const int* __restrict__ ext_cport,
                                                int pc, int cs, int g, int ldz, const int2* __restrict__ slot_rc,
                                                const uint16_t* __restrict__ Xprev, const uint16_t* __restrict__ Z,
                                                uint16_t* __restrict__ Xnext) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)n * ne) return;
    const int s = int(gid / ne), e = int(gid % ne);
    const int2 rc = slot_rc[s];
    const int q = ((rc.x / cs) % g) * g + (rc.y / cs) % g;
    uint16_t v = Z[size_t(s) * ldz + ext_zcol[e]];
    if (ext_child[e] == q) {
        const uint16_t d = Xprev[size_t(s) * pc + ext_cport[e]];
        v = d < v ? d : v;
    }
    Xnext[gid] = v;
}

void launch_q_lift(int n, int ne, const int* ext_zcol, const int* ext_child, const int* ext_cport, int pc, int cs,
                   int g, int ldz, const int2* slot_rc, const uint16_t* Xprev, const uint16_t* Z, uint16_t* Xnext,
                   hipStream_t st) {
    if (n == 0) return;
    const long long total = (long long)n * ne;
    hipLaunchKernelGGL(k_q_lift, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, n, ne, ext_zcol, ext_child,
                       ext_cport, pc, cs, g, ldz, slot_rc, Xprev, Z, Xnext);
}

}  // namespace htg
