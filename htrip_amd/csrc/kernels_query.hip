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
__device__ __forceinline__ uint32_t qadd2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(us2q, a),
                                                                      __builtin_bit_cast(us2q, b)));
}
__device__ __forceinline__ uint32_t qmin2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2q, a),
                                                                  __builtin_bit_cast(us2q, b)));
}

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

// Projection Z (see file comment), split over the seam rows: work item
// (W, q, first slot, count), column tile (256 columns, 4 per lane) and row
// chunk blockIdx.z.  Each wavefront keeps M x 4 packed accumulators and reads
// one 8-byte slice of a D_W row per seam row; the M viewpoints' X values for
// a chunk of rows sit in LDS duplicated into both u16 halves.  Partial minima
// go to Zp[chunk]; k_q_zred folds the chunks.
template <int M>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_q_z(int ld, int pc, const uint16_t* __restrict__ DU, size_t node_stride,
                                            const int* __restrict__ seam_list, const int* __restrict__ seam_off,
                                            const int* __restrict__ cols, int ncols, int ldp,
                                            const int4* __restrict__ items, int ksplit,
                                            const uint16_t* __restrict__ X, uint16_t* __restrict__ Zp,
                                            size_t part_stride) {
    constexpr int ACH = 64;  // seam rows staged per round
    constexpr int PF = 8;    // D rows in flight per lane
    // Rows past the item's count (v >= m) and past the chunk's end carry INF
    // in xs: min(acc, INF (+) d) leaves acc unchanged, so the loop runs
    // branch-free over full tiles.
    __shared__ __align__(16) uint32_t xs[ACH][M];  // row-major by seam row: one row's M values are contiguous
    __shared__ unsigned roff[ACH];
    const int4 it = items[blockIdx.x];
    const int W = it.x, q = it.y, s0 = it.z, m = it.w;
    const int tid = threadIdx.x;
    const int c0 = (blockIdx.y * 64 + tid) * 4;
    const bool active = c0 < ncols;
    // Projection columns come in aligned groups of 4 consecutive V_U ports,
    // so one 8-byte load per seam row feeds the lane's 4 columns.
    const int base = active ? cols[c0] : 0;
    const uint16_t* D = DU + size_t(W) * node_stride + base;
    const int* seam = seam_list + seam_off[q];
    const int ns = seam_off[q + 1] - seam_off[q];
    const int a_lo = int((long long)ns * blockIdx.z / ksplit), a_hi = int((long long)ns * (blockIdx.z + 1) / ksplit);
    uint32_t acc[M][2];
#pragma unroll
    for (int v = 0; v < M; ++v) acc[v][0] = acc[v][1] = kInf2;
    for (int a0 = a_lo; a0 < a_hi; a0 += ACH) {
        const int na = min(ACH, a_hi - a0);
        __syncthreads();
        const int sr = tid < na ? seam[a0 + tid] : seam[0];
        roff[tid] = unsigned(q * pc + sr) * unsigned(ld);
#pragma unroll
        for (int v = 0; v < M; ++v)
            xs[tid][v] = (v < m && tid < na) ? qdup(X[size_t(s0 + v) * pc + sr]) : kInf2;
        __syncthreads();
        if (!active) continue;
        for (int a = 0; a < na; a += PF) {
            uint2 dv[PF];
#pragma unroll
            for (int u = 0; u < PF; ++u) dv[u] = *reinterpret_cast<const uint2*>(D + roff[a + u]);
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const uint32_t* xr = xs[a + u];
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
    if (!active) return;
    uint16_t* P = Zp + size_t(blockIdx.z) * part_stride;
#pragma unroll
    for (int v = 0; v < M; ++v)
        if (v < m) *reinterpret_cast<uint2*>(P + size_t(s0 + v) * ldp + c0) = make_uint2(acc[v][0], acc[v][1]);
}

void launch_q_z(int M, int ld, int pc, const uint16_t* DU, size_t node_stride, const int* seam_list,
                const int* seam_off, const int* cols, int ncols, int ldp, const int4* items, int n_items, int ksplit,
                const uint16_t* X, uint16_t* Zp, size_t part_stride, hipStream_t st) {
    if (n_items == 0) return;
    dim3 grid(unsigned(n_items), unsigned((ncols + 255) / 256), unsigned(ksplit));
#define HTG_QZ(MM)                                                                                                  \
    hipLaunchKernelGGL((k_q_z<MM>), grid, dim3(64), 0, st, ld, pc, DU, node_stride, seam_list, seam_off, cols, ncols, \
                       ldp, items, ksplit, X, Zp, part_stride)
    if (M <= 1) HTG_QZ(1);
    else if (M <= 2) HTG_QZ(2);
    else if (M <= 4) HTG_QZ(4);
    else if (M <= 8) HTG_QZ(8);
    else HTG_QZ(16);
#undef HTG_QZ
}

// Fold the row chunks and lay Z out as contiguous per-child seam blocks
// (then the remaining exterior ports): Z[slot][pos] = min_c Zp[c][slot][src[pos]],
// INF where src[pos] < 0 (block padding).
__global__ __launch_bounds__(256) void k_q_zred(int n, int nz, int ldz, const int* __restrict__ src, int ldp,
                                                int ksplit, const uint16_t* __restrict__ Zp, size_t part_stride,
                                                uint16_t* __restrict__ Z) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)n * nz) return;
    const int s = int(gid / nz), pos = int(gid % nz);
    const int col = src[pos];
    uint16_t v = kInf;
    if (col >= 0) {
        const uint16_t* p = Zp + size_t(s) * ldp + col;
        for (int c = 0; c < ksplit; ++c) {
            const uint16_t x = p[size_t(c) * part_stride];
            v = x < v ? x : v;
        }
    }
    Z[size_t(s) * ldz + pos] = v;
}

void launch_q_zred(int n, int nz, int ldz, const int* src, int ldp, int ksplit, const uint16_t* Zp,
                   size_t part_stride, uint16_t* Z, hipStream_t st) {
    if (n == 0) return;
    const long long total = (long long)n * nz;
    hipLaunchKernelGGL(k_q_zred, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, n, nz, ldz, src, ldp, ksplit,
                       Zp, part_stride, Z);
}

// Meet: m_ij = min_{b in seam(q_j)} Z_i[b] (+) Z_j[b].  Every candidate is the
// length of a real walk (i -> b -> j inside W), and Z_j[b] <= X_j[b] on q_j's
// seam ports (D_W[b][b] = 0), so the candidate is never above the
// reference's X_i-projection (+) X_j meet: the minimum over levels is exact.
// One 256-thread workgroup computes a 16 x 16 tile of (i, j) pairs of one
// (query, W, q2) item, streaming the seam block through LDS as packed pairs.
__global__ __launch_bounds__(256) void k_q_meet(const int* __restrict__ tiles, const int4* __restrict__ groups,
                                                const int* __restrict__ glist, const int* __restrict__ zoff,
                                                const int* __restrict__ zlen, int ldz,
                                                const int* __restrict__ slot_i, const int* __restrict__ q_k,
                                                const long long* __restrict__ q_out, const uint16_t* __restrict__ Z,
                                                uint16_t* __restrict__ out) {
    constexpr int CH = 256;  // packed pairs per round (512 columns)
    __shared__ uint32_t A[16][CH + 1];
    __shared__ uint32_t Bm[16][CH + 1];
    const int* t = tiles + size_t(blockIdx.x) * 8;
    const int g = t[0], q2 = t[1], i0 = t[2], j0 = t[3], jend = t[4];
    const int4 grp = groups[g];
    const int* lst = glist + grp.z;
    const int n = grp.w;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int off = zoff[q2], len2 = zlen[q2] / 2;  // block length is even
    const int ni = min(16, n - i0), nj = min(16, jend - j0);
    // row pointers of this tile (packed-pair view of Z)
    const int r = tid >> 4, lane16 = tid & 15;
    const uint32_t* za = r < ni ? reinterpret_cast<const uint32_t*>(Z + size_t(lst[i0 + r]) * ldz + off) : nullptr;
    const uint32_t* zb = r < nj ? reinterpret_cast<const uint32_t*>(Z + size_t(lst[j0 + r]) * ldz + off) : nullptr;
    uint32_t acc = kInf2;
    for (int c0 = 0; c0 < len2; c0 += CH) {
        const int cn = min(CH, len2 - c0);
        __syncthreads();
        // 16 rows x CH pairs for each side; thread (r, lane16) loads its row's
        // columns lane16, lane16 + 16, ... (CH / 16 independent loads).
#pragma unroll
        for (int u = 0; u < CH / 16; ++u) {
            const int cc = lane16 + 16 * u;
            A[r][cc] = (za && cc < cn) ? za[c0 + cc] : kInf2;
            Bm[r][cc] = (zb && cc < cn) ? zb[c0 + cc] : kInf2;
        }
        __syncthreads();
#pragma unroll 8
        for (int cc = 0; cc < cn; ++cc) acc = qrelax2(acc, A[ti][cc], Bm[tj][cc]);
    }
    if (ti >= ni || tj >= nj) return;
    const int si = lst[i0 + ti], sj = lst[j0 + tj];
    if (si == sj) return;
    const uint16_t best = min(uint16_t(acc & 0xffffu), uint16_t(acc >> 16));
    const int Q = grp.x;
    uint16_t* o = out + q_out[Q] + size_t(slot_i[si]) * q_k[Q] + slot_i[sj];
    if (best < *o) *o = best;
}

void launch_q_meet(const int* tiles, int n_tiles, const int4* groups, const int* glist, const int* zoff,
                   const int* zlen, int ldz, const int* slot_i, const int* q_k, const long long* q_out,
                   const uint16_t* Z, uint16_t* out, hipStream_t st) {
    if (n_tiles == 0) return;
    hipLaunchKernelGGL(k_q_meet, dim3(unsigned(n_tiles)), dim3(256), 0, st, tiles, groups, glist, zoff, zlen, ldz,
                       slot_i, q_k, q_out, Z, out);
}

// Lift to W's exterior ports (src/minplus_oracle.cpp:1113-1154).
__global__ __launch_bounds__(256) void k_q_lift(int n, int ne, const int* __restrict__ ext_zpos,
                                                const int* __restrict__ ext_child, const int* __restrict__ ext_cport,
                                                int pc, int cs, int g, int ldz, const int2* __restrict__ slot_rc,
                                                const uint16_t* __restrict__ Xprev, const uint16_t* __restrict__ Z,
                                                uint16_t* __restrict__ Xnext) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long long)n * ne) return;
    const int s = int(gid / ne), e = int(gid % ne);
    const int2 rc = slot_rc[s];
    const int q = ((rc.x / cs) % g) * g + (rc.y / cs) % g;
    uint16_t v = Z[size_t(s) * ldz + ext_zpos[e]];
    if (ext_child[e] == q) {
        const uint16_t d = Xprev[size_t(s) * pc + ext_cport[e]];
        v = d < v ? d : v;
    }
    Xnext[gid] = v;
}

void launch_q_lift(int n, int ne, const int* ext_zpos, const int* ext_child, const int* ext_cport, int pc, int cs,
                   int g, int ldz, const int2* slot_rc, const uint16_t* Xprev, const uint16_t* Z, uint16_t* Xnext,
                   hipStream_t st) {
    if (n == 0) return;
    const long long total = (long long)n * ne;
    hipLaunchKernelGGL(k_q_lift, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, n, ne, ext_zpos, ext_child,
                       ext_cport, pc, cs, g, ldz, slot_rc, Xprev, Z, Xnext);
}

}  // namespace htg
