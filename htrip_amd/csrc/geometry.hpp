// Hierarchy geometry shared by host and device code.
//
// A map of side S = b * g^H is cut into (S/b)^2 leaf tiles of side b; level l
// (1..H) groups g x g tiles of level l-1 into one node.  Port order, V_U order
// and the exterior-port list are the reference's (src/leaf_tile.cpp:11-30,
// src/internal_node.cpp:13-56): every table index in this library means what
// it means in the reference, so tables compare element for element.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define HTG_HD __host__ __device__ __forceinline__

namespace htg {

constexpr uint16_t kInf = 60000;
constexpr uint32_t kInf2 = 0xEA60EA60u;  // (INF, INF) packed

// Position of port p on the perimeter of a side-s tile, relative to its
// top-left cell: top row, bottom row, left column, right column
// (src/leaf_tile.cpp:11-30).
HTG_HD void port_rc(int p, int s, int& r, int& c) {
    if (p < s) { r = 0; c = p; }
    else if (p < 2 * s) { r = s - 1; c = p - s; }
    else if (p < 3 * s - 2) { r = p - 2 * s + 1; c = 0; }
    else { r = p - (3 * s - 2) + 1; c = s - 1; }
}

// Inverse of port_rc for a cell given relative to the tile origin; -1 when the
// cell is not on the perimeter (src/internal_node.cpp:290-304).
HTG_HD int port_index(int r, int c, int s) {
    if (r == 0 && c >= 0 && c < s) return c;
    if (r == s - 1 && c >= 0 && c < s) return s + c;
    if (c == 0 && r > 0 && r < s - 1) return 2 * s + (r - 1);
    if (c == s - 1 && r > 0 && r < s - 1) return 3 * s - 2 + (r - 1);
    return -1;
}

HTG_HD int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Per-level sizes.  cs: child side, pc: ports per child, nV = g^2 * pc
// (|V_U|), s: node side, ne: exterior ports, nt: nodes per side.
struct LevelDims {
    int lvl, g, cs, pc, nV, ld, s, ne, nt;
};

inline LevelDims level_dims(int lvl, int b, int g, int S) {
    LevelDims d;
    d.lvl = lvl;
    d.g = g;
    d.cs = b;
    for (int l = 1; l < lvl; ++l) d.cs *= g;
    d.pc = 4 * d.cs - 4;
    d.nV = g * g * d.pc;
    d.ld = round_up(d.nV, 64);  // padded row stride of D_U (128-byte rows)
    d.s = d.cs * g;
    d.ne = 4 * d.s - 4;
    d.nt = S / d.s > 0 ? S / d.s : 1;
    return d;
}

}  // namespace htg
