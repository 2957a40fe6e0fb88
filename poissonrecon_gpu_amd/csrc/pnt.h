// Neighbours of a node from its sibling block's parent-neighbour row.
//
// Neighbour (dx, dy, dz) of child slot s = (sx, sy, sz) lies in parent
// neighbour ((sx + dx) >> 1) + 1 along x, at child bit (sx + dx) & 1 (the
// lookup computeNeighbor performs with LUTparent/LUTchild, main.cu:492-509).
// Along x that entry is sx for dx = -1, sx + 1 for dx = +1 and 1 for dx = 0,
// so 8 row entries serve the 27 neighbours: E[a][b][c] is entry
// (sx + a, sy + b, sz + c), and a dx = 0 neighbour takes a = 1 - sx (likewise
// for y and z).
#pragma once

#include <cstdint>

namespace prg {

// the 8 row entries slot s uses: e[4a + 2b + c] = entry (sx + a, sy + b, sz + c)
__device__ __forceinline__ void pnt_entries(const int32_t* __restrict__ row, int s, int32_t e[8]) {
    const int sx = (s >> 2) & 1, sy = (s >> 1) & 1, sz = s & 1;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) e[4 * a + 2 * b + c] = row[9 * (sx + a) + 3 * (sy + b) + (sz + c)];
}

// nb[j] = level-local index of neighbour j = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)
// of slot s from its 8 row entries, or -1 where the neighbour does not exist
__device__ __forceinline__ void pnt_expand(const int32_t e[8], int s, int32_t nb[27]) {
    const int sx = (s >> 2) & 1, sy = (s >> 1) & 1, sz = s & 1;
    int32_t E[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) E[a][b][c] = e[4 * a + 2 * b + c];
    int32_t F[3][2][2];  // x resolved: F[dx + 1][b][c]
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            F[0][b][c] = E[0][b][c];
            F[1][b][c] = sx ? E[0][b][c] : E[1][b][c];
            F[2][b][c] = E[1][b][c];
        }
    int32_t Gy[3][3][2];  // y resolved
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            Gy[a][0][c] = F[a][0][c];
            Gy[a][1][c] = sy ? F[a][0][c] : F[a][1][c];
            Gy[a][2][c] = F[a][1][c];
        }
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int jx = j / 9, jy = (j / 3) % 3, jz = j % 3;
        const int vx = sx + (jx - 1), vy = sy + (jy - 1), vz = sz + (jz - 1);
        const int32_t cb = jz == 0 ? Gy[jx][jy][0] : (jz == 2 ? Gy[jx][jy][1] : (sz ? Gy[jx][jy][0] : Gy[jx][jy][1]));
        nb[j] = cb < 0 ? -1 : cb + (((vx & 1) << 2) | ((vy & 1) << 1) | (vz & 1));
    }
}

__device__ __forceinline__ void pnt_neighbours(const int32_t* __restrict__ row, int s, int32_t nb[27]) {
    int32_t e[8];
    pnt_entries(row, s, e);
    pnt_expand(e, s, nb);
}

}  // namespace prg
