// Sibling-block tiles of one depth and their plans.
//
// A tile is kTile consecutive sibling blocks (kTileThreads rows, one per
// thread).  Its plan, built once per octree by k_tile_plan, lists the distinct
// sibling blocks the tile's 27-neighbourhoods touch (`ub`: each block's first
// row, level-local) and, per (block of the tile, parent-neighbour entry), that
// block's position in the list (`rw`, -1 where the parent neighbour does not
// exist).  Kernels that gather over the 27 neighbours of every row load each
// listed block once into LDS as a contiguous run of 8 values and read the
// neighbours from there.
#pragma once

#include <cstdint>

namespace prg {

constexpr int kTileThreads = 512;
constexpr int kTile = kTileThreads / 8;   // sibling blocks per tile
constexpr int kTileEntries = 27 * kTile;  // parent-neighbour entries of a tile

struct TilePlan {
    const int32_t* cnt;  // distinct blocks per tile
    const int32_t* ub;   // [tile][kTileEntries]: first row of each listed block
    const int16_t* rw;   // [tile][kTileEntries]: list position per (block, entry), or -1
};

// Neighbour j = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) of child slot s: the
// parent-neighbour entry it lies in and its child slot there.  Along x the
// neighbour's offset v = sx + dx in {-1, 0, 1, 2} lies in parent neighbour
// (v + 2) >> 1 at child bit v & 1 (as pnt_neighbours resolves it).
__device__ __forceinline__ void tile_neighbour(int s, int j, int& entry, int& child) {
    const int vx = ((s >> 2) & 1) + j / 9 - 1, vy = ((s >> 1) & 1) + (j / 3) % 3 - 1, vz = (s & 1) + j % 3 - 1;
    entry = 9 * ((vx + 2) >> 1) + 3 * ((vy + 2) >> 1) + ((vz + 2) >> 1);
    child = ((vx & 1) << 2) | ((vy & 1) << 1) | (vz & 1);
}

}  // namespace prg
