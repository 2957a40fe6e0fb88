// Device helpers shared by the marching-cubes kernels (leaf level and
// coarse-leaf subdivision): triangle and edge tables, Morton bit handling,
// basis evaluation and the implicit function over an ancestor chain.
#pragma once

#include "ctx.h"
#include "pnt.h"

namespace prg {
namespace {

#include "mc_triangle_table.inc"
__constant__ int8_t c_tri[256][16] = {PRG_MC_TRIANGLE_TABLE_DATA};
// edgeVertex (MarchingCubes.cuh:693-706): square-numbered corners of each edge kind
__constant__ int8_t c_edge_vertex[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 3}, {1, 2},
                                            {4, 7}, {5, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
// the same table as compile-time constants, for fully unrolled per-slot loops
__device__ constexpr int8_t c_edge_vertex_h[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 3}, {1, 2},
                                                      {4, 7}, {5, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};

__device__ __forceinline__ uint32_t compact3(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

// value() of baseFunctions[idx] (ConfirmedPPolynomial<3,4>, ConfirmedPPolynomial.cuh:78-91)
__device__ __forceinline__ float basis_value(const float* __restrict__ bf, int64_t idx, float x) {
    const float* f = bf + idx * 20;
    // the pieces up to the first start not below x are added, in order; all
    // 20 coefficients load together and a select replaces the early exit
    float c[20];
#pragma unroll
    for (int i = 0; i < 20; i++) c[i] = f[i];
    float res = 0.0f;
    bool on = true;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        on = on && (x > c[5 * i]);
        float temp = 1.0f, v = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v = v + temp * c[5 * i + 1 + j];
            temp = temp * x;
        }
        res = on ? res + v : res;
    }
    return res;
}

// Per-depth offsets of the node arrays and of the parent-neighbour tables.
struct Levels {
    int64_t base[kMaxDepth + 1];
    int64_t pnt_off[kMaxDepth + 1];
};

// Bits 3i of x spread from bits i (i = 0..9): the inverse of compact3.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x000003ffu;
    x = (x | (x << 16)) & 0xff0000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// D-level key of the depth-D cell with integer coordinates (x, y, z)
__device__ __forceinline__ uint32_t cell_key(uint32_t x, uint32_t y, uint32_t z) {
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

// Contribution of node nd (depth d >= 1, slot t in its depth) and its 27
// neighbours to f, added to val in the order j = 0..26, given the basis
// values of the node's cell offsets -1, 0, +1 along each axis (bx, by, bz);
// the neighbours come from the node's parent-neighbour row.
__device__ __forceinline__ float level_sum(int64_t t, int d, const float bx[3], const float by[3], const float bz[3],
                                           const int32_t* __restrict__ pnt, const Levels& L,
                                           const float* __restrict__ x, float val) {
    const int32_t* row = pnt + 27 * (L.pnt_off[d] + (t >> 3));
    const float* xd = x + L.base[d];
    int32_t nb[27];
    pnt_neighbours(row, int(t & 7), nb);
    // all 27 loads issued before the sums (a missing neighbour loads the
    // level's first value and its term is not added)
    float xv[27];
#pragma unroll
    for (int j = 0; j < 27; j++) xv[j] = xd[nb[j] < 0 ? 0 : nb[j]];
#pragma unroll
    for (int j = 0; j < 27; j++) {
        float tt = xv[j] * bx[j / 9];
        tt = tt * by[(j / 3) % 3];
        tt = tt * bz[j % 3];
        val = nb[j] < 0 ? val : val + tt;
    }
    return val;
}

// the basis values level_contrib evaluates for node nd of depth d at (px, py, pz)
__device__ __forceinline__ void node_basis(int64_t nd, int d, float px, float py, float pz,
                                           const uint32_t* __restrict__ key, const float* __restrict__ bf, int D,
                                           float bx[3], float by[3], float bz[3]) {
    const uint32_t k = key[nd];
    const int64_t fb = (int64_t(1) << d) - 1, width = int64_t(1) << d;
    const int64_t ox = compact3(k >> 2) >> (D - d), oy = compact3(k >> 1) >> (D - d), oz = compact3(k) >> (D - d);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int64_t u = ox + a - 1, v = oy + a - 1, w = oz + a - 1;
        bx[a] = (u >= 0 && u < width) ? basis_value(bf, fb + u, px) : 0.0f;
        by[a] = (v >= 0 && v < width) ? basis_value(bf, fb + v, py) : 0.0f;
        bz[a] = (w >= 0 && w < width) ? basis_value(bf, fb + w, pz) : 0.0f;
    }
}

// Contribution of node nd (depth d >= 1) and its 27 neighbours to f at
// (px, py, pz), added to val in the order j = 0..26.  Only three basis
// functions per axis are involved (offsets -1, 0, +1 of the node's cell), so 9
// evaluations serve the 27 products.
__device__ __forceinline__ float level_contrib(int64_t nd, int d, float px, float py, float pz,
                                               const uint32_t* __restrict__ key, const int32_t* __restrict__ pnt,
                                               const Levels& L, const float* __restrict__ x,
                                               const float* __restrict__ bf, int D, float val) {
    float bx[3], by[3], bz[3];
    node_basis(nd, d, px, py, pz, key, bf, D, bx, by, bz);
    return level_sum(nd - L.base[d], d, bx, by, bz, pnt, L, x, val);
}

// Basis values at the corners of the leaf level.  A corner's coordinate along
// one axis is its leaf's centre (offset ix at depth D) moved half a cell to
// side cx (0 low, 1 high), and the leaf's ancestor at depth d has offset
// ix >> (D - d), so every value level_contrib needs at a corner is
//   tab[((d 2^D + ix) 2 + cx) 4 + a] = B_{2^d - 1 + (ix >> (D - d)) + a - 1}(coordinate)
// (0 outside the depth's range; a = 0..2, slot 3 unused).  The same table
// serves all three axes: leaf centres and basis functions are the same per
// axis.  Depth 0 holds the root's basis in slot a = 1.  Built with the same
// float operations as leaf_center and basis_value, so the values are those
// the direct evaluation computes.
__device__ __forceinline__ void corner_basis_entry(const float* __restrict__ bf, int D, int d, int64_t ix, int cx,
                                                   float out[4]) {
    float c = 0.5f, w = 0.25f;
    for (int i = D - 1; i >= 0; --i) {
        c = ((ix >> i) & 1) ? c + w : c - w;
        w = w / 2;
    }
    const float hw = 1.0f / float(1 << (D + 1));
    const float p = c + float(2 * cx - 1) * hw;
    const int64_t fb = (int64_t(1) << d) - 1, width = int64_t(1) << d, o = ix >> (D - d);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int64_t u = o + a - 1;
        out[a] = (u >= 0 && u < width) ? basis_value(bf, fb + u, p) : 0.0f;
    }
    out[3] = 0.0f;
}

// f(pos) summed over the ancestor chain of node `nd` at depth `d` (node
// first), continuing from val; the root's only neighbour is itself
// (main.cu:803-808).  Values and summation order are those of the
// reference's loops (main.cu:1353-1374, 2277-2297).
__device__ __forceinline__ float chain_value(int64_t nd, int d, float px, float py, float pz,
                                             const uint32_t* __restrict__ key, const int32_t* __restrict__ pnt,
                                             const Levels& L, const int32_t* __restrict__ parent,
                                             const float* __restrict__ x, const float* __restrict__ bf, int D,
                                             float val) {
    while (d > 0) {
        val = level_contrib(nd, d, px, py, pz, key, pnt, L, x, bf, D, val);
        nd = parent[nd];
        d--;
    }
    float tt = x[0] * basis_value(bf, 0, px);
    tt = tt * basis_value(bf, 0, py);
    tt = tt * basis_value(bf, 0, pz);
    return val + tt;
}

// Coarse rows: for every node of depths 1 .. rows_depth, its 27 neighbours'
// solution values in the order j (0 where a neighbour does not exist: a +-0
// term added to a sum that started at +0 leaves it unchanged, exactly as
// skipping the term) and its parent index (float bits of the int32), 32
// floats per row from the first node of depth 1.
constexpr int kRowFloats = 32;

// level_sum from a coarse row
__device__ __forceinline__ float row_sum(const float* __restrict__ R, const float bx[3], const float by[3],
                                         const float bz[3], float val) {
    float xv[27];
#pragma unroll
    for (int j = 0; j < 27; j++) xv[j] = R[j];
#pragma unroll
    for (int j = 0; j < 27; j++) {
        float tt = xv[j] * bx[j / 9];
        tt = tt * by[(j / 3) % 3];
        tt = tt * bz[j % 3];
        val = val + tt;
    }
    return val;
}

// True when every active lane of the wave holds the same v.
__device__ __forceinline__ bool wave_uniform(int32_t v) {
    return __builtin_amdgcn_ballot_w64(v != __builtin_amdgcn_readfirstlane(v)) == 0;
}

// chain_value from a leaf-level node, for waves whose lanes all start at depth
// D on nodes in Morton order.  Such lanes soon share their ancestor; from that
// depth on the chain runs on one wave-uniform node index, so its neighbour
// rows, coefficients and parents come through the scalar cache once per wave
// instead of once per lane.  Values and summation order are chain_value's.
__device__ __forceinline__ float leaf_chain_value(int64_t nd, float px, float py, float pz,
                                                  const uint32_t* __restrict__ key, const int32_t* __restrict__ pnt,
                                                  const Levels& L, const int32_t* __restrict__ parent,
                                                  const float* __restrict__ x, const float* __restrict__ bf, int D,
                                                  const float* __restrict__ rows, int rows_depth) {
    float val = 0.0f;
    int d = D;
    for (; d > 0; d--) {
        if (wave_uniform(int32_t(nd))) break;
        if (d <= rows_depth) {
            // the node's row: its 27 neighbour values and its parent
            const float* R = rows + kRowFloats * (nd - L.base[1]);
            float bx[3], by[3], bz[3];
            node_basis(nd, d, px, py, pz, key, bf, D, bx, by, bz);
            val = row_sum(R, bx, by, bz, val);
            nd = __float_as_int(R[27]);
        } else {
            val = level_contrib(nd, d, px, py, pz, key, pnt, L, x, bf, D, val);
            nd = parent[nd];
        }
    }
    if (d > 0) {
        int32_t u = __builtin_amdgcn_readfirstlane(int32_t(nd));
        for (; d > 0; d--) {
            if (d <= rows_depth) {
                // one row per depth: neighbour values and the parent in one load
                const float* R = rows + kRowFloats * (int64_t(u) - L.base[1]);
                float bx[3], by[3], bz[3];
                node_basis(u, d, px, py, pz, key, bf, D, bx, by, bz);
                val = row_sum(R, bx, by, bz, val);
                u = __builtin_amdgcn_readfirstlane(__float_as_int(R[27]));
            } else {
                val = level_contrib(u, d, px, py, pz, key, pnt, L, x, bf, D, val);
                u = __builtin_amdgcn_readfirstlane(parent[u]);
            }
        }
    }
    float tt = x[0] * basis_value(bf, 0, px);
    tt = tt * basis_value(bf, 0, py);
    tt = tt * basis_value(bf, 0, pz);
    return val + tt;
}

// node centre (getNodeCenterAllDepth, main.cu:1403-1420) at the leaf level
__device__ __forceinline__ void leaf_center(uint32_t k, int D, float c[3]) {
    c[0] = c[1] = c[2] = 0.5f;
    float w = 0.25f;
    for (int i = D - 1; i >= 0; --i) {
        c[0] = ((k >> (3 * i + 2)) & 1u) ? c[0] + w : c[0] - w;
        c[1] = ((k >> (3 * i + 1)) & 1u) ? c[1] + w : c[1] - w;
        c[2] = ((k >> (3 * i)) & 1u) ? c[2] + w : c[2] - w;
        w = w / 2;
    }
}

// node centre (getNodeCenterAllDepth, main.cu:1403-1420) at depth d
__device__ __forceinline__ void node_center(uint32_t k, int d, int D, float c[3]) {
    c[0] = c[1] = c[2] = 0.5f;
    float w = 0.25f;
    for (int i = D - 1; i >= D - d; --i) {
        c[0] = ((k >> (3 * i + 2)) & 1u) ? c[0] + w : c[0] - w;
        c[1] = ((k >> (3 * i + 1)) & 1u) ? c[1] + w : c[1] - w;
        c[2] = ((k >> (3 * i)) & 1u) ? c[2] + w : c[2] - w;
        w = w / 2;
    }
}

__device__ __forceinline__ int tri_count(int cube) {
    int n = 0;
    while (n < 15 && c_tri[cube][n] >= 0) n++;
    return n / 3;
}

inline Levels levels_of(const prg_ctx* c) {

    Levels L{};
    for (int d = 0; d <= c->D; d++) {
        L.base[d] = c->bases[d];
        L.pnt_off[d] = c->pnt_off[d];
    }
    return L;
}

}  // namespace
}  // namespace prg
