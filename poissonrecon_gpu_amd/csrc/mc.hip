// Isovalue and leaf-level marching cubes (reference:
// calculatePointsImplicitFunctionValue main.cu:1334-1381, the MC topology
// kernels :1438-2082, computeVertexImplicitFunctionValue :2259-2326,
// generateVexNums/TriNums/IntersectionPoint/TrianglePos :2457-2757, driver
// :3502-3795).
//
// The implicit function at a position is the float sum, over the node's
// ancestor chain (leaf first) and their 27 neighbours (j = 0..26), of
// x_j * B_x(p_x) * B_y(p_y) * B_z(p_z) with the reference's per-index float
// basis polynomials.
//
// Topology uses the reference's ownership rule without its pre-arrays of
// 56-byte vertex records: a corner (edge) of the leaf level belongs to the
// node with the smallest key among the nodes sharing it; corner values are
// evaluated once, by the owner; mesh vertices are the owned edges with a sign
// change, numbered by an exclusive scan over (owner node, edge kind) - the
// order of the reference's EdgeArray - so triangle indices come out
// identical.
#include <cmath>

#include "ctx.h"
#include "mc_common.h"

namespace prg {
namespace {

constexpr int kBlock = 256;

// One row per node of depths 1 .. Dr (the layout kRowFloats describes).  The
// evaluation chains visit only ancestors of leaf-level slots, so a node
// without any below it (dnum 0) gets no row.
__global__ void __launch_bounds__(kBlock) k_coarse_rows(const int32_t* __restrict__ pnt, Levels L,
                                                        const int32_t* __restrict__ parent,
                                                        const int32_t* __restrict__ dnum,
                                                        const float* __restrict__ x, int Dr, int64_t n,
                                                        float* __restrict__ rows) {
    const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t i = L.base[1] + r;
    if (dnum[i] == 0) return;
    int d = 1;
    while (d < Dr && i >= L.base[d + 1]) d++;
    const int64_t t = i - L.base[d];
    int32_t nb[27];
    pnt_neighbours(pnt + 27 * (L.pnt_off[d] + (t >> 3)), int(t & 7), nb);
    const float* xd = x + L.base[d];
    float v[kRowFloats];
#pragma unroll
    for (int j = 0; j < 27; j++) v[j] = nb[j] < 0 ? 0.0f : xd[nb[j] < 0 ? 0 : nb[j]];
    v[27] = __int_as_float(parent[i]);
#pragma unroll
    for (int j = 28; j < kRowFloats; j++) v[j] = 0.0f;
    float4* o = reinterpret_cast<float4*>(rows + kRowFloats * r);
#pragma unroll
    for (int q = 0; q < kRowFloats / 4; q++) o[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

__global__ void __launch_bounds__(kBlock, 6) k_point_values(const float* __restrict__ px, const float* __restrict__ py,
                                                         const float* __restrict__ pz,
                                                         const int32_t* __restrict__ point_leaf, int64_t n,
                                                         int64_t baseD, const uint32_t* __restrict__ key,
                                                         const int32_t* __restrict__ pnt, Levels L,
                                                         const int32_t* __restrict__ parent,
                                                         const float* __restrict__ x, const float* __restrict__ bf,
                                                         int D, const float* __restrict__ rows, int rows_depth,
                                                         float* __restrict__ pvals, double* __restrict__ partial) {
    __shared__ double lds[kBlock / 64];
    const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double v = 0.0;
    if (r < n) {
        // points are in leaf order, so a wave's lanes share their coarse ancestors
        const float f = leaf_chain_value(baseD + point_leaf[r], px[r], py[r], pz[r], key, pnt, L, parent, x, bf, D,
                                         rows, rows_depth);
        pvals[r] = f;
        v = double(f);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; w++) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

// sum of n partials in a fixed order: contiguous runs per thread, then a tree
// ... and the isovalue sum / count (main.cu:3494-3496) in float
__global__ void __launch_bounds__(1024) k_sum_partials(const double* partial, int64_t n, int64_t count,
                                                       float* iso) {
    __shared__ double red[1024];
    const int tid = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = int64_t(tid) * per, e = b + per < n ? b + per : n;
    double s = 0.0;
    for (int64_t i = b; i < e; i++) s += partial[i];
    red[tid] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) *iso = float(red[0]) / float(count);
}

// corner owners (initVertexOwner, main.cu:1438-1498): corner j = x | y<<1 | z<<2.
// One thread per leaf slot: the slot's 26 neighbours and their keys are read
// once and serve its 8 corners.  Corner j is owned by the smallest key among
// the (up to 8) leaf slots sharing it; own[8 t + j] marks slot t's own
// corners and oref[8 t + j] = 8 * (owner slot) + the corner's number there.
// cmask[t] has bit j set for the corners slot t owns, ccount[t] counts them.
__global__ void __launch_bounds__(kBlock) k_corner_owner(const uint32_t* key, const int32_t* pntD, int64_t baseD,
                                                         int64_t MD, int32_t* cmask, int32_t* ccount, int32_t* oref) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int64_t i = baseD + t;
    int32_t nb[27];
    uint32_t kk[27];
    pnt_neighbours(pntD + 27 * (t >> 3), int(t & 7), nb);
#pragma unroll
    for (int dir = 0; dir < 27; dir++) {
        nb[dir] = nb[dir] < 0 ? -1 : int32_t(baseD + nb[dir]);
        kk[dir] = nb[dir] == -1 ? 0xffffffffu : key[nb[dir]];
    }
    int32_t ow[8], orf[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int cx = j & 1, cy = (j >> 1) & 1, cz = (j >> 2) & 1;
        uint32_t best = 0xffffffffu;
        int64_t owner = i;
        int jo = j;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int dx = a ? (2 * cx - 1) : 0, dy = b ? (2 * cy - 1) : 0, dz = e ? (2 * cz - 1) : 0;
                    const int dir = 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1);
                    if (nb[dir] == -1) continue;
                    if (kk[dir] < best) {
                        best = kk[dir];
                        owner = nb[dir];
                        jo = (cx ^ a) | ((cy ^ b) << 1) | ((cz ^ e) << 2);
                    }
                }
        ow[j] = (owner == i) ? 1 : 0;
        orf[j] = int32_t(8 * (owner - baseD) + jo);
    }
    int m = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) m |= ow[j] << j;
    cmask[t] = m;
    ccount[t] = __popc(m);
    int4* r4 = reinterpret_cast<int4*>(oref + 8 * t);
    r4[0] = make_int4(orf[0], orf[1], orf[2], orf[3]);
    r4[1] = make_int4(orf[4], orf[5], orf[6], orf[7]);
}

// owned corners, compacted: vlist[vid] = 8 * leaf slot + corner
// Owned corners are numbered in (slot, corner) order: cbase is the exclusive
// scan of the per-slot counts, and a corner's rank inside its slot counts the
// owned corners before it.
__device__ __forceinline__ int32_t corner_id(const int32_t* cbase, const int32_t* cmask, int32_t g) {
    const int32_t t = g >> 3;
    return cbase[t] + __popc(cmask[t] & ((1 << (g & 7)) - 1));
}

// owned corners, compacted: vlist[id] = 8 * leaf slot + corner
__global__ void __launch_bounds__(kBlock) k_corner_list(const int32_t* cmask, const int32_t* cbase, int64_t MD,
                                                        int32_t* vlist) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int m = cmask[t];
    int32_t r = cbase[t];
#pragma unroll
    for (int j = 0; j < 8; j++)
        if ((m >> j) & 1) vlist[r++] = int32_t(8 * t + j);
}

// the corner basis table (corner_basis_entry) for depths 0..D
__global__ void __launch_bounds__(kBlock) k_corner_basis(const float* bf, int D, float* tab) {
    const int64_t G = int64_t(1) << D;
    const int64_t g = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (g >= int64_t(D + 1) * G * 2) return;
    const int cx = int(g & 1);
    const int64_t ix = (g >> 1) % G;
    const int d = int((g >> 1) / G);
    float v[4];
    corner_basis_entry(bf, D, d, ix, cx, v);
    reinterpret_cast<float4*>(tab)[g] = make_float4(v[0], v[1], v[2], v[3]);
}

// f(corner) - iso for the owned corners (computeVertexImplicitFunctionValue,
// main.cu:2259-2326): the chain sum (chain_value) at the corner's position,
// with the basis values read from the corner basis table
__global__ void __launch_bounds__(kBlock, 5) k_corner_value(const int32_t* __restrict__ vlist, int64_t V,
                                                         const uint32_t* __restrict__ key,
                                                         const int32_t* __restrict__ pnt, Levels L,
                                                         const int32_t* __restrict__ parent,
                                                         const float* __restrict__ x, const float4* __restrict__ tab,
                                                         const float* __restrict__ rows, int rows_depth,
                                                         int64_t baseD, int D, const float* __restrict__ iso,
                                                         float* __restrict__ vval) {
    const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (v >= V) return;
    const int64_t g = vlist[v];
    const int j = int(g & 7);
    int64_t nd = baseD + (g >> 3);
    const uint32_t k = key[nd];
    const int64_t G = int64_t(1) << D;
    // table rows of this corner's three coordinates at depth 0
    const int64_t rx = 2 * int64_t(compact3(k >> 2)) + (j & 1);
    const int64_t ry = 2 * int64_t(compact3(k >> 1)) + ((j >> 1) & 1);
    const int64_t rz = 2 * int64_t(compact3(k)) + ((j >> 2) & 1);
    // corners are listed in leaf order: once a wave's lanes share their
    // ancestor, the rest of the chain runs on one wave-uniform node (rows and
    // coefficients through the scalar cache); values and order are unchanged
    float val = 0.0f;
    int d = D;
    for (; d > 0; d--) {
        if (wave_uniform(int32_t(nd))) break;
        const int64_t o = 2 * G * d;
        const float4 tx = tab[o + rx], ty = tab[o + ry], tz = tab[o + rz];
        const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
        if (d <= rows_depth) {
            // the node's row: its 27 neighbour values and its parent
            const float* R = rows + kRowFloats * (nd - L.base[1]);
            val = row_sum(R, bx, by, bz, val);
            nd = __float_as_int(R[27]);
        } else {
            val = level_sum(nd - L.base[d], d, bx, by, bz, pnt, L, x, val);
            nd = parent[nd];
        }
    }
    if (d > 0) {
        int32_t u = __builtin_amdgcn_readfirstlane(int32_t(nd));
        for (; d > 0; d--) {
            const int64_t o = 2 * G * d;
            const float4 tx = tab[o + rx], ty = tab[o + ry], tz = tab[o + rz];
            const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
            if (d <= rows_depth) {
                // one row per depth: neighbour values and the parent in one load
                const float* R = rows + kRowFloats * (int64_t(u) - L.base[1]);
                val = row_sum(R, bx, by, bz, val);
                u = __builtin_amdgcn_readfirstlane(__float_as_int(R[27]));
            } else {
                val = level_sum(int64_t(u) - L.base[d], d, bx, by, bz, pnt, L, x, val);
                u = __builtin_amdgcn_readfirstlane(parent[u]);
            }
        }
    }
    float tt = x[0] * tab[rx].y;
    tt = tt * tab[ry].y;
    tt = tt * tab[rz].y;
    vval[v] = (val + tt) - *iso;
}

// edge owners (initEdgeArray / maintainEdgeNodePointer, main.cu:1795-2082) and
// the sign-change test of the owner's corners (generateVexNums, :2457-2479).
// One thread per leaf slot.  Corner values are per corner, so an edge's
// sign-change test gives the same answer from every slot sharing it; a slot
// none of whose 12 edges changes sign flags nothing and skips the ownership
// search (its eref entries are never read: only edges of cubes with
// triangles are, and those change sign).
__device__ __forceinline__ void edge_owner_slot(int64_t t, const float cv[8], const uint32_t* __restrict__ key,
                                                const int32_t* __restrict__ pntD, int64_t baseD, int32_t* emask,
                                                int32_t* ecount, int32_t* eref) {
    const int64_t i = baseD + t;
    bool sc[12];
    bool any = false;
#pragma unroll
    for (int e = 0; e < 12; e++) {
        sc[e] = cv[c_edge_vertex_h[e][0]] * cv[c_edge_vertex_h[e][1]] <= 0;
        any = any || sc[e];
    }
    if (!any) {
        emask[t] = 0;
        ecount[t] = 0;
        return;
    }
    // the 27 neighbours and their keys (13 is the slot itself); corner
    // neighbours share no edge and are not looked at
    int32_t nb[27];
    uint32_t kk[27];
    pnt_neighbours(pntD + 27 * (t >> 3), int(t & 7), nb);
#pragma unroll
    for (int dir = 0; dir < 27; dir++) {
        const int ddx = dir / 9 - 1, ddy = (dir / 3) % 3 - 1, ddz = dir % 3 - 1;
        nb[dir] = (nb[dir] < 0 || (ddx != 0 && ddy != 0 && ddz != 0)) ? -1 : int32_t(baseD + nb[dir]);
        kk[dir] = nb[dir] == -1 ? 0xffffffffu : key[nb[dir]];
    }
    int32_t fl[12], er[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const int o = e >> 2;
        const int a1 = (o == 0) ? 1 : 0, a2 = (o == 2) ? 1 : 2;
        const int b1 = e & 1, b2 = (e >> 1) & 1;
        uint32_t best = 0xffffffffu;
        int64_t owner = i;
        int eo = e;
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 2; v++) {
                int dd[3] = {0, 0, 0};
                dd[a1] = u ? (2 * b1 - 1) : 0;
                dd[a2] = v ? (2 * b2 - 1) : 0;
                const int dir = 9 * (dd[0] + 1) + 3 * (dd[1] + 1) + (dd[2] + 1);
                if (nb[dir] == -1) continue;
                if (kk[dir] < best) {
                    best = kk[dir];
                    owner = nb[dir];
                    eo = (o << 2) | ((b2 ^ v) << 1) | (b1 ^ u);
                }
            }
        er[e] = int32_t(12 * (owner - baseD) + eo);
        fl[e] = (owner == i && sc[e]) ? 1 : 0;
    }
    int m = 0;
#pragma unroll
    for (int e = 0; e < 12; e++) m |= fl[e] << e;
    emask[t] = m;
    ecount[t] = __popc(m);
    int4* r4 = reinterpret_cast<int4*>(eref + 12 * t);
    r4[0] = make_int4(er[0], er[1], er[2], er[3]);
    r4[1] = make_int4(er[4], er[5], er[6], er[7]);
    r4[2] = make_int4(er[8], er[9], er[10], er[11]);
}

// Per leaf slot: its 8 corner values in square numbering (vertices[],
// main.cu:1659-1668, written for the vertex positions and the buffer read),
// the cube index and triangle count, and the edge owners (edge_owner_slot),
// all from the corner values in registers.
__global__ void __launch_bounds__(kBlock) k_corner_edges(const int32_t* __restrict__ oref,
                                                         const int32_t* __restrict__ cmask,
                                                         const int32_t* __restrict__ cbase,
                                                         const float* __restrict__ vval,
                                                         const uint32_t* __restrict__ key,
                                                         const int32_t* __restrict__ pntD, int64_t baseD, int64_t MD,
                                                         float* __restrict__ cval, uint8_t* __restrict__ cubes,
                                                         int32_t* __restrict__ ntri, int32_t* __restrict__ emask,
                                                         int32_t* __restrict__ ecount, int32_t* __restrict__ eref) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int4* o4 = reinterpret_cast<const int4*>(oref + 8 * t);
    const int4 oa = o4[0], ob = o4[1];
    const int32_t o[8] = {oa.x, oa.y, oa.z, oa.w, ob.x, ob.y, ob.z, ob.w};
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = (q >> 1) & 1, x = (q & 1) ^ y, z = q >> 2;
        v[q] = vval[corner_id(cbase, cmask, o[x | (y << 1) | (z << 2)])];
    }
    float4* c4 = reinterpret_cast<float4*>(cval + 8 * t);
    c4[0] = make_float4(v[0], v[1], v[2], v[3]);
    c4[1] = make_float4(v[4], v[5], v[6], v[7]);
    int cube = 0;
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (v[q] < 0) cube |= 1 << q;
    cubes[t] = uint8_t(cube);
    ntri[t] = tri_count(cube);
    edge_owner_slot(t, v, key, pntD, baseD, emask, ecount, eref);
}

// interpolatePoint (main.cu:2584-2597) for every owned edge with a sign
// change; one thread per leaf slot
// Mesh vertices are numbered in (slot, edge kind) order: ebase is the
// exclusive scan of the per-slot counts of owned sign-change edges.
__device__ __forceinline__ int32_t edge_id(const int32_t* ebase, const int32_t* emask, int32_t r) {
    const int32_t t = r / 12;
    return ebase[t] + __popc(emask[t] & ((1 << (r - 12 * t)) - 1));
}

__global__ void __launch_bounds__(kBlock) k_edge_vertex(const int32_t* emask, const int32_t* ebase, const uint32_t* key,
                                                        const float* cval, int64_t baseD, int64_t MD, int D,
                                                        float* mesh_v) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int m = emask[t];
    if (!m) return;
    int64_t v = ebase[t];
    float c[3];
    leaf_center(key[baseD + t], D, c);
    const float hw = 1.0f / float(1 << (D + 1));
#pragma unroll
    for (int e = 0; e < 12; e++) {
        if (!((m >> e) & 1)) continue;
        const int q1 = c_edge_vertex_h[e][0], q2 = c_edge_vertex_h[e][1];
        float p1[3], p2[3];
        {
            const int y = (q1 >> 1) & 1, x = (q1 & 1) ^ y, z = q1 >> 2;
            p1[0] = c[0] + float(2 * x - 1) * hw;
            p1[1] = c[1] + float(2 * y - 1) * hw;
            p1[2] = c[2] + float(2 * z - 1) * hw;
        }
        {
            const int y = (q2 >> 1) & 1, x = (q2 & 1) ^ y, z = q2 >> 2;
            p2[0] = c[0] + float(2 * x - 1) * hw;
            p2[1] = c[1] + float(2 * y - 1) * hw;
            p2[2] = c[2] + float(2 * z - 1) * hw;
        }
        const float f1 = cval[8 * t + q1], f2 = cval[8 * t + q2];
        const int dim = e >> 2;
        float out[3] = {p1[0], p1[1], p1[2]};
        const float pivot = f1 / (f1 - f2);
        const float another = 1 - pivot;
        const float s1 = p2[dim] * pivot;
        const float s2 = p1[dim] * another;
        out[dim] = s1 + s2;
        mesh_v[3 * v + 0] = out[0];
        mesh_v[3 * v + 1] = out[1];
        mesh_v[3 * v + 2] = out[2];
        v++;
    }
}

__global__ void __launch_bounds__(kBlock) k_tri_emit(const uint8_t* cubes, const int32_t* toff, const int32_t* eref,
                                                     const int32_t* emask, const int32_t* ebase, int64_t MD,
                                                     int32_t* mesh_t) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int cube = cubes[t];
    const int n = 3 * tri_count(cube);
    if (n == 0) return;
    const int64_t o = 3 * int64_t(toff[t]);
    // the 12 edge references and their vertex ids at once (a slot with
    // triangles has all 12 references), then the triangles' picks
    const int4* r4 = reinterpret_cast<const int4*>(eref + 12 * t);
    const int4 ra = r4[0], rb = r4[1], rc = r4[2];
    const int32_t r[12] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w, rc.x, rc.y, rc.z, rc.w};
    int32_t id[12];
#pragma unroll
    for (int e = 0; e < 12; e++) id[e] = edge_id(ebase, emask, r[e]);
    for (int k = 0; k < n; k++) {
        const int e = c_tri[cube][k];
        int32_t pick = id[0];
#pragma unroll
        for (int q = 1; q < 12; q++) pick = e == q ? id[q] : pick;
        mesh_t[o + k] = pick;
    }
}


}  // namespace

// Rows of depths 1 .. D - 1 (every depth above the leaf level) for the current solution (rebuilt when the
// solution or the octree changes); returns the deepest depth they cover (0:
// none).
int ensure_coarse_rows(prg_ctx* c, hipStream_t s) {
    const int D = c->D, Dr = D - 1;
    if (Dr < 1) return 0;
    if (c->rows_x == c->x_gen && c->rows_tree == c->tree_gen) return c->rows_depth;
    const int64_t n = c->bases[Dr + 1] - c->bases[1];
    float* rows = c->coarse_rows.reserve(size_t(kRowFloats) * std::max<int64_t>(1, n));
    TimedScope ts(c->kt, "coarse_rows", 128.0 * double(n));
    k_coarse_rows<<<grid_for(n, kBlock), kBlock, 0, s>>>(c->pnt.get(), levels_of(c), c->parent.get(),
                                                         c->dnum.get(), c->x.get(), Dr, n, rows);
    PRG_CHECK_LAUNCH();
    c->rows_depth = Dr;
    c->rows_x = c->x_gen;
    c->rows_tree = c->tree_gen;
    return Dr;
}

void compute_isovalue(prg_ctx* c) {
    ensure_tables(c);
    StageTimer st(c, PRG_STAGE_ISOVALUE);
    hipStream_t s = c->stream;
    const int64_t n = c->N;
    const unsigned nb = grid_for(n, kBlock);
    PoolBuf<double> part{&c->pool[76]};
    part.reserve(nb);
    {
        TimedScope ts(c->kt, "point_values", 24.0 * double(n));
        const int rd = ensure_coarse_rows(c, s);
        k_point_values<<<nb, kBlock, 0, s>>>(c->px.get(), c->py.get(), c->pz.get(), c->point_leaf.get(), n,
                                             c->bases[c->D], c->key.get(), c->pnt.get(), levels_of(c),
                                             c->parent.get(), c->x.get(), c->d_base_fn.get(), c->D,
                                             c->coarse_rows.get(), rd, c->pvals.reserve(n), part.get());
        PRG_CHECK_LAUNCH();
    }
    // the isovalue stays on the device (no host wait); prg_isovalue reads it back when asked
    k_sum_partials<<<1, 1024, 0, s>>>(part.get(), nb, n, c->iso_dev.reserve(1));
    PRG_CHECK_LAUNCH();
}

// pinned area of the owned-corner count (the last bytes of the context's
// pinned staging page, after the solver's areas)
int32_t* topo_count_host(prg_ctx* c) { return reinterpret_cast<int32_t*>(static_cast<uint8_t*>(c->cg_host) + 4092); }

void start_topology(prg_ctx* c) {
    ensure_tables(c);
    const bool timed = c->kt.enabled;
    hipStream_t s = timed ? c->stream : c->stream2;  // the plans have made stream2 wait for the tree
    Scratch& scr = timed ? c->scratch : c->scratch2;
    const int D = c->D;
    const int64_t MD = c->MD(), baseD = c->bases[D];
    KernelTimer& kt = c->kt;
    PoolBuf<int32_t> cmask{&c->pool[78]}, oref{&c->pool[79]}, cbase{&c->pool[80]}, vtot{&c->pool[81]},
        ccount{&c->pool[92]};
    PoolBuf<float> btab{&c->pool[91]};
    cmask.reserve(MD);
    ccount.reserve(MD);
    oref.reserve(8 * MD);
    cbase.reserve(MD);
    vtot.reserve(1);
    // stage time: events on the stream the topology runs on
    resolve_stage(c, PRG_STAGE_MC_TOPOLOGY);
    PRG_HIP(hipEventRecord(c->stage_ev[PRG_STAGE_MC_TOPOLOGY][0], s));
    {
        // per leaf slot: its key, its share of the block's parent-neighbour row
        // (13.5 B), the mask and count written, 8 owner references written
        TimedScope ts(kt, "corner_owner", (4.0 + 13.5 + 8.0 + 32.0) * double(MD));
        k_corner_owner<<<grid_for(MD, kBlock), kBlock, 0, s>>>(c->key.get(), c->pnt.get() + 27 * c->pnt_off[D], baseD,
                                                               MD, cmask.get(), ccount.get(), oref.get());
        PRG_CHECK_LAUNCH();
    }
    scan_exclusive(ccount.get(), cbase.get(), MD, vtot.get(), s, scr, kt);
    PRG_HIP(hipMemcpyAsync(topo_count_host(c), vtot.get(), 4, hipMemcpyDeviceToHost, s));
    const int64_t nbt = int64_t(D + 1) * (int64_t(1) << D) * 2;
    btab.reserve(4 * nbt);
    k_corner_basis<<<grid_for(nbt, kBlock), kBlock, 0, s>>>(c->d_base_fn.get(), D, btab.get());
    PRG_CHECK_LAUNCH();
    PRG_HIP(hipEventRecord(c->stage_ev[PRG_STAGE_MC_TOPOLOGY][1], s));
    c->stage_pending[PRG_STAGE_MC_TOPOLOGY] = true;
    PRG_HIP(hipEventRecord(c->topo_ev, s));
    c->topo_pending = true;
    c->topo_gen = c->tree_gen;
}

void extract_leaf_mesh(prg_ctx* c) {
    ensure_tables(c);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t MD = c->MD(), baseD = c->bases[D];
    KernelTimer& kt = c->kt;
    // per leaf slot: owned-corner mask and count, their scan (corner ids);
    // owned sign-change edge mask and count, their scan (vertex ids)
    PoolBuf<int32_t> cmask{&c->pool[78]}, oref{&c->pool[79]}, cbase{&c->pool[80]}, emask{&c->pool[82]},
        eref{&c->pool[83]}, ebase{&c->pool[84]}, etot{&c->pool[85]}, ntri{&c->pool[86]}, toff{&c->pool[87]},
        ttot{&c->pool[88]}, ecount{&c->pool[93]};
    PoolBuf<float> vval{&c->pool[89]}, btab{&c->pool[91]};
    float* CV = c->corner_vals.reserve(8 * MD);
    int32_t NV = 0, NT = 0;
    if (c->topo_gen != c->tree_gen) start_topology(c);
    if (c->topo_pending) {
        PRG_HIP(hipEventSynchronize(c->topo_ev));
        c->topo_corners = *topo_count_host(c);
        c->topo_pending = false;
    }
    const int32_t V = c->topo_corners;
    bool direct = false;
    {
        StageTimer st(c, PRG_STAGE_MC_EVAL);
        vval.reserve(V);
        PoolBuf<int32_t> vlist{&c->pool[90]};
        vlist.reserve(V);
        k_corner_list<<<grid_for(MD, kBlock), kBlock, 0, s>>>(cmask.get(), cbase.get(), MD, vlist.get());
        PRG_CHECK_LAUNCH();
        {
            TimedScope ts(kt, "corner_value", 24.0 * double(V));
            const int rd = ensure_coarse_rows(c, s);
            k_corner_value<<<grid_for(V, kBlock), kBlock, 0, s>>>(
                vlist.get(), V, c->key.get(), c->pnt.get(), levels_of(c), c->parent.get(), c->x.get(),
                reinterpret_cast<const float4*>(btab.get()), c->coarse_rows.get(), rd, baseD, D, c->iso_dev.get(),
                vval.get());
            PRG_CHECK_LAUNCH();
        }
        emask.reserve(MD);
        ecount.reserve(MD);
        eref.reserve(12 * MD);
        ebase.reserve(MD);
        etot.reserve(1);
        ntri.reserve(MD);
        toff.reserve(MD);
        ttot.reserve(1);
        PoolBuf<uint8_t> cubes{&c->pool[94]};
        cubes.reserve(MD);
        {
            // corner values gathered (8 floats), cube index, triangle count and edge owners per slot
            // per leaf slot: 8 owner references and 8 corner values read, its
            // row share (13.5 B); 8 corner values, cube, triangle count, edge
            // mask and count, 12 edge references written
            TimedScope ts(kt, "edge_owner", (32.0 + 32.0 + 13.5 + 32.0 + 1.0 + 4.0 + 8.0 + 48.0) * double(MD));
            k_corner_edges<<<grid_for(MD, kBlock), kBlock, 0, s>>>(oref.get(), cmask.get(), cbase.get(), vval.get(),
                                                                   c->key.get(), c->pnt.get() + 27 * c->pnt_off[D],
                                                                   baseD, MD, CV, cubes.get(), ntri.get(), emask.get(),
                                                                   ecount.get(), eref.get());
            PRG_CHECK_LAUNCH();
        }
        scan_exclusive2(ecount.get(), ebase.get(), etot.get(), ntri.get(), toff.get(), ttot.get(), MD, s, c->scratch,
                        kt);
        PRG_HIP(hipMemcpyAsync(&NV, etot.get(), 4, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipMemcpyAsync(&NT, ttot.get(), 4, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
        float* MV = c->mesh_v.reserve(size_t(3) * NV);
        int32_t* MT = c->mesh_t.reserve(size_t(3) * NT);
        // caller buffers that hold the mesh: the vertex copy (second stream)
        // runs while the triangles are emitted
        direct = c->out_v && c->out_t && 3 * int64_t(NV) <= c->out_vcap && 3 * int64_t(NT) <= c->out_tcap;
        k_edge_vertex<<<grid_for(MD, kBlock), kBlock, 0, s>>>(emask.get(), ebase.get(), c->key.get(), CV, baseD, MD,
                                                              D, MV);
        PRG_CHECK_LAUNCH();
        if (direct && NV) {
            PRG_HIP(hipEventRecord(c->mesh_ev, s));
            PRG_HIP(hipStreamWaitEvent(c->stream2, c->mesh_ev, 0));
            PRG_HIP(hipMemcpyAsync(c->out_v, MV, size_t(3) * NV * 4, hipMemcpyDeviceToHost, c->stream2));
            PRG_HIP(hipEventRecord(c->mesh_ev, c->stream2));
        }
        k_tri_emit<<<grid_for(MD, kBlock), kBlock, 0, s>>>(cubes.get(), toff.get(), eref.get(), emask.get(),
                                                           ebase.get(), MD, MT);
        PRG_CHECK_LAUNCH();
        if (!direct) PRG_HIP(hipStreamSynchronize(s));
    }
    c->mesh_nv = NV;
    c->mesh_nt = NT;
    c->out_copied = direct;
    if (direct && c->out_async) {
        // the subdivision follows on the context stream: the triangles go out
        // on the second stream too, behind the vertices, while it runs
        // (copy_mesh waits for mesh_ev)
        if (NT) {
            PRG_HIP(hipEventRecord(c->mesh_ev2, s));
            PRG_HIP(hipStreamWaitEvent(c->stream2, c->mesh_ev2, 0));
            PRG_HIP(hipMemcpyAsync(c->out_t, c->mesh_t.get(), size_t(3) * NT * 4, hipMemcpyDeviceToHost, c->stream2));
        }
        PRG_HIP(hipEventRecord(c->mesh_ev, c->stream2));
    } else if (direct) {
        StageTimer st(c, PRG_STAGE_MESH_COPY);
        if (NT) PRG_HIP(hipMemcpyAsync(c->out_t, c->mesh_t.get(), size_t(3) * NT * 4, hipMemcpyDeviceToHost, s));
        if (NV) PRG_HIP(hipStreamWaitEvent(s, c->mesh_ev, 0));
        PRG_HIP(hipStreamSynchronize(s));
    }
}

}  // namespace prg
