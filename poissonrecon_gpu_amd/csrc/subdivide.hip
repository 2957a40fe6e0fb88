// Coarse-leaf subdivision (reference: ProcessLeafNodesAtOtherDepth / ifSubdivide
// main.cu:2957-2992; the face array :2827-2955 and its surface marks
// :2734-2755; the virtual subtrees, their neighbours and marching cubes
// :1500-2222, 2328-2442, 2481-2825, 3114-3217; driver :3799-4561).
//
// A leaf above the leaf level is refined virtually down to depth D when its
// corner values change sign or one of its faces carries a leaf-level surface
// crossing.  The reference builds one array of virtual nodes per group (each
// root shallower than depth 3 alone, then all roots of one depth together),
// finds neighbours through the parents with the parent's child pointer
// redirected to the virtual root, and runs marching cubes over the virtual
// leaves, one mesh batch per group.
//
// Here no virtual tree is built.  A virtual leaf is (root, Morton offset inside
// the root); its key is the root's key with the offset in the low bits.  The
// combined tree of a group has exactly these facts:
//   - a depth-D cell is virtual iff its ancestor at the root depth r is the
//     root itself or (roots at depth >= 3) another root of depth r;
//   - a real node below depth r is reached from one of the root's 27 real
//     neighbours by its key bits; a coarse leaf (in the group or not) has no
//     real descendants.
// Corners and edges are owned by the smallest-key virtual leaf sharing them
// (real leaves never own, main.cu:1541-1545); vertex values sum the real
// neighbours of the owner's virtual ancestors (virtual coefficients are 0) and
// then the root's real ancestor chain (main.cu:2346-2380).  Leaves are
// numbered in root order, Morton order inside a root: the reference's batch
// order, so one scan numbers every batch's vertices and triangles.
//
// Work is split into bricks: Morton-contiguous blocks of up to 8^3 virtual
// leaves of one root (the whole root when it is smaller), one wavefront each.
// A wave computes the brick's (S+1)^3 corner values into LDS, each from its
// owner's point of view (the owner fixes the summation order), then the
// leaves' cubes and owned crossing edges from LDS.  Three passes:
//   count     every brick: vertex and triangle counts      (k_brick_count)
//   vertices  bricks with output: per-leaf edge masks, cube indices and
//             vertex offsets into small per-brick tables, and the vertices
//   triangles bricks with output: triangle entries, a vertex owned by a leaf
//             of another brick found through that brick's table
// Memory grows with the number of bricks (about 1 M for 506 M virtual
// leaves) and with the bricks that carry output, never with the virtual leaves.
#include <algorithm>
#include <vector>

#include "ctx.h"
#include "mc_common.h"
#include "tile.h"

namespace prg {
namespace {

constexpr int kBlockS = 256;
constexpr int kFinerDepth = 3;  // main.cu:3894: shallower roots go one at a time

// parentFaceKind (MarchingCubes.cuh:708-717), indexed [sonKey][face] literally
__constant__ int8_t c_parent_face[8][6] = {{0, -1, 2, -1, 4, -1}, {-1, 1, 2, -1, 4, -1}, {0, -1, -1, 3, 4, -1},
                                           {-1, 1, -1, 3, 4, -1}, {0, -1, 2, -1, -1, 5}, {-1, 1, 2, -1, -1, 5},
                                           {0, -1, -1, 3, -1, 5}, {0, 1, -1, 3, -1, 5}};
// faceEdges (MarchingCubes.cuh:734-741) as bit masks of edge kinds
constexpr uint16_t kFaceEdges[6] = {(1 << 4) | (1 << 6) | (1 << 8) | (1 << 10), (1 << 5) | (1 << 7) | (1 << 9) | (1 << 11),
                                    (1 << 0) | (1 << 2) | (1 << 8) | (1 << 9),  (1 << 1) | (1 << 3) | (1 << 10) | (1 << 11),
                                    (1 << 0) | (1 << 1) | (1 << 4) | (1 << 5),  (1 << 2) | (1 << 3) | (1 << 6) | (1 << 7)};
// Faces of a leaf cube touched by its triangles' edges, per cube
// configuration: bit j when an edge of face j (kFaceEdges) is an edge of one
// of the configuration's triangles.  Built at compile time from the
// marching-cubes triangle table.
struct CubeFaces {
    uint8_t f[256];
};
constexpr CubeFaces make_cube_faces() {
    constexpr int8_t tri[256][16] = {PRG_MC_TRIANGLE_TABLE_DATA};
    CubeFaces t{};
    for (int cube = 0; cube < 256; cube++) {
        uint32_t em = 0;
        for (int k = 0; k < 15 && tri[cube][k] >= 0; k++) em |= 1u << tri[cube][k];
        uint8_t f = 0;
        for (int j = 0; j < 6; j++)
            if (em & kFaceEdges[j]) f = uint8_t(f | (1u << j));
        t.f[cube] = f;
    }
    return t;
}
__constant__ CubeFaces c_cube_faces = make_cube_faces();
// childrenVertexKind (MarchingCubes.cuh:720-723), used literally
__constant__ int8_t c_children_vertex_kind[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// slots of the device counter array
enum { kCntCoarse = 0, kCntCorners = 1, kCntRoots = 2, kCntBricks = 3, kCntVerts = 4, kCntTris = 5, kCntActive = 6,
       kCntLeaves = 7, kCntErr = 8, kCntDense = 9, kCntDepBricks = 10, kCntDepPairs = 11, kCntWork = 12,
       kCntExact = 13, kCntSuper = 14, kCntList2 = 15, kCntWork2 = 16, kCntExactSorted = 17, kCntWork3 = 18,
       kCntWork4 = 19, kCntDenseSorted = 20, kCntList2Sorted = 21, kCntSlots = 32 };
static_assert(kCntSlots * 4 <= 128, "the pinned counter copy holds 32 ints");

__device__ __forceinline__ int depth_of(int64_t i, const Levels& L, int D) {
    int d = 0;
    while (d < D && i >= L.base[d + 1]) d++;
    return d;
}

__device__ __forceinline__ int nidx(int dx, int dy, int dz) { return 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1); }

// Neighbour j = nidx(dx, dy, dz) of node i at depth d, or -1 (computeNeighbor,
// main.cu:492-509): child LUTchild[s][j] of the parent's neighbour
// LUTparent[s][j], read from the sibling block's parent-neighbour row (the
// level-local child base of each of the parent's 27 neighbours).  Nodes of a
// depth come in sibling blocks of 8 in child order, so the slot s is the
// node's level-local index mod 8.  The root's only neighbour is itself.
__device__ __forceinline__ int32_t neighbour_of(const int32_t* __restrict__ pnt, const Levels& L, int d, int64_t i,
                                                int j) {
    if (d == 0) return j == 13 ? int32_t(i) : -1;
    const int64_t loc = i - L.base[d];
    int entry, child;
    tile_neighbour(int(loc & 7), j, entry, child);
    const int32_t cb = pnt[27 * (L.pnt_off[d] + (loc >> 3)) + entry];
    return cb < 0 ? -1 : int32_t(L.base[d] + cb + child);
}

// The face record holding face j of node X at depth d (initFaceArray /
// maintainFaceNodePointer, main.cu:2827-2955): owned by the smaller key of X
// and its neighbour across the face, numbered (owner, kind seen from the owner).
__device__ __forceinline__ int64_t face_rec(int64_t X, int d, int j, const uint32_t* __restrict__ key,
                                            const int32_t* __restrict__ pnt, const Levels& L) {
    const int o = j >> 1, dl = (j & 1) ? 1 : -1;
    const int32_t nb = neighbour_of(pnt, L, d, X, nidx(o == 0 ? dl : 0, o == 1 ? dl : 0, o == 2 ? dl : 0));
    if (nb != -1 && key[nb] < key[X]) return 6 * int64_t(nb) + (j ^ 1);
    return 6 * X + j;
}

// hasSurfaceIntersection marks from the leaf-level triangles (generateTrianglePos,
// main.cu:2734-2755): every face of a leaf-level node touched by one of its
// triangle edges, and the faces up the parent chain while the face record
// has a parent face.  The reference walks each chain from its leaf-level
// node; here the walk goes level by level.  A node's face j is visited when a
// chain reaches (node, j); the chain moves on to (parent, j) when the face's
// record has a parent face (hasParentFace, MarchingCubes.cuh:708-717, on the
// record's owner and kind).  So the leaf pass marks its own faces and ORs the
// continuing faces into the parent's byte of `up`; each coarser depth, after
// the finer one, does the same for the faces its byte collected.  The set of
// marked records is the reference's.
__device__ __forceinline__ void face_visit(int64_t Y, uint32_t faces, const uint32_t* __restrict__ key,
                                           const int32_t* __restrict__ pnt, const Levels& L,
                                           const int32_t* __restrict__ parent, int d, int D, uint8_t* inter,
                                           uint32_t* up) {
    const uint32_t ky = key[Y];
    uint32_t cont = 0;
    for (int j = 0; j < 6; j++) {
        if (!((faces >> j) & 1u)) continue;
        // the face record (face_rec): owned by the smaller key of Y and its
        // neighbour across the face, both at depth d
        const int o3 = j >> 1, dl = (j & 1) ? 1 : -1;
        const int32_t nb = neighbour_of(pnt, L, d, Y, nidx(o3 == 0 ? dl : 0, o3 == 1 ? dl : 0, o3 == 2 ? dl : 0));
        int64_t owner = Y;
        int kind = j;
        uint32_t ko = ky;
        if (nb != -1) {
            const uint32_t kn = key[nb];
            if (kn < ky) {
                owner = nb;
                kind = j ^ 1;
                ko = kn;
            }
        }
        inter[6 * owner + kind] = 1;
        // only the root (depth 0) has no parent
        if (d > 0 && c_parent_face[(ko >> (3 * (D - d))) & 7u][kind] != -1) cont |= 1u << j;
    }
    if (cont) {
        const int64_t P = parent[Y];
        atomicOr(up + (P >> 2), cont << (8 * (P & 3)));
    }
}

// (cubes: the leaf-level cube indices the marching cubes pass stored)
__global__ void __launch_bounds__(kBlockS) k_face_marks(const uint8_t* cubes, int64_t MD, int64_t baseD,
                                                        const uint32_t* key, const int32_t* pnt, Levels L,
                                                        const int32_t* parent, int D, uint8_t* inter, uint32_t* up) {
    const int64_t t = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (t >= MD) return;
    const uint32_t faces = c_cube_faces.f[cubes[t]];
    if (faces) face_visit(baseD + t, faces, key, pnt, L, parent, D, D, inter, up);
}

// the faces the chains carried up to the nodes of depth d (nodes base .. base + n)
__global__ void __launch_bounds__(kBlockS) k_face_marks_up(int64_t base, int64_t n, int d, const uint32_t* key,
                                                           const int32_t* pnt, Levels L, const int32_t* parent,
                                                           int D, uint8_t* inter, uint32_t* up) {
    const int64_t t = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (t >= n) return;
    const int64_t Y = base + t;
    const uint32_t faces = (up[Y >> 2] >> (8 * (Y & 3))) & 0xffu;
    if (faces) face_visit(Y, faces, key, pnt, L, parent, d, D, inter, up);
}

__global__ void __launch_bounds__(kBlockS) k_flag_coarse_leaves(const int32_t* child_base, int64_t n, int32_t* flag) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < n) flag[i] = (i >= 1 && child_base[i] < 0) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlockS) k_compact(const int32_t* flag, const int32_t* off, int64_t n, int32_t* out) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < n && flag[i]) out[off[i]] = int32_t(i);
}

// mark[list[i]] = 1 for the *n entries of a list of brick numbers below nb
__global__ void __launch_bounds__(kBlockS) k_mark_listed(const int32_t* list, const int32_t* n, int64_t nb,
                                                         int32_t* mark) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < nb && i < *n) mark[list[i]] = 1;
}

// Corner owners of the coarse leaves' corners among the same-depth nodes
// (initVertexOwner, main.cu:1438-1498); marks each (owner, corner) needed.
// ncl (device) coarse leaves in cl.
__global__ void __launch_bounds__(kBlockS) k_coarse_corner_owner(const int32_t* cl, const int32_t* ncl,
                                                                 const uint32_t* key, const int32_t* pnt, Levels L,
                                                                 int D, int32_t* oref, int32_t* need) {
    const int64_t g = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (g >= 8 * int64_t(*ncl)) return;
    const int64_t X = cl[g >> 3];
    const int dX = depth_of(X, L, D);
    const int j = int(g & 7);
    const int cx = j & 1, cy = (j >> 1) & 1, cz = (j >> 2) & 1;
    uint32_t best = key[X];
    int64_t owner = X;
    int jo = j;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int e = 0; e < 2; e++) {
                if (!(a | b | e)) continue;
                const int32_t nb = neighbour_of(pnt, L, dX, X, nidx(a ? 2 * cx - 1 : 0, b ? 2 * cy - 1 : 0, e ? 2 * cz - 1 : 0));
                if (nb == -1) continue;
                if (key[nb] < best) {
                    best = key[nb];
                    owner = nb;
                    jo = j ^ (a | (b << 1) | (e << 2));
                }
            }
    oref[g] = int32_t(8 * owner + jo);
    need[8 * owner + jo] = 1;
}

// computeVertexImplicitFunctionValue (main.cu:2259-2326) for a corner of a
// node above the leaf level: the owner's ancestor chain, then the descent
// through children[childrenVertexKind[corner]] down to D.
// chain_value for lanes that start on different nodes at different depths
// (the owners of coarse-leaf corners, in node order).  Each lane adds its
// depths d, d - 1, ..., 1 in that order, then the root, as chain_value does.
// The deepest lanes step first, so the wave's lanes reach each depth
// together; once every walking lane stands on the same node the rest of the
// chain runs on one wave-uniform index (its row and coefficients come through
// the scalar cache once per wave).  A node with leaf-level slots below it
// (dnum > 0) has a coarse row, whose sum is level_contrib's (row_sum); the
// others gather through the parent-neighbour rows.
__device__ __forceinline__ float coarse_chain_value(int64_t nd, int d, bool valid, float px, float py, float pz,
                                                    const uint32_t* __restrict__ key, const int32_t* __restrict__ pnt,
                                                    const Levels& L, const int32_t* __restrict__ parent,
                                                    const int32_t* __restrict__ dnum, const float* __restrict__ x,
                                                    const float* __restrict__ bf, int D, const float* __restrict__ rows,
                                                    int rows_depth) {
    float val = 0.0f;
    int dm = 0;  // the deepest walking lane's depth
    for (int t = kMaxDepth; t >= 1; t--)
        if (__builtin_amdgcn_ballot_w64(valid && d == t) != 0) {
            dm = t;
            break;
        }
    for (; dm > 0; dm--) {
        const bool walking = valid && d > 0;
        const uint64_t wl = __builtin_amdgcn_ballot_w64(walking);
        const int first = int(__builtin_ctzll(wl));  // (wl != 0 while dm > 0)
        const int64_t n0 = __builtin_amdgcn_readlane(int32_t(nd), first);
        if (__builtin_amdgcn_ballot_w64(walking && (d != dm || nd != n0)) == 0) break;
        if (walking && d == dm) {
            if (dm <= rows_depth && dnum[nd] > 0) {
                const float* R = rows + kRowFloats * (nd - L.base[1]);
                float bx[3], by[3], bz[3];
                node_basis(nd, dm, px, py, pz, key, bf, D, bx, by, bz);
                val = row_sum(R, bx, by, bz, val);
            } else {
                val = level_contrib(nd, dm, px, py, pz, key, pnt, L, x, bf, D, val);
            }
            nd = parent[nd];
            d--;
        }
    }
    if (dm > 0) {  // every walking lane on node u at depth dm
        const uint64_t wl = __builtin_amdgcn_ballot_w64(valid && d > 0);
        int32_t u = __builtin_amdgcn_readlane(int32_t(nd), int(__builtin_ctzll(wl)));
        for (; dm > 0; dm--) {
            const bool row = dm <= rows_depth && dnum[u] > 0;
            float bx[3], by[3], bz[3];
            if (valid) {
                if (row) {
                    node_basis(u, dm, px, py, pz, key, bf, D, bx, by, bz);
                    val = row_sum(rows + kRowFloats * (int64_t(u) - L.base[1]), bx, by, bz, val);
                } else {
                    val = level_contrib(u, dm, px, py, pz, key, pnt, L, x, bf, D, val);
                }
            }
            u = __builtin_amdgcn_readfirstlane(parent[u]);
        }
    }
    float tt = x[0] * basis_value(bf, 0, px);
    tt = tt * basis_value(bf, 0, py);
    tt = tt * basis_value(bf, 0, pz);
    return val + tt;
}

// (five waves per SIMD: 95 VGPRs without spills, 7% faster than four)
__global__ void __launch_bounds__(kBlockS, 5) k_coarse_corner_value(const int32_t* list, const int32_t* n_dev,
                                                                 const uint32_t* key, const int32_t* pnt, Levels L,
                                                                 const int32_t* parent, const int32_t* child_base,
                                                                 const int32_t* dnum, const float* x, const float* bf,
                                                                 const float* rows, int rows_depth, int D,
                                                                 const float* iso, float* out) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    const bool valid = i < *n_dev;
    if (__builtin_amdgcn_ballot_w64(valid) == 0) return;  // (wave-uniform)
    int64_t O = 0;
    int jo = 0, d = 0;
    float px = 0.5f, py = 0.5f, pz = 0.5f;
    if (valid) {
        O = list[i] >> 3;
        jo = list[i] & 7;
        d = depth_of(O, L, D);
        float c[3];
        node_center(key[O], d, D, c);
        const float hw = 1.0f / float(1 << (d + 1));
        px = c[0] + float(2 * (jo & 1) - 1) * hw;
        py = c[1] + float(2 * ((jo & 2) >> 1) - 1) * hw;
        pz = c[2] + float(2 * ((jo & 4) >> 2) - 1) * hw;
    }
    // (every lane of the wave takes part in the chain walk's wave votes)
    float f = coarse_chain_value(O, d, valid, px, py, pz, key, pnt, L, parent, dnum, x, bf, D, rows, rows_depth);
    if (!valid) return;
    const int ex = c_children_vertex_kind[jo];
    int64_t nd = O;
    for (int dd = d + 1; dd <= D; dd++) {
        const int32_t cb = child_base[nd];
        if (cb < 0) break;
        nd = cb + ex;
        f = level_contrib(nd, dd, px, py, pz, key, pnt, L, x, bf, D, f);
    }
    out[i] = f - *iso;
}

// hasTriangle || hasIntersection for each coarse leaf (ProcessLeafNodesAtOtherDepth
// + ifSubdivide, main.cu:2957-2992); 0 past the ncl coarse leaves, so a scan
// over all n slots counts the roots.
__global__ void __launch_bounds__(kBlockS) k_coarse_select(const int32_t* cl, const int32_t* ncl, int64_t n,
                                                           const int32_t* oref, const int32_t* nid, const float* cv,
                                                           const uint32_t* key, const int32_t* pnt, Levels L, int D,
                                                           const uint8_t* inter, int32_t* rflag) {
    const int64_t c = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (c >= n) return;
    if (c >= *ncl) {
        rflag[c] = 0;
        return;
    }
    const int64_t X = cl[c];
    // vertices[0] is the (0,0,0) corner in either numbering
    const float v0 = cv[nid[oref[8 * c]]];
    const float sign = (v0 < 0) ? -1.0f : 1.0f;
    bool tri = false;
    for (int j = 1; j < 8; j++)
        if (sign * cv[nid[oref[8 * c + j]]] < 0) tri = true;
    bool hit = false;
    const int dX = depth_of(X, L, D);
    for (int j = 0; j < 6; j++)
        if (inter[face_rec(X, dX, j, key, pnt, L)]) hit = true;
    rflag[c] = (tri || hit) ? 1 : 0;
}

// ---- roots and bricks ---------------------------------------------------------------
constexpr int kWave = 64;
constexpr int kBrickLog = 3;  // bricks of up to 8^3 virtual leaves
constexpr int kBrickLeaves = 1 << (3 * kBrickLog);
constexpr int kBrickPts = ((1 << kBrickLog) + 1) * ((1 << kBrickLog) + 1) * ((1 << kBrickLog) + 1);

__device__ __forceinline__ int brick_log(int D, int r) { return D - r < kBrickLog ? D - r : kBrickLog; }

// Per root (q < nroots): its real index, bricks and virtual leaves; zero past
// the roots, so scans over all n slots give the brick and leaf offsets.
__global__ void __launch_bounds__(kBlockS) k_root_info(const int32_t* cl, const int32_t* ridx, const int32_t* nroots,
                                                       int64_t n, Levels L, int D, int32_t* roots, int32_t* rbricks,
                                                       int32_t* rleaves) {
    const int64_t q = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (q >= n) return;
    if (q >= *nroots) {
        rbricks[q] = 0;
        rleaves[q] = 0;
        return;
    }
    const int32_t rho = cl[ridx[q]];
    const int r = depth_of(rho, L, D);
    roots[q] = rho;
    rbricks[q] = 1 << (3 * (D - r - brick_log(D, r)));
    rleaves[q] = 1 << (3 * (D - r));
}

__global__ void __launch_bounds__(kBlockS) k_root_of(const int32_t* roots, const int32_t* nroots, int32_t* root_of) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < *nroots) root_of[roots[i]] = int32_t(i);
}

struct SubArgs {
    const uint32_t* key;
    const int32_t* child_base;
    const int32_t* parent;
    const int32_t* pnt;
    Levels L;
    const float* x;
    const float* bf;
    const float4* btab;  // corner basis table (corner_basis_entry)
    const float* rows;   // coarse rows of depths 1 .. D - 1
    int D;
    const float* iso;
    const int32_t* roots;
    const int32_t* boff;     // first brick of each root (exclusive scan, nroots + 1 entries)
    const int32_t* root_of;  // per real node: its root number, or -1
    const int32_t* cnt;      // device counters
    const int32_t* bmap;     // root number of each brick
    const int32_t* rec;      // per root: its record (kRecWords words, root_record)
};

// Per-root record, built once per root and loaded by each of its bricks in
// one coalesced read: the root's 27 neighbours, their child bases and root
// numbers (group), its ancestor chain's neighbour values (27 per depth
// 0..r), the root's index, depth and key.
constexpr int kRecWords = 384;
constexpr int kRecNB = 0, kRecNBK = 27, kRecNBR = 54, kRecX = 81, kRecRho = 378, kRecR = 379, kRecKey = 380;

// A brick and its root (wave-uniform).
struct Brick {
    int32_t q, rho, r, kb;
    int32_t R[3];     // root coordinates at depth r
    int32_t R0[3];    // first depth-D cell of the root
    int32_t Sr;       // root side in depth-D cells
    int32_t org[3];   // first depth-D cell of the brick
    uint32_t bkey;    // key of the brick's first leaf
    bool deep;        // a neighbour of the root has children (real nodes below depth r)
    uint32_t kids;    // bit j: neighbour j (j != 13) has children
};

constexpr int kP = (1 << kBrickLog) + 1;  // grid points per brick side (at most)

// Blocks of the real nodes below the root that a brick's corner owners see:
// at most (D - 3 - r) blocks of 4^3 cells and blocks of 5^3, 7^3 and 11^3
// cells at the three finest depths.
constexpr int kXBlk = 2304;

// LDS of one brick's wave: the root's neighbourhood, the brick's basis
// tables and blocks (BrickCore), and for the evaluation its corner values.
// One basis table entry of the brick (the owner's three values along one
// axis): 12 bytes, the corner basis table's fourth float is never read.
struct BasisT {
    float x, y, z;
};

template <int kXBCap, int kTP = kP, bool kHasT = true>
struct BrickCoreT {
    static constexpr int kTPts = kTP;  // grid coordinates per axis (of T, or of the screen's ranges)
    // per depth d, axis a and grid coordinate i: the corner basis table entry
    // of the owner (own_cell) of grid coordinate org[a] + i, (d, a, i) at
    // d * 3 kP + a kP + i (the screens reduce the entries straight from the
    // global table and keep no T)
    // (a layout without blocks, kXBCap = 1, is for the mesh passes, which
    // evaluate nothing: no table and no chain)
    BasisT T[kHasT && kXBCap > 1 ? (kMaxDepth + 1) * 3 * kTP : 1];
    float X[kXBCap > 1 ? (kMaxDepth + 1) * 27 : 1];  // the root's chain: neighbour values at depths 0..r
    float XB[kXBCap];                  // per depth below the root: its block (xb_off)
    int32_t xb_off[kMaxDepth + 1];     // start of depth dd's block in XB, or -1 (no real node in it)
    int32_t NB[27], NBK[27], NBR[27];  // the root's neighbours, their child bases, their root numbers (group)
    float tmax[kMaxDepth + 1];         // brick_abs_bound: per depth, the largest |basis value|
    float xsum[kMaxDepth + 1];         // and the sum of |coefficients|
    // brick_interval_sign (the screens, which keep no T): per depth, axis and
    // block slot, the range of the weight over the brick's points
    float wlo[kHasT ? 1 : (kMaxDepth + 1) * 12];
    float whi[kHasT ? 1 : (kMaxDepth + 1) * 12];
};

using BrickCore = BrickCoreT<kXBlk>;
// The certificate screen's LDS: blocks up to kXBlkScreen floats (the depths
// where the brick lies in one cell, and the first finer one); bricks that
// need more go to the exact certificate, whose LDS holds every block.
constexpr int kXBlkScreen = 512;
using BrickScreenLds = BrickCoreT<kXBlkScreen, kP, false>;
// The super-brick screen's: 2x2x2 bricks, 17 grid coordinates per axis.
constexpr int kPSuper = 2 * (1 << kBrickLog) + 1;
using BrickSuperLds = BrickCoreT<kXBlkScreen, kPSuper, false>;

// The vertex and triangle passes read the corner values the count pass
// stored (the count pass keeps them over its blocks): no blocks.
struct BrickMeshLds : BrickCoreT<1> {
    float G[kBrickPts];  // f - iso at the brick's corners: (i, j, k) at (k kP + j) kP + i
};

__device__ __forceinline__ void decode_cell(uint32_t k, int32_t c[3]) {
    c[0] = int32_t(compact3(k >> 2));
    c[1] = int32_t(compact3(k >> 1));
    c[2] = int32_t(compact3(k));
}

// The record of root q (one wave per root).
__global__ void __launch_bounds__(kWave) k_root_records(SubArgs A, int32_t* rec) {
    const int D = A.D, lane = threadIdx.x;
    const int32_t q = blockIdx.x;
    const int32_t rho = A.roots[q];
    const int r = depth_of(rho, A.L, D);
    int32_t* R = rec + int64_t(q) * kRecWords;
    if (lane < 27) {
        const int32_t nb = neighbour_of(A.pnt, A.L, r, rho, lane);
        R[kRecNB + lane] = nb;
        R[kRecNBK + lane] = nb < 0 ? -1 : A.child_base[nb];
        R[kRecNBR + lane] = (nb < 0 || r < kFinerDepth) ? -1 : A.root_of[nb];
        R[kRecX + 27 * r + lane] = __float_as_int(nb < 0 ? 0.0f : A.x[nb]);
    }
    // ancestors' rows (their 27 neighbour values, then the parent), depths r-1 .. 1
    int32_t a = A.parent[rho];
    for (int d = r - 1; d >= 1; d--) {
        const float* Row = A.rows + kRowFloats * (int64_t(a) - A.L.base[1]);
        if (lane < 27) R[kRecX + 27 * d + lane] = __float_as_int(Row[lane]);
        a = __builtin_amdgcn_readfirstlane(__float_as_int(Row[27]));
    }
    if (lane == 0) {
        R[kRecX] = __float_as_int(A.x[0]);
        R[kRecRho] = rho;
        R[kRecR] = r;
        R[kRecKey] = int32_t(A.key[rho]);
    }
}

// bmap[b] = the root of brick b (binary search of the brick offsets).
__global__ void __launch_bounds__(kBlockS) k_brick_map(const int32_t* boff, const int32_t* cnt, int64_t nb,
                                                       int32_t* bmap) {
    const int64_t b = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (b >= nb) return;
    int32_t lo = 0, hi = cnt[kCntRoots] - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    bmap[b] = lo;
}

// Loads brick b's root record into LDS (with the chain's neighbour values
// when kChain) and sets the brick's geometry.
// q_loaded: the root whose record S already holds (the previous brick's, in a
// kernel that walks consecutive bricks), or -1.  up = 1 sets up the
// super-brick of the 2x2x2 bricks that start at brick b.
template <bool kChain, class L>
__device__ __forceinline__ void brick_setup(const SubArgs& A, int32_t b, Brick& B, L& S, int32_t q_loaded = -1,
                                            int up = 0) {
    const int D = A.D, lane = threadIdx.x;
    B.q = __builtin_amdgcn_readfirstlane(A.bmap[b]);
    const int32_t* R = A.rec + int64_t(B.q) * kRecWords;
    B.rho = __builtin_amdgcn_readfirstlane(R[kRecRho]);
    B.r = __builtin_amdgcn_readfirstlane(R[kRecR]);
    const uint32_t rkey = uint32_t(__builtin_amdgcn_readfirstlane(R[kRecKey]));
    if (B.q != q_loaded) {
        for (int w = lane; w < kRecX; w += kWave) {
            const int32_t v = R[w];
            if (w < kRecNBK) S.NB[w] = v;
            else if (w < kRecNBR) S.NBK[w - kRecNBK] = v;
            else S.NBR[w - kRecNBR] = v;
        }
        if constexpr (kChain) {
            static_assert(sizeof(S.X) == (kMaxDepth + 1) * 27 * sizeof(float), "the layout holds the chain");
            for (int w = lane; w < 27 * (B.r + 1); w += kWave) S.X[w] = __int_as_float(R[kRecX + w]);
        }
    }
    B.kb = brick_log(D, B.r) + up;  // up = 1: the 2x2x2 bricks from brick b (a multiple of 8 in the root)
    int32_t rc[3];
    decode_cell(rkey, rc);
    B.Sr = 1 << (D - B.r);
    for (int a = 0; a < 3; a++) {
        B.R[a] = rc[a] >> (D - B.r);
        B.R0[a] = B.R[a] << (D - B.r);
    }
    B.bkey = rkey | (uint32_t(b - A.boff[B.q]) << (3 * (B.kb - up)));
    decode_cell(B.bkey, B.org);
    __syncthreads();
    const uint64_t kids = __builtin_amdgcn_ballot_w64(lane < 27 && lane != 13 && S.NBK[lane < 27 ? lane : 0] >= 0);
    B.kids = uint32_t(kids);
    B.deep = B.kids != 0;
}

// Whether the depth-D cell c (inside the root's 3x3x3 neighbourhood) is a
// virtual leaf of the brick's group; *q receives its root's neighbour slot
// (13 for the brick's own root).
template <class L>
__device__ __forceinline__ bool cell_virtual(const Brick& B, const L& S, int D, const int32_t c[3], int* slot) {
    const int W = 1 << D;
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= W || c[1] >= W || c[2] >= W) return false;
    const int sh = D - B.r;
    const int ax = (c[0] >> sh) - B.R[0], ay = (c[1] >> sh) - B.R[1], az = (c[2] >> sh) - B.R[2];
    const int j = nidx(ax, ay, az);
    *slot = j;
    if (j == 13) return true;
    return S.NBR[j] >= 0;
}

// Width in cells of the block at depth dd: the owners' ancestors (the cells
// the brick spans, and one more below for owners in the lower brick) +-1.
__device__ __forceinline__ int block_width(int D, int kb, int dd) {
    const int e = dd - (D - kb);
    return (e > 0 ? (1 << e) : 1) + 3;
}

// Along one axis of grid coordinate g, the owner cell for a root whose owners
// are all its own leaves: the smaller of the two cells around g that lies in
// the root (Morton keys grow with each coordinate).
__device__ __forceinline__ int32_t own_cell(int32_t g, int32_t R0) { return g == R0 ? g : g - 1; }

// Loads the brick's basis table entries (S.T) for owners inside its root.
// The smallest-key leaf around a grid point, when it is a leaf of the
// brick's own root, is own_cell along every axis, so every corner the
// brick's root owns reads its three basis values per depth from S.T.
template <class L>
__device__ __forceinline__ void brick_tables(const SubArgs& A, const Brick& B, L& S) {
    const int D = A.D, P = (1 << B.kb) + 1;
    constexpr int TP = L::kTPts;
    for (int e = threadIdx.x; e < (D + 1) * TP; e += kWave) {
        const int d = e / TP, i = e % TP;
        if (i >= P) continue;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {  // constant indices keep B in registers
            const int32_t g = B.org[ax] + i, o = own_cell(g, B.R0[ax]);
            const float4 t = A.btab[((int64_t(d) << D) + o) * 2 + (g - o)];
            S.T[3 * TP * d + TP * ax + i] = BasisT{t.x, t.y, t.z};
        }
    }
}

// Fills the brick's blocks (depths r+1..D): each cell's coefficient.  A
// depth's block is needed when it reaches outside the root into a
// neighbour with children; the parents of a block's cells lie in the block
// one depth up, so the depths that need one are r+1 .. some dmax, and the
// blocks are filled level by level: the cells of depth r+1 from the
// neighbours' child bases, each deeper cell from its parent's node (one
// child-base load per cell).  A block holds its cells' node indices until
// the next depth has read them, then their coefficients.  A depth whose
// block holds no real node gets none.
template <class L>
__device__ __forceinline__ bool brick_blocks(const SubArgs& A, const Brick& B, L& S) {
    constexpr int kCap = int(sizeof(S.XB) / sizeof(float));
    const int D = A.D, lane = threadIdx.x;
    int off[kMaxDepth + 1];  // uniform: constant indices below keep it in scalar registers
    int total = 0, dmax = B.r;
#pragma unroll
    for (int dd = 0; dd <= kMaxDepth; dd++) {
        off[dd] = -1;
        if (dd > D || dd <= B.r || !B.deep || dmax != dd - 1) continue;
        const int W = block_width(D, B.kb, dd), sh = D - dd, up = dd - B.r;
        // the root slots the block touches: filled only if one of them has children
        uint32_t in[3];  // per axis, bit t: offset t - 1 is touched
        for (int a = 0; a < 3; a++) {
            const int32_t lo = (B.org[a] >> sh) - 2;
            const int slo = (lo >> up) - B.R[a], shi = ((lo + W - 1) >> up) - B.R[a];
            in[a] = (slo <= -1 ? 1u : 0u) | (slo <= 0 && shi >= 0 ? 2u : 0u) | (shi >= 1 ? 4u : 0u);
        }
        uint32_t touched = 0;
#pragma unroll
        for (int j = 0; j < 27; j++)
            if (((in[0] >> (j / 9)) & (in[1] >> ((j / 3) % 3)) & (in[2] >> (j % 3))) & 1u) touched |= 1u << j;
        if (!(touched & B.kids)) continue;
        if (total + W * W * W > kCap) return false;  // (uniform) the caller's LDS cannot hold the blocks
        off[dd] = total;
        total += W * W * W;
        dmax = dd;
    }
    int32_t* NI = reinterpret_cast<int32_t*>(S.XB);
    uint32_t nz = 0;  // depths where this lane found a real node
#pragma unroll
    for (int dd = 1; dd <= kMaxDepth + 1; dd++) {
        if (dd <= kMaxDepth && dd > B.r && dd <= dmax) {
            // node indices of depth dd
            const int W = block_width(D, B.kb, dd), sh = D - dd, n = W * W * W;
            const uint32_t rw = (65536u + uint32_t(W) - 1u) / uint32_t(W);
            const int32_t lo[3] = {(B.org[0] >> sh) - 2, (B.org[1] >> sh) - 2, (B.org[2] >> sh) - 2};
            const int32_t width = 1 << dd;
            const int Wp = dd > B.r + 1 ? block_width(D, B.kb, dd - 1) : 0;
            const int32_t plo[3] = {(B.org[0] >> (sh + 1)) - 2, (B.org[1] >> (sh + 1)) - 2, (B.org[2] >> (sh + 1)) - 2};
            for (int e = lane; e < n; e += kWave) {
                // l / W and l / W^2 for l < W^3 <= 1331 by a 16-bit reciprocal (exact: l / W
                // has a fractional part of at most (W-1)/W, the reciprocal's excess adds < 0.03)
                const int q1 = int((uint32_t(e) * rw) >> 16), q2 = int((uint32_t(q1) * rw) >> 16);
                const int32_t c[3] = {lo[0] + q2, lo[1] + (q1 - q2 * W), lo[2] + (e - q1 * W)};
                const int oct = ((c[0] & 1) << 2) | ((c[1] & 1) << 1) | (c[2] & 1);
                int32_t node = -1;
                if (c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < width && c[1] < width && c[2] < width) {
                    if (dd == B.r + 1) {
                        const int ox = (c[0] >> 1) - B.R[0], oy = (c[1] >> 1) - B.R[1], oz = (c[2] >> 1) - B.R[2];
                        const int32_t kids = (ox | oy | oz) ? S.NBK[nidx(ox, oy, oz)] : -1;
                        node = kids < 0 ? -1 : kids + oct;
                    } else {
                        const int32_t pe =
                            (((c[0] >> 1) - plo[0]) * Wp + ((c[1] >> 1) - plo[1])) * Wp + ((c[2] >> 1) - plo[2]);
                        const int32_t par = NI[off[dd - 1] + pe];
                        const int32_t cb = par < 0 ? -1 : A.child_base[par];
                        node = cb < 0 ? -1 : cb + oct;
                    }
                }
                NI[off[dd] + e] = node;
            }
        }
        // coefficients of depth dd - 1, whose node indices depth dd has read
        const int dc = dd - 1;
        if (dc > B.r && dc <= dmax) {
            const int W = block_width(D, B.kb, dc), n = W * W * W;
            for (int e = lane; e < n; e += kWave) {
                const int32_t node = NI[off[dc] + e];
                const float xv = node < 0 ? 0.0f : A.x[node];
                S.XB[off[dc] + e] = xv;
                if (xv != 0.0f) nz |= 1u << dc;
            }
        }
        if (dd <= dmax + 1) __syncthreads();  // this depth's node indices are read by the next
    }
#pragma unroll
    for (int dd = 0; dd <= kMaxDepth; dd++) {
        const bool any = __builtin_amdgcn_ballot_w64((nz >> dd) & 1u) != 0;
        if (lane == 0) S.xb_off[dd] = any ? off[dd] : -1;
    }
    __syncthreads();
    return true;
}

// deep_sum for an owner leaf o of the brick's own root (the owner of grid
// point org + gi), reading the blocks and S.T.  Every one of the 27 terms is
// added: a neighbour that is virtual, absent or outside the domain
// contributes +-0 (its coefficient or its basis value is 0), which leaves a
// float sum that started at +0 unchanged, so the result is deep_sum's bit
// for bit.
template <class L>
__device__ __forceinline__ float deep_sum_lds(const SubArgs& A, const Brick& B, const L& S, const int32_t o[3], const int gi[3]) {
    const int D = A.D;
    float f = 0.0f;
    for (int dd = D; dd > B.r; dd--) {
        const int slot = S.xb_off[dd];
        if (slot < 0) continue;
        const int up = dd - B.r, sh = D - dd;
        const int32_t ax = o[0] >> sh, ay = o[1] >> sh, az = o[2] >> sh;
        const int32_t ext = (1 << up) - 1;
        const int32_t lx = B.R[0] << up, ly = B.R[1] << up, lz = B.R[2] << up;
        if (ax - 1 >= lx && ax + 1 <= lx + ext && ay - 1 >= ly && ay + 1 <= ly + ext && az - 1 >= lz &&
            az + 1 <= lz + ext)
            continue;
        const int W = block_width(D, B.kb, dd);
        const int bx0 = ax - 1 - ((B.org[0] >> sh) - 2), by0 = ay - 1 - ((B.org[1] >> sh) - 2),
                  bz0 = az - 1 - ((B.org[2] >> sh) - 2);
        const BasisT tx = S.T[3 * kP * dd + gi[0]];
        const BasisT ty = S.T[3 * kP * dd + kP + gi[1]];
        const BasisT tz = S.T[3 * kP * dd + 2 * kP + gi[2]];
        const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
        const float* Xb = S.XB + slot;
#pragma unroll
        for (int k = 0; k < 27; k++) {
            const float xv = Xb[((bx0 + k / 9) * W + (by0 + (k / 3) % 3)) * W + (bz0 + k % 3)];
            float tt = xv * bx[k / 9];
            tt = tt * by[(k / 3) % 3];
            tt = tt * bz[k % 3];
            f = f + tt;
        }
    }
    return f;
}

// Sum over the real nodes below depth r (virtual coefficients are 0) for the
// corner of owner leaf o on sides s, in the reference's order: depth D up to
// r + 1, neighbours j = 0..26 of the owner's virtual ancestor
// (computeSubdivideVertexImplicitFunctionValue, main.cu:2346-2372).  R is the
// owner's root at depth r; nb(j) its neighbour j and kids(j) that neighbour's
// child base.
template <class NbFn, class KidsFn>
__device__ __forceinline__ float deep_sum(const SubArgs& A, int r, const int32_t R[3], const int32_t o[3], const int s[3], NbFn nb,
                          KidsFn kids) {
    const int D = A.D;
    float f = 0.0f;
    for (int dd = D; dd > r; dd--) {
        const int up = dd - r, sh = D - dd;
        const int32_t ax = o[0] >> sh, ay = o[1] >> sh, az = o[2] >> sh;
        const int32_t ext = (1 << up) - 1;
        const int32_t lx = R[0] << up, ly = R[1] << up, lz = R[2] << up;
        // all 27 neighbours inside the root: virtual, nothing to add
        if (ax - 1 >= lx && ax + 1 <= lx + ext && ay - 1 >= ly && ay + 1 <= ly + ext && az - 1 >= lz &&
            az + 1 <= lz + ext)
            continue;
        const float4 tx = A.btab[((int64_t(dd) << D) + o[0]) * 2 + s[0]];
        const float4 ty = A.btab[((int64_t(dd) << D) + o[1]) * 2 + s[1]];
        const float4 tz = A.btab[((int64_t(dd) << D) + o[2]) * 2 + s[2]];
        const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
        const int32_t width = 1 << dd;
#pragma unroll
        for (int k = 0; k < 27; k++) {
            const int32_t nx = ax + (k / 9 - 1), ny = ay + ((k / 3) % 3 - 1), nz = az + (k % 3 - 1);
            if (nx < 0 || ny < 0 || nz < 0 || nx >= width || ny >= width || nz >= width) continue;
            const int cx = (nx >> up) - R[0], cy = (ny >> up) - R[1], cz = (nz >> up) - R[2];
            if (!(cx | cy | cz)) continue;  // inside the root: virtual
            const int jn = nidx(cx, cy, cz);
            if (kids(jn) < 0) continue;  // absent, or a coarse leaf: no real node below
            int32_t cur = nb(jn);
            for (int l = r + 1; l <= dd && cur >= 0; l++) {
                const int32_t cb = A.child_base[cur];
                const int s2 = dd - l;
                cur = cb < 0 ? -1 : cb + ((((nx >> s2) & 1) << 2) | (((ny >> s2) & 1) << 1) | ((nz >> s2) & 1));
            }
            if (cur < 0) continue;
            float tt = A.x[cur] * bx[k / 9];
            tt = tt * by[(k / 3) % 3];
            tt = tt * bz[k % 3];
            f = f + tt;
        }
    }
    return f;
}

// The root's chain from LDS (chain_value's order: depth r down to 1, then the
// root of the tree), continuing from f, for the owner of grid point org + gi
// in the brick's own root.
template <class L>
__device__ __forceinline__ float chain_lds(const L& S, int r, const int gi[3], float f) {
    for (int d = r; d >= 1; d--) {
        const BasisT tx = S.T[3 * kP * d + gi[0]];
        const BasisT ty = S.T[3 * kP * d + kP + gi[1]];
        const BasisT tz = S.T[3 * kP * d + 2 * kP + gi[2]];
        const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
        const float* X = S.X + 27 * d;
#pragma unroll
        for (int j = 0; j < 27; j++) {
            float tt = X[j] * bx[j / 9];
            tt = tt * by[(j / 3) % 3];
            tt = tt * bz[j % 3];
            f = f + tt;
        }
    }
    const float b0x = S.T[gi[0]].y, b0y = S.T[kP + gi[1]].y, b0z = S.T[2 * kP + gi[2]].y;
    float tt = S.X[0] * b0x;
    tt = tt * b0y;
    tt = tt * b0z;
    return f + tt;
}

// The owner of grid point g (depth-D corner coordinates) of the brick's root
// cell: the smallest-key virtual leaf of the group around g.  Returns its
// root's neighbour slot (13: the brick's own root) and its cell in o.
template <class L>
__device__ __forceinline__ int corner_owner(const Brick& B, const L& S, int D, const int32_t g[3],
                                            int32_t o[3]) {
    uint32_t best = 0xffffffffu;
    int slot = 13;
    for (int t = 0; t < 8; t++) {
        const int32_t c[3] = {g[0] - 1 + (t & 1), g[1] - 1 + ((t >> 1) & 1), g[2] - 1 + (t >> 2)};
        int sl;
        if (!cell_virtual(B, S, D, c, &sl)) continue;
        const uint32_t k = cell_key(uint32_t(c[0]), uint32_t(c[1]), uint32_t(c[2]));
        if (k < best) {
            best = k;
            slot = sl;
            for (int a = 0; a < 3; a++) o[a] = c[a];
        }
    }
    return slot;
}

// f - iso at grid point g (depth-D corner coordinates) of the brick's root cell,
// computed by its owner: the smallest-key virtual leaf around it.
template <class L>
__device__ __forceinline__ float grid_value(const SubArgs& A, const Brick& B, const L& S, const int32_t g[3]) {
    const int D = A.D;
    int32_t o[3];
    int side[3];
    int slot = 13;
    const bool interior = g[0] > B.R0[0] && g[0] < B.R0[0] + B.Sr && g[1] > B.R0[1] && g[1] < B.R0[1] + B.Sr &&
                          g[2] > B.R0[2] && g[2] < B.R0[2] + B.Sr;
    if (interior) {
        // all 8 cells around are leaves of this root; (g - 1) has the smallest key
        for (int a = 0; a < 3; a++) {
            o[a] = g[a] - 1;
            side[a] = 1;
        }
    } else {
        slot = corner_owner(B, S, D, g, o);
        for (int a = 0; a < 3; a++) side[a] = g[a] - o[a];
    }
    if (slot == 13) {
        // o is own_cell along every axis: S.T holds its basis values
        const int gi[3] = {g[0] - B.org[0], g[1] - B.org[1], g[2] - B.org[2]};
        const float f = B.deep ? deep_sum_lds(A, B, S, o, gi) : 0.0f;
        return chain_lds(S, B.r, gi, f) - *A.iso;
    }
    // owned by a leaf of another root of the group: that root's record (its
    // neighbours, their child bases, its chain's values: what the neighbour
    // planes and chain_value would read), with the owner's basis values from
    // the corner basis table
    const int32_t* R2 = A.rec + int64_t(S.NBR[slot]) * kRecWords;
    const int32_t RC[3] = {B.R[0] + slot / 9 - 1, B.R[1] + (slot / 3) % 3 - 1, B.R[2] + slot % 3 - 1};
    float f = deep_sum(
        A, B.r, RC, o, side, [&](int j) { return R2[kRecNB + j]; }, [&](int j) { return R2[kRecNBK + j]; });
    for (int d = B.r; d >= 1; d--) {
        const float4 tx = A.btab[((int64_t(d) << D) + o[0]) * 2 + side[0]];
        const float4 ty = A.btab[((int64_t(d) << D) + o[1]) * 2 + side[1]];
        const float4 tz = A.btab[((int64_t(d) << D) + o[2]) * 2 + side[2]];
        const float bx[3] = {tx.x, tx.y, tx.z}, by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
        const int32_t* X = R2 + kRecX + 27 * d;
#pragma unroll
        for (int j = 0; j < 27; j++) {
            float tt = __int_as_float(X[j]) * bx[j / 9];
            tt = tt * by[(j / 3) % 3];
            tt = tt * bz[j % 3];
            f = f + tt;
        }
    }
    const float b0x = A.btab[o[0] * 2 + side[0]].y, b0y = A.btab[o[1] * 2 + side[1]].y, b0z = A.btab[o[2] * 2 + side[2]].y;
    float tt = __int_as_float(R2[kRecX]) * b0x;
    tt = tt * b0y;
    tt = tt * b0z;
    return f + tt - *A.iso;
}

// The brick's corner values (grid_value; 0 at the points a smaller brick
// does not have), stored over its blocks, which nothing reads once the values
// are known: each lane holds its values in registers until every lane's
// block reads are done.  The blocks' LDS is then the corner values' (G), so
// the evaluation needs no LDS beyond the certificate's.  The values also go
// to gv (kBrickPts floats).
constexpr int kValRounds = (kBrickPts + kWave - 1) / kWave;

template <class L>
__device__ __forceinline__ float* brick_values(const SubArgs& A, const Brick& B, L& S, float* gv) {
    static_assert(sizeof(S.XB) >= kBrickPts * sizeof(float), "the corner values fit over the blocks");
    const int P = (1 << B.kb) + 1;
    float held[kValRounds];
#pragma unroll
    for (int t = 0; t < kValRounds; t++) held[t] = 0.0f;
    for (int t = 0; t < kValRounds; t++) {
        const int p = threadIdx.x + kWave * t;
        const int i = p % kP, j = (p / kP) % kP, k = p / (kP * kP);
        float v = 0.0f;
        if (p < kBrickPts && i < P && j < P && k < P) {
            const int32_t g[3] = {B.org[0] + i, B.org[1] + j, B.org[2] + k};
            v = grid_value(A, B, S, g);
        }
        // (a register queue: constant indices keep held out of scratch memory)
#pragma unroll
        for (int q = 0; q + 1 < kValRounds; q++) held[q] = held[q + 1];
        held[kValRounds - 1] = v;
    }
    __syncthreads();  // the blocks' last reads
    float* G = S.XB;
#pragma unroll
    for (int t = 0; t < kValRounds; t++) {
        const int p = threadIdx.x + kWave * t;
        if (p < kBrickPts) {
            G[p] = held[t];
            gv[p] = held[t];
        }
    }
    __syncthreads();
    return G;
}

// Owner of edge e of leaf v among the virtual leaves sharing it
// (initSubdivideEdgeArray, main.cu:1867-2019): its cell and the edge kind seen
// from it.
template <class L>
__device__ __forceinline__ void edge_owner(const Brick& B, const L& S, int D, const int32_t v[3], int e,
                                           int32_t ow[3], int* eo) {
    const int o = e >> 2;
    const int a1 = (o == 0) ? 1 : 0, a2 = (o == 2) ? 1 : 2;
    const int b1 = e & 1, b2 = (e >> 1) & 1;
    uint32_t best = cell_key(uint32_t(v[0]), uint32_t(v[1]), uint32_t(v[2]));
    for (int a = 0; a < 3; a++) ow[a] = v[a];
    *eo = e;
    for (int u = 0; u < 2; u++)
        for (int w = 0; w < 2; w++) {
            if (!(u | w)) continue;
            int32_t n[3] = {v[0], v[1], v[2]};
            if (u) n[a1] += 2 * b1 - 1;
            if (w) n[a2] += 2 * b2 - 1;
            int sl;
            if (!cell_virtual(B, S, D, n, &sl)) continue;
            const uint32_t kk = cell_key(uint32_t(n[0]), uint32_t(n[1]), uint32_t(n[2]));
            if (kk < best) {
                best = kk;
                for (int a = 0; a < 3; a++) ow[a] = n[a];
                *eo = (o << 2) | ((b2 ^ w) << 1) | (b1 ^ u);
            }
        }
}

// Cube index and owned crossing edges of leaf l of the brick
// (generateSubdivideVexNums, main.cu:2481-2531; generateTriNums, :2562-2582).
template <class L>
__device__ __forceinline__ void leaf_cube(const Brick& B, const L& S, const float* G, int D, int l, int32_t v[3],
                                          float cv[8], int* cube, int* emask) {
    const int lx = int(compact3(uint32_t(l) >> 2)), ly = int(compact3(uint32_t(l) >> 1)), lz = int(compact3(uint32_t(l)));
    v[0] = B.org[0] + lx;
    v[1] = B.org[1] + ly;
    v[2] = B.org[2] + lz;
    int cb = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = (q >> 1) & 1, xx = (q & 1) ^ y, z = q >> 2;
        cv[q] = G[((lz + z) * kP + (ly + y)) * kP + (lx + xx)];
        if (cv[q] < 0) cb |= 1 << q;
    }
    *cube = cb;
    // a leaf not on its root's boundary owns exactly its edges on the high side
    // of both other axes (kinds 3, 7, 11): its 26 neighbours are all in the root
    bool inner = true;
    for (int a = 0; a < 3; a++) inner = inner && v[a] > B.R0[a] && v[a] < B.R0[a] + B.Sr - 1;
    int m = 0;
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const float f1 = cv[c_edge_vertex_h[e][0]], f2 = cv[c_edge_vertex_h[e][1]];
        if (!(f1 * f2 <= 0)) continue;
        bool own;
        if (inner) {
            own = (e & 3) == 3;
        } else {
            int32_t ow[3];
            int eo;
            edge_owner(B, S, D, v, e, ow, &eo);
            own = ow[0] == v[0] && ow[1] == v[1] && ow[2] == v[2];
        }
        if (own) m |= 1 << e;
    }
    *emask = m;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ int wave_excl_scan(int v) {
    int inc = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(inc, o, kWave);
        if (int(threadIdx.x) >= o) inc += t;
    }
    return inc - v;
}

// The separable contraction of one depth's coefficients over z and y for
// the lane's weights: Y[m] = sum_n wy[n] sum_l wz[l] X[m][n][l].
// contract27: the root's chain at a depth 1..r, whose 27 neighbours all lie
// around the owners' common ancestor (3 slots per axis, weights (t.x, t.y,
// t.z)).  contract64: a 4^3 block (cells A - 2 .. A + 1 around the brick's
// ancestor A) with 4-slot weights.
__device__ __forceinline__ void contract27(const float* X27, BasisT ty, BasisT tz, float Y[3]) {
    const float wy[3] = {ty.x, ty.y, ty.z}, wz[3] = {tz.x, tz.y, tz.z};
    float xv[27];
#pragma unroll
    for (int q = 0; q < 27; q++) xv[q] = X27[q];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        float y = 0.0f;
#pragma unroll
        for (int n = 0; n < 3; n++) {
            float z = 0.0f;
#pragma unroll
            for (int l = 0; l < 3; l++) z = __builtin_fmaf(xv[9 * m + 3 * n + l], wz[l], z);
            y = __builtin_fmaf(wy[n], z, y);
        }
        Y[m] = y;
    }
}

__device__ __forceinline__ void contract64(const float* X64, const float wy[4], const float wz[4], float Y[4]) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
        float xv[16];
#pragma unroll
        for (int q = 0; q < 16; q++) xv[q] = X64[16 * m + q];
        float y = 0.0f;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            float z = 0.0f;
#pragma unroll
            for (int l = 0; l < 4; l++) z = __builtin_fmaf(xv[4 * n + l], wz[l], z);
            y = __builtin_fmaf(wy[n], z, y);
        }
        Y[m] = y;
    }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ float wave_fsum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// An upper bound of A(g) = sum over depths and the owner's 27 neighbours of
// |x_k b_x b_y b_z| for every corner g of the brick owned inside its root:
// per depth, the sum of |x| over the depth's block (which holds every
// owner's 27 neighbours) times the cube of the largest |basis value| among
// the brick's table entries at that depth.  The float sums of non-negative
// terms are within 0.1% of their exact values.
template <class L>
__device__ __forceinline__ float brick_abs_bound(const Brick& B, L& S, int D) {
    const int lane = threadIdx.x, P = (1 << B.kb) + 1, dlast = D - B.kb;
    if (lane <= kMaxDepth) S.tmax[lane] = S.xsum[lane] = 0.0f;
    __syncthreads();
    // |basis values| (non-negative floats order as their bit patterns)
    constexpr int TP = L::kTPts;
    for (int e = lane; e < (D + 1) * 3 * TP; e += kWave) {
        const int d = e / (3 * TP), i = e % TP;
        if (i >= P || (d > B.r && S.xb_off[d] < 0)) continue;
        const BasisT t = S.T[e];
        const float m = fmaxf(fabsf(t.x), fmaxf(fabsf(t.y), fabsf(t.z)));
        atomicMax(reinterpret_cast<int*>(&S.tmax[d]), __float_as_int(m));
    }
    // |coefficients|: the chain, then the blocks
    for (int e = lane; e < 27 * B.r; e += kWave) atomicAdd(&S.xsum[1 + e / 27], fabsf(S.X[27 + e]));
    if (lane == 0) atomicAdd(&S.xsum[0], fabsf(S.X[0]));
    for (int d = B.r + 1; d <= D; d++) {
        if (S.xb_off[d] < 0) continue;
        const int W = d <= dlast ? 4 : block_width(D, B.kb, d), n = W * W * W;
        float sx = 0.0f;
        for (int e = lane; e < n; e += kWave) sx += fabsf(S.XB[S.xb_off[d] + e]);
        if (sx != 0.0f) atomicAdd(&S.xsum[d], sx);
    }
    __syncthreads();
    float bound = 0.0f;
#pragma unroll
    for (int d = 0; d <= kMaxDepth; d++) {
        const float tm = S.tmax[d];
        if (d <= D) bound += S.xsum[d] * (tm * tm * tm);
    }
    return 1.001f * bound;
}

// brick_abs_bound for the screens, and the ranges brick_interval_sign needs,
// from the corner basis table in global memory (the screens keep no T): one
// lane per (depth, axis) reads the owner's entry for each of the brick's grid
// coordinates (brick_tables' entries), takes the largest |basis value| (for
// tmax) and, at depths d <= D - kb, each block slot's weight range over the
// points (wlo, whi).  The figures are those brick_abs_bound and the range
// pass of the T path compute.
template <class L>
__device__ __forceinline__ float brick_screen_bound(const SubArgs& A, const Brick& B, L& S) {
    static_assert(sizeof(S.wlo) == (kMaxDepth + 1) * 12 * sizeof(float), "the screens' LDS holds the slot ranges");
    const int D = A.D, lane = threadIdx.x, P = (1 << B.kb) + 1, dlast = D - B.kb;
    if (lane <= kMaxDepth) S.tmax[lane] = S.xsum[lane] = 0.0f;
    __syncthreads();
    if (lane < 3 * (D + 1)) {
        const int d = lane / 3, a = lane % 3;
        if (d <= B.r || S.xb_off[d] >= 0) {
            // (arithmetic selects: an index by a would put B in scratch memory)
            const int32_t og = (a == 0) * B.org[0] + (a == 1) * B.org[1] + (a == 2) * B.org[2];
            const int32_t r0 = (a == 0) * B.R0[0] + (a == 1) * B.R0[1] + (a == 2) * B.R0[2];
            float mx = 0.0f, lo[4], hi[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                lo[m] = 3.4e38f;
                hi[m] = -3.4e38f;
            }
#pragma unroll
            for (int i = 0; i < L::kTPts; i++) {
                if (i < P) {
                    const int32_t g = og + i, o = own_cell(g, r0);
                    const float4 t = A.btab[((int64_t(d) << D) + o) * 2 + (g - o)];
                    mx = fmaxf(mx, fmaxf(fabsf(t.x), fmaxf(fabsf(t.y), fabsf(t.z))));
                    // the owner's ancestor is the brick's (slots 1..3) or the one below (0..2)
                    const bool below = d > B.r && (o >> (D - d)) != (og >> (D - d));
                    const float w0 = below ? t.x : 0.0f, w1 = below ? t.y : t.x, w2 = below ? t.z : t.y,
                                w3 = below ? 0.0f : t.z;
                    lo[0] = fminf(lo[0], w0), hi[0] = fmaxf(hi[0], w0);
                    lo[1] = fminf(lo[1], w1), hi[1] = fmaxf(hi[1], w1);
                    lo[2] = fminf(lo[2], w2), hi[2] = fmaxf(hi[2], w2);
                    lo[3] = fminf(lo[3], w3), hi[3] = fmaxf(hi[3], w3);
                }
            }
            // |basis values| (non-negative floats order as their bit patterns)
            atomicMax(reinterpret_cast<int*>(&S.tmax[d]), __float_as_int(mx));
            if (d <= dlast) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    S.wlo[12 * d + 4 * a + m] = lo[m];
                    S.whi[12 * d + 4 * a + m] = hi[m];
                }
            }
        }
    }
    // |coefficients|: the chain, then the blocks (as brick_abs_bound)
    for (int e = lane; e < 27 * B.r; e += kWave) atomicAdd(&S.xsum[1 + e / 27], fabsf(S.X[27 + e]));
    if (lane == 0) atomicAdd(&S.xsum[0], fabsf(S.X[0]));
    for (int d = B.r + 1; d <= D; d++) {
        if (S.xb_off[d] < 0) continue;
        const int W = d <= dlast ? 4 : block_width(D, B.kb, d), n = W * W * W;
        float sx = 0.0f;
        for (int e = lane; e < n; e += kWave) sx += fabsf(S.XB[S.xb_off[d] + e]);
        if (sx != 0.0f) atomicAdd(&S.xsum[d], sx);
    }
    __syncthreads();
    float bound = 0.0f;
#pragma unroll
    for (int d = 0; d <= kMaxDepth; d++) {
        const float tm = S.tmax[d];
        if (d <= D) bound += S.xsum[d] * (tm * tm * tm);
    }
    return 1.001f * bound;
}

// The basis weights over the 4 block slots along axis ax for grid coordinate
// org[ax] + i at depth d <= D - kb: the owner's three values, placed at its
// ancestor's offset (-1 or 0) from the brick's ancestor.
template <class L>
__device__ __forceinline__ void axis_weights4(const Brick& B, const L& S, int D, int d, int ax, int i,
                                              float w[4]) {
    const int32_t g = B.org[ax] + i, o = own_cell(g, B.R0[ax]);
    const int delta = (o >> (D - d)) - (B.org[ax] >> (D - d));
    const BasisT t = S.T[3 * L::kTPts * d + L::kTPts * ax + i];
    if (delta == 0) {
        w[0] = 0.0f;
        w[1] = t.x;
        w[2] = t.y;
        w[3] = t.z;
    } else {
        w[0] = t.x;
        w[1] = t.y;
        w[2] = t.z;
        w[3] = 0.0f;
    }
}

// A sign for every corner the brick's root owns from one interval for the
// whole brick.  With slot weights (zero outside the owner's three per axis),
// a depth's sum is sum over the block's slots (m, n, l) of
// X[m][n][l] w_x(i)[m] w_y(j)[n] w_z(k)[l] exactly; each weight lies in its
// slot's range over the brick's points, so each term lies in X times the
// product of three ranges.  The finer depths, where the brick spans several
// cells, are bounded by sum |x| times the cube of their largest basis value
// (brick_abs_bound's per-depth figures).  The float evaluation differs from
// the exact sum by at most 351 2^-24 A(g) (the certificate's analysis below)
// and the interval's own float sums by less than 40 2^-24 of the same bound:
// the brick is certified when the interval clears iso by 700 2^-24 times
// the bound of A (plus the rounding of the subtraction).
template <class L>
__device__ __forceinline__ int brick_interval_sign(const SubArgs& A, const Brick& B, L& S, float abound) {
    const int D = A.D, lane = threadIdx.x, dlast = D - B.kb;
    // (the slot ranges wlo, whi come from brick_screen_bound, which ended with a barrier)
    const int m = lane >> 4, n = (lane >> 2) & 3, l = lane & 3;
    float lo = 0.0f, hi = 0.0f;
#pragma unroll
    for (int d = 0; d <= kMaxDepth; d++) {
        float x = 0.0f;
        if (d == 0) {
            x = (m == 2 && n == 2 && l == 2) ? S.X[0] : 0.0f;
        } else if (d <= B.r) {
            x = (m && n && l) ? S.X[27 * d + 9 * (m - 1) + 3 * (n - 1) + (l - 1)] : 0.0f;
        } else if (d <= dlast) {
            const int o = S.xb_off[d];
            x = o < 0 ? 0.0f : S.XB[o + lane];
        }
        if (d <= dlast) {
            const float a1 = S.wlo[12 * d + m], a2 = S.whi[12 * d + m];
            const float b1 = S.wlo[12 * d + 4 + n], b2 = S.whi[12 * d + 4 + n];
            const float c1 = S.wlo[12 * d + 8 + l], c2 = S.whi[12 * d + 8 + l];
            const float u1 = fminf(fminf(a1 * b1, a1 * b2), fminf(a2 * b1, a2 * b2));
            const float u2 = fmaxf(fmaxf(a1 * b1, a1 * b2), fmaxf(a2 * b1, a2 * b2));
            const float p1 = fminf(fminf(u1 * c1, u1 * c2), fminf(u2 * c1, u2 * c2));
            const float p2 = fmaxf(fmaxf(u1 * c1, u1 * c2), fmaxf(u2 * c1, u2 * c2));
            if (x != 0.0f) {
                lo += x >= 0.0f ? x * p1 : x * p2;
                hi += x >= 0.0f ? x * p2 : x * p1;
            }
        }
    }
    lo = wave_fsum(lo);
    hi = wave_fsum(hi);
    float fine = 0.0f;
    for (int d = dlast + 1; d <= D; d++) {
        if (S.xb_off[d] < 0) continue;
        const float t = S.tmax[d];
        fine += S.xsum[d] * (t * t * t);
    }
    lo -= 1.001f * fine;
    hi += 1.001f * fine;
    const float iso = *A.iso;
    const float margin = 700.0f * 5.9604645e-8f * abound + 1.2e-7f * (fabsf(lo) + fabsf(hi) + 2.0f * fabsf(iso)) + 1e-15f;
    if (lo - iso > margin) return 1;
    if (hi - iso < -margin) return -1;
    return 0;
}

// The offsets (bit nidx(delta)) of the bricks holding the owners of the
// brick's corners that leaves of other roots of the group own.
template <class L>
__device__ __forceinline__ uint32_t brick_group_deps(const Brick& B, const L& S, int D) {
    const int Sb = 1 << B.kb, P = Sb + 1;
    uint32_t deps = 0;
    for (int p = threadIdx.x; p < kBrickPts; p += kWave) {
        const int i = p % kP, j = (p / kP) % kP, k = p / (kP * kP);
        if (i >= P || j >= P || k >= P) continue;
        const int32_t g[3] = {B.org[0] + i, B.org[1] + j, B.org[2] + k};
        bool interior = true;
        for (int a = 0; a < 3; a++) interior = interior && g[a] > B.R0[a] && g[a] < B.R0[a] + B.Sr;
        if (interior) continue;
        int32_t o[3];
        if (corner_owner(B, S, D, g, o) == 13) continue;
        int dv[3];
        for (int a = 0; a < 3; a++) dv[a] = o[a] < B.org[a] ? -1 : (o[a] >= B.org[a] + Sb ? 1 : 0);
        deps |= 1u << nidx(dv[0], dv[1], dv[2]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) deps |= uint32_t(__shfl_xor(int(deps), o, kWave));
    return deps;
}

// Proves, when it can, that every corner value of the brick has one strict
// sign, so the brick has no crossing edge and no triangle.  For each depth
// the sum S(g) = sum_k x_k b_x b_y b_z over the corner owner's 27 neighbours
// is evaluated from the same float coefficients and basis table entries the
// evaluation reads, as a separable contraction: each lane owns the column of
// grid points (*, j, k), contracts the depth's coefficients over z and y
// with its own weights, then over x for the nine points of its column.  At
// the chain depths 1..r every owner has the root's ancestor, whose 27
// neighbours are the chain; at depths r+1..D-kb (the whole brick in one cell)
// the owners' ancestors are that cell or the one below it, which a 4x4x4
// block covers; at the finer depths, where the brick spans several cells, the
// owner's 27 terms are read from the brick's block point by point.  The float
// value the evaluation computes, a sum of n <= 27 (D+1) + 1 products of
// three roundings each, differs from S(g) by at most (n+3) 2^-24 A(g) with
// A(g) = sum_k |x_k b_x b_y b_z| (recursive summation bound); the
// contraction's own rounding adds less than 60 2^-24 A(g).  A(g) is bounded
// for the whole brick by brick_abs_bound.  The brick is certified when
// |S(g) - iso| exceeds 560 2^-24 times that bound (plus the rounding of the
// subtraction) with one sign at every corner its root owns.  A corner on the
// root's boundary may be owned by a leaf of another root of the group: its
// value is the one that root's brick holding the owner computes for its own
// corner (the same owner, table entries and order), so the brick records
// that brick's offset instead (deps) and the sign is compared with that
// brick's certificate afterwards (k_brick_deps).
struct Certificate {
    int sign;       // +1 / -1: every corner the brick's root owns has this sign; 0: not proven
    uint32_t deps;  // bit nidx(delta): a corner owned by a leaf of another root, in the brick at offset delta
};

// Whether the brick touches a face, edge or corner its root shares with
// another root of the group (whose leaves may own corners of the brick).
template <class L>
__device__ __forceinline__ bool brick_in_group(const Brick& B, const L& S) {
    const int Sb = 1 << B.kb;
    bool lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = B.org[a] == B.R0[a];
        hi[a] = B.org[a] + Sb == B.R0[a] + B.Sr;
    }
    bool group = false;
    for (int j = 0; j < 27; j++) {
        if (j == 13 || S.NBR[j] < 0) continue;
        const int dv[3] = {j / 9 - 1, (j / 3) % 3 - 1, j % 3 - 1};
        bool touches = true;
        for (int a = 0; a < 3; a++)
            if (dv[a] != 0) touches = touches && (dv[a] < 0 ? lo[a] : hi[a]);
        group = group || touches;
    }
    return group;
}

// The interval certificate (brick_interval_sign) with the brick's
// dependencies on other roots' bricks.
template <class L>
__device__ __forceinline__ Certificate brick_certify_interval(const SubArgs& A, const Brick& B, L& S) {
    Certificate c;
    c.sign = brick_interval_sign(A, B, S, brick_screen_bound(A, B, S));
    c.deps = (c.sign != 0 && brick_in_group(B, S)) ? brick_group_deps(B, S, A.D) : 0;
    return c;
}

template <class L>
__device__ __forceinline__ Certificate brick_certify(const SubArgs& A, const Brick& B, L& S) {
    const int D = A.D, kb = B.kb, Sb = 1 << kb, P = Sb + 1;
    const int lane = threadIdx.x;
    const bool group = brick_in_group(B, S);
    const float iso = *A.iso;
    const int dlast = D - kb;  // deepest depth whose cell holds the whole brick
    const float marginA = 560.0f * 5.9604645e-8f * brick_abs_bound(B, S, D) + 1e-15f;
    bool ok = true, pos = false, neg = false;
    uint32_t deps = 0;
    for (int pp = lane; pp < kP * kP; pp += kWave) {
        const int j = pp % kP, k = pp / kP;
        if (j >= P || k >= P) continue;
        float v[kP];
#pragma unroll
        for (int i = 0; i < kP; i++) v[i] = 0.0f;
        // (points i >= P of a smaller brick are accumulated from unset table
        // entries and never checked)
        {  // depth 0: the tree's root, one coefficient
            const float y = (S.X[0] * S.T[2 * kP + k].y) * S.T[kP + j].y;
#pragma unroll
            for (int i = 0; i < kP; i++) v[i] = __builtin_fmaf(S.T[i].y, y, v[i]);
        }
        for (int d = 1; d <= B.r; d++) {  // the chain: all owners share their ancestor
            float Y[3];
            contract27(S.X + 27 * d, S.T[3 * kP * d + kP + j], S.T[3 * kP * d + 2 * kP + k], Y);
#pragma unroll
            for (int i = 0; i < kP; i++) {
                const BasisT t = S.T[3 * kP * d + i];
                v[i] = __builtin_fmaf(t.x, Y[0], v[i]);
                v[i] = __builtin_fmaf(t.y, Y[1], v[i]);
                v[i] = __builtin_fmaf(t.z, Y[2], v[i]);
            }
        }
        for (int d = B.r + 1; d <= dlast; d++) {  // 4^3 blocks of real nodes
            if (S.xb_off[d] < 0) continue;
            float w1[4], w2[4], Y[4];
            axis_weights4(B, S, D, d, 1, j, w1);
            axis_weights4(B, S, D, d, 2, k, w2);
            contract64(S.XB + S.xb_off[d], w1, w2, Y);
#pragma unroll
            for (int i = 0; i < kP; i++) {
                float w0[4];
                axis_weights4(B, S, D, d, 0, i, w0);
#pragma unroll
                for (int m = 0; m < 4; m++) v[i] = __builtin_fmaf(w0[m], Y[m], v[i]);
            }
        }
        // the finer depths, where the brick spans several cells: per x slot m
        // of the block, Y_m = sum over the lane's 3 y and 3 z slots, then
        // each point adds its owner's three x slots' Y (the slots a point's
        // owner sees follow from grid coordinates alone: uniform per i)
        for (int d = dlast + 1; d <= D; d++) {
            const int slot = S.xb_off[d];
            if (slot < 0) continue;
            const int sh = D - d, W = block_width(D, kb, d);
            const int b1 = (own_cell(B.org[1] + j, B.R0[1]) >> sh) - 1 - ((B.org[1] >> sh) - 2);
            const int b2 = (own_cell(B.org[2] + k, B.R0[2]) >> sh) - 1 - ((B.org[2] >> sh) - 2);
            const BasisT ty = S.T[3 * kP * d + kP + j], tz = S.T[3 * kP * d + 2 * kP + k];
            const float by[3] = {ty.x, ty.y, ty.z}, bz[3] = {tz.x, tz.y, tz.z};
            int b0[kP];
            float bx[kP][3];
#pragma unroll
            for (int i = 0; i < kP; i++) {
                b0[i] = (own_cell(B.org[0] + i, B.R0[0]) >> sh) - 1 - ((B.org[0] >> sh) - 2);
                const BasisT tx = S.T[3 * kP * d + i];
                bx[i][0] = tx.x;
                bx[i][1] = tx.y;
                bx[i][2] = tx.z;
            }
#pragma unroll
            for (int m = 0; m < 11; m++) {  // W <= 11
                if (m >= W || m < b0[0] || m > b0[P - 1] + 2) continue;  // (uniform) a slot no point sees
                const float* Xm = S.XB + slot + m * W * W;
                float y = 0.0f;
#pragma unroll
                for (int n = 0; n < 3; n++) {
                    float z = 0.0f;
#pragma unroll
                    for (int l = 0; l < 3; l++) z = __builtin_fmaf(Xm[(b1 + n) * W + (b2 + l)], bz[l], z);
                    y = __builtin_fmaf(by[n], z, y);
                }
#pragma unroll
                for (int i = 0; i < kP; i++) {
                    const int q = m - b0[i];  // uniform
                    if (i < P && q >= 0 && q <= 2) v[i] = __builtin_fmaf(q == 0 ? bx[i][0] : q == 1 ? bx[i][1] : bx[i][2], y, v[i]);
                }
            }
        }
        const bool jk_inner = B.org[1] + j > B.R0[1] && B.org[1] + j < B.R0[1] + B.Sr && B.org[2] + k > B.R0[2] &&
                              B.org[2] + k < B.R0[2] + B.Sr;
#pragma unroll
        for (int i = 0; i < kP; i++) {
            if (i >= P) continue;
            const int32_t gx = B.org[0] + i;
            if (group && !(jk_inner && gx > B.R0[0] && gx < B.R0[0] + B.Sr)) {
                // on the root's boundary: is the owner a leaf of another root?
                const int32_t g[3] = {gx, B.org[1] + j, B.org[2] + k};
                int32_t o[3];
                if (corner_owner(B, S, D, g, o) != 13) {
                    int dv[3];
                    for (int a = 0; a < 3; a++) dv[a] = o[a] < B.org[a] ? -1 : (o[a] >= B.org[a] + Sb ? 1 : 0);
                    deps |= 1u << nidx(dv[0], dv[1], dv[2]);
                    continue;
                }
            }
            const float val = v[i] - iso;
            const float margin = marginA + 1.2e-7f * (fabsf(v[i]) + fabsf(iso));
            ok = ok && fabsf(val) > margin;
            pos = pos || val > 0;
            neg = neg || val < 0;
        }
    }
    const bool lane_good = ok && !(pos && neg);
    const uint64_t good = __builtin_amdgcn_ballot_w64(lane_good);
    const uint64_t anypos = __builtin_amdgcn_ballot_w64(pos), anyneg = __builtin_amdgcn_ballot_w64(neg);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) deps |= uint32_t(__shfl_xor(int(deps), o, kWave));
    Certificate c;
    c.sign = (good == ~uint64_t(0) && !(anypos && anyneg)) ? (anypos ? 1 : -1) : 0;
    c.deps = deps;
    return c;
}

// Records a brick's certificate: its sign, and either its dependencies
// (dbricks: (b, first, count); dpairs: the bricks it depends on) or, when
// not certified, its place in `list`.
template <class L>
__device__ __forceinline__ void record_certificate(const SubArgs& A, const Brick& B, const L& S, int32_t b,
                                                   const Certificate& c, int8_t* sign, int32_t* list, int list_slot,
                                                   int32_t* dbricks, int32_t* dpairs, int32_t* cnt) {
    const int lane = threadIdx.x, D = A.D;
    int32_t first = 0;
    if (lane == 0) {
        sign[b] = int8_t(c.sign);
        if (c.sign == 0) {
            list[atomicAdd(cnt + list_slot, 1)] = b;
        } else if (c.deps) {
            const int n = __popc(c.deps);
            first = atomicAdd(cnt + kCntDepPairs, n);
            const int32_t e = atomicAdd(cnt + kCntDepBricks, 1);
            dbricks[3 * e] = b;
            dbricks[3 * e + 1] = first;
            dbricks[3 * e + 2] = n;
        }
    }
    first = __shfl(first, 0, kWave);
    if (c.sign != 0 && lane < 27 && ((c.deps >> lane) & 1u)) {
        // the brick at offset delta(lane), in the root of the group that holds it
        const int Sb = 1 << B.kb, sh = D - B.r;
        const int dv[3] = {lane / 9 - 1, (lane / 3) % 3 - 1, lane % 3 - 1};
        int32_t cc[3];
        int ro[3];
        for (int a = 0; a < 3; a++) {
            cc[a] = B.org[a] + dv[a] * Sb;
            ro[a] = (cc[a] >> sh) - B.R[a];
        }
        const int slot = nidx(ro[0], ro[1], ro[2]);
        const int32_t q2 = slot == 13 ? B.q : S.NBR[slot];
        const uint32_t rmask = 3 * sh >= 32 ? 0xffffffffu : ((1u << (3 * sh)) - 1u);
        const int32_t b2 =
            q2 < 0 ? -1
                   : A.boff[q2] + int32_t((cell_key(uint32_t(cc[0]), uint32_t(cc[1]), uint32_t(cc[2])) & rmask) >>
                                          (3 * B.kb));
        dpairs[first + __popc(c.deps & ((1u << lane) - 1u))] = b2;
    }
}

// Pass 0a: the interval certificate for the super-bricks (2x2x2 bricks) of
// the roots with at least 8 bricks (a fixed grid takes chunks of consecutive
// super-bricks from a counter).  A certified super-brick settles its 8
// bricks (counts 0, its sign); the bricks of the others, and of super-bricks
// that touch another root of the group (whose leaves may own corners) or
// whose blocks the LDS cannot hold, are listed for the brick screen.
__global__ void __launch_bounds__(kWave, 5) k_brick_screen_super(SubArgs A, const int32_t* smap, const int32_t* soff,
                                                              int32_t* bnv, int32_t* bnt, int8_t* sign,
                                                              int32_t* list2, int32_t* cnt, int certify) {
    __shared__ BrickSuperLds S;
    const int lane = threadIdx.x;
    const int32_t ns = cnt[kCntSuper];
    constexpr int kChunk = 8;
    int32_t q_loaded = -1, sb = 0, s_end = 0;
    for (;;) {
        if (sb == s_end) {
            int32_t c0 = 0;
            if (lane == 0) c0 = atomicAdd(cnt + kCntWork, kChunk);
            c0 = __shfl(c0, 0, kWave);
            if (c0 >= ns) break;
            sb = c0;
            s_end = std::min(ns, c0 + kChunk);
        }
        const int32_t q = smap[sb];
        const int32_t b0 = A.boff[q] + 8 * (sb - soff[q]);
        Brick B;
        __syncthreads();  // the previous super-brick's LDS reads are done
        brick_setup<true>(A, b0, B, S, q_loaded, 1);
        q_loaded = B.q;
        int sgn = 0;
        if (certify && !brick_in_group(B, S) && brick_blocks(A, B, S))
            sgn = brick_interval_sign(A, B, S, brick_screen_bound(A, B, S));
        if (sgn != 0) {
            if (lane < 8) {
                bnv[b0 + lane] = 0;
                bnt[b0 + lane] = 0;
                sign[b0 + lane] = int8_t(sgn);
            }
        } else {
            int32_t at = 0;
            if (lane == 0) at = atomicAdd(cnt + kCntList2, 8);
            at = __shfl(at, 0, kWave);
            if (lane < 8) list2[at + lane] = b0 + lane;
        }
        sb++;
    }
}

// The bricks of the roots with fewer than 8 bricks, for the brick screen.
__global__ void __launch_bounds__(kBlockS) k_list_small_root_bricks(const int32_t* bmap, const int32_t* rbricks,
                                                                    int64_t nb, int32_t* list2, int32_t* cnt) {
    const int64_t b = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (b >= nb) return;
    if (rbricks[bmap[b]] < 8) list2[atomicAdd(cnt + kCntList2, 1)] = int32_t(b);
}

// Per root: its super-brick count (a root of at least 8 bricks has a
// multiple of 8, in 2x2x2 blocks of Morton-consecutive bricks).
__global__ void __launch_bounds__(kBlockS) k_super_count(const int32_t* rbricks, const int32_t* cnt, int64_t nr_max,
                                                         int32_t* nsup) {
    const int64_t q = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (q >= nr_max) return;
    int32_t n = 0;
    if (q < cnt[kCntRoots]) n = rbricks[q] >= 8 ? rbricks[q] / 8 : 0;
    nsup[q] = n;
}

// smap[s] = the root of super-brick s (binary search of the super-brick offsets).
__global__ void __launch_bounds__(kBlockS) k_super_map(const int32_t* soff, const int32_t* cnt, int64_t ns_max,
                                                       int32_t* smap) {
    const int64_t sb = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (sb >= ns_max || sb >= cnt[kCntSuper]) return;
    int32_t lo = 0, hi = cnt[kCntRoots] - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (soff[mid] <= sb) lo = mid;
        else hi = mid - 1;
    }
    smap[sb] = lo;
}

// Pass 0b: the interval certificate for the listed bricks (a fixed grid takes
// chunks of the list from a counter).  Every listed brick's counts start at
// 0; a certified brick keeps them unless a brick it depends on has another
// sign (k_brick_deps); the others are listed for the exact certificate
// (k_brick_exact).
__global__ void __launch_bounds__(kWave, 5) k_brick_screen(SubArgs A, const int32_t* list2, int32_t* bnv, int32_t* bnt,
                                                        int8_t* sign, int32_t* xlist, int32_t* dbricks,
                                                        int32_t* dpairs, int32_t* cnt, int certify) {
    __shared__ BrickScreenLds S;
    const int lane = threadIdx.x;
    const int32_t nl = cnt[kCntList2Sorted];
    constexpr int kChunk = 16;  // consecutive entries per grab: mostly one root, whose record stays in LDS
    int32_t q_loaded = -1, it = 0, it_end = 0;
    for (;;) {
        if (it == it_end) {
            int32_t c0 = 0;
            if (lane == 0) c0 = atomicAdd(cnt + kCntWork2, kChunk);
            c0 = __shfl(c0, 0, kWave);
            if (c0 >= nl) break;
            it = c0;
            it_end = std::min(nl, c0 + kChunk);
        }
        const int32_t b = list2[it];
        Brick B;
        __syncthreads();  // the previous brick's LDS reads are done
        brick_setup<true>(A, b, B, S, q_loaded);
        q_loaded = B.q;
        Certificate c;
        c.sign = 0;
        c.deps = 0;
        if (brick_blocks(A, B, S) && certify) c = brick_certify_interval(A, B, S);  // else: to the exact certificate
        if (lane == 0) {
            bnv[b] = 0;
            bnt[b] = 0;
        }
        record_certificate(A, B, S, b, c, sign, xlist, kCntExact, dbricks, dpairs, cnt);
        it++;
    }
}

// Pass 0b: the exact certificate (brick_certify) for the bricks the interval
// left (a fixed grid walks the list); the bricks it does not certify are
// listed for the evaluation.
// The list is in brick order (root-major), and a resident grid takes chunks
// of consecutive entries from a counter, so consecutive bricks mostly share
// their root and its record stays in LDS.  Compiled for three waves per SIMD
// (168 VGPRs, a few spilled; the LDS allows eleven workgroups per CU): the
// waves wait on memory about half their time, and two per SIMD hid less.
__global__ void __launch_bounds__(kWave, 3) k_brick_exact(SubArgs A, const int32_t* xlist, int8_t* sign, int32_t* list,
                                                       int32_t* dbricks, int32_t* dpairs, int32_t* cnt, int certify) {
    __shared__ BrickCore S;
    const int lane = threadIdx.x;
    const int32_t n = cnt[kCntExactSorted];
    // chunks of up to 8 entries, at least about four per workgroup
    const int32_t chunk = std::max(1, std::min(8, n / (4 * int32_t(gridDim.x))));
    int32_t q_loaded = -1, it = 0, it_end = 0;
    for (;;) {
        if (it == it_end) {
            int32_t c0 = 0;
            if (lane == 0) c0 = atomicAdd(cnt + kCntWork3, chunk);
            c0 = __shfl(c0, 0, kWave);
            if (c0 >= n) break;
            it = c0;
            it_end = std::min(n, c0 + chunk);
        }
        const int32_t b = xlist[it++];
        Brick B;
        __syncthreads();  // the previous brick's LDS reads are done
        brick_setup<true>(A, b, B, S, q_loaded);
        q_loaded = B.q;
        brick_tables(A, B, S);
        brick_blocks(A, B, S);
        Certificate c;
        c.sign = 0;
        c.deps = 0;
        if (certify) c = brick_certify(A, B, S);
        record_certificate(A, B, S, b, c, sign, list, kCntDense, dbricks, dpairs, cnt);
    }
}

// The certified bricks with corners owned by other roots: certified only when
// every brick they depend on carries the same sign; the others are listed
// for the evaluation.
__global__ void __launch_bounds__(kBlockS) k_brick_deps(const int32_t* dbricks, const int32_t* dpairs,
                                                        const int8_t* sign, int32_t* list, int32_t* cnt) {
    const int32_t e = blockIdx.x * kBlockS + threadIdx.x;
    if (e >= cnt[kCntDepBricks]) return;
    const int32_t b = dbricks[3 * e], first = dbricks[3 * e + 1], n = dbricks[3 * e + 2];
    const int8_t sb = sign[b];
    bool ok = true;
    for (int t = 0; t < n; t++) {
        const int32_t b2 = dpairs[first + t];
        ok = ok && b2 >= 0 && sign[b2] == sb;
    }
    if (!ok) list[atomicAdd(cnt + kCntDense, 1)] = b;
}

// Pass 1: vertex and triangle counts of the bricks the certificate left
// (a fixed grid walks the list).
// The corner values go to gvals (kBrickPts per listed brick) and the
// brick's place in the list to didx, for the vertex pass; in LDS they take
// the place of the blocks (brick_values), so the pass needs the
// certificate's LDS only.
// The list is in brick order and a resident grid takes chunks of consecutive
// entries from a counter (kCntWork4), so the root's record mostly stays in LDS.
__global__ void __launch_bounds__(kWave) k_brick_count(SubArgs A, const int32_t* list, int32_t* cnt,
                                                       int32_t* bnv, int32_t* bnt, float* gvals, int32_t* didx) {
    __shared__ BrickCore S;
    const int32_t n = cnt[kCntDenseSorted];
    // chunks of up to 8 entries, at least about four per workgroup (a short
    // list in long chunks leaves some workgroups with twice the others' work)
    const int32_t chunk = std::max(1, std::min(8, n / (4 * int32_t(gridDim.x))));
    int32_t q_loaded = -1, it = 0, it_end = 0;
    for (;;) {
        if (it == it_end) {
            int32_t c0 = 0;
            if (threadIdx.x == 0) c0 = atomicAdd(cnt + kCntWork4, chunk);
            c0 = __shfl(c0, 0, kWave);
            if (c0 >= n) break;
            it = c0;
            it_end = std::min(n, c0 + chunk);
        }
        Brick B;
        const int32_t b = list[it];
        __syncthreads();  // the previous brick's LDS reads are done
        brick_setup<true>(A, b, B, S, q_loaded);
        q_loaded = B.q;
        brick_tables(A, B, S);
        brick_blocks(A, B, S);
        const float* G = brick_values(A, B, S, gvals + int64_t(it) * kBrickPts);
        if (threadIdx.x == 0) didx[b] = it;
        int nv = 0, nt = 0;
        const int L = 1 << (3 * B.kb);
        for (int l = threadIdx.x; l < L; l += kWave) {
            int32_t v[3];
            float cv[8];
            int cube, em;
            leaf_cube(B, S, G, A.D, l, v, cv, &cube, &em);
            nv += __popc(uint32_t(em));
            nt += tri_count(cube);
        }
        nv = wave_sum(nv);
        nt = wave_sum(nt);
        if (threadIdx.x == 0) {
            bnv[b] = nv;
            bnt[b] = nt;
        }
        it++;
    }
}

__global__ void __launch_bounds__(kBlockS) k_brick_active(const int32_t* bnv, const int32_t* bnt, int64_t n,
                                                          int32_t* flag) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < n) flag[i] = (bnv[i] | bnt[i]) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlockS) k_brick_compact(const int32_t* flag, const int32_t* off, int64_t n,
                                                           int32_t* alist, int32_t* amap) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) {
        alist[off[i]] = int32_t(i);
        amap[i] = off[i];
    } else {
        amap[i] = -1;
    }
}

// Pass 2 (bricks with output): per-leaf edge masks, cube indices and vertex
// offsets into the brick's tables; the vertices (interpolatePoint and
// generateSubdivideIntersectionPoint, main.cu:2584-2597, 2629-2697).
__global__ void __launch_bounds__(kWave) k_brick_vertices(SubArgs A, const int32_t* alist, const int32_t* bvoff,
                                                          uint16_t* lofs_tab, uint16_t* emask_tab, uint8_t* cube_tab,
                                                          const float* gvals, const int32_t* didx, float* out_v) {
    __shared__ BrickMeshLds S;
    Brick B;
    const int32_t ai = blockIdx.x;
    const int32_t b = alist[ai];
    brick_setup<false>(A, b, B, S);
    {  // the corner values the count pass computed
        const float* G = gvals + int64_t(didx[b]) * kBrickPts;
        for (int p = threadIdx.x; p < kBrickPts; p += kWave) S.G[p] = G[p];
        __syncthreads();
    }
    const int D = A.D;
    const int L = 1 << (3 * B.kb);
    const int per = (L + kWave - 1) / kWave;
    const int l0 = min(L, int(threadIdx.x) * per), l1 = min(L, l0 + per);
    int cnt = 0;
    for (int l = l0; l < l1; l++) {
        int32_t v[3];
        float cv[8];
        int cube, em;
        leaf_cube(B, S, S.G, D, l, v, cv, &cube, &em);
        emask_tab[int64_t(ai) * kBrickLeaves + l] = uint16_t(em);
        cube_tab[int64_t(ai) * kBrickLeaves + l] = uint8_t(cube);
        cnt += __popc(uint32_t(em));
    }
    int run = wave_excl_scan(cnt);
    const float hw = 1.0f / float(1 << (D + 1));
    for (int l = l0; l < l1; l++) {
        lofs_tab[int64_t(ai) * kBrickLeaves + l] = uint16_t(run);
        const int em = emask_tab[int64_t(ai) * kBrickLeaves + l];
        if (!em) continue;
        int32_t v[3];
        float cv[8];
        int cube, em2;
        leaf_cube(B, S, S.G, D, l, v, cv, &cube, &em2);
        float c[3];
        node_center(B.bkey | uint32_t(l), D, D, c);
        for (int e = 0; e < 12; e++) {
            if (!(em & (1 << e))) continue;
            const int q1 = c_edge_vertex[e][0], q2 = c_edge_vertex[e][1];
            float p1[3], p2[3];
            {
                const int y = (q1 >> 1) & 1, xx = (q1 & 1) ^ y, z = q1 >> 2;
                p1[0] = c[0] + float(2 * xx - 1) * hw;
                p1[1] = c[1] + float(2 * y - 1) * hw;
                p1[2] = c[2] + float(2 * z - 1) * hw;
            }
            {
                const int y = (q2 >> 1) & 1, xx = (q2 & 1) ^ y, z = q2 >> 2;
                p2[0] = c[0] + float(2 * xx - 1) * hw;
                p2[1] = c[1] + float(2 * y - 1) * hw;
                p2[2] = c[2] + float(2 * z - 1) * hw;
            }
            const float f1 = cv[q1], f2 = cv[q2];
            const int dim = e >> 2;
            float o3[3] = {p1[0], p1[1], p1[2]};
            const float pivot = f1 / (f1 - f2);
            const float another = 1 - pivot;
            const float s1 = p2[dim] * pivot;
            const float s2 = p1[dim] * another;
            o3[dim] = s1 + s2;
            const int64_t id = int64_t(bvoff[b]) + run++;
            out_v[3 * id + 0] = o3[0];
            out_v[3 * id + 1] = o3[1];
            out_v[3 * id + 2] = o3[2];
        }
    }
}

// Pass 3 (bricks with output): triangle entries (generateSubdivideTrianglePos,
// main.cu:2759-2825).  A vertex owned by a leaf of another brick (of this
// root or of another root of the group) is found through that brick's tables.
__global__ void __launch_bounds__(kWave) k_brick_triangles(SubArgs A, const int32_t* alist, const int32_t* amap,
                                                           const int32_t* bvoff, const int32_t* btoff,
                                                           const uint16_t* lofs_tab, const uint16_t* emask_tab,
                                                           const uint8_t* cube_tab, int64_t vbase, int32_t* out_t,
                                                           int32_t* err) {
    __shared__ BrickCoreT<1> S;  // the root's neighbours only
    Brick B;
    const int32_t ai = blockIdx.x;
    const int32_t b = alist[ai];
    brick_setup<false>(A, b, B, S);
    const int D = A.D;
    const int L = 1 << (3 * B.kb);
    const int per = (L + kWave - 1) / kWave;
    const int l0 = min(L, int(threadIdx.x) * per), l1 = min(L, l0 + per);
    int cnt = 0;
    for (int l = l0; l < l1; l++) cnt += tri_count(cube_tab[int64_t(ai) * kBrickLeaves + l]);
    int64_t o = 3 * (int64_t(btoff[b]) + wave_excl_scan(cnt));
    const int sh = 3 * (D - B.r);
    const uint32_t rmask = sh >= 32 ? 0xffffffffu : ((1u << sh) - 1u);
    for (int l = l0; l < l1; l++) {
        const int cube = cube_tab[int64_t(ai) * kBrickLeaves + l];
        const int n = 3 * tri_count(cube);
        if (!n) continue;
        const int32_t v[3] = {B.org[0] + int32_t(compact3(uint32_t(l) >> 2)),
                              B.org[1] + int32_t(compact3(uint32_t(l) >> 1)), B.org[2] + int32_t(compact3(uint32_t(l)))};
        for (int k = 0; k < n; k++) {
            const int e = c_tri[cube][k];
            int32_t ow[3];
            int eo;
            edge_owner(B, S, D, v, e, ow, &eo);
            int sl = 13;
            cell_virtual(B, S, D, ow, &sl);
            const int32_t q2 = sl == 13 ? B.q : S.NBR[sl];
            const uint32_t mort = cell_key(uint32_t(ow[0]), uint32_t(ow[1]), uint32_t(ow[2])) & rmask;
            const int32_t bi = q2 < 0 ? 0 : A.boff[q2] + int32_t(mort >> (3 * B.kb));
            const int32_t lu = int32_t(mort & ((1u << (3 * B.kb)) - 1u));
            const int32_t au = q2 < 0 ? -1 : amap[bi];
            int32_t id = 0;
            if (au < 0) {
                err[0] = 1;  // a crossing edge whose owner brick has no output: inconsistent counts
            } else {
                const int64_t t = int64_t(au) * kBrickLeaves + lu;
                id = int32_t(vbase + bvoff[bi] + lofs_tab[t] + __popc(uint32_t(emask_tab[t]) & ((1u << eo) - 1u)));
            }
            out_t[o + k] = id;
        }
        o += n;
    }
}

}  // namespace

// the coarse rows of the current solution (mc.hip)
int ensure_coarse_rows(prg_ctx* c, hipStream_t s);

void extract_subdivision(prg_ctx* c) {
    ensure_tables(c);
    StageTimer st(c, PRG_STAGE_SUBDIVIDE);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M, MD = c->MD(), baseD = c->bases[D];
    KernelTimer& kt = c->kt;
    const Levels Lv = levels_of(c);
    c->sub_nv = c->sub_nt = 0;
    c->sub_roots = c->sub_leaves = 0;
    c->sub_bricks = c->sub_exact = c->sub_evaluated = 0;
    if (baseD <= 1) return;
    if (8 * baseD >= (int64_t(1) << 31)) throw Error{PRG_E_ARG, "too many nodes above the leaf level for subdivision"};
    const int rd = ensure_coarse_rows(c, s);
    if (rd < D - 1) throw Error{PRG_E_STATE, "coarse rows missing"};

    // pool slots 97..139
    PoolBuf<uint8_t> inter{&c->pool[97]};
    PoolBuf<int32_t> cnt{&c->pool[98]}, flag{&c->pool[99]}, off{&c->pool[100]}, cl{&c->pool[101]},
        oref{&c->pool[102]}, need{&c->pool[103]}, nid{&c->pool[104]}, clist{&c->pool[105]}, rflag{&c->pool[107]},
        roff{&c->pool[108]}, ridx{&c->pool[109]}, roots{&c->pool[110]}, rbricks{&c->pool[111]},
        boff{&c->pool[112]}, rleaves{&c->pool[113]}, loff{&c->pool[114]}, root_of{&c->pool[115]};
    PoolBuf<float> cv{&c->pool[106]};
    int32_t* C = cnt.reserve(kCntSlots);
    PRG_HIP(hipMemsetAsync(C, 0, kCntSlots * 4, s));

    // faces touched by the leaf-level surface, and up the parent chains
    inter.reserve(6 * M);
    PRG_HIP(hipMemsetAsync(inter.get(), 0, size_t(6 * M), s));
    {
        PoolBuf<uint32_t> up{&c->pool[140]};
        up.reserve((M + 3) / 4);
        PRG_HIP(hipMemsetAsync(up.get(), 0, size_t((M + 3) / 4) * 4, s));
        // bytes: the cube index of each leaf-level node and the face byte of
        // each coarser node, read once (the face visits' gathers come on top)
        TimedScope ts(kt, "face_marks", double(M));
        // the leaf-level cube indices of the marching cubes pass that ran just before
        PoolBuf<uint8_t> cubes{&c->pool[94]};
        k_face_marks<<<grid_for(MD, kBlockS), kBlockS, 0, s>>>(cubes.get(), MD, baseD, c->key.get(), c->pnt.get(),
                                                                Lv, c->parent.get(), D, inter.get(), up.get());
        PRG_CHECK_LAUNCH();
        for (int d = D - 1; d >= 0; d--) {
            k_face_marks_up<<<grid_for(c->counts[d], kBlockS), kBlockS, 0, s>>>(
                c->bases[d], c->counts[d], d, c->key.get(), c->pnt.get(), Lv, c->parent.get(), D, inter.get(),
                up.get());
            PRG_CHECK_LAUNCH();
        }
    }
    // coarse leaves (counts stay on the device; launches are sized by upper bounds)
    flag.reserve(baseD);
    off.reserve(baseD);
    cl.reserve(baseD);
    k_flag_coarse_leaves<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(c->child_base.get(), baseD, flag.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(flag.get(), off.get(), baseD, C + kCntCoarse, s, c->scratch, kt);
    k_compact<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(flag.get(), off.get(), baseD, cl.get());
    PRG_CHECK_LAUNCH();
    // their corner values
    oref.reserve(8 * baseD);
    need.reserve(8 * baseD);
    nid.reserve(8 * baseD);
    clist.reserve(8 * baseD);
    cv.reserve(8 * baseD);
    PRG_HIP(hipMemsetAsync(need.get(), 0, size_t(8 * baseD) * 4, s));
    k_coarse_corner_owner<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(cl.get(), C + kCntCoarse, c->key.get(),
                                                                            c->pnt.get(), Lv, D, oref.get(), need.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(need.get(), nid.get(), 8 * baseD, C + kCntCorners, s, c->scratch, kt);
    k_compact<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(need.get(), nid.get(), 8 * baseD, clist.get());
    PRG_CHECK_LAUNCH();
    {
        TimedScope ts(kt, "coarse_corner_value", 24.0 * double(baseD));
        k_coarse_corner_value<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(
            clist.get(), C + kCntCorners, c->key.get(), c->pnt.get(), Lv, c->parent.get(), c->child_base.get(),
            c->dnum.get(), c->x.get(), c->d_base_fn.get(), c->coarse_rows.get(), rd, D, c->iso_dev.get(), cv.get());
        PRG_CHECK_LAUNCH();
    }
    // roots, their bricks and virtual leaves
    rflag.reserve(baseD);
    roff.reserve(baseD);
    ridx.reserve(baseD);
    roots.reserve(baseD);
    rbricks.reserve(baseD);
    boff.reserve(baseD);
    rleaves.reserve(baseD);
    loff.reserve(baseD);
    k_coarse_select<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(cl.get(), C + kCntCoarse, baseD, oref.get(),
                                                                  nid.get(), cv.get(), c->key.get(), c->pnt.get(), Lv,
                                                                  D, inter.get(), rflag.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(rflag.get(), roff.get(), baseD, C + kCntRoots, s, c->scratch, kt);
    k_compact<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(rflag.get(), roff.get(), baseD, ridx.get());
    PRG_CHECK_LAUNCH();
    k_root_info<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(cl.get(), ridx.get(), C + kCntRoots, baseD, Lv, D,
                                                              roots.get(), rbricks.get(), rleaves.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive2(rbricks.get(), boff.get(), C + kCntBricks, rleaves.get(), loff.get(), C + kCntLeaves, baseD, s,
                    c->scratch, kt);
    int32_t* H = c->sub_host;
    PRG_HIP(hipMemcpyAsync(H, C, kCntSlots * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    const int64_t nr = H[kCntRoots], nb = H[kCntBricks];
    c->sub_roots = nr;
    c->sub_leaves = H[kCntLeaves];
    if (nr == 0) return;
    if (nb <= 0 || H[kCntLeaves] <= 0) throw Error{PRG_E_HIP, "subdivision: inconsistent brick counts"};
    root_of.reserve(M);
    PRG_HIP(hipMemsetAsync(root_of.get(), 0xff, size_t(M) * 4, s));
    k_root_of<<<grid_for(nr, kBlockS), kBlockS, 0, s>>>(roots.get(), C + kCntRoots, root_of.get());
    PRG_CHECK_LAUNCH();

    PoolBuf<float> btab{&c->pool[91]};
    SubArgs A{c->key.get(),
                    c->child_base.get(), c->parent.get(), c->pnt.get(),
                    Lv,                c->x.get(),      c->d_base_fn.get(),
                    reinterpret_cast<const float4*>(btab.get()), c->coarse_rows.get(), D,
                    c->iso_dev.get(),  roots.get(),     boff.get(),
                    root_of.get(),     C,               nullptr,
                    nullptr};

    // per-root records and the brick -> root map
    PoolBuf<int32_t> rec{&c->pool[127]}, bmap{&c->pool[128]};
    rec.reserve(size_t(nr) * kRecWords);
    bmap.reserve(size_t(nb));
    A.rec = rec.get();
    A.bmap = bmap.get();
    k_root_records<<<unsigned(nr), kWave, 0, s>>>(A, rec.get());
    PRG_CHECK_LAUNCH();
    k_brick_map<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(boff.get(), C, nb, bmap.get());
    PRG_CHECK_LAUNCH();

    // pass 1: counts per brick
    PoolBuf<int32_t> bnv{&c->pool[116]}, bnt{&c->pool[117]}, bvoff{&c->pool[118]}, btoff{&c->pool[119]},
        aflag{&c->pool[120]}, aoff{&c->pool[121]}, alist{&c->pool[122]}, amap{&c->pool[123]};
    bnv.reserve(nb);
    bnt.reserve(nb);
    bvoff.reserve(nb);
    btoff.reserve(nb);
    aflag.reserve(nb);
    aoff.reserve(nb);
    alist.reserve(nb);
    amap.reserve(nb);
    // pass 0: certificates; pass 1: counts of the bricks left
    PoolBuf<int32_t> dlist{&c->pool[129]}, dbricks{&c->pool[130]}, dpairs{&c->pool[131]}, xlist{&c->pool[133]};
    PoolBuf<int8_t> bsign{&c->pool[132]};
    dlist.reserve(nb);
    xlist.reserve(nb);
    dbricks.reserve(3 * nb);
    dpairs.reserve(26 * nb);
    bsign.reserve(nb);
    {
        TimedScope ts(kt, "subdivide_screen", 0.0);  // latency- and issue-bound: no byte figure
        // super-bricks of the roots with at least 8 bricks, then the brick
        // screen over the bricks they leave and those of the smaller roots
        PoolBuf<int32_t> nsup{&c->pool[136]}, soff{&c->pool[137]}, smap{&c->pool[138]}, list2{&c->pool[139]};
        nsup.reserve(nr);
        soff.reserve(nr);
        smap.reserve(std::max<int64_t>(1, nb / 8));
        list2.reserve(nb);
        k_super_count<<<grid_for(nr, kBlockS), kBlockS, 0, s>>>(rbricks.get(), C, nr, nsup.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive(nsup.get(), soff.get(), nr, C + kCntSuper, s, c->scratch, kt);
        k_super_map<<<grid_for(std::max<int64_t>(1, nb / 8), kBlockS), kBlockS, 0, s>>>(soff.get(), C,
                                                                                      std::max<int64_t>(1, nb / 8),
                                                                                      smap.get());
        PRG_CHECK_LAUNCH();
        k_list_small_root_bricks<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(bmap.get(), rbricks.get(), nb,
                                                                            list2.get(), C);
        PRG_CHECK_LAUNCH();
        // resident grids: every workgroup walks its share
        int per_cu = 0;
        PRG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_brick_screen_super, kWave, 0));
        int64_t grid = std::min<int64_t>(std::max<int64_t>(1, nb / 8), int64_t(std::max(1, per_cu)) * std::max(1, c->num_cu));
        k_brick_screen_super<<<unsigned(grid), kWave, 0, s>>>(A, smap.get(), soff.get(), bnv.get(), bnt.get(),
                                                              bsign.get(), list2.get(), C, int(c->sub_certify));
        PRG_CHECK_LAUNCH();
        // the lists the passes append to, each put in brick order (marks,
        // scan, compact) before the next pass walks it
        PoolBuf<int32_t> xmark{&c->pool[141]}, xoff{&c->pool[142]}, xsorted{&c->pool[143]};
        xmark.reserve(nb);
        xoff.reserve(nb);
        xsorted.reserve(nb);
        PRG_HIP(hipMemsetAsync(xmark.get(), 0, size_t(nb) * 4, s));
        k_mark_listed<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(list2.get(), C + kCntList2, nb, xmark.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive(xmark.get(), xoff.get(), nb, C + kCntList2Sorted, s, c->scratch, kt);
        k_compact<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(xmark.get(), xoff.get(), nb, list2.get());
        PRG_CHECK_LAUNCH();
        PRG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_brick_screen, kWave, 0));
        grid = std::min<int64_t>(nb, int64_t(std::max(1, per_cu)) * std::max(1, c->num_cu));
        k_brick_screen<<<unsigned(grid), kWave, 0, s>>>(A, list2.get(), bnv.get(), bnt.get(), bsign.get(), xlist.get(),
                                                        dbricks.get(), dpairs.get(), C, int(c->sub_certify));
        PRG_CHECK_LAUNCH();
        // the exact certificate's list
        PRG_HIP(hipMemsetAsync(xmark.get(), 0, size_t(nb) * 4, s));
        k_mark_listed<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(xlist.get(), C + kCntExact, nb, xmark.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive(xmark.get(), xoff.get(), nb, C + kCntExactSorted, s, c->scratch, kt);
        k_compact<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(xmark.get(), xoff.get(), nb, xsorted.get());
        PRG_CHECK_LAUNCH();
        PRG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_brick_exact, kWave, 0));
        grid = std::min<int64_t>(nb, int64_t(std::max(1, per_cu)) * std::max(1, c->num_cu));
        k_brick_exact<<<unsigned(grid), kWave, 0, s>>>(A, xsorted.get(), bsign.get(), dlist.get(), dbricks.get(),
                                                       dpairs.get(), C, int(c->sub_certify));
        PRG_CHECK_LAUNCH();
        k_brick_deps<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(dbricks.get(), dpairs.get(), bsign.get(),
                                                                dlist.get(), C);
        PRG_CHECK_LAUNCH();
        // the evaluation's list in brick order, in place of the unordered one
        PRG_HIP(hipMemsetAsync(xmark.get(), 0, size_t(nb) * 4, s));
        k_mark_listed<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(dlist.get(), C + kCntDense, nb, xmark.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive(xmark.get(), xoff.get(), nb, C + kCntDenseSorted, s, c->scratch, kt);
        k_compact<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(xmark.get(), xoff.get(), nb, dlist.get());
        PRG_CHECK_LAUNCH();
    }
    PRG_HIP(hipMemcpyAsync(H, C, kCntSlots * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    const int64_t ndense = H[kCntDense];
    PoolBuf<float> gvals{&c->pool[134]};
    PoolBuf<int32_t> didx{&c->pool[135]};
    gvals.reserve(size_t(std::max<int64_t>(1, ndense)) * kBrickPts);
    didx.reserve(nb);
    {
        TimedScope ts(kt, "subdivide_count", 0.0);
        int per_cu = 0;
        PRG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_brick_count, kWave, 0));
        const int64_t grid = std::min<int64_t>(ndense, int64_t(std::max(1, per_cu)) * std::max(1, c->num_cu));
        if (grid > 0) {
            k_brick_count<<<unsigned(grid), kWave, 0, s>>>(A, dlist.get(), C, bnv.get(), bnt.get(), gvals.get(),
                                                           didx.get());
            PRG_CHECK_LAUNCH();
        }
    }
    scan_exclusive2(bnv.get(), bvoff.get(), C + kCntVerts, bnt.get(), btoff.get(), C + kCntTris, nb, s, c->scratch,
                    kt);
    k_brick_active<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(bnv.get(), bnt.get(), nb, aflag.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(aflag.get(), aoff.get(), nb, C + kCntActive, s, c->scratch, kt);
    k_brick_compact<<<grid_for(nb, kBlockS), kBlockS, 0, s>>>(aflag.get(), aoff.get(), nb, alist.get(), amap.get());
    PRG_CHECK_LAUNCH();
    PRG_HIP(hipMemcpyAsync(H, C, kCntSlots * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    const int64_t snv = H[kCntVerts], snt = H[kCntTris], nact = H[kCntActive];
    c->sub_bricks = nb;
    c->sub_exact = H[kCntExact];
    c->sub_evaluated = H[kCntDense];
    if (snv < 0 || snt < 0 || nact < 0) throw Error{PRG_E_NOMEM, "subdivision mesh exceeds int32 counts"};
    if (c->mesh_nv + snv >= (int64_t(1) << 31)) throw Error{PRG_E_NOMEM, "mesh vertex count exceeds int32 indices"};
    if (nact == 0) return;

    // passes 2 and 3 over the bricks with output
    PoolBuf<uint16_t> lofs_tab{&c->pool[124]}, emask_tab{&c->pool[125]};
    PoolBuf<uint8_t> cube_tab{&c->pool[126]};
    lofs_tab.reserve(size_t(nact) * kBrickLeaves);
    emask_tab.reserve(size_t(nact) * kBrickLeaves);
    cube_tab.reserve(size_t(nact) * kBrickLeaves);
    float* SV = c->sub_v.reserve(size_t(3) * snv + 3);
    int32_t* ST = c->sub_t.reserve(size_t(3) * snt + 3);
    {
        TimedScope ts(kt, "subdivide_vertices", 12.0 * double(snv));
        k_brick_vertices<<<unsigned(nact), kWave, 0, s>>>(A, alist.get(), bvoff.get(), lofs_tab.get(),
                                                          emask_tab.get(), cube_tab.get(), gvals.get(), didx.get(), SV);
        PRG_CHECK_LAUNCH();
    }
    {
        TimedScope ts(kt, "subdivide_triangles", 12.0 * double(snt));
        k_brick_triangles<<<unsigned(nact), kWave, 0, s>>>(A, alist.get(), amap.get(), bvoff.get(), btoff.get(),
                                                           lofs_tab.get(), emask_tab.get(), cube_tab.get(),
                                                           c->mesh_nv, ST, C + kCntErr);
        PRG_CHECK_LAUNCH();
    }
    PRG_HIP(hipMemcpyAsync(H + kCntErr, C + kCntErr, 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    if (H[kCntErr]) throw Error{PRG_E_HIP, "subdivision: a triangle edge has no owner vertex"};
    c->sub_nv = snv;
    c->sub_nt = snt;
}

}  // namespace prg
