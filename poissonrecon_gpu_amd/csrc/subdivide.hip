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
// Here every virtual leaf of every group is handled in one pass, one thread
// per leaf, without building the virtual tree.  A virtual leaf is (root,
// Morton offset inside the root); its key is the root's key with the offset in
// the low bits.  The combined tree of a group has exactly these facts:
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
#include <algorithm>
#include <vector>

#include "ctx.h"
#include "mc_common.h"

namespace prg {
namespace {

constexpr int kBlockS = 256;
constexpr int kFinerDepth = 3;           // main.cu:3894: shallower roots go one at a time
constexpr int64_t kMaxVirtualLeaves = int64_t(1) << 28;

// parentFaceKind (MarchingCubes.cuh:708-717), indexed [sonKey][face] literally
__constant__ int8_t c_parent_face[8][6] = {{0, -1, 2, -1, 4, -1}, {-1, 1, 2, -1, 4, -1}, {0, -1, -1, 3, 4, -1},
                                           {-1, 1, -1, 3, 4, -1}, {0, -1, 2, -1, -1, 5}, {-1, 1, 2, -1, -1, 5},
                                           {0, -1, -1, 3, -1, 5}, {0, 1, -1, 3, -1, 5}};
// faceEdges (MarchingCubes.cuh:734-741) as bit masks of edge kinds
__constant__ uint16_t c_face_edges[6] = {(1 << 4) | (1 << 6) | (1 << 8) | (1 << 10), (1 << 5) | (1 << 7) | (1 << 9) | (1 << 11),
                                         (1 << 0) | (1 << 2) | (1 << 8) | (1 << 9),  (1 << 1) | (1 << 3) | (1 << 10) | (1 << 11),
                                         (1 << 0) | (1 << 1) | (1 << 4) | (1 << 5),  (1 << 2) | (1 << 3) | (1 << 6) | (1 << 7)};
// childrenVertexKind (MarchingCubes.cuh:720-723), used literally
__constant__ int8_t c_children_vertex_kind[8] = {0, 1, 3, 2, 4, 5, 7, 6};

__device__ __forceinline__ int depth_of(int64_t i, const Levels& L, int D) {
    int d = 0;
    while (d < D && i >= L.base[d + 1]) d++;
    return d;
}

__device__ __forceinline__ int nidx(int dx, int dy, int dz) { return 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1); }

// The face record holding face j of node X (initFaceArray / maintainFaceNodePointer,
// main.cu:2827-2955): owned by the smaller key of X and its neighbour across
// the face, numbered (owner, kind seen from the owner).
__device__ __forceinline__ int64_t face_rec(int64_t X, int j, const uint32_t* __restrict__ key,
                                            const int32_t* __restrict__ neigh, int64_t M) {
    const int o = j >> 1, dl = (j & 1) ? 1 : -1;
    const int jj = nidx(o == 0 ? dl : 0, o == 1 ? dl : 0, o == 2 ? dl : 0);
    const int32_t nb = neigh[int64_t(jj) * M + X];
    if (nb != -1 && key[nb] < key[X]) return 6 * int64_t(nb) + (j ^ 1);
    return 6 * X + j;
}

__device__ __forceinline__ bool has_parent_face(int64_t rec, const uint32_t* __restrict__ key,
                                                const int32_t* __restrict__ parent, const Levels& L, int D) {
    const int64_t o = rec / 6;
    const int kind = int(rec % 6);
    if (parent[o] == -1) return false;
    const int d = depth_of(o, L, D);
    const int son = int((key[o] >> (3 * (D - d))) & 7u);
    return c_parent_face[son][kind] != -1;
}

// hasSurfaceIntersection marks from the leaf-level triangles (generateTrianglePos,
// main.cu:2734-2755): every face of a leaf-level node touched by one of its
// triangle edges, and the faces up the parent chain while the face record
// has a parent face.
__global__ void __launch_bounds__(kBlockS) k_face_marks(const float* cval, int64_t MD, int64_t baseD,
                                                        const uint32_t* key, const int32_t* neigh, int64_t M,
                                                        const int32_t* parent, Levels L, int D, uint8_t* inter) {
    const int64_t t = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (t >= MD) return;
    const int cube = cube_of(cval, t);
    const int n = 3 * tri_count(cube);
    if (!n) return;
    uint32_t em = 0;
    for (int k = 0; k < n; k++) em |= 1u << c_tri[cube][k];
    const int64_t X = baseD + t;
    for (int j = 0; j < 6; j++) {
        if (!(em & c_face_edges[j])) continue;
        int64_t rec = face_rec(X, j, key, neigh, M);
        inter[rec] = 1;
        int64_t pnode = parent[X];
        while (has_parent_face(rec, key, parent, L, D)) {
            rec = face_rec(pnode, j, key, neigh, M);
            pnode = parent[pnode];
            inter[rec] = 1;
        }
    }
}

__global__ void __launch_bounds__(kBlockS) k_flag_coarse_leaves(const int32_t* child_base, int64_t n, int32_t* flag) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < n) flag[i] = (i >= 1 && child_base[i] < 0) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlockS) k_compact(const int32_t* flag, const int32_t* off, int64_t n, int32_t* out) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < n && flag[i]) out[off[i]] = int32_t(i);
}

// Corner owners of the coarse leaves' corners among the same-depth nodes
// (initVertexOwner, main.cu:1438-1498); marks each (owner, corner) needed.
__global__ void __launch_bounds__(kBlockS) k_coarse_corner_owner(const int32_t* cl, int64_t ncl, const uint32_t* key,
                                                                 const int32_t* neigh, int64_t M, int32_t* oref,
                                                                 int32_t* need) {
    const int64_t g = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (g >= 8 * ncl) return;
    const int64_t X = cl[g >> 3];
    const int j = int(g & 7);
    const int cx = j & 1, cy = (j >> 1) & 1, cz = (j >> 2) & 1;
    uint32_t best = key[X];
    int64_t owner = X;
    int jo = j;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int e = 0; e < 2; e++) {
                if (!(a | b | e)) continue;
                const int32_t nb = neigh[int64_t(nidx(a ? 2 * cx - 1 : 0, b ? 2 * cy - 1 : 0, e ? 2 * cz - 1 : 0)) * M + X];
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
__global__ void __launch_bounds__(kBlockS) k_coarse_corner_value(const int32_t* list, int64_t n, const uint32_t* key,
                                                                 const int32_t* pnt, Levels L, const int32_t* parent,
                                                                 const int32_t* child_base, const float* x,
                                                                 const float* bf, int D, const float* iso, float* out) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i >= n) return;
    const int64_t O = list[i] >> 3;
    const int jo = list[i] & 7;
    const int d = depth_of(O, L, D);
    float c[3];
    node_center(key[O], d, D, c);
    const float hw = 1.0f / float(1 << (d + 1));
    const float px = c[0] + float(2 * (jo & 1) - 1) * hw;
    const float py = c[1] + float(2 * ((jo & 2) >> 1) - 1) * hw;
    const float pz = c[2] + float(2 * ((jo & 4) >> 2) - 1) * hw;
    float f = implicit_value(O, d, px, py, pz, key, pnt, L, parent, x, bf, D);
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
// + ifSubdivide, main.cu:2957-2992)
__global__ void __launch_bounds__(kBlockS) k_coarse_select(const int32_t* cl, int64_t ncl, const int32_t* oref,
                                                           const int32_t* nid, const float* cv, const uint32_t* key,
                                                           const int32_t* neigh, int64_t M, const uint8_t* inter,
                                                           int32_t* rflag) {
    const int64_t c = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (c >= ncl) return;
    const int64_t X = cl[c];
    // vertices[0] is the (0,0,0) corner in either numbering
    const float v0 = cv[nid[oref[8 * c]]];
    const float sign = (v0 < 0) ? -1.0f : 1.0f;
    bool tri = false;
    for (int j = 1; j < 8; j++)
        if (sign * cv[nid[oref[8 * c + j]]] < 0) tri = true;
    bool hit = false;
    for (int j = 0; j < 6; j++)
        if (inter[face_rec(X, j, key, neigh, M)]) hit = true;
    rflag[c] = (tri || hit) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlockS) k_root_of(const int32_t* roots, int64_t nr, int32_t* root_of) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i < nr) root_of[roots[i]] = int32_t(i);
}

// ---- virtual leaves ------------------------------------------------------------------
struct Groups {
    const int64_t* loff;     // first virtual leaf of each root, nr + 1 entries
    const int32_t* roots;    // real index of each root, in node order
    const int8_t* rdepth;    // depth of each root
    const int32_t* root_of;  // per real node: its root number or -1
    int64_t nr;
};

struct VLeaf {
    int64_t ri;
    int32_t rho;
    int r;
    uint32_t key;
    int32_t cx, cy, cz;  // depth-D cell coordinates
    int32_t rx, ry, rz;  // the root's depth-r coordinates
};

__device__ __forceinline__ VLeaf decode(int64_t v, const Groups& G, const uint32_t* __restrict__ key, int D) {
    int64_t lo = 0, hi = G.nr - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (G.loff[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    VLeaf L;
    L.ri = lo;
    L.rho = G.roots[lo];
    L.r = G.rdepth[lo];
    L.key = key[L.rho] | uint32_t(v - G.loff[lo]);
    L.cx = int32_t(compact3(L.key >> 2));
    L.cy = int32_t(compact3(L.key >> 1));
    L.cz = int32_t(compact3(L.key));
    L.rx = L.cx >> (D - L.r);
    L.ry = L.cy >> (D - L.r);
    L.rz = L.cz >> (D - L.r);
    return L;
}

// Virtual leaf number of the depth-D cell (nx, ny, nz) in the group of leaf L,
// or -1 if that cell is not a virtual leaf of the group.
__device__ __forceinline__ int64_t virt_index(const VLeaf& L, int32_t nx, int32_t ny, int32_t nz, const Groups& G,
                                              const int32_t* __restrict__ neigh, int64_t M, int D) {
    const int32_t W = 1 << D;
    if (nx < 0 || ny < 0 || nz < 0 || nx >= W || ny >= W || nz >= W) return -1;
    const int sh = D - L.r;
    const int ax = (nx >> sh) - L.rx, ay = (ny >> sh) - L.ry, az = (nz >> sh) - L.rz;
    const uint32_t low = cell_key(uint32_t(nx), uint32_t(ny), uint32_t(nz)) & ((1u << (3 * sh)) - 1u);
    if (!(ax | ay | az)) return G.loff[L.ri] + low;
    if (L.r < kFinerDepth) return -1;
    const int32_t a = neigh[int64_t(nidx(ax, ay, az)) * M + L.rho];
    if (a < 0) return -1;
    const int32_t q = G.root_of[a];
    if (q < 0) return -1;
    return G.loff[q] + low;
}

// initSubdivideVertexOwner (main.cu:1500-1568): corners of leaf v it owns
__global__ void __launch_bounds__(kBlockS) k_vcorner_own(int64_t nv, Groups G, const uint32_t* key,
                                                         const int32_t* neigh, int64_t M, int D, uint8_t* cmask,
                                                         int32_t* ccnt) {
    const int64_t v = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (v >= nv) return;
    const VLeaf L = decode(v, G, key, D);
    int mask = 0;
    for (int j = 0; j < 8; j++) {
        const int sx = (j & 1) ? 1 : -1, sy = (j & 2) ? 1 : -1, sz = (j & 4) ? 1 : -1;
        bool own = true;
        for (int q = 1; q < 8 && own; q++) {
            const int32_t nx = L.cx + ((q & 1) ? sx : 0), ny = L.cy + ((q & 2) ? sy : 0), nz = L.cz + ((q & 4) ? sz : 0);
            if (virt_index(L, nx, ny, nz, G, neigh, M, D) < 0) continue;
            if (cell_key(uint32_t(nx), uint32_t(ny), uint32_t(nz)) < L.key) own = false;
        }
        if (own) mask |= 1 << j;
    }
    cmask[v] = uint8_t(mask);
    ccnt[v] = __popc(mask);
}

__global__ void __launch_bounds__(kBlockS) k_vcorner_list(int64_t nv, const uint8_t* cmask, const int32_t* coff,
                                                          int64_t* clist) {
    const int64_t v = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (v >= nv) return;
    const int m = cmask[v];
    int rank = 0;
    for (int j = 0; j < 8; j++)
        if (m & (1 << j)) clist[coff[v] + rank++] = 8 * v + j;
}

// computeSubdivideVertexImplicitFunctionValue (main.cu:2328-2442) for each
// owned corner: the real neighbours of the owner's virtual ancestors (depth D
// up to r + 1), then the root's chain through its real neighbours and real
// ancestors.
__global__ void __launch_bounds__(kBlockS) k_vcorner_value(const int64_t* clist, int64_t n, Groups G,
                                                           const uint32_t* key, const int32_t* neigh, int64_t M,
                                                           const int32_t* child_base, const int32_t* pnt, Levels Lv,
                                                           const int32_t* parent, const float* x, const float* bf,
                                                           int D, const float* iso, float* vval) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i >= n) return;
    const int64_t v = clist[i] >> 3;
    const int j = int(clist[i] & 7);
    const VLeaf L = decode(v, G, key, D);
    float c[3];
    node_center(L.key, D, D, c);
    const float hw = 1.0f / float(1 << (D + 1));
    const float px = c[0] + float(2 * (j & 1) - 1) * hw;
    const float py = c[1] + float(2 * ((j & 2) >> 1) - 1) * hw;
    const float pz = c[2] + float(2 * ((j & 4) >> 2) - 1) * hw;
    float f = 0.0f;
    for (int dd = D; dd > L.r; dd--) {
        const int32_t axo = L.cx >> (D - dd), ayo = L.cy >> (D - dd), azo = L.cz >> (D - dd);
        const int64_t fb = (int64_t(1) << dd) - 1;
        const int32_t width = 1 << dd;
        float bx[3], by[3], bz[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int32_t u = axo + a - 1, w = ayo + a - 1, z = azo + a - 1;
            bx[a] = (u >= 0 && u < width) ? basis_value(bf, fb + u, px) : 0.0f;
            by[a] = (w >= 0 && w < width) ? basis_value(bf, fb + w, py) : 0.0f;
            bz[a] = (z >= 0 && z < width) ? basis_value(bf, fb + z, pz) : 0.0f;
        }
        const int up = dd - L.r;
        for (int k = 0; k < 27; k++) {
            const int32_t nx = axo + (k / 9 - 1), ny = ayo + ((k / 3) % 3 - 1), nz = azo + (k % 3 - 1);
            if (nx < 0 || ny < 0 || nz < 0 || nx >= width || ny >= width || nz >= width) continue;
            const int ax = (nx >> up) - L.rx, ay = (ny >> up) - L.ry, az = (nz >> up) - L.rz;
            if (!(ax | ay | az)) continue;  // inside the root: virtual, coefficient 0
            int32_t cur = neigh[int64_t(nidx(ax, ay, az)) * M + L.rho];
            for (int l = L.r + 1; l <= dd && cur >= 0; l++) {
                const int32_t cb = child_base[cur];
                const int sh = dd - l;
                cur = cb < 0 ? -1 : cb + ((((nx >> sh) & 1) << 2) | (((ny >> sh) & 1) << 1) | ((nz >> sh) & 1));
            }
            if (cur < 0) continue;  // absent, or below a coarse leaf (virtual or none)
            float tt = x[cur] * bx[k / 9];
            tt = tt * by[(k / 3) % 3];
            tt = tt * bz[k % 3];
            f = f + tt;
        }
    }
    f = chain_value(L.rho, L.r, px, py, pz, key, pnt, Lv, parent, x, bf, D, f);
    vval[i] = f - *iso;
}

// Corner values of each virtual leaf in square numbering (vertices[],
// maintainSubdivideVertexNodePointer main.cu:1683-1740), from the owners' records.
__global__ void __launch_bounds__(kBlockS) k_vcorner_square(int64_t nv, Groups G, const uint32_t* key,
                                                            const int32_t* neigh, int64_t M, int D,
                                                            const uint8_t* cmask, const int32_t* coff,
                                                            const float* vval, float* cval) {
    const int64_t v = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (v >= nv) return;
    const VLeaf L = decode(v, G, key, D);
    for (int q = 0; q < 8; q++) {
        const int y = (q >> 1) & 1, xx = (q & 1) ^ y, z = q >> 2;
        const int j = xx | (y << 1) | (z << 2);
        const int sx = xx ? 1 : -1, sy = y ? 1 : -1, sz = z ? 1 : -1;
        uint32_t best = L.key;
        int64_t ov = v;
        int jo = j;
        for (int m = 1; m < 8; m++) {
            const int32_t nx = L.cx + ((m & 1) ? sx : 0), ny = L.cy + ((m & 2) ? sy : 0), nz = L.cz + ((m & 4) ? sz : 0);
            const int64_t vi = virt_index(L, nx, ny, nz, G, neigh, M, D);
            if (vi < 0) continue;
            const uint32_t kk = cell_key(uint32_t(nx), uint32_t(ny), uint32_t(nz));
            if (kk < best) {
                best = kk;
                ov = vi;
                jo = j ^ m;
            }
        }
        const int idx = coff[ov] + __popc(uint32_t(cmask[ov]) & ((1u << jo) - 1u));
        cval[8 * v + q] = vval[idx];
    }
}

// Owner of edge e of leaf L among the virtual leaves sharing it
// (initSubdivideEdgeArray, main.cu:1867-2019); returns its number and the
// edge kind seen from it.
__device__ __forceinline__ int64_t edge_owner(const VLeaf& L, int64_t v, int e, const Groups& G,
                                              const int32_t* __restrict__ neigh, int64_t M, int D, int* eo) {
    const int o = e >> 2;
    const int a1 = (o == 0) ? 1 : 0, a2 = (o == 2) ? 1 : 2;
    const int b1 = e & 1, b2 = (e >> 1) & 1;
    uint32_t best = L.key;
    int64_t ov = v;
    *eo = e;
    for (int u = 0; u < 2; u++)
        for (int w = 0; w < 2; w++) {
            if (!(u | w)) continue;
            int dd[3] = {0, 0, 0};
            dd[a1] = u ? (2 * b1 - 1) : 0;
            dd[a2] = w ? (2 * b2 - 1) : 0;
            const int32_t nx = L.cx + dd[0], ny = L.cy + dd[1], nz = L.cz + dd[2];
            const int64_t vi = virt_index(L, nx, ny, nz, G, neigh, M, D);
            if (vi < 0) continue;
            const uint32_t kk = cell_key(uint32_t(nx), uint32_t(ny), uint32_t(nz));
            if (kk < best) {
                best = kk;
                ov = vi;
                *eo = (o << 2) | ((b2 ^ w) << 1) | (b1 ^ u);
            }
        }
    return ov;
}

// owned edges with a sign change (generateSubdivideVexNums, main.cu:2481-2531)
// and triangle counts (generateTriNums, main.cu:2562-2582)
__global__ void __launch_bounds__(kBlockS) k_vedges(int64_t nv, Groups G, const uint32_t* key, const int32_t* neigh,
                                                    int64_t M, int D, const float* cval, uint16_t* emask,
                                                    int32_t* ecnt, int32_t* tcnt) {
    const int64_t v = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (v >= nv) return;
    const VLeaf L = decode(v, G, key, D);
    int mask = 0;
    for (int e = 0; e < 12; e++) {
        int eo;
        if (edge_owner(L, v, e, G, neigh, M, D, &eo) != v) continue;
        const float f1 = cval[8 * v + c_edge_vertex[e][0]];
        const float f2 = cval[8 * v + c_edge_vertex[e][1]];
        if (f1 * f2 <= 0) mask |= 1 << e;
    }
    emask[v] = uint16_t(mask);
    ecnt[v] = __popc(mask);
    tcnt[v] = tri_count(cube_of(cval, v));
}

// interpolatePoint (main.cu:2584-2597) per owned crossing edge and the
// triangles of each leaf (generateSubdivideTrianglePos, main.cu:2759-2825)
__global__ void __launch_bounds__(kBlockS) k_vemit(int64_t nv, Groups G, const uint32_t* key, const int32_t* neigh,
                                                   int64_t M, int D, const float* cval, const uint16_t* emask,
                                                   const int32_t* eoff, const int32_t* toff, int64_t vbase,
                                                   float* out_v, int32_t* out_t) {
    const int64_t v = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (v >= nv) return;
    const VLeaf L = decode(v, G, key, D);
    const int mine = emask[v];
    if (mine) {
        float c[3];
        node_center(L.key, D, D, c);
        const float hw = 1.0f / float(1 << (D + 1));
        int rank = 0;
        for (int e = 0; e < 12; e++) {
            if (!(mine & (1 << e))) continue;
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
            const float f1 = cval[8 * v + q1], f2 = cval[8 * v + q2];
            const int dim = e >> 2;
            float o3[3] = {p1[0], p1[1], p1[2]};
            const float pivot = f1 / (f1 - f2);
            const float another = 1 - pivot;
            const float s1 = p2[dim] * pivot;
            const float s2 = p1[dim] * another;
            o3[dim] = s1 + s2;
            const int64_t id = eoff[v] + rank++;
            out_v[3 * id + 0] = o3[0];
            out_v[3 * id + 1] = o3[1];
            out_v[3 * id + 2] = o3[2];
        }
    }
    const int cube = cube_of(cval, v);
    const int n = 3 * tri_count(cube);
    const int64_t o = 3 * int64_t(toff[v]);
    for (int k = 0; k < n; k++) {
        const int e = c_tri[cube][k];
        int eo;
        const int64_t ov = edge_owner(L, v, e, G, neigh, M, D, &eo);
        out_t[o + k] = int32_t(vbase + eoff[ov] + __popc(uint32_t(emask[ov]) & ((1u << eo) - 1u)));
    }
}

int32_t read_total(const DBuf<int32_t>& tot, hipStream_t s) {
    int32_t v = 0;
    PRG_HIP(hipMemcpyAsync(&v, tot.get(), 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

void extract_subdivision(prg_ctx* c) {
    ensure_neighbours(c);
    ensure_tables(c);
    StageTimer st(c, PRG_STAGE_SUBDIVIDE);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M, MD = c->MD(), baseD = c->bases[D];
    KernelTimer& kt = c->kt;
    const Levels Lv = levels_of(c);
    c->sub_nv = c->sub_nt = 0;
    c->sub_roots = c->sub_leaves = 0;
    if (baseD <= 1) return;
    if (8 * baseD >= (int64_t(1) << 31)) throw Error{PRG_E_ARG, "too many nodes above the leaf level for subdivision"};

    // faces touched by the leaf-level surface
    DBuf<uint8_t> inter;
    inter.reserve(6 * M);
    PRG_HIP(hipMemsetAsync(inter.get(), 0, size_t(6 * M), s));
    {
        TimedScope ts(kt, "face_marks", 32.0 * double(MD));
        k_face_marks<<<grid_for(MD, kBlockS), kBlockS, 0, s>>>(c->corner_vals.get(), MD, baseD, c->key.get(),
                                                                c->neigh.get(), M, c->parent.get(), Lv, D, inter.get());
        PRG_CHECK_LAUNCH();
    }
    // coarse leaves
    DBuf<int32_t> flag, off, tot, cl;
    flag.reserve(baseD);
    off.reserve(baseD);
    tot.reserve(1);
    k_flag_coarse_leaves<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(c->child_base.get(), baseD, flag.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(flag.get(), off.get(), baseD, tot.get(), s, c->scratch, kt);
    const int64_t ncl = read_total(tot, s);
    if (ncl == 0) return;
    cl.reserve(ncl);
    k_compact<<<grid_for(baseD, kBlockS), kBlockS, 0, s>>>(flag.get(), off.get(), baseD, cl.get());
    PRG_CHECK_LAUNCH();
    // their corner values
    DBuf<int32_t> oref, need, nid, clist;
    oref.reserve(8 * ncl);
    need.reserve(8 * baseD);
    nid.reserve(8 * baseD);
    PRG_HIP(hipMemsetAsync(need.get(), 0, size_t(8 * baseD) * 4, s));
    k_coarse_corner_owner<<<grid_for(8 * ncl, kBlockS), kBlockS, 0, s>>>(cl.get(), ncl, c->key.get(), c->neigh.get(),
                                                                          M, oref.get(), need.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(need.get(), nid.get(), 8 * baseD, tot.get(), s, c->scratch, kt);
    const int64_t ncv = read_total(tot, s);
    clist.reserve(ncv);
    k_compact<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(need.get(), nid.get(), 8 * baseD, clist.get());
    PRG_CHECK_LAUNCH();
    DBuf<float> cv;
    cv.reserve(ncv);
    {
        TimedScope ts(kt, "coarse_corner_value", 24.0 * double(ncv));
        k_coarse_corner_value<<<grid_for(ncv, kBlockS), kBlockS, 0, s>>>(
            clist.get(), ncv, c->key.get(), c->pnt.get(), Lv, c->parent.get(), c->child_base.get(), c->x.get(),
            c->d_base_fn.get(), D, c->iso_dev.get(), cv.get());
        PRG_CHECK_LAUNCH();
    }
    // roots
    DBuf<int32_t> rflag, roff, roots;
    rflag.reserve(ncl);
    roff.reserve(ncl);
    k_coarse_select<<<grid_for(ncl, kBlockS), kBlockS, 0, s>>>(cl.get(), ncl, oref.get(), nid.get(), cv.get(),
                                                                c->key.get(), c->neigh.get(), M, inter.get(),
                                                                rflag.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(rflag.get(), roff.get(), ncl, tot.get(), s, c->scratch, kt);
    const int64_t nr = read_total(tot, s);
    if (nr == 0) return;
    DBuf<int32_t> ridx;  // positions in cl
    ridx.reserve(nr);
    k_compact<<<grid_for(ncl, kBlockS), kBlockS, 0, s>>>(rflag.get(), roff.get(), ncl, ridx.get());
    PRG_CHECK_LAUNCH();
    std::vector<int32_t> hpos(nr), hcl(ncl);
    PRG_HIP(hipMemcpyAsync(hpos.data(), ridx.get(), size_t(nr) * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipMemcpyAsync(hcl.data(), cl.get(), size_t(ncl) * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> hroots(nr);
    std::vector<int8_t> hdepth(nr);
    std::vector<int64_t> hloff(nr + 1, 0);
    for (int64_t i = 0; i < nr; i++) {
        hroots[i] = hcl[hpos[i]];
        int d = 0;
        while (d < D && hroots[i] >= c->bases[d + 1]) d++;
        hdepth[i] = int8_t(d);
        hloff[i + 1] = hloff[i] + (int64_t(1) << (3 * (D - d)));
    }
    const int64_t nv = hloff[nr];
    if (nv > kMaxVirtualLeaves)
        throw Error{PRG_E_NOMEM, "coarse-leaf subdivision needs " + std::to_string(nv) +
                                     " virtual leaves, more than the supported 2^28"};
    c->sub_roots = nr;
    c->sub_leaves = nv;
    DBuf<int32_t> droots, root_of;
    DBuf<int8_t> ddepth;
    DBuf<int64_t> dloff;
    PRG_HIP(hipMemcpyAsync(droots.reserve(nr), hroots.data(), size_t(nr) * 4, hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(ddepth.reserve(nr), hdepth.data(), size_t(nr), hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(dloff.reserve(nr + 1), hloff.data(), size_t(nr + 1) * 8, hipMemcpyHostToDevice, s));
    root_of.reserve(M);
    PRG_HIP(hipMemsetAsync(root_of.get(), 0xff, size_t(M) * 4, s));
    k_root_of<<<grid_for(nr, kBlockS), kBlockS, 0, s>>>(droots.get(), nr, root_of.get());
    PRG_CHECK_LAUNCH();
    Groups G{dloff.get(), droots.get(), ddepth.get(), root_of.get(), nr};

    // marching cubes over the virtual leaves
    DBuf<uint8_t> cmask;
    DBuf<int32_t> ccnt, coff;
    cmask.reserve(nv);
    ccnt.reserve(nv);
    coff.reserve(nv);
    const unsigned gv = grid_for(nv, kBlockS);
    k_vcorner_own<<<gv, kBlockS, 0, s>>>(nv, G, c->key.get(), c->neigh.get(), M, D, cmask.get(), ccnt.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(ccnt.get(), coff.get(), nv, tot.get(), s, c->scratch, kt);
    const int64_t nvc = read_total(tot, s);
    DBuf<int64_t> vlist;
    DBuf<float> vval, cval;
    vlist.reserve(nvc);
    vval.reserve(nvc);
    cval.reserve(8 * nv);
    k_vcorner_list<<<gv, kBlockS, 0, s>>>(nv, cmask.get(), coff.get(), vlist.get());
    PRG_CHECK_LAUNCH();
    {
        TimedScope ts(kt, "subdivide_corner_value", 24.0 * double(nvc));
        k_vcorner_value<<<grid_for(nvc, kBlockS), kBlockS, 0, s>>>(
            vlist.get(), nvc, G, c->key.get(), c->neigh.get(), M, c->child_base.get(), c->pnt.get(), Lv,
            c->parent.get(), c->x.get(), c->d_base_fn.get(), D, c->iso_dev.get(), vval.get());
        PRG_CHECK_LAUNCH();
    }
    k_vcorner_square<<<gv, kBlockS, 0, s>>>(nv, G, c->key.get(), c->neigh.get(), M, D, cmask.get(), coff.get(),
                                            vval.get(), cval.get());
    PRG_CHECK_LAUNCH();
    DBuf<uint16_t> emask;
    DBuf<int32_t> ecnt, eoff, tcnt, toff, ttot;
    emask.reserve(nv);
    ecnt.reserve(nv);
    eoff.reserve(nv);
    tcnt.reserve(nv);
    toff.reserve(nv);
    ttot.reserve(1);
    k_vedges<<<gv, kBlockS, 0, s>>>(nv, G, c->key.get(), c->neigh.get(), M, D, cval.get(), emask.get(), ecnt.get(),
                                    tcnt.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(ecnt.get(), eoff.get(), nv, tot.get(), s, c->scratch, kt);
    scan_exclusive(tcnt.get(), toff.get(), nv, ttot.get(), s, c->scratch, kt);
    const int64_t snv = read_total(tot, s);
    const int64_t snt = read_total(ttot, s);
    if (c->mesh_nv + snv >= (int64_t(1) << 31)) throw Error{PRG_E_NOMEM, "mesh vertex count exceeds int32 indices"};
    float* SV = c->sub_v.reserve(size_t(3) * snv + 3);
    int32_t* ST = c->sub_t.reserve(size_t(3) * snt + 3);
    k_vemit<<<gv, kBlockS, 0, s>>>(nv, G, c->key.get(), c->neigh.get(), M, D, cval.get(), emask.get(), eoff.get(),
                                   toff.get(), c->mesh_nv, SV, ST);
    PRG_CHECK_LAUNCH();
    PRG_HIP(hipStreamSynchronize(s));
    c->sub_nv = snv;
    c->sub_nt = snt;
}

}  // namespace prg
