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

namespace prg {
namespace {

constexpr int kBlockS = 256;
constexpr int kFinerDepth = 3;  // main.cu:3894: shallower roots go one at a time

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

// slots of the device counter array
enum { kCntCoarse = 0, kCntCorners = 1, kCntRoots = 2, kCntBricks = 3, kCntVerts = 4, kCntTris = 5, kCntActive = 6,
       kCntLeaves = 7, kCntErr = 8, kCntSlots = 16 };

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
// ncl (device) coarse leaves in cl.
__global__ void __launch_bounds__(kBlockS) k_coarse_corner_owner(const int32_t* cl, const int32_t* ncl,
                                                                 const uint32_t* key, const int32_t* neigh, int64_t M,
                                                                 int32_t* oref, int32_t* need) {
    const int64_t g = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (g >= 8 * int64_t(*ncl)) return;
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
__global__ void __launch_bounds__(kBlockS) k_coarse_corner_value(const int32_t* list, const int32_t* n_dev,
                                                                 const uint32_t* key, const int32_t* pnt, Levels L,
                                                                 const int32_t* parent, const int32_t* child_base,
                                                                 const float* x, const float* bf, int D,
                                                                 const float* iso, float* out) {
    const int64_t i = int64_t(blockIdx.x) * kBlockS + threadIdx.x;
    if (i >= *n_dev) return;
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
// + ifSubdivide, main.cu:2957-2992); 0 past the ncl coarse leaves, so a scan
// over all n slots counts the roots.
__global__ void __launch_bounds__(kBlockS) k_coarse_select(const int32_t* cl, const int32_t* ncl, int64_t n,
                                                           const int32_t* oref, const int32_t* nid, const float* cv,
                                                           const uint32_t* key, const int32_t* neigh, int64_t M,
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
    for (int j = 0; j < 6; j++)
        if (inter[face_rec(X, j, key, neigh, M)]) hit = true;
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

// Largest |x| of each depth (one workgroup per depth).
__global__ void __launch_bounds__(1024) k_depth_absmax(const float* x, Levels L, int D, int64_t M, float* out) {
    __shared__ float red[1024];
    const int d = blockIdx.x;
    const int64_t b = L.base[d], e = d < D ? L.base[d + 1] : M;
    float m = 0.0f;
    for (int64_t i = b + threadIdx.x; i < e; i += 1024) m = fmaxf(m, fabsf(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[d] = red[0];
}

// Largest |value| of each depth's entries of the corner basis table.
__global__ void __launch_bounds__(1024) k_table_absmax(const float4* tab, int D, float* out) {
    __shared__ float red[1024];
    const int d = blockIdx.x;
    const int64_t n = int64_t(2) << D;  // (ix, side) entries per depth
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < n; i += 1024) {
        const float4 t = tab[int64_t(d) * n + i];
        m = fmaxf(m, fmaxf(fabsf(t.x), fmaxf(fabsf(t.y), fabsf(t.z))));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[d] = red[0];
}

struct SubArgs {
    const uint32_t* key;
    const int32_t* neigh;  // 27 planes
    int64_t M;
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
    const float* xmax;       // per depth: largest |x|
    const float* bmax;       // per depth: largest |basis value| in the corner basis table
};

// A brick and its root (wave-uniform).
struct Brick {
    int32_t q, rho, r, kb;
    int32_t R[3];     // root coordinates at depth r
    int32_t R0[3];    // first depth-D cell of the root
    int32_t Sr;       // root side in depth-D cells
    int32_t org[3];   // first depth-D cell of the brick
    uint32_t bkey;    // key of the brick's first leaf
    bool deep;        // a neighbour of the root has children (real nodes below depth r)
};

constexpr int kP = (1 << kBrickLog) + 1;  // grid points per brick side (at most)

// Scratch of the brick certificate (brick_certify): per depth, the
// coefficients of a 4x4x4 block of nodes and the separable partial sums.
struct CertLds {
    float w[3][kP][4], wa[3][kP][4];  // per axis and grid coordinate: basis weights over the block (and |.|)
    float X4[64], X4a[64];
    float Z[16][kP], Za[16][kP];
    float Y[4][kP][kP], Ya[4][kP][kP];
};

// Blocks of the real nodes below the root that a brick's corner owners see:
// at most (D - 3 - r) blocks of 4^3 cells and blocks of 5^3, 7^3 and 11^3
// cells at the three finest depths.
constexpr int kXBlk = 2304;

// LDS of one brick's wave
struct BrickLds {
    union {
        float G[kBrickPts];  // f - iso at the brick's corners (grid points)
        CertLds cert;
    };
    float X[(kMaxDepth + 1) * 27];     // the root's chain: neighbour values at depths 0..r
    float XB[kXBlk];                   // per depth below the root: its block (xb_off)
    int32_t xb_off[kMaxDepth + 1];     // start of depth dd's block in XB, or -1 (no real node in it)
    int32_t NB[27], NBK[27], NBR[27];  // the root's neighbours, their child bases, their root numbers (group)
};

__device__ __forceinline__ void decode_cell(uint32_t k, int32_t c[3]) {
    c[0] = int32_t(compact3(k >> 2));
    c[1] = int32_t(compact3(k >> 1));
    c[2] = int32_t(compact3(k));
}

// Finds brick b's root (binary search of the brick offsets) and loads the
// root's neighbours and ancestor chain into LDS; `chain` = 0 skips the chain.
__device__ void brick_setup(const SubArgs& A, int32_t b, Brick& B, BrickLds& S, bool chain) {
    const int D = A.D;
    int32_t lo = 0, hi = A.cnt[kCntRoots] - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (A.boff[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    B.q = lo;
    B.rho = A.roots[lo];
    B.r = depth_of(B.rho, A.L, D);
    B.kb = brick_log(D, B.r);
    const uint32_t rkey = A.key[B.rho];
    int32_t rc[3];
    decode_cell(rkey, rc);
    B.Sr = 1 << (D - B.r);
    for (int a = 0; a < 3; a++) {
        B.R[a] = rc[a] >> (D - B.r);
        B.R0[a] = B.R[a] << (D - B.r);
    }
    B.bkey = rkey | (uint32_t(b - A.boff[lo]) << (3 * B.kb));
    decode_cell(B.bkey, B.org);
    const int lane = threadIdx.x;
    if (lane < 27) {
        const int32_t nb = A.neigh[int64_t(lane) * A.M + B.rho];
        S.NB[lane] = nb;
        S.NBK[lane] = nb < 0 ? -1 : A.child_base[nb];
        S.NBR[lane] = (nb < 0 || B.r < kFinerDepth) ? -1 : A.root_of[nb];
        if (chain) S.X[27 * B.r + lane] = nb < 0 ? 0.0f : A.x[nb];
    }
    if (chain) {
        // ancestors' rows (their 27 neighbour values, then the parent), depths r-1 .. 1
        int32_t a = A.parent[B.rho];
        for (int d = B.r - 1; d >= 1; d--) {
            const float* R = A.rows + kRowFloats * (int64_t(a) - A.L.base[1]);
            if (lane < 27) S.X[27 * d + lane] = R[lane];
            a = __builtin_amdgcn_readfirstlane(__float_as_int(R[27]));
        }
        if (lane == 0) S.X[0] = A.x[0];
    }
    __syncthreads();
    bool deep = false;
    for (int j = 0; j < 27; j++)
        if (j != 13 && S.NBK[j] >= 0) deep = true;
    B.deep = deep;
}

// Whether the depth-D cell c (inside the root's 3x3x3 neighbourhood) is a
// virtual leaf of the brick's group; *q receives its root's neighbour slot
// (13 for the brick's own root).
__device__ __forceinline__ bool cell_virtual(const Brick& B, const BrickLds& S, int D, const int32_t c[3], int* slot) {
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

// Coefficient of the depth-dd cell c for corners of the brick's root: the
// real node reached from the root's neighbour by c's key bits, or 0 (inside
// the root: virtual; absent; below a coarse leaf).
__device__ __forceinline__ float deep_coef(const SubArgs& A, const Brick& B, const BrickLds& S, int dd,
                                           const int32_t c[3]) {
    const int32_t W = 1 << dd;
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= W || c[1] >= W || c[2] >= W) return 0.0f;
    const int up = dd - B.r;
    const int ox = (c[0] >> up) - B.R[0], oy = (c[1] >> up) - B.R[1], oz = (c[2] >> up) - B.R[2];
    if (!(ox | oy | oz)) return 0.0f;
    const int jn = nidx(ox, oy, oz);
    int32_t cur = S.NBK[jn] < 0 ? -1 : S.NB[jn];
    for (int l = B.r + 1; l <= dd && cur >= 0; l++) {
        const int32_t cb = A.child_base[cur];
        const int s2 = dd - l;
        cur = cb < 0 ? -1 : cb + ((((c[0] >> s2) & 1) << 2) | (((c[1] >> s2) & 1) << 1) | ((c[2] >> s2) & 1));
    }
    return cur >= 0 ? A.x[cur] : 0.0f;
}

// Fills the brick's blocks (depths r+1..D): each cell's coefficient, one
// walk down the tree per cell instead of one per corner and neighbour.  A
// depth whose block lies inside the root, or holds no real node, gets none.
__device__ void brick_blocks(const SubArgs& A, const Brick& B, BrickLds& S) {
    const int D = A.D, lane = threadIdx.x;
    int off = 0;
    for (int dd = B.r + 1; dd <= D; dd++) {
        const int W = block_width(D, B.kb, dd), sh = D - dd, up = dd - B.r;
        int32_t lo[3];
        bool inside = true;
        for (int a = 0; a < 3; a++) {
            lo[a] = (B.org[a] >> sh) - 2;
            const int32_t rlo = B.R[a] << up, rhi = rlo + (1 << up) - 1;
            inside = inside && lo[a] >= rlo && lo[a] + W - 1 <= rhi;
        }
        int slot = -1;
        if (!inside && B.deep) {
            const int n = W * W * W;
            bool any = false;
            for (int e = lane; e < n; e += kWave) {
                const int32_t c[3] = {lo[0] + e / (W * W), lo[1] + (e / W) % W, lo[2] + e % W};
                const float xv = deep_coef(A, B, S, dd, c);
                S.XB[off + e] = xv;
                any = any || xv != 0.0f;
            }
            if (__builtin_amdgcn_ballot_w64(any)) {
                slot = off;
                off += n;
            }
        }
        if (lane == 0) S.xb_off[dd] = slot;
    }
    __syncthreads();
}

// deep_sum for an owner leaf o of the brick's own root, reading the blocks.
// Every one of the 27 terms is added: a neighbour that is virtual, absent or
// outside the domain contributes +-0 (its coefficient or its basis value is
// 0), which leaves a float sum that started at +0 unchanged, so the result
// is deep_sum's bit for bit.
__device__ float deep_sum_lds(const SubArgs& A, const Brick& B, const BrickLds& S, const int32_t o[3], const int s[3]) {
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
        const float4 tx = A.btab[((int64_t(dd) << D) + o[0]) * 2 + s[0]];
        const float4 ty = A.btab[((int64_t(dd) << D) + o[1]) * 2 + s[1]];
        const float4 tz = A.btab[((int64_t(dd) << D) + o[2]) * 2 + s[2]];
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
__device__ float deep_sum(const SubArgs& A, int r, const int32_t R[3], const int32_t o[3], const int s[3], NbFn nb,
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
// root of the tree), continuing from f.
__device__ __forceinline__ float chain_lds(const SubArgs& A, const BrickLds& S, int r, const int32_t o[3],
                                           const int s[3], float f) {
    const int D = A.D;
    for (int d = r; d >= 1; d--) {
        const float4 tx = A.btab[((int64_t(d) << D) + o[0]) * 2 + s[0]];
        const float4 ty = A.btab[((int64_t(d) << D) + o[1]) * 2 + s[1]];
        const float4 tz = A.btab[((int64_t(d) << D) + o[2]) * 2 + s[2]];
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
    const float b0x = A.btab[o[0] * 2 + s[0]].y, b0y = A.btab[o[1] * 2 + s[1]].y, b0z = A.btab[o[2] * 2 + s[2]].y;
    float tt = S.X[0] * b0x;
    tt = tt * b0y;
    tt = tt * b0z;
    return f + tt;
}

// f - iso at grid point g (depth-D corner coordinates) of the brick's root cell,
// computed by its owner: the smallest-key virtual leaf around it.
__device__ float grid_value(const SubArgs& A, const Brick& B, const BrickLds& S, const int32_t g[3]) {
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
        uint32_t best = 0xffffffffu;
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
        for (int a = 0; a < 3; a++) side[a] = g[a] - o[a];
    }
    if (slot == 13) {
        const float f = B.deep ? deep_sum_lds(A, B, S, o, side) : 0.0f;
        return chain_lds(A, S, B.r, o, side, f) - *A.iso;
    }
    // owned by a leaf of another root of the group: that root's neighbours and chain
    const int32_t rho2 = S.NB[slot];
    const int32_t R2[3] = {B.R[0] + slot / 9 - 1, B.R[1] + (slot / 3) % 3 - 1, B.R[2] + slot % 3 - 1};
    const float f = deep_sum(
        A, B.r, R2, o, side, [&](int j) { return A.neigh[int64_t(j) * A.M + rho2]; },
        [&](int j) {
            const int32_t a = A.neigh[int64_t(j) * A.M + rho2];
            return a < 0 ? -1 : A.child_base[a];
        });
    float c[3];
    node_center(cell_key(uint32_t(o[0]), uint32_t(o[1]), uint32_t(o[2])), D, D, c);
    const float hw = 1.0f / float(1 << (D + 1));
    const float px = c[0] + float(2 * side[0] - 1) * hw;
    const float py = c[1] + float(2 * side[1] - 1) * hw;
    const float pz = c[2] + float(2 * side[2] - 1) * hw;
    return chain_value(rho2, B.r, px, py, pz, A.key, A.pnt, A.L, A.parent, A.x, A.bf, D, f) - *A.iso;
}

__device__ void brick_values(const SubArgs& A, const Brick& B, BrickLds& S) {
    const int P = (1 << B.kb) + 1;
    for (int p = threadIdx.x; p < P * P * P; p += kWave) {
        const int32_t g[3] = {B.org[0] + p % P, B.org[1] + (p / P) % P, B.org[2] + p / (P * P)};
        S.G[p] = grid_value(A, B, S, g);
    }
    __syncthreads();
}

// Owner of edge e of leaf v among the virtual leaves sharing it
// (initSubdivideEdgeArray, main.cu:1867-2019): its cell and the edge kind seen
// from it.
__device__ __forceinline__ void edge_owner(const Brick& B, const BrickLds& S, int D, const int32_t v[3], int e,
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
__device__ __forceinline__ void leaf_cube(const Brick& B, const BrickLds& S, int D, int l, int32_t v[3], float cv[8],
                                          int* cube, int* emask) {
    const int P = (1 << B.kb) + 1;
    const int lx = int(compact3(uint32_t(l) >> 2)), ly = int(compact3(uint32_t(l) >> 1)), lz = int(compact3(uint32_t(l)));
    v[0] = B.org[0] + lx;
    v[1] = B.org[1] + ly;
    v[2] = B.org[2] + lz;
    int cb = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = (q >> 1) & 1, xx = (q & 1) ^ y, z = q >> 2;
        cv[q] = S.G[((lz + z) * P + (ly + y)) * P + (lx + xx)];
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

// Owner cell along one axis of grid coordinate g for a root whose owners are
// all its own leaves: the smaller of the two cells around g that lies in the
// root (Morton keys grow with each coordinate).
__device__ __forceinline__ int32_t own_cell(int32_t g, int32_t R0) { return g == R0 ? g : g - 1; }

// Proves, when it can, that every corner value of the brick has one strict
// sign, so the brick has no crossing edge and no triangle.  For each depth
// whose nodes are fixed across the brick (the chain depths 0..r and the
// depths r+1..D-kb, where the whole brick lies in one cell) the sum
// S(g) = sum_k x_k b_x b_y b_z over the corner owner's 27 neighbours is
// evaluated exactly as a separable contraction of a 4x4x4 block of
// coefficients with per-axis weights taken from the same float basis table
// the evaluation reads (the fourth layer holds the neighbours of owners in the
// lower brick), together with A(g) = sum |x_k b_x b_y b_z|.  The float value
// the evaluation computes, a sum of n <= 27 (D+1) + 1 products of three
// roundings each, differs from S(g) by at most (n+3) 2^-24 A(g) (recursive
// summation bound) plus the rounding of this contraction; each finer depth
// contributes at most U(g): the sum of |x| over the 27 neighbours the owner's
// ancestor sees there (from the brick's blocks) times the depth's largest
// basis value cubed.  The brick is certified when
// |S(g) - iso| exceeds 420 2^-24 (A + U) + U with one sign at every corner.
// Bricks whose corners may be owned by another root of the group are not.
__device__ bool brick_certify(const SubArgs& A, const Brick& B, BrickLds& S) {
    const int D = A.D, kb = B.kb, Sb = 1 << kb, P = Sb + 1;
    const int lane = threadIdx.x;
    CertLds& L = S.cert;
    bool lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = B.org[a] == B.R0[a];
        hi[a] = B.org[a] + Sb == B.R0[a] + B.Sr;
    }
    for (int j = 0; j < 27; j++) {
        if (j == 13 || S.NBR[j] < 0) continue;
        const int dv[3] = {j / 9 - 1, (j / 3) % 3 - 1, j % 3 - 1};
        bool touches = true;
        for (int a = 0; a < 3; a++)
            if (dv[a] != 0) touches = touches && (dv[a] < 0 ? lo[a] : hi[a]);
        if (touches) return false;
    }
    constexpr int kPer = (kBrickPts + kWave - 1) / kWave;
    float sacc[kPer], aacc[kPer];
#pragma unroll
    for (int t = 0; t < kPer; t++) sacc[t] = aacc[t] = 0.0f;
    const int dlast = D - kb;  // deepest depth whose cell holds the whole brick
    for (int d = 0; d <= dlast; d++) {
        // the block's coefficients: cells A - 2 .. A + 1 per axis around the
        // brick's ancestor A at depth d
        float xv = 0.0f;
        {
            const int mx = lane >> 4, my = (lane >> 2) & 3, mz = lane & 3;
            if (d <= B.r) {
                if (mx >= 1 && my >= 1 && mz >= 1) {
                    if (d == 0) xv = (mx == 2 && my == 2 && mz == 2) ? S.X[0] : 0.0f;
                    else xv = S.X[27 * d + 9 * (mx - 1) + 3 * (my - 1) + (mz - 1)];
                }
            } else if (S.xb_off[d] >= 0) {
                xv = S.XB[S.xb_off[d] + lane];  // a 4^3 block, cells A - 2 .. A + 1
            }
        }
        if (__builtin_amdgcn_ballot_w64(xv != 0.0f) == 0) continue;  // nothing at this depth
        L.X4[lane] = xv;
        L.X4a[lane] = fabsf(xv);
        // per-axis weights: the owner's three basis values, placed at its
        // ancestor's offset inside the block
        if (lane < 3 * P) {
            const int ax = lane / P, i = lane % P;
            const int32_t g = B.org[ax] + i;
            const int32_t o = own_cell(g, B.R0[ax]);
            const int side = g - o;
            const int delta = (o >> (D - d)) - (B.org[ax] >> (D - d));  // -1 or 0
            const float4 t = A.btab[((int64_t(d) << D) + o) * 2 + side];
            const float tv[3] = {t.x, t.y, t.z};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int a = m - delta - 1;
                const float v = (a >= 0 && a <= 2) ? tv[a] : 0.0f;
                L.w[ax][i][m] = v;
                L.wa[ax][i][m] = fabsf(v);
            }
        }
        __syncthreads();
        for (int e = lane; e < 16 * P; e += kWave) {
            const int mn = e / P, k = e % P;
            float z = 0.0f, za = 0.0f;
#pragma unroll
            for (int l = 0; l < 4; l++) {
                z = __builtin_fmaf(L.X4[4 * mn + l], L.w[2][k][l], z);
                za = __builtin_fmaf(L.X4a[4 * mn + l], L.wa[2][k][l], za);
            }
            L.Z[mn][k] = z;
            L.Za[mn][k] = za;
        }
        __syncthreads();
        for (int e = lane; e < 4 * P * P; e += kWave) {
            const int m = e / (P * P), j = (e / P) % P, k = e % P;
            float y = 0.0f, ya = 0.0f;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                y = __builtin_fmaf(L.w[1][j][n], L.Z[4 * m + n][k], y);
                ya = __builtin_fmaf(L.wa[1][j][n], L.Za[4 * m + n][k], ya);
            }
            L.Y[m][j][k] = y;
            L.Ya[m][j][k] = ya;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kPer; t++) {
            const int p = lane + kWave * t;
            if (p < P * P * P) {
                const int i = p % P, j = (p / P) % P, k = p / (P * P);
                float v = 0.0f, va = 0.0f;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    v = __builtin_fmaf(L.w[0][i][m], L.Y[m][j][k], v);
                    va = __builtin_fmaf(L.wa[0][i][m], L.Ya[m][j][k], va);
                }
                sacc[t] += v;
                aacc[t] += va;
            }
        }
        __syncthreads();
    }
    const float iso = *A.iso;
    bool ok = true, pos = false, neg = false;
#pragma unroll
    for (int t = 0; t < kPer; t++) {
        const int p = lane + kWave * t;
        if (p >= P * P * P) continue;
        const int32_t g[3] = {B.org[0] + p % P, B.org[1] + (p / P) % P, B.org[2] + p / (P * P)};
        // the finer depths, where the owner's ancestor touches a root face
        // whose neighbours have children
        float U = 0.0f;
        for (int d = max(dlast + 1, B.r + 1); d <= D; d++) {
            const int slot = S.xb_off[d];
            if (slot < 0) continue;
            // the 27 coefficients the owner's ancestor sees at this depth
            const int sh = D - d, W = block_width(D, kb, d);
            int b0[3];
            for (int a = 0; a < 3; a++) b0[a] = (own_cell(g[a], B.R0[a]) >> sh) - 1 - ((B.org[a] >> sh) - 2);
            float sx = 0.0f;
            for (int k = 0; k < 27; k++)
                sx += fabsf(S.XB[slot + ((b0[0] + k / 9) * W + (b0[1] + (k / 3) % 3)) * W + (b0[2] + k % 3)]);
            const float bm = A.bmax[d];
            U += 1.0001f * sx * (bm * bm * bm);
        }
        const float v = sacc[t] - iso;
        const float margin = 420.0f * 5.9604645e-8f * (aacc[t] + U) + U + 1.2e-7f * (fabsf(sacc[t]) + fabsf(iso)) + 1e-15f;
        ok = ok && fabsf(v) > margin;
        pos = pos || v > 0;
        neg = neg || v < 0;
    }
    const bool lane_good = ok && !(pos && neg);
    const uint64_t good = __builtin_amdgcn_ballot_w64(lane_good);
    const uint64_t anypos = __builtin_amdgcn_ballot_w64(pos), anyneg = __builtin_amdgcn_ballot_w64(neg);
    return good == ~uint64_t(0) && !(anypos && anyneg);
}

// Pass 1: vertex and triangle counts of every brick.
__global__ void __launch_bounds__(kWave) k_brick_count(SubArgs A, int32_t* bnv, int32_t* bnt) {
    __shared__ BrickLds S;
    Brick B;
    const int32_t b = blockIdx.x;
    brick_setup(A, b, B, S, true);
    brick_blocks(A, B, S);
    if (brick_certify(A, B, S)) {
        if (threadIdx.x == 0) bnv[b] = bnt[b] = 0;
        return;
    }
    __syncthreads();
    brick_values(A, B, S);
    int nv = 0, nt = 0;
    const int L = 1 << (3 * B.kb);
    for (int l = threadIdx.x; l < L; l += kWave) {
        int32_t v[3];
        float cv[8];
        int cube, em;
        leaf_cube(B, S, A.D, l, v, cv, &cube, &em);
        nv += __popc(uint32_t(em));
        nt += tri_count(cube);
    }
    nv = wave_sum(nv);
    nt = wave_sum(nt);
    if (threadIdx.x == 0) {
        bnv[b] = nv;
        bnt[b] = nt;
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
                                                          float* out_v) {
    __shared__ BrickLds S;
    Brick B;
    const int32_t ai = blockIdx.x;
    const int32_t b = alist[ai];
    brick_setup(A, b, B, S, true);
    brick_blocks(A, B, S);
    brick_values(A, B, S);
    const int D = A.D;
    const int L = 1 << (3 * B.kb);
    const int per = (L + kWave - 1) / kWave;
    const int l0 = min(L, int(threadIdx.x) * per), l1 = min(L, l0 + per);
    int cnt = 0;
    for (int l = l0; l < l1; l++) {
        int32_t v[3];
        float cv[8];
        int cube, em;
        leaf_cube(B, S, D, l, v, cv, &cube, &em);
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
        leaf_cube(B, S, D, l, v, cv, &cube, &em2);
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
    __shared__ BrickLds S;
    Brick B;
    const int32_t ai = blockIdx.x;
    const int32_t b = alist[ai];
    brick_setup(A, b, B, S, false);
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
    const int rd = ensure_coarse_rows(c, s);
    if (rd < D - 1) throw Error{PRG_E_STATE, "coarse rows missing"};

    // pool slots 97..127
    PoolBuf<uint8_t> inter{&c->pool[97]};
    PoolBuf<int32_t> cnt{&c->pool[98]}, flag{&c->pool[99]}, off{&c->pool[100]}, cl{&c->pool[101]},
        oref{&c->pool[102]}, need{&c->pool[103]}, nid{&c->pool[104]}, clist{&c->pool[105]}, rflag{&c->pool[107]},
        roff{&c->pool[108]}, ridx{&c->pool[109]}, roots{&c->pool[110]}, rbricks{&c->pool[111]},
        boff{&c->pool[112]}, rleaves{&c->pool[113]}, loff{&c->pool[114]}, root_of{&c->pool[115]};
    PoolBuf<float> cv{&c->pool[106]};
    int32_t* C = cnt.reserve(kCntSlots);
    PRG_HIP(hipMemsetAsync(C, 0, kCntSlots * 4, s));

    // faces touched by the leaf-level surface
    inter.reserve(6 * M);
    PRG_HIP(hipMemsetAsync(inter.get(), 0, size_t(6 * M), s));
    {
        TimedScope ts(kt, "face_marks", 32.0 * double(MD));
        k_face_marks<<<grid_for(MD, kBlockS), kBlockS, 0, s>>>(c->corner_vals.get(), MD, baseD, c->key.get(),
                                                                c->neigh.get(), M, c->parent.get(), Lv, D, inter.get());
        PRG_CHECK_LAUNCH();
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
                                                                            c->neigh.get(), M, oref.get(), need.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(need.get(), nid.get(), 8 * baseD, C + kCntCorners, s, c->scratch, kt);
    k_compact<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(need.get(), nid.get(), 8 * baseD, clist.get());
    PRG_CHECK_LAUNCH();
    {
        TimedScope ts(kt, "coarse_corner_value", 24.0 * double(baseD));
        k_coarse_corner_value<<<grid_for(8 * baseD, kBlockS), kBlockS, 0, s>>>(
            clist.get(), C + kCntCorners, c->key.get(), c->pnt.get(), Lv, c->parent.get(), c->child_base.get(),
            c->x.get(), c->d_base_fn.get(), D, c->iso_dev.get(), cv.get());
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
                                                                  nid.get(), cv.get(), c->key.get(), c->neigh.get(), M,
                                                                  inter.get(), rflag.get());
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

    PoolBuf<float> btab{&c->pool[91]}, bounds{&c->pool[127]};
    float* XB = bounds.reserve(2 * (D + 1));
    k_depth_absmax<<<D + 1, 1024, 0, s>>>(c->x.get(), Lv, D, M, XB);
    PRG_CHECK_LAUNCH();
    k_table_absmax<<<D + 1, 1024, 0, s>>>(reinterpret_cast<const float4*>(btab.get()), D, XB + D + 1);
    PRG_CHECK_LAUNCH();
    const SubArgs A{c->key.get(),      c->neigh.get(),  M,
                    c->child_base.get(), c->parent.get(), c->pnt.get(),
                    Lv,                c->x.get(),      c->d_base_fn.get(),
                    reinterpret_cast<const float4*>(btab.get()), c->coarse_rows.get(), D,
                    c->iso_dev.get(),  roots.get(),     boff.get(),
                    root_of.get(),     C,               XB,
                    XB + D + 1};

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
    {
        TimedScope ts(kt, "subdivide_count", 4.0 * double(c->sub_leaves));
        k_brick_count<<<unsigned(nb), kWave, 0, s>>>(A, bnv.get(), bnt.get());
        PRG_CHECK_LAUNCH();
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
                                                          emask_tab.get(), cube_tab.get(), SV);
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
