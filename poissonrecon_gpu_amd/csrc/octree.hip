// Octree construction on the GPU (reference: pipelineBuildNodeArray,
// main.cu:511-841).
//
// The reference builds each coarser level with two 8M-slot hash tables and
// global atomics per depth.  Here every level comes from sorted keys alone:
//   1. Morton keys of the normalised points (encodePoint, main.cu:132-164) and
//      a stable radix sort of (key, index) - the order of thrust::sort_by_key
//      on key<<32|index (main.cu:596-603);
//   2. run heads + scan give the unique leaves (pidx = run start, pnum = run length);
//   3. bottom-up, the non-empty nodes of depth d are the blocks of depth d+1:
//      parent-boundary flags + scan give each non-empty node its block, so its
//      slot in the sibling-complete level array is 8*block + child slot; the
//      aggregates (pnum, dnum) are differences of consecutive run starts
//      because points and leaf-level slots are contiguous per subtree;
//   4. a per-slot pass fills empty slots (key, running pidx/didx, parent) as
//      updateEmptyNodeInfo (main.cu:433-490) intends;
//   5. neighbours top-down with the parent/child lookup of computeNeighbor
//      (main.cu:492-509).
#include <algorithm>
#include <cmath>

#include "ctx.h"
#include "pnt.h"

namespace prg {
namespace {

constexpr int kBlock = 256;

// LUTparent / LUTchild (main.cu:80-99), derived in build order: child slot s,
// neighbour j -> parent neighbour slot and child slot inside it.
struct NeighLut {
    int8_t parent[8][27];
    int8_t child[8][27];
};
__constant__ NeighLut c_lut;

NeighLut make_lut() {
    NeighLut l;
    for (int s = 0; s < 8; s++)
        for (int j = 0; j < 27; j++) {
            int cc[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
            int dd[3] = {j / 9 - 1, (j / 3) % 3 - 1, j % 3 - 1};
            int pj = 0, cs = 0;
            for (int a = 0; a < 3; a++) {
                int v = cc[a] + dd[a];
                int po = v < 0 ? -1 : (v > 1 ? 1 : 0);
                pj = pj * 3 + (po + 1);
                cs = (cs << 1) | (v - 2 * po);
            }
            l.parent[s][j] = int8_t(pj);
            l.child[s][j] = int8_t(cs);
        }
    return l;
}

// ---- normalisation (main.cu:530-571) -------------------------------------------
__global__ void __launch_bounds__(kBlock) k_bbox_partial(const float* xyz, int64_t n, float* part) {
    __shared__ float smin[3][kBlock], smax[3][kBlock];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        for (int a = 0; a < 3; a++) {
            float v = xyz[3 * i + a];
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    for (int a = 0; a < 3; a++) {
        smin[a][threadIdx.x] = mn[a];
        smax[a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int a = 0; a < 3; a++) {
                smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + o]);
                smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) {
            part[6 * blockIdx.x + a] = smin[a][0];
            part[6 * blockIdx.x + 3 + a] = smax[a][0];
        }
}

// Scale and centre from the block minima / maxima, in the reference's float
// operations (main.cu:539-545): tr = (centre x, y, z, scale).
__global__ void __launch_bounds__(1024) k_bbox_final(const float* part, int nb, float* tr) {
    // minima and maxima are exact in any order: one partial per thread, then a tree
    __shared__ float red[6][1024];
    const int tid = threadIdx.x;
    for (int a = 0; a < 6; a++) red[a][tid] = tid < nb ? part[6 * tid + a] : part[a];
    for (int b = tid + 1024; b < nb; b += 1024)
        for (int a = 0; a < 3; a++) {
            red[a][tid] = fminf(red[a][tid], part[6 * b + a]);
            red[3 + a][tid] = fmaxf(red[3 + a][tid], part[6 * b + 3 + a]);
        }
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (tid < o)
            for (int a = 0; a < 3; a++) {
                red[a][tid] = fminf(red[a][tid], red[a][tid + o]);
                red[3 + a][tid] = fmaxf(red[3 + a][tid], red[3 + a][tid + o]);
            }
        __syncthreads();
    }
    if (tid != 0) return;
    float mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = red[a][0];
        mx[a] = red[3 + a][0];
    }
    float scale = 1.0f, center[3];
    for (int a = 0; a < 3; a++) {
        const float ext = mx[a] - mn[a];
        if (!a || scale < ext) scale = ext;
        center[a] = (mx[a] + mn[a]) / 2;
    }
    scale = scale * 1.25f;
    for (int a = 0; a < 3; a++) tr[a] = center[a] - scale / 2;
    tr[3] = scale;
}

// one 32-byte record per point (x, y, z, nx, ny, nz, 0, 0), so the sort's
// permutation later gathers each point with two 16-byte loads
__global__ void __launch_bounds__(kBlock) k_normalize(const float* xyz, const float* nxyz, int64_t n,
                                                      const float* tr, float nscale, float4* stage) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const float cx = tr[0], cy = tr[1], cz = tr[2], scale = tr[3];
    const float px = (xyz[3 * i + 0] - cx) / scale;
    const float py = (xyz[3 * i + 1] - cy) / scale;
    const float pz = (xyz[3 * i + 2] - cz) / scale;
    const float a = nxyz[3 * i + 0], b = nxyz[3 * i + 1], c = nxyz[3 * i + 2];
    float sq = a * a;
    sq = sq + b * b;
    sq = sq + c * c;
    float len = float(sqrt(double(sq)));  // Length(): double sqrt of a float sum (Geometry.inl:29-33)
    if (len > 1e-6f) len = 1.0f / len;
    len = len * nscale;
    stage[2 * i] = make_float4(px, py, pz, a * len);
    stage[2 * i + 1] = make_float4(b * len, c * len, 0.0f, 0.0f);
}

// ---- keys and sort ----------------------------------------------------------------
__device__ inline uint32_t encode_point(float x, float y, float z, int D) {
    uint32_t key = 0;
    float c0 = 0.5f, c1 = 0.5f, c2 = 0.5f, w = 0.25f;
    for (int i = D - 1; i >= 0; --i) {
        if (x > c0) {
            key |= 1u << (3 * i + 2);
            c0 += w;
        } else {
            c0 -= w;
        }
        if (y > c1) {
            key |= 1u << (3 * i + 1);
            c1 += w;
        } else {
            c1 -= w;
        }
        if (z > c2) {
            key |= 1u << (3 * i);
            c2 += w;
        } else {
            c2 -= w;
        }
        w *= 0.5f;
    }
    return key;
}

__global__ void __launch_bounds__(kBlock) k_morton(const float4* stage, int64_t n, int D, uint32_t* key,
                                                   uint32_t* val) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = stage[2 * i];
    key[i] = encode_point(p.x, p.y, p.z, D);
    val[i] = uint32_t(i);
}

// point r of the output is staged record perm[r] (record r without a permutation)
__global__ void __launch_bounds__(kBlock) k_gather_points(const uint32_t* perm, int64_t n, const float4* stage,
                                                          float* px, float* py, float* pz, float* nx, float* ny,
                                                          float* nz) {
    const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t i = perm ? int64_t(perm[r]) : r;
    const float4 a = stage[2 * i], b = stage[2 * i + 1];
    px[r] = a.x;
    py[r] = a.y;
    pz[r] = a.z;
    nx[r] = a.w;
    ny[r] = b.x;
    nz[r] = b.y;
}

// run heads of the sorted keys (generateMark, main.cu:195-211)
__global__ void __launch_bounds__(kBlock) k_heads(const uint32_t* key, int64_t n, int shift, int32_t* flag) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || (key[i] >> shift) != (key[i - 1] >> shift)) ? 1 : 0;
}

// the same over the first *count of ub entries (0 beyond)
__global__ void __launch_bounds__(kBlock) k_unique(const uint32_t* key, const int32_t* flag, const int32_t* rank,
                                                   int64_t n, uint32_t* uk, int32_t* upidx) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n || !flag[i]) return;
    uk[rank[i]] = key[i];
    upidx[rank[i]] = int32_t(i);
}

// exclusive head count -> index of the run containing each element
__global__ void __launch_bounds__(kBlock) k_excl_to_incl(int32_t* out, const int32_t* excl, const int32_t* flag,
                                                         int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) out[i] = excl[i] + flag[i] - 1;
}

__device__ __forceinline__ void level_init_slot(int64_t base, int64_t s, bool leaf, int32_t* pnum, int32_t* dnum,
                                                int32_t* child_base) {
    pnum[base + s] = 0;
    dnum[base + s] = leaf ? 1 : 0;
    child_base[base + s] = -1;
}

__global__ void __launch_bounds__(kBlock) k_level_init(int64_t base, int64_t count, bool leaf, int32_t* pnum,
                                                       int32_t* dnum, int32_t* child_base) {
    const int64_t s = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (s >= count) return;
    level_init_slot(base, s, leaf, pnum, dnum, child_base);
}

// Scatter the non-empty nodes of depth d into their slots.
__device__ __forceinline__ void level_scatter_node(int64_t q, const uint32_t* nk, const int32_t* np,
                                                   const int32_t* ndi, const int32_t* pos, int64_t ne,
                                                   int64_t npoints, int64_t nleaf_slots, bool leaf, int64_t base,
                                                   int64_t child_level_base, uint32_t* key, int32_t* pidx,
                                                   int32_t* pnum, int32_t* didx, int32_t* dnum,
                                                   int32_t* child_base) {
    const int64_t g = base + pos[q];
    key[g] = nk[q];
    pidx[g] = np[q];
    pnum[g] = int32_t(((q + 1 < ne) ? int64_t(np[q + 1]) : npoints) - np[q]);
    if (leaf) {
        didx[g] = pos[q];
        dnum[g] = 1;
    } else {
        didx[g] = ndi[q];
        dnum[g] = int32_t(((q + 1 < ne) ? int64_t(ndi[q + 1]) : nleaf_slots) - ndi[q]);
        child_base[g] = int32_t(child_level_base + 8 * q);
    }
}

__global__ void __launch_bounds__(kBlock) k_level_scatter(const uint32_t* nk, const int32_t* np, const int32_t* ndi,
                                                          const int32_t* pos, int64_t ne, int64_t npoints,
                                                          int64_t nleaf_slots, bool leaf, int64_t base,
                                                          int64_t child_level_base, uint32_t* key, int32_t* pidx,
                                                          int32_t* pnum, int32_t* didx, int32_t* dnum,
                                                          int32_t* child_base) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= ne) return;
    level_scatter_node(q, nk, np, ndi, pos, ne, npoints, nleaf_slots, leaf, base, child_level_base, key, pidx, pnum,
                       didx, dnum, child_base);
}

// Every slot of depth d: key, running pidx / didx inside the sibling block and
// the parent index (updateEmptyNodeInfo's intent, main.cu:433-490).
__device__ __forceinline__ void level_fill_slot(int64_t s, const uint32_t* pk, const int32_t* pp,
                                                const int32_t* pdi, const int32_t* ppos, int D, int d, int64_t base,
                                                int64_t parent_base, uint32_t* key, int32_t* pidx,
                                                const int32_t* pnum, int32_t* didx, const int32_t* dnum,
                                                int32_t* parent) {
    const int64_t b = s >> 3;
    const int j = int(s & 7);
    const int64_t g = base + s;
    key[g] = pk[b] | (uint32_t(j) << (3 * (D - d)));
    int32_t run_p = pp[b], run_d = pdi[b];
    for (int k = 0; k < j; k++) {
        run_p += pnum[base + 8 * b + k];
        run_d += dnum[base + 8 * b + k];
    }
    pidx[g] = run_p;
    didx[g] = (d != D) ? run_d : int32_t(s);  // leaf level: every slot is its own first leaf slot
    parent[g] = int32_t(parent_base + ppos[b]);
}

__global__ void __launch_bounds__(kBlock) k_level_fill(const uint32_t* pk, const int32_t* pp, const int32_t* pdi,
                                                       const int32_t* ppos, int64_t count, int D, int d, int64_t base,
                                                       int64_t parent_base, uint32_t* key, int32_t* pidx,
                                                       const int32_t* pnum, int32_t* didx, const int32_t* dnum,
                                                       int32_t* parent) {
    const int64_t s = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (s >= count) return;
    level_fill_slot(s, pk, pp, pdi, ppos, D, d, base, parent_base, key, pidx, pnum, didx, dnum, parent);
}

__global__ void __launch_bounds__(kBlock) k_point_leaf(const int32_t* rank, const int32_t* pos_leaf, int64_t n,
                                                       int32_t* point_leaf) {
    const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= n) return;
    point_leaf[r] = pos_leaf[rank[r]];
}

// computeNeighbor (main.cu:492-509): neighbour j of a depth-d node is child
// LUTchild[s][j] of its parent's neighbour LUTparent[s][j], when that exists.
__global__ void __launch_bounds__(kBlock) k_neighbors(const uint32_t* key, const int32_t* parent,
                                                      const int32_t* child_base, int32_t* neigh, int64_t M,
                                                      int64_t base, int64_t count, int D, int d) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= count) return;
    const int64_t i = base + t;
    const int s = int((key[i] >> (3 * (D - d))) & 7u);
    const int64_t p = parent[i];
    // The parent's rows (depth d - 1) and the rows written here (depth d) are
    // disjoint: all loads are issued before the first store, so the 27
    // lookups overlap instead of waiting on each other.
    int32_t np[27], cb[27];
#pragma unroll
    for (int j = 0; j < 27; j++) np[j] = neigh[int64_t(c_lut.parent[s][j]) * M + p];
#pragma unroll
    for (int j = 0; j < 27; j++) cb[j] = np[j] != -1 ? child_base[np[j]] : -1;
#pragma unroll
    for (int j = 0; j < 27; j++) neigh[int64_t(j) * M + i] = cb[j] != -1 ? cb[j] + c_lut.child[s][j] : -1;
}

// Parent-neighbour table of one depth: for sibling block b, the level-local
// child base of each of the parent's 27 neighbours (-1 where the neighbour
// does not exist or is empty).  Neighbour j of slot s in the block is
// pnt[b][LUTparent[s][j]] + LUTchild[s][j] - the same lookup computeNeighbor
// performs (main.cu:492-509), kept per block so the solver reads 13.5 bytes
// per node instead of 108.  Built top-down: the parent's neighbours come from
// the parent depth's own table (the root's only neighbour is itself,
// main.cu:803-808), so no 27-plane neighbour table is needed.  One thread per
// block.
__device__ __forceinline__ void pnt_block(int64_t b, const int32_t* parent, const int32_t* child_base,
                                          const int32_t* pnt_prev, int64_t base_prev, int64_t base, int32_t* pnt) {
    const int64_t tp = parent[base + 8 * b] - base_prev;
    int32_t nbp[27];
    if (!pnt_prev) {
#pragma unroll
        for (int j = 0; j < 27; j++) nbp[j] = (j == 13) ? 0 : -1;
    } else {
        pnt_neighbours(pnt_prev + 27 * (tp >> 3), int(tp & 7), nbp);
    }
    int32_t cb[27];
#pragma unroll
    for (int j = 0; j < 27; j++) cb[j] = nbp[j] < 0 ? -1 : child_base[base_prev + nbp[j]];
    int32_t* row = pnt + 27 * b;
#pragma unroll
    for (int j = 0; j < 27; j++) row[j] = cb[j] == -1 ? -1 : int32_t(cb[j] - base);
}

__global__ void __launch_bounds__(kBlock) k_parent_neighbor_table(const int32_t* parent, const int32_t* child_base,
                                                                  const int32_t* pnt_prev, int64_t base_prev,
                                                                  int64_t base, int64_t nblocks, int32_t* pnt) {
    const int64_t b = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (b >= nblocks) return;
    pnt_block(b, parent, child_base, pnt_prev, base_prev, base, pnt);
}

// ---- the bottom-up step of a large depth in three launches ---------------------------
// The non-empty nodes of depth d (compact, key order) give their slot in the
// depth-d level array and the compact arrays of depth d - 1: the parent-run
// heads are counted per tile, the tile counts scanned (their total is the
// parent depth's non-empty count), and the apply pass recomputes the heads,
// scans them inside the tile and writes each node's slot and, per head, its
// parent's key, first point and first leaf-level slot.
constexpr int kLvlBlock = 256;
constexpr int kLvlItems = 8;
constexpr int kLvlTile = kLvlBlock * kLvlItems;

__device__ __forceinline__ int32_t head_flag(const uint32_t* nk, int64_t i, int64_t n, int hs) {
    return (i < n && (i == 0 || (nk[i] >> hs) != (nk[i - 1] >> hs))) ? 1 : 0;
}

// exclusive scan of one value per thread over the workgroup, and its total
__device__ __forceinline__ int32_t lvl_block_scan(int32_t v, int32_t* lds, int32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int32_t before = 0, tot = 0;
    for (int w = 0; w < int(blockDim.x >> 6); w++) {
        const int32_t t = lds[w];
        if (w < wave) before += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

__global__ void __launch_bounds__(kLvlBlock) k_lvl_reduce(const uint32_t* nk, const int32_t* ne, int hs,
                                                          int32_t* sums) {
    __shared__ int32_t lds[kLvlBlock / 64];
    const int64_t n = *ne, base = int64_t(blockIdx.x) * kLvlTile;
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kLvlItems; k++) acc += head_flag(nk, base + int64_t(k) * kLvlBlock + threadIdx.x, n, hs);
    int32_t tot;
    (void)lvl_block_scan(acc, lds, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of the tile counts in place, total to *ne_parent
__global__ void __launch_bounds__(1024) k_lvl_sums(int32_t* sums, int64_t tiles, int32_t* ne_parent) {
    __shared__ int32_t lds[16];
    int32_t carry = 0;
    for (int64_t base = 0; base < tiles; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int32_t v = i < tiles ? sums[i] : 0;
        int32_t tot;
        const int32_t ex = lvl_block_scan(v, lds, tot);
        if (i < tiles) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *ne_parent = carry;
}

__global__ void __launch_bounds__(kLvlBlock) k_lvl_apply(const uint32_t* nk, const int32_t* np, const int32_t* ndi,
                                                         const int32_t* ne, int hs, int D, int d, const int32_t* sums,
                                                         int32_t* pos, uint32_t* pk, int32_t* pp, int32_t* pdi) {
    __shared__ int32_t lds[kLvlBlock / 64];
    const int64_t n = *ne, base = int64_t(blockIdx.x) * kLvlTile + int64_t(threadIdx.x) * kLvlItems;
    const int sh = 3 * (D - d);
    int32_t f[kLvlItems];
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kLvlItems; k++) {
        f[k] = head_flag(nk, base + k, n, hs);
        acc += f[k];
    }
    int32_t tot;
    int32_t ex = lvl_block_scan(acc, lds, tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kLvlItems; k++) {
        const int64_t i = base + k;
        if (i < n) {
            const int32_t b = ex + f[k] - 1;  // parent block: heads up to i, inclusive, minus one
            const uint32_t key = nk[i];
            pos[i] = 8 * b + int32_t((key >> sh) & 7u);
            if (f[k]) {
                pk[b] = key & ~(7u << sh);
                pp[b] = np[i];
                pdi[b] = (d == D) ? 8 * b : ndi[i];
            }
        }
        ex += f[k];
    }
}

// ---- small depths in one workgroup -------------------------------------------------
// The coarse depths hold few nodes; one 1024-thread workgroup runs their whole
// per-depth sequence in a single launch instead of several launches per depth.
constexpr int kSmallThreads = 1024;
constexpr int64_t kSmallLevelSlots = 4096;  // slots of a depth built by k_levels_small
constexpr int64_t kSmallLevelUp = 32768;     // bound on the non-empty nodes of a depth k_levels_up_small takes

struct LevelArrays {
    uint32_t* nk[kMaxDepth + 1];  // compact non-empty nodes of each depth: keys,
    int32_t* np[kMaxDepth + 1];   // first sorted point,
    int32_t* ndi[kMaxDepth + 1];  // first leaf-level slot,
    int32_t* pos[kMaxDepth + 1];  // and slot in the depth's level array
};

// The bottom-up step (k_lvl_reduce, k_lvl_sums, k_lvl_apply) for depths
// ds .. 1: per depth, chunks of the compact nodes with a workgroup scan of
// the parent-run heads carried across chunks.
__global__ void __launch_bounds__(kSmallThreads) k_levels_up_small(LevelArrays A, int32_t* ne, int ds, int D) {
    __shared__ int32_t sc[kSmallThreads];
    __shared__ int32_t carry;
    const int tid = threadIdx.x;
    for (int d = ds; d >= 1; --d) {
        const int64_t n = ne[d];
        const int hs = 3 * (D - d + 1), sh = 3 * (D - d);
        const uint32_t* nk = A.nk[d];
        if (tid == 0) carry = 0;
        __syncthreads();
        for (int64_t c0 = 0; c0 < n; c0 += kSmallThreads) {
            const int64_t i = c0 + tid;
            int32_t f = 0;
            uint32_t k = 0;
            if (i < n) {
                k = nk[i];
                f = (i == 0 || (k >> hs) != (nk[i - 1] >> hs)) ? 1 : 0;
            }
            sc[tid] = f;
            __syncthreads();
            for (int o = 1; o < kSmallThreads; o <<= 1) {
                const int32_t v = tid >= o ? sc[tid - o] : 0;
                __syncthreads();
                sc[tid] += v;
                __syncthreads();
            }
            const int32_t b = carry + sc[tid] - 1;  // parent block of node i
            if (i < n) {
                A.pos[d][i] = 8 * b + int32_t((k >> sh) & 7u);
                if (f) {
                    A.nk[d - 1][b] = k & ~(7u << sh);
                    A.np[d - 1][b] = A.np[d][i];
                    A.ndi[d - 1][b] = (d == D) ? 8 * b : A.ndi[d][i];
                }
            }
            __syncthreads();
            if (tid == kSmallThreads - 1) carry = b + 1;
            __syncthreads();
        }
        if (tid == 0) ne[d - 1] = carry;
        __syncthreads();
    }
}

struct SmallLevels {
    int64_t base[kMaxDepth + 1], count[kMaxDepth + 1], ne[kMaxDepth + 1], pnt_off[kMaxDepth + 2];
};

// Root, then for depths 1 .. df: slot defaults and non-empty nodes (k_level_init,
// k_level_scatter), every slot's fields (k_level_fill) and the parent-neighbour
// rows (k_parent_neighbor_table, top-down).
__global__ void __launch_bounds__(kSmallThreads) k_levels_small(LevelArrays A, SmallLevels S, int df, int D,
                                                                int64_t M, int64_t npoints, int64_t MD,
                                                                uint32_t* key, int32_t* pidx, int32_t* pnum,
                                                                int32_t* parent, int32_t* child_base, int32_t* didx,
                                                                int32_t* dnum, int32_t* pnt) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        key[0] = 0;
        pidx[0] = 0;
        pnum[0] = int32_t(npoints);
        parent[0] = -1;
        child_base[0] = int32_t(S.base[1]);
        didx[0] = 0;
        dnum[0] = int32_t(MD);
    }
    for (int d = 1; d <= df; d++)
        for (int64_t s = tid; s < S.count[d]; s += kSmallThreads) level_init_slot(S.base[d], s, d == D, pnum, dnum, child_base);
    __syncthreads();
    for (int d = 1; d <= df; d++)
        for (int64_t q = tid; q < S.ne[d]; q += kSmallThreads)
            level_scatter_node(q, A.nk[d], A.np[d], A.ndi[d], A.pos[d], S.ne[d], npoints, MD, d == D, S.base[d],
                               d == D ? 0 : S.base[d + 1], key, pidx, pnum, didx, dnum, child_base);
    __syncthreads();
    for (int d = 1; d <= df; d++)
        for (int64_t s = tid; s < S.count[d]; s += kSmallThreads)
            level_fill_slot(s, A.nk[d - 1], A.np[d - 1], A.ndi[d - 1], A.pos[d - 1], D, d, S.base[d], S.base[d - 1],
                            key, pidx, pnum, didx, dnum, parent);
    __syncthreads();
    for (int d = 1; d <= df; d++) {
        for (int64_t b = tid; b < S.count[d] / 8; b += kSmallThreads)
            pnt_block(b, parent, child_base, d == 1 ? nullptr : pnt + 27 * S.pnt_off[d - 1], S.base[d - 1],
                      S.base[d], pnt + 27 * S.pnt_off[d]);
        __syncthreads();
    }
}

}  // namespace

void normalize_points(prg_ctx* c, const float* xyz, const float* nxyz, int64_t n, bool on_device) {
    hipStream_t s = c->stream;
    PoolBuf<float> hxyz{&c->pool[0]}, hn{&c->pool[1]};
    const float* dxyz = xyz;
    const float* dn = nxyz;
    if (!on_device) {
        PRG_HIP(hipMemcpyAsync(hxyz.reserve(size_t(3 * n)), xyz, size_t(3 * n) * 4, hipMemcpyHostToDevice, s));
        PRG_HIP(hipMemcpyAsync(hn.reserve(size_t(3 * n)), nxyz, size_t(3 * n) * 4, hipMemcpyHostToDevice, s));
        dxyz = hxyz.get();
        dn = hn.get();
    }
    const int blocks = int(std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
    PoolBuf<float> part{&c->pool[2]};
    part.reserve(size_t(6 * blocks));
    {
        TimedScope ts(c->kt, "bbox_partial", 12.0 * double(n));
        k_bbox_partial<<<blocks, kBlock, 0, s>>>(dxyz, n, part.get());
        PRG_CHECK_LAUNCH();
    }
    {
        k_bbox_final<<<1, 1024, 0, s>>>(part.get(), blocks, c->transform_dev.reserve(4));
        PRG_CHECK_LAUNCH();
    }
    c->N = n;
    {
        TimedScope ts(c->kt, "normalize", 48.0 * double(n));
        k_normalize<<<grid_for(n, kBlock), kBlock, 0, s>>>(dxyz, dn, n, c->transform_dev.get(), float(2 << c->D),
                                                           reinterpret_cast<float4*>(c->pstage.reserve(8 * n)));
        PRG_CHECK_LAUNCH();
    }
    c->points_sorted = false;
    // the host copy of (centre, scale) arrives with the next synchronisation
    PRG_HIP(hipMemcpyAsync(c->transform_host, c->transform_dev.get(), 16, hipMemcpyDeviceToHost, s));
    c->transform_pending = true;
    if (!on_device) PRG_HIP(hipStreamSynchronize(s));  // the staging copies must finish before the caller reuses its arrays
}

void unpack_points(prg_ctx* c) {
    if (c->points_sorted) return;
    const int64_t n = c->N;
    k_gather_points<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(
        nullptr, n, reinterpret_cast<const float4*>(c->pstage.get()), c->px.reserve(n), c->py.reserve(n),
        c->pz.reserve(n), c->nx.reserve(n), c->ny.reserve(n), c->nz.reserve(n));
    PRG_CHECK_LAUNCH();
}

// The 27-plane neighbour table (plane j holds neighbour j of every node), for
// callers that read it (PRG_BUF_NEIGHS and its view); the pipeline's own
// kernels take neighbours from the parent-neighbour rows.  The root's only
// neighbour is itself (main.cu:803-808).
namespace {
__global__ void k_neighbors_root(int32_t* neigh, int64_t M) {
    const int j = threadIdx.x;
    if (j < 27) neigh[int64_t(j) * M] = (j == 13) ? 0 : -1;
}
}  // namespace

void ensure_neighbours(prg_ctx* c) {
    if (c->neigh_ready) return;
    c->neigh.reserve(size_t(27) * c->M);
    k_neighbors_root<<<1, 64, 0, c->stream>>>(c->neigh.get(), c->M);
    PRG_CHECK_LAUNCH();
    for (int d = 1; d <= c->D; d++) {
        TimedScope ts(c->kt, "neighbors", 128.0 * double(c->counts[d]));
        k_neighbors<<<grid_for(c->counts[d], kBlock), kBlock, 0, c->stream>>>(
            c->key.get(), c->parent.get(), c->child_base.get(), c->neigh.get(), c->M, c->bases[d], c->counts[d], c->D, d);
        PRG_CHECK_LAUNCH();
    }
    c->neigh_ready = true;
}

void build_octree(prg_ctx* c) {
    static bool lut_ready = false;
    if (!lut_ready) {
        NeighLut l = make_lut();
        PRG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_lut), &l, sizeof(l)));
        lut_ready = true;
    }
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t n = c->N;
    KernelTimer& kt = c->kt;

    // 1. keys + stable sort
    PoolBuf<uint32_t> key{&c->pool[3]}, val{&c->pool[4]}, key2{&c->pool[5]}, val2{&c->pool[6]};
    key.reserve(n);
    val.reserve(n);
    key2.reserve(n);
    val2.reserve(n);
    {
        TimedScope ts(kt, "morton", 20.0 * double(n));
        k_morton<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<const float4*>(c->pstage.get()), n, D,
                                                         key.get(), val.get());
        PRG_CHECK_LAUNCH();
    }
    radix_sort_pairs(key.get(), val.get(), key2.get(), val2.get(), n, 3 * D, s, c->scratch, kt);
    {
        TimedScope ts(kt, "gather_points", 52.0 * double(n));
        k_gather_points<<<grid_for(n, kBlock), kBlock, 0, s>>>(
            val.get(), n, reinterpret_cast<const float4*>(c->pstage.get()), c->px.reserve(n), c->py.reserve(n),
            c->pz.reserve(n), c->nx.reserve(n), c->ny.reserve(n), c->nz.reserve(n));
        PRG_CHECK_LAUNCH();
    }
    c->points_sorted = true;

    // 2. unique leaves
    PoolBuf<int32_t> flag{&c->pool[7]}, rank{&c->pool[8]}, tot{&c->pool[9]};
    flag.reserve(n);
    rank.reserve(n);
    tot.reserve(1);
    {
        TimedScope ts(kt, "heads", 12.0 * double(n));
        k_heads<<<grid_for(n, kBlock), kBlock, 0, s>>>(key.get(), n, 0, flag.get());
        PRG_CHECK_LAUNCH();
    }
    PoolBuf<int32_t> ne_dev{&c->pool[14]};  // non-empty nodes per depth
    ne_dev.reserve(D + 1);
    scan_exclusive(flag.get(), rank.get(), n, ne_dev.get() + D, s, c->scratch, kt);
    // upper bounds of the non-empty node counts, min(n, 8^d): the unique
    // leaves number at most n, and the counts stay on the device until the
    // levels are built
    const int64_t U = std::max<int64_t>(1, n);
    std::vector<int64_t> ub(D + 1);
    for (int d = 0; d <= D; d++) ub[d] = std::min<int64_t>(U, int64_t(1) << (3 * d));

    // per-depth compact arrays of non-empty nodes
    std::vector<std::unique_ptr<PoolBuf<uint32_t>>> NK(D + 1);
    std::vector<std::unique_ptr<PoolBuf<int32_t>>> NP(D + 1), NDI(D + 1), POS(D + 1);
    std::vector<int64_t> NE(D + 1, 0);
    for (int d = 0; d <= D; d++) {
        NK[d].reset(new PoolBuf<uint32_t>{&c->pool[16 + 4 * d]});
        NP[d].reset(new PoolBuf<int32_t>{&c->pool[17 + 4 * d]});
        NDI[d].reset(new PoolBuf<int32_t>{&c->pool[18 + 4 * d]});
        POS[d].reset(new PoolBuf<int32_t>{&c->pool[19 + 4 * d]});
    }
    NE[D] = U;
    NK[D]->reserve(U);
    NP[D]->reserve(U);
    NDI[D]->reserve(U);
    {
        TimedScope ts(kt, "unique", 16.0 * double(n));
        k_unique<<<grid_for(n, kBlock), kBlock, 0, s>>>(key.get(), flag.get(), rank.get(), n, NK[D]->get(),
                                                         NP[D]->get());
        PRG_CHECK_LAUNCH();
    }
    PoolBuf<int32_t> prank{&c->pool[10]};  // unique-leaf rank of every sorted point
    prank.reserve(n);

    // 3. levels bottom-up: sizes stay on the device (upper bounds size the
    // launches); one read-back after the loop
    std::vector<int64_t> counts(D + 1, 0);
    // depths ds .. 1 (at most kSmallLevelUp non-empty nodes each) in one launch
    int ds = 0;
    while (ds < D && ub[ds + 1] <= kSmallLevelUp) ds++;
    PoolBuf<int32_t> lsums{&c->pool[95]};
    for (int d = D; d > ds; --d) {
        NK[d - 1]->reserve(ub[d - 1]);
        NP[d - 1]->reserve(ub[d - 1]);
        NDI[d - 1]->reserve(ub[d - 1]);
        POS[d]->reserve(ub[d]);
        const int64_t tiles = (ub[d] + kLvlTile - 1) / kLvlTile;
        int32_t* sums = lsums.reserve(tiles);
        const int hs = 3 * (D - d + 1);
        TimedScope ts(kt, "level_up", 28.0 * double(ub[d]));
        k_lvl_reduce<<<unsigned(tiles), kLvlBlock, 0, s>>>(NK[d]->get(), ne_dev.get() + d, hs, sums);
        PRG_CHECK_LAUNCH();
        k_lvl_sums<<<1, 1024, 0, s>>>(sums, tiles, ne_dev.get() + d - 1);
        PRG_CHECK_LAUNCH();
        k_lvl_apply<<<unsigned(tiles), kLvlBlock, 0, s>>>(NK[d]->get(), NP[d]->get(), NDI[d]->get(),
                                                           ne_dev.get() + d, hs, D, d, sums, POS[d]->get(),
                                                           NK[d - 1]->get(), NP[d - 1]->get(), NDI[d - 1]->get());
        PRG_CHECK_LAUNCH();
    }
    LevelArrays LA{};
    for (int d = ds; d >= 1; --d) {
        NK[d - 1]->reserve(ub[d - 1]);
        NP[d - 1]->reserve(ub[d - 1]);
        NDI[d - 1]->reserve(ub[d - 1]);
        POS[d]->reserve(ub[d]);
    }
    POS[0]->reserve(1);
    for (int d = 0; d <= D; d++) {
        LA.nk[d] = NK[d]->get();
        LA.np[d] = NP[d]->get();
        LA.ndi[d] = NDI[d]->get();
        LA.pos[d] = POS[d]->get();
    }
    if (ds >= 1) {
        TimedScope ts(kt, "level_up_small", 28.0 * double(ub[ds]));
        k_levels_up_small<<<1, kSmallThreads, 0, s>>>(LA, ne_dev.get(), ds, D);
        PRG_CHECK_LAUNCH();
    }
    {
        std::vector<int32_t> hne(D + 1);
        PRG_HIP(hipMemcpyAsync(hne.data(), ne_dev.get(), size_t(D + 1) * 4, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
        for (int d = 0; d <= D; d++) NE[d] = hne[d];
        if (c->transform_pending) {
            for (int a = 0; a < 3; a++) c->center[a] = c->transform_host[a];
            c->scale = c->transform_host[3];
            c->transform_pending = false;
        }
        for (int d = 1; d <= D; d++) counts[d] = 8 * NE[d - 1];
    }
    counts[0] = 1;
    // root compact arrays: key 0, pidx 0, didx 0, slot 0
    PRG_HIP(hipMemsetAsync(POS[0]->get(), 0, 4, s));

    std::vector<int64_t> bases(D + 1, 0);
    for (int d = 1; d <= D; d++) bases[d] = bases[d - 1] + counts[d - 1];
    const int64_t M = bases[D] + counts[D];
    if (M >= (int64_t(1) << 31)) throw Error{PRG_E_ARG, "octree exceeds 2^31 nodes"};
    c->counts = counts;
    c->bases = bases;
    c->M = M;
    uint32_t* K = c->key.reserve(M);
    int32_t* PI = c->pidx.reserve(M);
    int32_t* PN = c->pnum.reserve(M);
    int32_t* PA = c->parent.reserve(M);
    int32_t* CB = c->child_base.reserve(M);
    int32_t* DI = c->didx.reserve(M);
    int32_t* DN = c->dnum.reserve(M);
    const int64_t MD = counts[D];
    c->pnt_off.assign(D + 2, 0);
    for (int d = 1; d <= D; d++) c->pnt_off[d + 1] = c->pnt_off[d] + counts[d] / 8;
    int32_t* PT = c->pnt.reserve(size_t(27) * std::max<int64_t>(1, c->pnt_off[D + 1]));
    // depths 1 .. df (at most kSmallLevelSlots slots each): root, levels and
    // parent-neighbour rows in one launch
    int df = 0;
    while (df < D && counts[df + 1] <= kSmallLevelSlots) df++;
    {
        SmallLevels SL{};
        for (int d = 0; d <= D; d++) {
            SL.base[d] = bases[d];
            SL.count[d] = counts[d];
            SL.ne[d] = NE[d];
        }
        for (int d = 0; d <= D + 1; d++) SL.pnt_off[d] = c->pnt_off[d];
        TimedScope ts(kt, "levels_small", 44.0 * double(bases[df] + counts[df]));
        k_levels_small<<<1, kSmallThreads, 0, s>>>(LA, SL, df, D, M, n, MD, K, PI, PN, PA, CB, DI, DN, PT);
        PRG_CHECK_LAUNCH();
    }
    {
        TimedScope ts(kt, "levels_fill", 44.0 * double(M));
        for (int d = df + 1; d <= D; d++) {
            const bool leaf = (d == D);
            k_level_init<<<grid_for(counts[d], kBlock), kBlock, 0, s>>>(bases[d], counts[d], leaf, PN, DN, CB);
            PRG_CHECK_LAUNCH();
            k_level_scatter<<<grid_for(NE[d], kBlock), kBlock, 0, s>>>(
                NK[d]->get(), NP[d]->get(), NDI[d]->get(), POS[d]->get(), NE[d], n, MD, leaf, bases[d],
                leaf ? 0 : bases[d + 1], K, PI, PN, DI, DN, CB);
            PRG_CHECK_LAUNCH();
        }
        for (int d = df + 1; d <= D; d++) {
            k_level_fill<<<grid_for(counts[d], kBlock), kBlock, 0, s>>>(
                NK[d - 1]->get(), NP[d - 1]->get(), NDI[d - 1]->get(), POS[d - 1]->get(), counts[d], D, d,
                bases[d], bases[d - 1], K, PI, PN, DI, DN, PA);
            PRG_CHECK_LAUNCH();
        }
    }
    // point -> leaf-level slot: rank of the run containing each sorted point
    {
        TimedScope ts(kt, "point_leaf", 12.0 * double(n));
        k_excl_to_incl<<<grid_for(n, kBlock), kBlock, 0, s>>>(prank.get(), rank.get(), flag.get(), n);
        PRG_CHECK_LAUNCH();
        k_point_leaf<<<grid_for(n, kBlock), kBlock, 0, s>>>(prank.get(), POS[D]->get(), n, c->point_leaf.reserve(n));
        PRG_CHECK_LAUNCH();
    }
    // 4. parent-neighbour tables (depths 1..D), top-down; the 27-plane
    // neighbour table (depths 1..D) is built on demand
    c->neigh_ready = false;
    for (int d = df + 1; d <= D; d++) {
        const int64_t nb = counts[d] / 8;
        TimedScope ts(kt, "parent_neighbor_table", 120.0 * double(nb));
        k_parent_neighbor_table<<<grid_for(nb, kBlock), kBlock, 0, s>>>(
            PA, CB, d == 1 ? nullptr : PT + 27 * c->pnt_off[d - 1], bases[d - 1], bases[d], nb, PT + 27 * c->pnt_off[d]);
        PRG_CHECK_LAUNCH();
    }
}

}  // namespace prg
