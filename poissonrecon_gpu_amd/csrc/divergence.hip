// Vector field and divergence (reference: computeVectorField main.cu:913-964,
// computeEncodedFinerNodesDivergence :1007-1060, the per-node coarse loop
// :1062-1141 / :3417-3459).
//
// Vector field: one thread per leaf-level slot gathers the points of its 27
// neighbours; each (slot, point) weight shifts the depth-D B-spline by the
// point coordinate and evaluates it at the slot centre, in the reference's
// float order (ConfirmedPPolynomial.cuh:37-91).
//
// Divergence b_i = sum over the leaf-level slots k of i's 27 neighbours'
// subtrees of float(V_k . (w(ux), w(uy), w(uz))), accumulated in double.  The
// reference runs one thread per node (finer depths) and a host loop with
// three cudaMallocs per node (coarser depths).  Here:
//   - leaf level: one thread per node, 27 terms;
//   - coarser depths: per-node subtree sums G[a][j] in one streaming pass
//     (wave-level reductions over chunks of a node's leaves, chunk partials
//     added in order), then b_i gathers 27 of them.  No global atomics; the
//     result is deterministic.
#include <cmath>
#include <map>
#include <mutex>

#include <algorithm>

#include "ctx.h"
#include "pnt.h"
#include "tile.h"

namespace prg {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 2048;  // terms per wave-chunk
// listed blocks per tile the LDS-staged leaf kernels hold (a larger tile takes
// their per-row path)
constexpr int kTileBlocks = 320;

struct BaseMD {
    float v[16];  // 4 pieces x (start, c0, c1, c2)
};
struct Bases {
    int64_t b[kMaxDepth + 2];
};

// Bits 3l (l = 0..9) of x gathered into bits l.
__device__ __forceinline__ uint32_t compact3(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

// value(BaseFunctionMaxDepth.shift(t), x) (ConfirmedPPolynomial.cuh:37-51,78-91)
__device__ __forceinline__ float shifted_value(const BaseMD& bm, float t, float x) {
    float res = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float start = bm.v[4 * i] + t;
        if (!(x > start)) break;
        float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
        const float c0 = bm.v[4 * i + 1], c1 = bm.v[4 * i + 2], c2 = bm.v[4 * i + 3];
        // k = 0
        q0 = q0 + c0 * 1.0f;
        // k = 1: j = 1, 0
        {
            float temp = 1.0f;
            q1 = q1 + c1 * temp;
            temp = temp * ((-t) * 1.0f);
            temp = temp / 1.0f;
            q0 = q0 + c1 * temp;
        }
        // k = 2: j = 2, 1, 0
        {
            float temp = 1.0f;
            q2 = q2 + c2 * temp;
            temp = temp * ((-t) * 2.0f);
            temp = temp / 1.0f;
            q1 = q1 + c2 * temp;
            temp = temp * ((-t) * 1.0f);
            temp = temp / 2.0f;
            q0 = q0 + c2 * temp;
        }
        // evaluation of the piece: v = 1*q0 + x*q1 + (x*x)*q2
        float tp = 1.0f, v = 0.0f;
        v = v + tp * q0;
        tp = tp * x;
        v = v + tp * q1;
        tp = tp * x;
        v = v + tp * q2;
        res = res + v;
    }
    return res;
}

// centre coordinate of the depth-D cell with offset o along one axis, as the
// reference computes it (BinaryNode::CenterAndWidth through double)
__device__ __forceinline__ float cell_centre(int64_t o, float width) {
    return float((0.5 + double(o)) * double(width));
}

// Per sorted point q and axis a, the weights of the three cells around its
// leaf along a: W[9 q + 3 a + k] = value(BaseFunctionMaxDepth.shift(p_a), c)
// with c the centre of the leaf's offset + k - 1.  Every slot that gathers q
// (the leaf's 27 neighbours) needs one of them per axis; computing them once
// per point instead of once per (slot, point) pair does the same float
// operations.
__global__ void __launch_bounds__(kBlock) k_point_weights(const uint32_t* key, const int32_t* point_leaf,
                                                          const float* px, const float* py, const float* pz,
                                                          int64_t n, int64_t baseD, int D, BaseMD bm, float* W) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n) return;
    const uint32_t k = key[baseD + point_leaf[q]];
    const float width = float(1.0 / double(1 << D));
    const int64_t o[3] = {compact3(k >> 2), compact3(k >> 1), compact3(k)};
    const float p[3] = {px[q], py[q], pz[q]};
    float w[9];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int kk = 0; kk < 3; kk++) w[3 * a + kk] = shifted_value(bm, p[a], cell_centre(o[a] + kk - 1, width));
#pragma unroll
    for (int i = 0; i < 9; i++) W[9 * q + i] = w[i];
}

// V_t = sum over t's neighbours j (in order) and their points (in order) of
// w * n_q, w = wx * wy * wz (computeVectorField, main.cu:913-964).  Slot t
// sits at offset -delta_j from neighbour j's leaf, so its weights are the
// point's W entries k = 1 - delta_j.
__device__ __forceinline__ void vector_field_row(const int32_t* __restrict__ pntD, const int32_t* __restrict__ pidx,
                                                 const int32_t* __restrict__ pnum, const float* __restrict__ W,
                                                 const float* __restrict__ nx, const float* __restrict__ ny,
                                                 const float* __restrict__ nz, int64_t baseD, int64_t t,
                                                 float* __restrict__ vfx, float* __restrict__ vfy,
                                                 float* __restrict__ vfz) {
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    int32_t nbr[27];
    pnt_neighbours(pntD + 27 * (t >> 3), int(t & 7), nbr);
    // the 27 point ranges load together (a missing neighbour has none)
    int32_t p0[27], pe[27];
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int64_t nb = baseD + (nbr[j] < 0 ? 0 : nbr[j]);
        p0[j] = pidx[nb];
        pe[j] = p0[j] + (nbr[j] < 0 ? 0 : pnum[nb]);
    }
    auto add = [&](float fx, float fy, float fz, float ux, float uy, float uz) {
        float w = fx * fy;
        w = w * fz;
        vx = vx + w * ux;
        vy = vy + w * uy;
        vz = vz + w * uz;
    };
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int kx = 2 - j / 9, ky = 2 - (j / 3) % 3, kz = 2 - j % 3;  // 1 - delta
        int32_t q = p0[j];
        // two points per round trip, added in order
        for (; q + 1 < pe[j]; q += 2) {
            const float* wa = W + 9 * int64_t(q);
            const float* wb = wa + 9;
            const float ax = wa[kx], ay = wa[3 + ky], az = wa[6 + kz];
            const float bx = wb[kx], by = wb[3 + ky], bz = wb[6 + kz];
            const float nax = nx[q], nay = ny[q], naz = nz[q];
            const float nbx = nx[q + 1], nby = ny[q + 1], nbz = nz[q + 1];
            add(ax, ay, az, nax, nay, naz);
            add(bx, by, bz, nbx, nby, nbz);
        }
        if (q < pe[j]) {
            const float* wa = W + 9 * int64_t(q);
            add(wa[kx], wa[3 + ky], wa[6 + kz], nx[q], ny[q], nz[q]);
        }
    }
    vfx[t] = 0.0f + vx;
    vfy[t] = 0.0f + vy;
    vfz[t] = 0.0f + vz;
}

__global__ void __launch_bounds__(kBlock) k_vector_field(const int32_t* pntD, const int32_t* pidx,
                                                         const int32_t* pnum, const float* W, const float* nx,
                                                         const float* ny, const float* nz, int64_t baseD, int64_t MD,
                                                         float* vfx, float* vfy, float* vfz) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    vector_field_row(pntD, pidx, pnum, W, nx, ny, nz, baseD, t, vfx, vfy, vfz);
}

// k_vector_field through the tile plan.  Per tile: the point ranges of the
// listed blocks' cells (first point, and their offsets in LDS by a scan), then
// the points themselves (9 weights and the normal, one array per component)
// staged in LDS once; every node then walks its 27 neighbours and their
// points in the same order as k_vector_field.  A tile with more than
// kVfPoints points reads them from global memory instead, and a tile with
// more than kTileBlocks listed blocks takes the per-row path (the choices are
// per workgroup).
constexpr int kVfPoints = 768;

__global__ void __launch_bounds__(kTileThreads, 6) k_vector_field_tiles(TilePlan P, int64_t nblk,
                                                                     const int32_t* __restrict__ pntD,
                                                                     const int32_t* __restrict__ pidx,
                                                                     const int32_t* __restrict__ pnum,
                                                                     const float* __restrict__ W,
                                                                     const float* __restrict__ nx,
                                                                     const float* __restrict__ ny,
                                                                     const float* __restrict__ nz, int64_t baseD,
                                                                     float* __restrict__ vfx, float* __restrict__ vfy,
                                                                     float* __restrict__ vfz, int32_t lim_blocks,
                                                                     int32_t lim_points) {
    // A sibling block's points are contiguous in key order, cell after cell,
    // from the block's first point (the first cell's pidx, also when empty):
    // a cell's first point is bfirst[block] + coff[cell] - coff[8 block].
    __shared__ int32_t bfirst[kTileBlocks];        // first point of each listed block
    __shared__ int32_t coff[8 * kTileBlocks + 1];  // LDS offsets of the cells' points
    __shared__ float pw[12 * kVfPoints];           // 9 weights and the normal, one array per component
    __shared__ int16_t rid[kTileEntries];
    __shared__ int32_t wtot[kTileThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tile = blockIdx.x;
    const int32_t ncell = 8 * P.cnt[tile];
    if (ncell > 8 * lim_blocks) {
        const int64_t t = tile * kTileThreads + tid;
        if (t < 8 * nblk) vector_field_row(pntD, pidx, pnum, W, nx, ny, nz, baseD, t, vfx, vfy, vfz);
        return;
    }
    const int16_t* rw = P.rw + tile * kTileEntries;
    const int32_t* ub = P.ub + tile * kTileEntries;
    for (int e = tid; e < kTileEntries; e += kTileThreads) rid[e] = rw[e];
    // cells [c0, c1) of this thread: first points and counts, then an
    // exclusive scan of the counts over the tile
    const int per = (ncell + kTileThreads - 1) / kTileThreads;
    const int c0 = min(ncell, tid * per), c1 = min(ncell, c0 + per);
    int32_t own = 0;
    for (int c = c0; c < c1; c++) {
        const int64_t gi = baseD + ub[c >> 3] + (c & 7);
        if ((c & 7) == 0) bfirst[c >> 3] = pidx[gi];
        own += pnum[gi];
    }
    int32_t inc = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    int32_t base = 0;
    for (int w = 0; w < wave; w++) base += wtot[w];
    int32_t run = base + inc - own;
    for (int c = c0; c < c1; c++) {
        coff[c] = run;
        run += pnum[baseD + ub[c >> 3] + (c & 7)];
    }
    if (tid == kTileThreads - 1) coff[ncell] = run;
    __syncthreads();
    const int32_t total = coff[ncell];
    const bool staged = total <= lim_points;
    if (staged) {
        for (int32_t i = tid; i < total; i += kTileThreads) {
            int lo = 0, hi = ncell;  // the last cell whose offset is <= i
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (coff[mid] <= i)
                    lo = mid;
                else
                    hi = mid;
            }
            const int64_t q = int64_t(bfirst[lo >> 3]) + (i - coff[lo & ~7]);
#pragma unroll
            for (int k = 0; k < 9; k++) pw[k * kVfPoints + i] = W[9 * q + k];
            pw[9 * kVfPoints + i] = nx[q];
            pw[10 * kVfPoints + i] = ny[q];
            pw[11 * kVfPoints + i] = nz[q];
        }
    }
    __syncthreads();
    const int g = tid >> 3, s = tid & 7;
    const int64_t blk = tile * kTile + g;
    if (blk >= nblk) return;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        int entry, child;
        tile_neighbour(s, j, entry, child);
        const int id = rid[27 * g + entry];
        if (id < 0) continue;
        const int cell = 8 * id + child;
        const int32_t o = coff[cell], e = coff[cell + 1];
        const int kx = 2 - j / 9, ky = 2 - (j / 3) % 3, kz = 2 - j % 3;  // 1 - delta
        if (staged) {
            for (int32_t q = o; q < e; q++) {
                const float fx = pw[kx * kVfPoints + q], fy = pw[(3 + ky) * kVfPoints + q],
                            fz = pw[(6 + kz) * kVfPoints + q];
                float w = fx * fy;
                w = w * fz;
                vx = vx + w * pw[9 * kVfPoints + q];
                vy = vy + w * pw[10 * kVfPoints + q];
                vz = vz + w * pw[11 * kVfPoints + q];
            }
        } else {
            const int64_t p0 = int64_t(bfirst[cell >> 3]) + (o - coff[cell & ~7]);
            for (int64_t q = p0; q < p0 + (e - o); q++) {
                const float* wq = W + 9 * q;
                const float fx = wq[kx], fy = wq[3 + ky], fz = wq[6 + kz];
                float w = fx * fy;
                w = w * fz;
                vx = vx + w * nx[q];
                vy = vy + w * ny[q];
                vz = vz + w * nz[q];
            }
        }
    }
    const int64_t t = 8 * blk + s;
    vfx[t] = 0.0f + vx;
    vfy[t] = 0.0f + vy;
    vfz[t] = 0.0f + vz;
}

// DotProduct (main.cu:966-972) in float
__device__ __forceinline__ float fdot(float vx, float vy, float vz, float ux, float uy, float uz) {
    float r = 0.0f;
    r = r + vx * ux;
    r = r + vy * uy;
    r = r + vz * uz;
    return r;
}

// leaf level: 27 terms per node, the neighbours from the leaf level's
// parent-neighbour rows
__device__ __forceinline__ void div_leaf_row(const int32_t* __restrict__ pntD, const float* __restrict__ vfx,
                                             const float* __restrict__ vfy, const float* __restrict__ vfz,
                                             const float* __restrict__ T, int64_t baseD, int64_t t,
                                             float* __restrict__ div) {
    int32_t nb[27];
    pnt_neighbours(pntD + 27 * (t >> 3), int(t & 7), nb);
    // all loads issued before the sum; a missing neighbour reads slot 0 and
    // its term is not added
    float ux[27], uy[27], uz[27];
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int64_t l = nb[j] < 0 ? 0 : nb[j];
        ux[j] = vfx[l];
        uy[j] = vfy[l];
        uz[j] = vfz[l];
    }
    double val = 0.0;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const float r = fdot(ux[j], uy[j], uz[j], T[j / 9], T[(j / 3) % 3], T[j % 3]);
        val = nb[j] < 0 ? val : val + double(r);
    }
    div[baseD + t] = float(val);
}

__global__ void __launch_bounds__(kBlock) k_div_leaf(const int32_t* pntD, const float* vfx, const float* vfy,
                                                     const float* vfz, const float* T, int64_t baseD, int64_t MD,
                                                     float* div) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    div_leaf_row(pntD, vfx, vfy, vfz, T, baseD, t, div);
}

// The leaf level through the tile plan: a tile's distinct neighbour blocks
// are loaded once into LDS (three components), then each node's 27 terms are
// summed from LDS in the order j = 0..26 (the terms and their order are
// k_div_leaf's).  A tile with more than kTileBlocks listed blocks takes the
// per-row path.
__global__ void __launch_bounds__(kTileThreads) k_div_leaf_tiles(TilePlan P, int64_t nblk,
                                                                 const int32_t* __restrict__ pntD,
                                                                 const float* __restrict__ vfx,
                                                                 const float* __restrict__ vfy,
                                                                 const float* __restrict__ vfz,
                                                                 const float* __restrict__ T, int64_t baseD,
                                                                 float* __restrict__ div, int32_t lim_blocks) {
    constexpr int stride = 8 * kTileBlocks;
    __shared__ float cells[3 * stride];
    __shared__ int16_t rid[kTileEntries];
    const int tid = threadIdx.x;
    const int64_t tile = blockIdx.x;
    const int32_t cnt = P.cnt[tile];
    if (cnt > lim_blocks) {
        const int64_t t = tile * kTileThreads + tid;
        if (t < 8 * nblk) div_leaf_row(pntD, vfx, vfy, vfz, T, baseD, t, div);
        return;
    }
    const int16_t* rw = P.rw + tile * kTileEntries;
    const int32_t* ub = P.ub + tile * kTileEntries;
    for (int e = tid; e < kTileEntries; e += kTileThreads) rid[e] = rw[e];
    for (int u = tid; u < 8 * cnt; u += kTileThreads) {
        const int64_t idx = int64_t(ub[u >> 3]) + (u & 7);
        cells[u] = vfx[idx];
        cells[stride + u] = vfy[idx];
        cells[2 * stride + u] = vfz[idx];
    }
    __syncthreads();
    const int g = tid >> 3, s = tid & 7;
    const int64_t blk = tile * kTile + g;
    if (blk >= nblk) return;
    double val = 0.0;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        int entry, child;
        tile_neighbour(s, j, entry, child);
        const int id = rid[27 * g + entry];
        const int l = id < 0 ? 0 : 8 * id + child;
        const float r = fdot(cells[l], cells[stride + l], cells[2 * stride + l], T[j / 9], T[(j / 3) % 3], T[j % 3]);
        val = id < 0 ? val : val + double(r);
    }
    div[baseD + 8 * blk + s] = float(val);
}

// The 27 terms fdot(v, (wx[j / 9], wy[(j / 3) % 3], wz[j % 3])) from the 9
// distinct products: each term is still ((0 + vx wx) + vy wy) + vz wz in
// float, the operations fdot performs.
__device__ __forceinline__ void terms27(float vx, float vy, float vz, const float wx[3], const float wy[3],
                                       const float wz[3], float r[27]) {
    float sx[3], sxy[3][3], pz[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        sx[a] = 0.0f + vx * wx[a];
        pz[a] = vz * wz[a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) sxy[a][b] = sx[a] + vy * wy[b];
#pragma unroll
    for (int j = 0; j < 27; j++) r[j] = sxy[j / 9][(j / 3) % 3] + pz[j % 3];
}

// Sums of 32 per-lane doubles over the 64 lanes of a wave in a fixed order
// (a reduce-scatter: each halving step exchanges half of the values still
// held), 32 shuffles instead of 6 per value.  Returns the total of entry
// q(lane) = 16 b5 + 8 b4 + 4 b3 + 2 b2 + b1 (b_i = bit i of the lane); lanes
// L and L ^ 1 hold the same entry.
__device__ __forceinline__ double wave_reduce_scatter32(double v[32], int lane) {
    const bool h5 = (lane >> 5) & 1, h4 = (lane >> 4) & 1, h3 = (lane >> 3) & 1, h2 = (lane >> 2) & 1,
               h1 = (lane >> 1) & 1;
    double k16[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const double give = h5 ? v[i] : v[16 + i];
        const double keep = h5 ? v[16 + i] : v[i];
        k16[i] = keep + __shfl_xor(give, 32, 64);
    }
    double k8[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double give = h4 ? k16[i] : k16[8 + i];
        const double keep = h4 ? k16[8 + i] : k16[i];
        k8[i] = keep + __shfl_xor(give, 16, 64);
    }
    double k4[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double give = h3 ? k8[i] : k8[4 + i];
        const double keep = h3 ? k8[4 + i] : k8[i];
        k4[i] = keep + __shfl_xor(give, 8, 64);
    }
    double k2[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double give = h2 ? k4[i] : k4[2 + i];
        const double keep = h2 ? k4[2 + i] : k4[i];
        k2[i] = keep + __shfl_xor(give, 4, 64);
    }
    const double give = h1 ? k2[0] : k2[1];
    const double keep = h1 ? k2[1] : k2[0];
    const double k1 = keep + __shfl_xor(give, 2, 64);
    // lanes L and L ^ 1 hold the same entry's two halves; the lower lane's
    // part comes first in the sum
    const double other = __shfl_xor(k1, 1, 64);
    return (lane & 1) ? other + k1 : k1 + other;
}

__device__ __forceinline__ int depth_of(const Bases& B, int D, int64_t i) {
    int d = 0;
    while (d < D && i >= B.b[d + 1]) d++;
    return d;
}

// ---- coarser depths: per-node subtree sums G, then a 27-term gather ---------------
// For a node a of depth d and a relative position j (a = i + delta_j seen from
// the node i being assembled), the reference's terms involving a's subtree are
//   floatdot(V_k, (w[(dx+1)s + rx_k], w[(dy+1)s + ry_k], w[(dz+1)s + rz_k]))
// with r_k the leaf's offset inside a's cell and s = 2^(D-d) - they depend on
// a's own leaves and on j only.  G[a][j] sums them over a's leaf range (one
// streaming pass per depth, 27 accumulators per lane), and
// b_i = sum_j G[neighbour_j(i)][j].  The float terms are the reference's;
// only the grouping of the double sum differs.

// chunks of a node's own leaf range
__global__ void __launch_bounds__(kBlock) k_g_chunks(const int32_t* dnum, int64_t n, int64_t first_skipped,
                                                     int32_t* chunks, int32_t* multi) {
    const int64_t a = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (a >= n) return;
    // nodes from first_skipped on (the depth just above the leaf level) are summed by k_g_parent_sums
    const int32_t c = a >= first_skipped ? 0 : (dnum[a] + kChunk - 1) / kChunk;
    chunks[a] = c;
    multi[a] = (c > 1) ? c : 0;
}

// node of each chunk: the last node whose chunk offset is <= c (the
// exclusive scan of chunk counts is non-decreasing; nodes without chunks
// share their successor's offset, so the last one is the chunk's owner)
__global__ void __launch_bounds__(kBlock) k_g_chunk_map(const int32_t* chunk_off, int64_t n, const int32_t* d_nchunks,
                                                        int32_t* chunk_node) {
    const int64_t c = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= *d_nchunks) return;
    int64_t lo = 0, hi = n;  // invariant: chunk_off[lo] <= c, and hi is past the answer
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (chunk_off[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    chunk_node[c] = int32_t(lo);
}

// the 27 terms of one leaf slot for a node whose cell spans sc leaf cells
// (offsets inside it from the slot's key, weights from the depth's table T)
__device__ __forceinline__ void leaf_terms(uint32_t kl, float vx, float vy, float vz, const float* T, int sc,
                                           uint32_t rmask, float r[27]) {
    const uint32_t rx = compact3(kl >> 2) & rmask, ry = compact3(kl >> 1) & rmask, rz = compact3(kl) & rmask;
    const float wx[3] = {T[rx], T[sc + rx], T[2 * sc + rx]};
    const float wy[3] = {T[ry], T[sc + ry], T[2 * sc + ry]};
    const float wz[3] = {T[rz], T[sc + rz], T[2 * sc + rz]};
    terms27(vx, vy, vz, wx, wy, wz, r);
}

// the coarse depths' tables (depths 0 .. kMaxDepth - 2: 3 * 2^(D - d) floats
// each) held in LDS by every k_g_sums block
constexpr int kCoarseTableFloats = 3 * ((2 << kMaxDepth) - 4);

// one wave per chunk: 27 double accumulators per lane, cross-lane reduction.
// The tables of depths 0 .. D-2 (table[0 .. tcount)) are staged in LDS.
__global__ void __launch_bounds__(kBlock) k_g_sums(const uint32_t* key, const int32_t* didx, const int32_t* dnum,
                                                   const float* vfx, const float* vfy, const float* vfz,
                                                   const float* table, int32_t tcount, const int64_t* table_off,
                                                   const int32_t* chunk_node, const int32_t* chunk_off,
                                                   const int32_t* chunks, const int32_t* multi_off,
                                                   const int32_t* d_nchunks, Bases B, int D, double* G,
                                                   int64_t gstride, double* partial) {
    __shared__ float tab[kCoarseTableFloats];
    for (int i = threadIdx.x; i < tcount; i += kBlock) tab[i] = table[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t nchunks = *d_nchunks;
    // the chunk, its node and their ranges are wave-uniform (scalar loads)
    const int64_t wave0 = __builtin_amdgcn_readfirstlane(int32_t((blockIdx.x * kBlock + threadIdx.x) >> 6));
    const int64_t nwaves = (int64_t(gridDim.x) * kBlock) >> 6;
    const int64_t baseD = B.b[D];
    for (int64_t c = wave0; c < nchunks; c += nwaves) {
        const int64_t a = chunk_node[c];
        const int64_t ci = c - chunk_off[a];
        const int64_t k0 = int64_t(didx[a]) + ci * kChunk;
        const int64_t k1 = min(k0 + kChunk, int64_t(didx[a]) + dnum[a]);
        const int d = depth_of(B, D, a);
        const int sc = 1 << (D - d);
        const uint32_t rmask = uint32_t(sc - 1);
        const float* T = tab + table_off[d];
        double acc[32];
#pragma unroll
        for (int j = 0; j < 32; j++) acc[j] = 0.0;
        // full steps of 128 slots: two per lane, k and k + 64, both loaded
        // before either's terms are computed; then the remaining slots one per
        // lane.  Each lane adds its slots in increasing order.
        int64_t kb = k0;
        for (; kb + 128 <= k1; kb += 128) {
            const int64_t k = kb + lane;
            const uint32_t kla = key[baseD + k], klb = key[baseD + k + 64];
            const float vxa = vfx[k], vya = vfy[k], vza = vfz[k];
            const float vxb = vfx[k + 64], vyb = vfy[k + 64], vzb = vfz[k + 64];
            float r[27];
            leaf_terms(kla, vxa, vya, vza, T, sc, rmask, r);
#pragma unroll
            for (int j = 0; j < 27; j++) acc[j] += double(r[j]);
            leaf_terms(klb, vxb, vyb, vzb, T, sc, rmask, r);
#pragma unroll
            for (int j = 0; j < 27; j++) acc[j] += double(r[j]);
        }
        for (int64_t k = kb + lane; k < k1; k += 64) {
            float r[27];
            leaf_terms(key[baseD + k], vfx[k], vfy[k], vfz[k], T, sc, rmask, r);
#pragma unroll
            for (int j = 0; j < 27; j++) acc[j] += double(r[j]);
        }
        const double tot = wave_reduce_scatter32(acc, lane);
        const int q = 16 * ((lane >> 5) & 1) + 8 * ((lane >> 4) & 1) + 4 * ((lane >> 3) & 1) + 2 * ((lane >> 2) & 1) +
                      ((lane >> 1) & 1);
        if (!(lane & 1) && q < 27) {
            if (chunks[a] == 1)
                G[q * gstride + a] = tot;
            else
                partial[27 * (int64_t(multi_off[a]) + ci) + q] = tot;
        }
    }
}

// nodes spanning several chunks: one workgroup per such node adds its chunk
// partials (fixed per-thread strides, fixed tree) for all 27 positions
__global__ void __launch_bounds__(kBlock) k_g_flag_multi(const int32_t* chunks, int64_t n, int32_t* flag) {
    const int64_t a = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (a < n) flag[a] = chunks[a] > 1 ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_g_list_multi(const int32_t* flag, const int32_t* pos, int64_t n,
                                                         int32_t* list) {
    const int64_t a = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (a < n && flag[a]) list[pos[a]] = int32_t(a);
}

__global__ void __launch_bounds__(kBlock) k_g_finish(const int32_t* list, const int32_t* d_count,
                                                     const int32_t* chunks, const int32_t* multi_off,
                                                     const double* partial, double* G, int64_t gstride) {
    __shared__ double red[kBlock / 64][27];
    const int64_t count = *d_count;  // the grid is sized by an upper bound and strides over the list
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t l = blockIdx.x; l < count; l += gridDim.x) {
        const int64_t a = list[l];
        const int32_t c = chunks[a];
        const double* P = partial + 27 * int64_t(multi_off[a]);
        double acc[27];
#pragma unroll
        for (int j = 0; j < 27; j++) acc[j] = 0.0;
        for (int32_t i = threadIdx.x; i < c; i += kBlock)
#pragma unroll
            for (int j = 0; j < 27; j++) acc[j] += P[27 * int64_t(i) + j];
#pragma unroll
        for (int j = 0; j < 27; j++) {
            double v = acc[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if (lane == 0) red[wave][j] = v;
        }
        __syncthreads();
        if (threadIdx.x < 27) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; w++) v += red[w][threadIdx.x];
            G[threadIdx.x * gstride + a] = v;
        }
        __syncthreads();  // red is reused by the next node
    }
}

// Depth D-1: every node's leaf range is its own sibling block of 8 leaf-level
// slots (or empty), so one lane per slot computes its 27 terms and the 8 lanes
// of a block add them in a fixed order (a reduce-scatter: 28 double
// exchanges instead of the 81 of a full butterfly).
__global__ void __launch_bounds__(kBlock) k_g_parent_sums(const uint32_t* key, const int32_t* parent,
                                                          const float* vfx, const float* vfy, const float* vfz,
                                                          const float* T, int64_t baseD, int64_t MD, double* G,
                                                          int64_t gstride) {
    const int64_t k = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= MD) return;  // MD is a multiple of 8 and kBlock of 8: blocks never straddle
    const uint32_t kl = key[baseD + k];
    const int rx = (kl >> 2) & 1, ry = (kl >> 1) & 1, rz = kl & 1;  // offset inside the parent's cell (s = 2)
    const float wx[3] = {T[rx], T[2 + rx], T[4 + rx]};
    const float wy[3] = {T[ry], T[2 + ry], T[4 + ry]};
    const float wz[3] = {T[rz], T[2 + rz], T[4 + rz]};
    float r[27];
    terms27(vfx[k], vfy[k], vfz[k], wx, wy, wz, r);
    // reduce-scatter over the block's 8 lanes (xor 4, 2, 1): each level
    // exchanges half of the values still held; lane s ends with positions
    // j0 .. j0 + 3, j0 = 16 b2 + 8 b1 + 4 b0 (b_i = bit i of s), its own part
    // first in every sum
    double v[32];
#pragma unroll
    for (int j = 0; j < 32; j++) v[j] = j < 27 ? double(r[j]) : 0.0;
    const int s = int(k & 7);
    const bool h2 = (s >> 2) & 1, h1 = (s >> 1) & 1, h0 = s & 1;
    double k16[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const double give = h2 ? v[i] : v[16 + i];
        const double keep = h2 ? v[16 + i] : v[i];
        k16[i] = keep + __shfl_xor(give, 4, 64);
    }
    double k8[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double give = h1 ? k16[i] : k16[8 + i];
        const double keep = h1 ? k16[8 + i] : k16[i];
        k8[i] = keep + __shfl_xor(give, 2, 64);
    }
    double k4[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double give = h0 ? k8[i] : k8[4 + i];
        const double keep = h0 ? k8[4 + i] : k8[i];
        k4[i] = keep + __shfl_xor(give, 1, 64);
    }
    const int64_t a = parent[baseD + k];
    const int j0 = 16 * h2 + 8 * h1 + 4 * h0;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (j0 + i < 27) G[(j0 + i) * gstride + a] = k4[i];
}

// b_i = sum_j G[neighbour_j(i)][j] for the nodes above the leaf level; the
// neighbours come from the parent-neighbour rows (depth >= 1), the root's
// only neighbour is itself.
// G holds every position of a node with leaf-level slots below it (dnum > 0);
// a node without any contributes no terms and its G is neither written nor
// read.
__global__ void __launch_bounds__(kBlock, 6) k_div_from_g(const int32_t* pnt, const int64_t* pnt_off, Bases B, int D,
                                                       const int32_t* dnum, const double* G, int64_t gstride,
                                                       int64_t n, float* div) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    double val = 0.0;
    if (i == 0) {
        if (dnum[0] > 0) val += G[13 * gstride + 0];
    } else {
        const int d = depth_of(B, D, i);
        const int64_t t = i - B.b[d];
        int32_t nbr[27];
        pnt_neighbours(pnt + 27 * (pnt_off[d] + (t >> 3)), int(t & 7), nbr);
        // loads issued nine at a time; G of a node without leaf-level slots
        // is never written, and its value is read but not added
#pragma unroll
        for (int j0 = 0; j0 < 27; j0 += 9) {
            int32_t dn[9];
            double g[9];
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const int j = j0 + q;
                const int64_t nb = B.b[d] + (nbr[j] < 0 ? 0 : nbr[j]);
                dn[q] = dnum[nb];
                g[q] = G[j * gstride + nb];
            }
#pragma unroll
            for (int q = 0; q < 9; q++) val = (nbr[j0 + q] < 0 || dn[q] == 0) ? val : val + g[q];
        }
    }
    div[i] = float(val);
}

}  // namespace

// Function tables depend on the depth only: built once per depth per process
// and shared (read-only) by every context.
std::shared_ptr<const FsTables> fs_tables_of(int D) {
    static std::mutex m;
    static std::map<int, std::shared_ptr<const FsTables>> cache;
    std::lock_guard<std::mutex> lock(m);
    auto it = cache.find(D);
    if (it != cache.end()) return it->second;
    auto t = std::make_shared<FsTables>();
    build_fs_tables(D, t.get());
    cache[D] = t;
    return t;
}

void ensure_tables(prg_ctx* c) {
    if (c->fs && c->fs->D == c->D) return;
    std::shared_ptr<const FsTables> t = fs_tables_of(c->D);
    hipStream_t s = c->stream;
    PRG_HIP(hipMemcpyAsync(c->d_base_fn.reserve(t->base_fn.size()), t->base_fn.data(), t->base_fn.size() * 4,
                           hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(c->d_stencil.reserve(t->stencil.size()), t->stencil.data(), t->stencil.size() * 4,
                           hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(c->d_div_table.reserve(t->div_table.size()), t->div_table.data(),
                           t->div_table.size() * 4, hipMemcpyHostToDevice, s));
    c->fs = std::move(t);  // the cached host arrays outlive the copies
}

void compute_divergence(prg_ctx* c) {
    ensure_tables(c);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M, MD = c->MD(), baseD = c->bases[D];
    BaseMD bm;
    for (int k = 0; k < 16; k++) bm.v[k] = c->fs->base_maxdepth[k];
    // LDS capacities of the tile kernels; a smaller limit (a test setting)
    // sends more tiles down the per-row and global-memory paths
    const int32_t lim_blocks = c->tile_limit > 0 ? std::min(c->tile_limit, kTileBlocks) : kTileBlocks;
    const int32_t lim_points = c->tile_limit > 0 ? std::min(c->tile_limit, kVfPoints) : kVfPoints;
    float* VX = c->vfx.reserve(MD);
    float* VY = c->vfy.reserve(MD);
    float* VZ = c->vfz.reserve(MD);
    {
        StageTimer st(c, PRG_STAGE_VECTOR_FIELD);
        PoolBuf<float> W{&c->pool[75]};
        W.reserve(size_t(9) * c->N);
        {
            TimedScope ts(c->kt, "point_weights", 52.0 * double(c->N));
            k_point_weights<<<grid_for(c->N, kBlock), kBlock, 0, s>>>(c->key.get(), c->point_leaf.get(), c->px.get(),
                                                                      c->py.get(), c->pz.get(), c->N, baseD, D, bm,
                                                                      W.get());
            PRG_CHECK_LAUNCH();
        }
        TimedScope ts(c->kt, "vector_field", 12.0 * double(MD) + 116.0 * double(MD) + 24.0 * double(c->N));
        TilePlan P{};
        if (tile_plan_on(c, D, &P, s)) {
            const int64_t nblk = MD / 8, ntiles = (nblk + kTile - 1) / kTile;
            k_vector_field_tiles<<<unsigned(ntiles), kTileThreads, 0, s>>>(
                P, nblk, c->pnt.get() + 27 * c->pnt_off[D], c->pidx.get(), c->pnum.get(), W.get(), c->nx.get(),
                c->ny.get(), c->nz.get(), baseD, VX, VY, VZ, lim_blocks, lim_points);
        } else {
            k_vector_field<<<grid_for(MD, kBlock), kBlock, 0, s>>>(c->pnt.get() + 27 * c->pnt_off[D], c->pidx.get(),
                                                                   c->pnum.get(), W.get(), c->nx.get(), c->ny.get(),
                                                                   c->nz.get(), baseD, MD, VX, VY, VZ);
        }
        PRG_CHECK_LAUNCH();
    }
    StageTimer st(c, PRG_STAGE_DIVERGENCE);
    float* DV = c->div.reserve(M);
    PoolBuf<int64_t> toff{&c->pool[60]};
    PRG_HIP(hipMemcpyAsync(toff.reserve(c->fs->div_off.size()), c->fs->div_off.data(), c->fs->div_off.size() * 8,
                           hipMemcpyHostToDevice, s));
    {
        TimedScope ts(c->kt, "div_leaf", 124.0 * double(MD));
        TilePlan P{};
        const float* T = c->d_div_table.get() + c->fs->div_off[D];
        if (tile_plan_on(c, D, &P, s)) {
            const int64_t nblk = MD / 8, ntiles = (nblk + kTile - 1) / kTile;
            k_div_leaf_tiles<<<unsigned(ntiles), kTileThreads, 0, s>>>(P, nblk, c->pnt.get() + 27 * c->pnt_off[D], VX,
                                                                       VY, VZ, T, baseD, DV, lim_blocks);
        } else {
            k_div_leaf<<<grid_for(MD, kBlock), kBlock, 0, s>>>(c->pnt.get() + 27 * c->pnt_off[D], VX, VY, VZ, T,
                                                               baseD, MD, DV);
        }
        PRG_CHECK_LAUNCH();
    }
    const int64_t n = baseD;  // nodes of depths 0 .. D-1
    if (n > 0) {
        PoolBuf<int32_t> chunks{&c->pool[61]}, multi{&c->pool[62]}, coff{&c->pool[63]}, moff{&c->pool[64]},
            cnode{&c->pool[65]}, ctot{&c->pool[66]}, mtot{&c->pool[67]};
        PoolBuf<double> G{&c->pool[68]}, partial{&c->pool[69]};
        chunks.reserve(n);
        multi.reserve(n);
        coff.reserve(n);
        moff.reserve(n);
        ctot.reserve(1);
        mtot.reserve(1);
        G.reserve(size_t(27) * n);  // written for every node with dnum > 0; only those are read
        const int64_t first_skipped = D >= 1 ? c->bases[D - 1] : n;
        if (D >= 1) {
            TimedScope ts(c->kt, "div_parent_sums", 16.0 * double(MD) + 216.0 * double(c->counts[D - 1]));
            k_g_parent_sums<<<grid_for(MD, kBlock), kBlock, 0, s>>>(c->key.get(), c->parent.get(), VX, VY, VZ,
                                                                   c->d_div_table.get() + c->fs->div_off[D - 1],
                                                                   baseD, MD, G.get(), n);
            PRG_CHECK_LAUNCH();
        }
        k_g_chunks<<<grid_for(n, kBlock), kBlock, 0, s>>>(c->dnum.get(), n, first_skipped, chunks.get(), multi.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive2(chunks.get(), coff.get(), ctot.get(), multi.get(), moff.get(), mtot.get(), n, s, c->scratch,
                        c->kt);
        // Sizes from upper bounds, the counts stay on the device (no host
        // round trip): a node's chunks number at most 1 + dnum / kChunk and
        // every depth's dnum add up to the leaf-level slot count, so
        // chunks <= n + (D - 1) (MD / kChunk + 1); a node with several chunks
        // has dnum > kChunk, so those chunks number at most
        // 2 (D - 1) (MD / kChunk + 1), and such nodes half that.
        const int64_t per_depth = MD / kChunk + 1;
        const int64_t nchunks = n + std::max(1, D - 1) * per_depth;
        const int64_t nmulti = 2 * std::max(1, D - 1) * per_depth;
        const int64_t nm = std::max(1, D - 1) * per_depth;
        cnode.reserve(nchunks);
        partial.reserve(size_t(27) * nmulti);
        k_g_chunk_map<<<grid_for(nchunks, kBlock), kBlock, 0, s>>>(coff.get(), n, ctot.get(), cnode.get());
        PRG_CHECK_LAUNCH();
        Bases B{};
        for (int d = 0; d <= D; d++) B.b[d] = c->bases[d];
        B.b[D + 1] = M;
        const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(nchunks, 16384));
        // depths 0 .. D-2 come first in the table, 3 * 2^(D - d) floats each
        const int32_t tcount = D >= 2 ? int32_t(c->fs->div_off[D - 1]) : 0;
        if (tcount > kCoarseTableFloats) throw Error{PRG_E_ARG, "divergence: coarse tables exceed their LDS area"};
        {
            // every leaf slot read once per coarser depth (key + vector field), G written once
            TimedScope ts(c->kt, "div_subtree_sums", 16.0 * double(MD) * double(D) + 216.0 * double(n));
            k_g_sums<<<grid_for(waves * 64, kBlock), kBlock, 0, s>>>(
                c->key.get(), c->didx.get(), c->dnum.get(), VX, VY, VZ, c->d_div_table.get(), tcount, toff.get(), cnode.get(),
                coff.get(), chunks.get(), moff.get(), ctot.get(), B, D, G.get(), n, partial.get());
            PRG_CHECK_LAUNCH();
        }
        {
            PoolBuf<int32_t> mflag{&c->pool[70]}, mpos{&c->pool[71]}, mcnt{&c->pool[72]}, mlist{&c->pool[73]};
            mflag.reserve(n);
            mpos.reserve(n);
            mcnt.reserve(1);
            k_g_flag_multi<<<grid_for(n, kBlock), kBlock, 0, s>>>(chunks.get(), n, mflag.get());
            PRG_CHECK_LAUNCH();
            scan_exclusive(mflag.get(), mpos.get(), n, mcnt.get(), s, c->scratch, c->kt);
            mlist.reserve(nm);
            k_g_list_multi<<<grid_for(n, kBlock), kBlock, 0, s>>>(mflag.get(), mpos.get(), n, mlist.get());
            PRG_CHECK_LAUNCH();
            TimedScope ts(c->kt, "div_subtree_finish", 216.0 * double(nm));
            const unsigned fgrid = unsigned(std::min<int64_t>(nm, 2048));
            k_g_finish<<<fgrid, kBlock, 0, s>>>(mlist.get(), mcnt.get(), chunks.get(), moff.get(), partial.get(),
                                                G.get(), n);
            PRG_CHECK_LAUNCH();
        }
        {
            PoolBuf<int64_t> poff{&c->pool[74]};
            PRG_HIP(hipMemcpyAsync(poff.reserve(D + 2), c->pnt_off.data(), size_t(D + 2) * 8, hipMemcpyHostToDevice, s));
            TimedScope ts(c->kt, "div_from_subtree_sums", 13.5 * double(n) + 216.0 * double(n) + 4.0 * double(n));
            k_div_from_g<<<grid_for(n, kBlock), kBlock, 0, s>>>(c->pnt.get(), poff.get(), B, D, c->dnum.get(), G.get(),
                                                               n, n, DV);
            PRG_CHECK_LAUNCH();
        }
    }
}

}  // namespace prg
