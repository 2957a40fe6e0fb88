// Per-depth Laplacian solves (reference: LaplacianIteration main.cu:1223-1332,
// gpuConjugateGradient CG_CUDA.cuh:256-324).
//
// The reference assembles a CSR matrix per depth in managed memory and runs
// a cooperative-groups CG with six grid syncs per iteration.  Its matrix is a
// constant 27-point stencil per depth (the dot tables depend on relative
// position only), so the SpMV here is matrix-free: y_i = sum_j L_d[j] *
// x[neighbour_j(i)] over existing neighbours, in the CSR row order j = 0..26
// and the float arithmetic of gpuSpMV.  Vectors and SpMV are FP32; every dot
// product and CG scalar is FP64; r = b - A x is recomputed every `refresh`
// iterations.
//
// Large depths (> kSmallRows nodes), two kernels per iteration:
//   k_cg_spmv    forms p_k = r + beta p_{k-1} on the fly for every row and
//                neighbour it touches (p double-buffered, so the direction
//                update needs no pass of its own), q = A p_k, FP64 partials of
//                p.q;
//   k_cg_update  every workgroup sums the p.q partials in a fixed order and
//                forms alpha; x += alpha p, r -= alpha q, FP64 partials of r.r
//                (summed the same way by the next SpMV for beta and the stop
//                test).
// Neighbours come from the parent-neighbour table (13.5 bytes per row instead
// of 27 neighbour indices); the SpMV stages each sibling block's 4x4x4
// neighbourhood in LDS.  Kernels return at once when the stop flag is
// set, so iterations are queued in batches with one host sync per batch.
//
// Small depths (<= kSmallRows nodes) run the whole CG inside one 1024-thread
// workgroup each, all small depths in a single launch on a second stream, so
// they overlap the large-depth solves (the per-depth systems are independent).
#include <algorithm>
#include <cmath>

#include "../../include/prg_unet.h"
#include "ctx.h"
#include "pnt.h"
#include "tile.h"

namespace prg {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kSmallBlock = 1024;
constexpr int64_t kSmallRows = 2048;
// Algorithmic bytes per row of the kernels that gather through the
// parent-neighbour table: the 27 entries of a row's 8-row sibling block (13.5 B
// per row), plus the vectors each reads and writes once per row.  The SpMV
// reads the tile plans instead (spmv_plan_bytes).  (The reference's
// representation, 27 int32 neighbour indices + p_i read + (Ap)_i written,
// would be 116 B per SpMV row.)
constexpr double kPntRowBytes = 13.5;
constexpr double kResidualRowBytes = kPntRowBytes + 12.0;  // x, b read; r written

struct Stencil {
    float L[27];
};

// Scalars of one chain's CG, written only by workgroup 0 of a kernel and read
// by later kernels (kernel boundaries order them) and by the host.
struct CgState {
    double bb;          // b.b
    double rr;          // r.r at the last stop test
    double rr_hist[2];  // rr_k in slot k & 1
    int32_t done, iters;
    int32_t restarted;  // a warm start with ||b - A x0|| > ||b|| was replaced by x0 = 0
    int32_t pad;
};

// One depth of a chain: its rows, its workgroup ranges in the chain's SpMV and
// vector launches, its tile plan, vectors (p double-buffered), FP64 partials
// and CG state, and its stencil.  A chain's kernels serve all its depths in
// one launch; a workgroup finds its depth from the ranges (wave-uniform).
struct Seg {
    int64_t n;
    int32_t wg0, nwg;    // SpMV workgroups [wg0, wg0 + nwg)
    int32_t vwg0, nvwg;  // vector-kernel workgroups [vwg0, vwg0 + nvwg)
    const int32_t* cnt;  // tile plan
    const int32_t* ub;
    const int16_t* rw;
    float *R, *P0, *P1, *Q, *XS;
    double *rr, *pq, *bb;  // r.r (two slots of kMaxWG), p.q, b.b partials
    CgState* st;
    float L[27];
};

__device__ __forceinline__ int seg_of(const Seg* __restrict__ segs, int nseg, int b, bool vec) {
    int k = 0;
    for (int i = 1; i < nseg; i++)
        if (b >= (vec ? segs[i].vwg0 : segs[i].wg0)) k = i;
    return k;
}

// y_t = sum_j L[j] * v[neighbour_j(t)] over existing neighbours with a
// nonzero coefficient (gpuSpMV order, float); the neighbours come from the
// row of t's sibling block (8 entries serve the 27)
__device__ __forceinline__ float spmv_pnt(const int32_t* __restrict__ pnt, int64_t t, const float* __restrict__ v,
                                          const Stencil& S) {
    int32_t nb[27];
    pnt_neighbours(pnt + 27 * (t >> 3), int(t & 7), nb);
    // the 27 loads go out together (a missing neighbour reads row 0); terms
    // of missing neighbours or zero coefficients are not added
    float vv[27];
#pragma unroll
    for (int j = 0; j < 27; j++) vv[j] = v[nb[j] < 0 ? 0 : nb[j]];
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const float prod = S.L[j] * vv[j];
        out = (nb[j] < 0 || S.L[j] == 0.0f) ? out : out + prod;
    }
    return out;
}

// Block-wide sum of a double in a fixed order; every thread gets the total.
template <int BS>
__device__ __forceinline__ double block_allsum(double v, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; w++) t += lds[w];
    return t;
}

// Sum of the np partials a previous kernel wrote, in a fixed order (fixed
// per-thread strides, fixed tree): every workgroup that calls it gets the
// same value, so the CG scalars need no cross-workgroup handshake.
template <int BS>
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int np, double* lds) {
    double s = 0.0;
    for (int b = threadIdx.x; b < np; b += BS) s += part[b];
    return block_allsum<BS>(s, lds);
}

constexpr int kVecThreads = 512;  // init / update / refresh / true residual
constexpr int kMaxWG = 1024;       // workgroups of a depth in the vector kernels (its partial count)
constexpr int kMaxSpmvWG = 2048;   // workgroups of a depth in the SpMV (its partial count)
constexpr int kSpmvThreads = kTileThreads;  // one row per lane, kTile sibling blocks per tile
constexpr int kHashSlots = 2048;          // > kTileEntries: linear probing always finds a slot
static_assert(kHashSlots > kTileEntries, "hash must hold every entry of a tile");

// r = b - A x0, p = r, xs = x0 (the solve works on the 16-byte-aligned copy
// xs); partials of r.r (slot 0) and b.b.  zero_x: x0 = 0 is known, and A x0
// is +0 (a float sum of signed zeros started at +0), so r = b - 0 without the
// SpMV.
__global__ void __launch_bounds__(kVecThreads) k_cg_init(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                         const float* x, int32_t zero_x, float* xs, float* r,
                                                         float* p, double* rr_part, double* bb_part, CgState* st) {
    __shared__ double lds[kVecThreads / 64];
    double bb = 0.0, rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kVecThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kVecThreads) {
        const float ax = zero_x ? 0.0f : spmv_pnt(pnt, t, x, S);
        const float rv = b[t] - ax;
        xs[t] = zero_x ? 0.0f : x[t];
        r[t] = rv;
        p[t] = rv;
        bb += double(b[t]) * double(b[t]);
        rr += double(rv) * double(rv);
    }
    const double sbb = block_allsum<kVecThreads>(bb, lds);
    const double srr = block_allsum<kVecThreads>(rr, lds);
    if (threadIdx.x == 0) {
        bb_part[blockIdx.x] = sbb;
        rr_part[blockIdx.x] = srr;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->done = 0;
        st->iters = 0;
        st->restarted = 0;
    }
}

// A warm start worse than none (||b - A x0|| > ||b||, from k_cg_init's
// partials, summed in a fixed order) is replaced by x0 = 0: xs = 0, r = p = b,
// and each workgroup's r.r partial becomes its b.b partial (the same grid as
// k_cg_init, so the same rows per workgroup: the state k_cg_init leaves for
// x0 = 0, bit for bit).  Every workgroup decides from the same sums; the
// partials are only rewritten by k_warm_reset_apply after all have decided.
__global__ void __launch_bounds__(kVecThreads) k_warm_reset_decide(const double* rr_part, const double* bb_part,
                                                                   int32_t np, CgState* st) {
    __shared__ double lds[kVecThreads / 64];
    const double rr = sum_partials<kVecThreads>(rr_part, np, lds);
    const double bb = sum_partials<kVecThreads>(bb_part, np, lds);
    if (threadIdx.x == 0) st->restarted = rr > bb ? 1 : 0;
}

__global__ void __launch_bounds__(kVecThreads) k_warm_reset_apply(int64_t n, const float* b, float* xs, float* r,
                                                                  float* p, double* rr_part, const double* bb_part,
                                                                  const CgState* st) {
    if (!st->restarted) return;
    for (int64_t t = int64_t(blockIdx.x) * kVecThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kVecThreads) {
        xs[t] = 0.0f;
        r[t] = b[t];
        p[t] = b[t];
    }
    if (threadIdx.x == 0) rr_part[blockIdx.x] = bb_part[blockIdx.x];
}

// Tile plan of one depth (once per octree): for each tile of kTile sibling
// blocks, the distinct sibling blocks its rows touch (`ubase`, their first
// rows) and, per (block, parent-neighbour entry), the slot of that block in
// the list (`rowid`, -1 if absent).  The parent-neighbour entries of
// neighbouring rows repeat (about 2 distinct blocks per block), so the SpMV
// loads each distinct block once per tile as full 32-byte segments.
__global__ void __launch_bounds__(kSpmvThreads) k_tile_plan(const int32_t* pnt, int64_t nblk, int32_t* count,
                                                            int32_t* ubase, int16_t* rowid) {
    __shared__ int32_t hkey[kHashSlots];
    __shared__ int16_t hslot[kHashSlots];
    __shared__ int16_t code[kTileEntries];
    __shared__ int32_t ucount;
    const int tid = threadIdx.x;
    const int64_t tile = blockIdx.x, b0 = tile * kTile;
    for (int h = tid; h < kHashSlots; h += kSpmvThreads) hkey[h] = -1;
    if (tid == 0) ucount = 0;
    __syncthreads();
    for (int e = tid; e < kTileEntries; e += kSpmvThreads) {
        int32_t v = -1;
        if (b0 + e / 27 < nblk) v = pnt[27 * b0 + e];
        int c = -1;
        if (v >= 0) {
            const int32_t key = v >> 3;
            uint32_t h = (uint32_t(key) * 2654435761u) >> 21;
            while (true) {
                const int32_t prev = atomicCAS(&hkey[h], -1, key);
                if (prev == -1) {
                    const int id = atomicAdd(&ucount, 1);
                    hslot[h] = int16_t(id);
                    ubase[tile * kTileEntries + id] = v;
                    break;
                }
                if (prev == key) break;
                h = (h + 1) & (kHashSlots - 1);
            }
            c = int(h);
        }
        code[e] = int16_t(c);
    }
    __syncthreads();
    for (int e = tid; e < kTileEntries; e += kSpmvThreads) {
        const int c = code[e];
        rowid[tile * kTileEntries + e] = c >= 0 ? hslot[c] : int16_t(-1);
    }
    if (tid == 0) count[tile] = ucount;
}

// Largest tile count and the sum of the tile counts of every depth (one
// workgroup per depth; the per-tile counts of depth d are count[off[d] ..
// off[d + 1]))
struct PlanOffsets {
    int64_t off[kMaxDepth + 2];
};
__global__ void __launch_bounds__(1024) k_plan_max(const int32_t* count, PlanOffsets P, int32_t* maxcount,
                                                   long long* sumcount) {
    __shared__ int32_t red[1024 / 64];
    __shared__ long long reds[1024 / 64];
    const int d = blockIdx.x;
    int32_t m = 0;
    long long sm = 0;
    for (int64_t t = P.off[d] + threadIdx.x; t < P.off[d + 1]; t += 1024) {
        m = max(m, count[t]);
        sm += count[t];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = max(m, __shfl_down(m, o, 64));
        sm += __shfl_down(sm, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = m;
        reds[threadIdx.x >> 6] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t r = 0;
        long long rs = 0;
        for (int w = 0; w < 1024 / 64; w++) {
            r = max(r, red[w]);
            rs += reds[w];
        }
        maxcount[d] = r;
        sumcount[d] = rs;
    }
}

// CG iteration `it` (1-based), first kernel.  Every workgroup sums the r.r
// partials of the previous iteration (rr_{it-1}) and applies the stop test
// (iteration cap, rr == 0, fixed count, rr <= tol^2 b.b); then q = A p_it over
// tiles, with p_it = r + beta p_{it-1} formed as the cells reach LDS (it == 1:
// p_1 = r is already in pnew), each distinct sibling block of a tile loaded
// once into LDS (see k_tile_plan), and the rows summed from LDS in the order
// j = 0..26.  Writes this workgroup's partial of p.q.
//
// The tile loop is software-pipelined two tiles deep: while a tile is summed,
// the raw cell values of the workgroup's next tile and the plan (count, row
// ids, block bases) of the one after are in flight in registers.  K = cells per thread covers the
// depth's largest tile (8 * count <= K * kSpmvThreads).
// Instances with at most kSpmvSmallK cells per thread fit 64 VGPRs: eight
// waves per SIMD, four workgroups per CU when LDS allows; the larger ones run
// two or three (the launch asks the runtime's occupancy).
constexpr int kSpmvSmallK = 5;

// Raw buffer descriptor of `bytes` bytes at p (the gfx9 data-format word);
// loads at offsets beyond it return 0 and stores there are dropped.  Ranges
// are capped just below 4 GiB so that kBufOob always lies outside (a depth
// holds fewer than 2^30 - 2 rows).  The SpMV's vector and plan traffic goes
// through these rather than through the Seg's pointers: those are generic,
// and generic (flat) loads count against the LDS wait counter too, so every
// LDS wait of the tile loop would also wait for the next tile's loads.
constexpr int kBufWord3 = 0x00020000;
constexpr uint32_t kBufOob = 0xFFFFFFFCu;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint64_t bytes) {
    const uint32_t nr = bytes < 0xFFFFFFF8ull ? uint32_t(bytes) : 0xFFFFFFF8u;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(nr), kBufWord3);
}

template <int K>
__global__ void __launch_bounds__(kSpmvThreads, K <= kSpmvSmallK ? 8 : (K <= 8 ? 4 : 2)) k_cg_spmv(const Seg* __restrict__ segs,
                                                                                 int32_t nseg, int32_t it,
                                                          float tol, int32_t fixed, int32_t max_iters) {
    static_assert(K * kSpmvThreads <= 8 * kTileEntries, "cells per thread beyond a tile's plan");
    constexpr int kRidWords = kTileEntries / 2;  // row ids as 32-bit words
    extern __shared__ float cellv[];
    __shared__ int32_t rid32[kRidWords];
    // each block's 4x4x4 neighbourhood, (x, y, z) = (c >> 4, (c >> 2) & 3, c & 3),
    // block g's at grid + 272 (g >> 2) + 64 (g & 3) + {0, 2, 8, 10}[g & 3]: the
    // four blocks of a half-wave (g >> 2 fixed) in windows that do not
    // overlap.  Their lanes' rows sit at c0 in {0, 1, 4, 5, 16, 17, 20, 21},
    // and the four shifts move those sets onto disjoint banks (the group
    // stride, 16 mod 32, maps the sets onto each other), so a stencil read
    // (ds_read2, constant offsets) meets one lane per bank; the fill writes
    // meet at most 2.
    __shared__ float grid[(kTile / 4) * 272];
    __shared__ double lds[kSpmvThreads / 64];
    const int16_t* rid = reinterpret_cast<const int16_t*>(rid32);
    const Seg& sg = segs[seg_of(segs, nseg, blockIdx.x, false)];
    CgState* st = sg.st;
    if (st->done) return;
    const int64_t n = sg.n;
    // p_k and q of the tile's own rows (it == 1: p_1 is already in place, the
    // empty range drops those stores)
    const auto rs_pnew = buf_rsrc(((it - 1) & 1) ? sg.P1 : sg.P0, it == 1 ? 0 : uint64_t(n) * sizeof(float));
    const auto rs_q = buf_rsrc(sg.Q, uint64_t(n) * sizeof(float));
    const double* rr_part = sg.rr + ((it - 1) & 1) * kMaxWG;
    const double* bb_part = sg.bb;
    const int32_t np = sg.nvwg;
    double* pq_part = sg.pq;
    const int bid = int(blockIdx.x) - sg.wg0, nblocks = sg.nwg;
    const int k1 = it - 1;
    const bool first = (it == 1);
    const int tid = threadIdx.x;
    const int g = tid >> 3, s = tid & 7;
    const int64_t nblk = n >> 3;
    const int64_t ntiles = (nblk + kTile - 1) / kTile;
    // Tile schedule.  Workgroups are dealt to the 8 XCDs in turn, so the
    // segment's workgroups bid = x, x + G, x + 2G, ... (G = min(8, nblocks))
    // run on one XCD.  Group x takes one contiguous range of tiles, in
    // proportion to its members, and its members stride through it together:
    // member l takes tiles tbeg + l, tbeg + l + m, ... (m members).  At any
    // moment an XCD works on a window of m consecutive tiles, whose
    // neighbourhoods overlap in that XCD's L2.
    // (Integer division runs on the vector unit; readfirstlane keeps the
    // results in scalar registers, where the buffer descriptors need them.)
    const int G = nblocks < 8 ? nblocks : 8;
    const int x = __builtin_amdgcn_readfirstlane(bid % G), l = __builtin_amdgcn_readfirstlane(bid / G);
    const int q8 = __builtin_amdgcn_readfirstlane(nblocks / G), r8 = nblocks - q8 * G;
    const int before = x * q8 + (x < r8 ? x : r8);  // workgroups of the groups before x
    const int m = q8 + (x < r8 ? 1 : 0);
    // tiles split evenly over the workgroups (a depth has fewer than 2^31 tiles)
    const int32_t per = __builtin_amdgcn_readfirstlane(int32_t(ntiles) / nblocks);
    const int32_t rem = int32_t(ntiles) - per * nblocks;
    const int64_t gbeg = int64_t(before) * per + (before < rem ? before : rem);
    const int64_t gend = int64_t(before + m) * per + (before + m < rem ? before + m : rem);
    const int64_t tbeg = gbeg + l, tstep = m, tend = gend;
    // this lane's row sits at (1 + sx, 1 + sy, 1 + sz) of its block's grid
    const int c0 = 16 * ((s >> 2) & 1) + 4 * ((s >> 1) & 1) + (s & 1);
    float* const gbase = grid + 272 * (g >> 2) + 64 * (g & 3) + ((g & 1) ? 2 : 0) + ((g & 2) ? 8 : 0);
    // plan of one tile: cell count, row-id words, block base of each of this thread's cells
    struct Plan {
        int32_t cnt;
        int32_t rw[2];
        int32_t ub[K];
    };
    // Plan and cell loads go through buffer descriptors: a 32-bit offset from
    // a uniform base (one address register per load), and offsets past the
    // descriptor's range read 0 without touching memory: the row-id words
    // beyond the plan, the cells beyond the tile's count, and every plan entry
    // of a tile past this workgroup's range (count 0, so none of its cells is
    // loaded either).  No load sits behind a branch, so the waits before the
    // values are used count exactly the loads and stores issued after them.
    // (the range check covers the vector offset only, so the tile's offset goes there)
    const auto rs_cnt = buf_rsrc(sg.cnt + gbeg, uint64_t(gend - gbeg) * sizeof(int32_t));
    auto load_plan = [&](int64_t tile, Plan& P) {
        const bool in = tile < tend;
        P.cnt = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rs_cnt, uint32_t(4 * (tile - gbeg)), 0, 0));
        const auto rw = buf_rsrc(sg.rw + tile * kTileEntries, in ? kTileEntries * sizeof(int16_t) : 0);
        P.rw[0] = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rw, 4 * tid, 0, 0));
        P.rw[1] = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rw, 4 * (tid + kSpmvThreads), 0, 0));
        const auto ub = buf_rsrc(sg.ub + tile * kTileEntries, in ? kTileEntries * sizeof(int32_t) : 0);
#pragma unroll
        for (int k = 0; k < K; k++)  // entries beyond the count are masked later
            P.ub[k] = int32_t(__builtin_amdgcn_raw_buffer_load_b32(ub, 4 * ((tid >> 3) + k * (kSpmvThreads / 8)), 0, 0));
    };
    // raw r and p_(k-1) of the tile's cells; p_k = r + beta p_(k-1) is formed
    // when the values reach LDS, so nothing waits for these loads before the
    // next tile.  it == 1: p_1 = r (k_cg_init copies r into p), so the second
    // load reads r again and the value is taken as it is.
    const auto rs_r = buf_rsrc(sg.R, uint64_t(n) * sizeof(float));
    const auto rs_p = buf_rsrc(first ? sg.R : ((it & 1) ? sg.P1 : sg.P0), uint64_t(n) * sizeof(float));
    auto load_cells = [&](const Plan& P, float vr[K], float vp[K]) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int u = tid + k * kSpmvThreads;
            const uint32_t off = u < 8 * P.cnt ? 4u * (uint32_t(P.ub[k]) + uint32_t(u & 7)) : kBufOob;
            vr[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_r, off, 0, 0));
            vp[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_p, off, 0, 0));
        }
    };
    Plan pa{}, pb{};
    float va[K], vpa[K];
    load_plan(tbeg, pa);
    // (the first plan is in flight while the partial sums load and reduce)
    const double rr = sum_partials<kSpmvThreads>(rr_part, np, lds);
    const double bb = (it == 1) ? sum_partials<kSpmvThreads>(bb_part, np, lds) : st->bb;
    const double tol2 = double(tol) * double(tol) * bb;
    const bool stop = (bb <= 0.0) || (k1 >= max_iters) || (rr == 0.0) || (fixed > 0 ? (k1 >= fixed) : (rr <= tol2));
    const float bf = (it == 1) ? 0.0f : float(rr / st->rr_hist[(k1 - 1) & 1]);
    if (bid == 0 && threadIdx.x == 0) {
        if (it == 1) st->bb = bb;
        st->rr_hist[k1 & 1] = rr;
        st->rr = rr;
        if (stop) {
            st->done = 1;
            st->iters = k1;
        }
    }
    if (stop) return;
    load_cells(pa, va, vpa);
    load_plan(tbeg + tstep, pb);
    double pq = 0.0;
    for (int64_t tile = tbeg; tile < tend; tile += tstep) {
        // this tile: values in va, plan in pa; the next one: plan in pb
        rid32[tid] = pa.rw[0];
        if (tid + kSpmvThreads < kRidWords) rid32[tid + kSpmvThreads] = pa.rw[1];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int u = tid + k * kSpmvThreads;
            if (u < 8 * pa.cnt) {
                const float tp = bf * vpa[k];
                cellv[u] = first ? va[k] : va[k] + tp;
            }
        }
        load_cells(pb, va, vpa);  // free again: the next tile's values
        pa = pb;
        load_plan(tile + 2 * tstep, pb);
        __syncthreads();
        // lanes of block g fill its 4x4x4 grid: cell (x, y, z) lies in parent
        // neighbour (x >> 1, y >> 1, z >> 1) at child slot (x & 1, y & 1, z & 1)
        // of the grid whose origin is the block's (-1, -1, -1) cell
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int cc = s + 8 * m;
            const int ax = (cc >> 4) + 1, ay = ((cc >> 2) & 3) + 1, az = (cc & 3) + 1;
            const int id = rid[27 * g + 9 * (ax >> 1) + 3 * (ay >> 1) + (az >> 1)];
            gbase[cc] = id < 0 ? 0.0f : cellv[8 * id + (((ax & 1) << 2) | ((ay & 1) << 1) | (az & 1))];
        }
        __syncthreads();
        // blocks past the last one have an empty plan (all-zero grids); their
        // stores fall outside the descriptors and are dropped, so the stores
        // are issued without a branch and the loads of the next tiles are
        // waited for by count
        const int64_t blk = tile * kTile + g;
        {
            // absent neighbours hold +0: adding L*0 leaves a float sum
            // unchanged, so the result equals the sum over present terms
            const float* gc = gbase + c0;
            float out = 0.0f;
#pragma unroll
            for (int j = 0; j < 27; j++) {
                const float prod = sg.L[j] * gc[16 * (j / 9) + 4 * ((j / 3) % 3) + (j % 3)];
                out = out + prod;
            }
            const uint32_t off = blk < nblk ? 4u * uint32_t(8 * blk + s) : kBufOob;
            const float pt = gc[16 + 4 + 1];
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(pt), rs_pnew, off, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(out), rs_q, off, 0, 0);
            const double pqt = pq + double(pt) * double(out);
            pq = blk < nblk ? pqt : pq;
        }
        __syncthreads();
    }
    const double sum = block_allsum<kSpmvThreads>(pq, lds);
    if (tid == 0) pq_part[bid] = sum;
}

// CG iteration `it`, second kernel: alpha = rr_{it-1} / p.q (every workgroup
// sums the p.q partials; p.q <= 0 after the residual vanished stops the
// solve), x += alpha p, r -= alpha q on 16-byte-aligned vectors four rows
// per lane, and this workgroup's partial of r.r (slot it & 1) unless r is
// recomputed next.
__global__ void __launch_bounds__(kVecThreads) k_cg_update(const Seg* __restrict__ segs, int32_t nseg, int32_t it,
                                                           int32_t refresh) {
    __shared__ double lds[kVecThreads / 64];
    const Seg& sg = segs[seg_of(segs, nseg, blockIdx.x, true)];
    CgState* st = sg.st;
    if (st->done) return;
    const int64_t n = sg.n;
    const float* p = ((it - 1) & 1) ? sg.P1 : sg.P0;
    const float* q = sg.Q;
    float* x = sg.XS;
    float* r = sg.R;
    double* rr_part = sg.rr + (it & 1) * kMaxWG;
    const bool refresh_next = (it % refresh) == 0;
    const int bid = int(blockIdx.x) - sg.vwg0, nblocks = sg.nvwg;
    const double PQ = sum_partials<kVecThreads>(sg.pq, sg.nwg, lds);
    if (!(PQ > 0.0)) {
        if (bid == 0 && threadIdx.x == 0) {
            st->done = 1;
            st->iters = it - 1;
        }
        return;
    }
    const float a = float(st->rr_hist[(it - 1) & 1] / PQ);
    double rr = 0.0;
    const int64_t n4 = n >> 2;
    const float4* p4 = reinterpret_cast<const float4*>(p);
    const float4* q4 = reinterpret_cast<const float4*>(q);
    float4* x4 = reinterpret_cast<float4*>(x);
    float4* r4 = reinterpret_cast<float4*>(r);
    for (int64_t t = int64_t(bid) * kVecThreads + threadIdx.x; t < n4; t += int64_t(nblocks) * kVecThreads) {
        const float4 pv = p4[t], qv = q4[t];
        float4 xv = x4[t], rv = r4[t];
        xv.x = xv.x + a * pv.x;
        xv.y = xv.y + a * pv.y;
        xv.z = xv.z + a * pv.z;
        xv.w = xv.w + a * pv.w;
        rv.x = rv.x - a * qv.x;
        rv.y = rv.y - a * qv.y;
        rv.z = rv.z - a * qv.z;
        rv.w = rv.w - a * qv.w;
        x4[t] = xv;
        r4[t] = rv;
        rr += double(rv.x) * double(rv.x);
        rr += double(rv.y) * double(rv.y);
        rr += double(rv.z) * double(rv.z);
        rr += double(rv.w) * double(rv.w);
    }
    for (int64_t t = 4 * n4 + int64_t(bid) * kVecThreads + threadIdx.x; t < n; t += int64_t(nblocks) * kVecThreads) {
        const float t1 = a * p[t];
        x[t] = x[t] + t1;
        const float t2 = a * q[t];
        const float rv = r[t] - t2;
        r[t] = rv;
        rr += double(rv) * double(rv);
    }
    if (!refresh_next) {
        const double sum = block_allsum<kVecThreads>(rr, lds);
        if (threadIdx.x == 0) rr_part[bid] = sum;
    }
    if (bid == 0 && threadIdx.x == 0) st->iters = it;
}

// r = b - A x (every `refresh` iterations) and its r.r partial (slot it & 1)
__global__ void __launch_bounds__(kVecThreads) k_cg_refresh(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                            const float* x, float* r, double* rr_part, CgState* st) {
    __shared__ double lds[kVecThreads / 64];
    if (st->done) return;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kVecThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kVecThreads) {
        const float ax = spmv_pnt(pnt, t, x, S);
        const float rv = b[t] - ax;
        r[t] = rv;
        rr += double(rv) * double(rv);
    }
    const double sum = block_allsum<kVecThreads>(rr, lds);
    if (threadIdx.x == 0) rr_part[blockIdx.x] = sum;
}

// ||b - A x||^2 of the final iterate, one FP64 partial per workgroup (FP32
// SpMV, FP64 sum of squares).
__global__ void __launch_bounds__(kVecThreads) k_true_residual(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                               const float* x, double* partial) {
    __shared__ double lds[kVecThreads / 64];
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kVecThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kVecThreads) {
        const float ax = spmv_pnt(pnt, t, x, S);
        const double rv = double(b[t]) - double(ax);
        rr += rv * rv;
    }
    const double sum = block_allsum<kVecThreads>(rr, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = sum;
}

// One workgroup: ||b - A x|| / ||b|| from the true-residual partials (nt of
// them), the recursive residual of the last iteration (its r.r partials, np
// of them, in slot iters & 1) and the iteration count.
__global__ void __launch_bounds__(kVecThreads) k_true_residual_finish(const double* partial, int32_t nt,
                                                                      const double* rr_slots, int32_t np,
                                                                      int32_t slot_stride, const CgState* st,
                                                                      double* out) {
    __shared__ double lds[kVecThreads / 64];
    const double RR = sum_partials<kVecThreads>(partial, nt, lds);
    const int32_t iters = st->iters;
    const double rec = sum_partials<kVecThreads>(rr_slots + int64_t(iters & 1) * slot_stride, np, lds);
    if (threadIdx.x == 0) {
        const double bb = st->bb;
        out[0] = (bb > 0.0) ? sqrt(RR / bb) : 0.0;
        out[1] = (bb > 0.0) ? sqrt(rec / bb) : 0.0;
        out[2] = double(iters);
        out[3] = double(st->restarted);
    }
}

// ---- small depths: one workgroup runs a whole CG ---------------------------------
struct SmallJobs {
    int32_t count;
    int32_t depth[kMaxDepth + 1];
    int64_t base[kMaxDepth + 1];
    int64_t n[kMaxDepth + 1];
    int64_t pnt_row[kMaxDepth + 1];  // first parent-neighbour row of the depth (-1 at the root)
};

// Row t of a small depth from v in LDS: neighbours from the row's 8
// parent-neighbour entries (depth >= 1; packed two int16 per word, each in
// -1 .. kSmallRows - 1; the root's only neighbour is itself), the terms and
// order of spmv_pnt.
__device__ __forceinline__ float spmv_small(const uint32_t ep[4], int s, bool root, const float* v, const float* L) {
    if (root) return (L[13] == 0.0f) ? 0.0f : 0.0f + L[13] * v[0];
    int32_t e[8];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        e[2 * m] = int32_t(int16_t(ep[m] & 0xffffu));
        e[2 * m + 1] = int32_t(int16_t(ep[m] >> 16));
    }
    int32_t nb[27];
    pnt_expand(e, s, nb);
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const float prod = L[j] * v[nb[j] < 0 ? 0 : nb[j]];
        out = (nb[j] < 0 || L[j] == 0.0f) ? out : out + prod;
    }
    return out;
}

// One workgroup runs a small depth's whole CG.  Its rows (at most
// kSmallRows) sit kSmallRows / kSmallBlock per thread: x, r, b, q in
// registers with each row's 8 parent-neighbour entries, and the vector the
// SpMV reads (p, or x for a residual) in LDS.  The arithmetic, the stop test
// and the reduction order (each thread's rows in order, then the workgroup
// sum) are the large depths'.
__global__ void __launch_bounds__(kSmallBlock) k_cg_small(SmallJobs jobs, const float* stencils, const int32_t* pnt_all,
                                                          const float* div, float* xall, int32_t warm_depth, float tol,
                                                          int32_t fixed, int32_t max_iters, int32_t refresh,
                                                          double* out) {
    constexpr int R = int(kSmallRows / kSmallBlock);
    static_assert(kSmallRows <= 32768, "parent-neighbour entries are packed as int16");
    __shared__ double lds[kSmallBlock / 64];
    __shared__ float L[27];
    __shared__ float sv[kSmallRows];  // p, or x while a residual is formed
    __shared__ float ps[kSmallRows];  // p of this thread's rows while sv holds x
    const int tid = threadIdx.x;
    const int job = blockIdx.x;
    const int d = jobs.depth[job];
    const int64_t base = jobs.base[job], n = jobs.n[job];
    const bool root = jobs.pnt_row[job] < 0;
    const int32_t* pnt = root ? nullptr : pnt_all + 27 * jobs.pnt_row[job];
    if (tid < 27) L[tid] = stencils[27 * d + tid];
    const float* bg = div + base;
    float* xg = xall + base;
    float x[R], r[R], b[R], q[R];
    uint32_t ep[R][4];
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t t = tid + int64_t(k) * kSmallBlock;
        const bool ok = t < n;
        b[k] = ok ? bg[t] : 0.0f;
        x[k] = (ok && d == warm_depth) ? xg[t] : 0.0f;  // zero start unless warm-started
        int32_t e[8];
        if (ok && !root) {
            pnt_entries(pnt + 27 * (t >> 3), int(t & 7), e);
        } else {
#pragma unroll
            for (int m = 0; m < 8; m++) e[m] = -1;
        }
#pragma unroll
        for (int m = 0; m < 4; m++) ep[k][m] = (uint32_t(e[2 * m]) & 0xffffu) | (uint32_t(e[2 * m + 1]) << 16);
        if (ok) sv[t] = x[k];
    }
    __syncthreads();  // x and L in LDS
    double bb = 0.0, rr = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t t = tid + int64_t(k) * kSmallBlock;
        if (t < n) {
            r[k] = b[k] - spmv_small(ep[k], int(t & 7), root, sv, L);
            bb += double(b[k]) * double(b[k]);
            rr += double(r[k]) * double(r[k]);
        }
    }
    bb = block_allsum<kSmallBlock>(bb, lds);  // every read of sv above is complete
    rr = block_allsum<kSmallBlock>(rr, lds);
    // a warm start worse than none (||b - A x0|| > ||b||) is replaced by x0 = 0:
    // r = b - (+0) = b and r.r = b.b, the state of a zero start
    const bool restarted = d == warm_depth && rr > bb;
    if (restarted) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            x[k] = 0.0f;
            r[k] = b[k];
        }
        rr = bb;
    }
    const double tol2 = double(tol) * double(tol) * bb;
    int it = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t t = tid + int64_t(k) * kSmallBlock;
        if (t < n) sv[t] = r[k];
    }
    while (bb > 0.0 && rr != 0.0 && it < max_iters && (fixed > 0 ? it < fixed : rr > tol2)) {
        __syncthreads();  // p complete in LDS
        double pq = 0.0;
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t t = tid + int64_t(k) * kSmallBlock;
            if (t < n) {
                q[k] = spmv_small(ep[k], int(t & 7), root, sv, L);
                pq += double(sv[t]) * double(q[k]);
            }
        }
        pq = block_allsum<kSmallBlock>(pq, lds);
        if (!(pq > 0.0)) break;
        const float af = float(rr / pq);
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t t = tid + int64_t(k) * kSmallBlock;
            if (t < n) {
                const float pk = sv[t];
                const float t1 = af * pk;
                x[k] = x[k] + t1;
                const float t2 = af * q[k];
                r[k] = r[k] - t2;
                ps[t] = pk;
            }
        }
        it++;
        if (refresh > 0 && it % refresh == 0) {
            __syncthreads();  // every row's p has been read from sv
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int64_t t = tid + int64_t(k) * kSmallBlock;
                if (t < n) sv[t] = x[k];
            }
            __syncthreads();  // x complete in LDS
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int64_t t = tid + int64_t(k) * kSmallBlock;
                if (t < n) r[k] = b[k] - spmv_small(ep[k], int(t & 7), root, sv, L);
            }
        }
        double rr2 = 0.0;
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t t = tid + int64_t(k) * kSmallBlock;
            if (t < n) rr2 += double(r[k]) * double(r[k]);
        }
        rr2 = block_allsum<kSmallBlock>(rr2, lds);  // every read of sv above is complete
        const float bf = float(rr2 / rr);
        rr = rr2;
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t t = tid + int64_t(k) * kSmallBlock;
            if (t < n) {
                const float tp = bf * ps[t];
                sv[t] = r[k] + tp;
            }
        }
    }
    __syncthreads();  // the last SpMV's reads of sv are complete
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t t = tid + int64_t(k) * kSmallBlock;
        if (t < n) sv[t] = x[k];
    }
    __syncthreads();
    double tr = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t t = tid + int64_t(k) * kSmallBlock;
        if (t < n) {
            const double rv = double(b[k]) - double(spmv_small(ep[k], int(t & 7), root, sv, L));
            tr += rv * rv;
            xg[t] = x[k];
        }
    }
    tr = block_allsum<kSmallBlock>(tr, lds);
    if (threadIdx.x == 0) {
        out[4 * d + 0] = (bb > 0.0) ? sqrt(tr / bb) : 0.0;
        out[4 * d + 1] = (bb > 0.0) ? sqrt(rr / bb) : 0.0;
        out[4 * d + 2] = double(it);
        out[4 * d + 3] = restarted ? 1.0 : 0.0;
    }
}

}  // namespace

namespace {

// A chain: a set of large depths solved together on one stream.  One SpMV and
// one update launch per iteration serve all of them (the per-depth systems are
// independent; each depth has its own scalars, partials and stop flag).
struct Chain {
    hipStream_t s = nullptr;
    std::vector<int> depths;
    int seg0 = 0;                // its first entry in the solve's segment table
    int32_t nwg = 0, nvwg = 0;   // workgroups of its SpMV and vector launches
    int inst = 0;                // SpMV instance (cells per thread)
    size_t lds = 0;              // SpMV dynamic LDS
    int issued = 0;              // iterations queued
    int batch = 24;
    bool waiting = false;
    bool ended = false;          // true residuals queued after the latest batch
    CgState* st = nullptr;       // device states, one per depth
    CgState* hst = nullptr;      // pinned host copies
    hipEvent_t ev = nullptr;     // marks the state copy
    std::vector<std::pair<int, int>> launched;
};

constexpr int kFirstBatch = 24;      // iterations queued before a chain's first state read-back
constexpr int kMaxSegs = 12;         // depths per solve above kSmallRows rows (D <= 10 has at most 10)
constexpr size_t kHostStates = 1024;  // pinned read-back area: states first, then the segment table
constexpr size_t kHostPlanMax = 3840;  // then the plans' per-depth largest tile counts (kMaxDepth + 1 int32)
constexpr size_t kHostPlanSum = 3888;  // and their per-depth sums of tile counts (kMaxDepth + 1 int64)
static_assert(kHostPlanMax + (kMaxDepth + 1) * 4 <= kHostPlanSum && kHostPlanSum % 8 == 0, "plan read-back layout");

int32_t* plan_maxcount_host(prg_ctx* c) {
    return reinterpret_cast<int32_t*>(static_cast<uint8_t*>(c->cg_host) + kHostPlanMax);
}
long long* plan_sumcount_host(prg_ctx* c) {
    return reinterpret_cast<long long*>(static_cast<uint8_t*>(c->cg_host) + kHostPlanSum);
}

// Bytes one SpMV launch must move for depth d in this layout: every tile's
// row-id words and count, the block bases of the distinct blocks the tiles
// list, r and p_(k-1) read and p_k and q written per row.
double spmv_plan_bytes(const prg_ctx* c, int d) {
    const int64_t nt = c->plan_tile_off[d + 1] - c->plan_tile_off[d];
    return double(nt) * (kTileEntries * sizeof(int16_t) + sizeof(int32_t)) + 4.0 * double(c->plan_cells[d]) +
           16.0 * double(c->counts[d]);
}

// SpMV instance: cells per thread covering the largest tile of the chain's
// depths.  A chain's launch is one round of workgroups (as many per CU as
// stay resident), shared among its depths by tile count, each pipelining a
// run of tiles.
int spmv_instance(int kcells) { return kcells <= 2 ? 0 : kcells <= 4 ? 1 : kcells <= 5 ? 4 : kcells <= 8 ? 2 : 3; }

// resident workgroups per CU of an instance with `lds` bytes of cell values
int spmv_wg_per_cu(int inst, size_t lds) {
    const void* fn = inst == 0   ? reinterpret_cast<const void*>(&k_cg_spmv<2>)
                     : inst == 1 ? reinterpret_cast<const void*>(&k_cg_spmv<4>)
                     : inst == 2 ? reinterpret_cast<const void*>(&k_cg_spmv<8>)
                     : inst == 4 ? reinterpret_cast<const void*>(&k_cg_spmv<5>)
                                 : reinterpret_cast<const void*>(&k_cg_spmv<27>);
    int n = 0;
    PRG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kSpmvThreads, lds));
    return std::max(1, n);
}

void launch_spmv(int inst, int32_t nwg, size_t lds, hipStream_t s, const Seg* segs, int32_t nseg, int32_t it,
                 float tol, int32_t fixed, int32_t max_iters) {
#define PRG_SPMV(KC) k_cg_spmv<KC><<<nwg, kSpmvThreads, lds, s>>>(segs, nseg, it, tol, fixed, max_iters)
    switch (inst) {
        case 0: PRG_SPMV(2); break;
        case 1: PRG_SPMV(4); break;
        case 2: PRG_SPMV(8); break;
        case 4: PRG_SPMV(5); break;
        default: PRG_SPMV(27); break;
    }
#undef PRG_SPMV
    PRG_CHECK_LAUNCH();
}

int64_t tiles_of(int64_t n) { return (n / 8 + kTile - 1) / kTile; }

}  // namespace

// Tile plans of every depth above kSmallRows rows, once per octree: on the
// second stream after the build (on the context stream when kernels are
// timed), with the per-depth largest tile count copied to pinned memory.
void start_plans(prg_ctx* c) {
    hipStream_t s = c->kt.enabled ? c->stream : c->stream2;
    const int D = c->D;
    int64_t off = 0;
    for (int d = 0; d <= D; d++) {
        c->plan_tile_off[d] = off;
        if (c->counts[d] > kSmallRows) off += tiles_of(c->counts[d]);
    }
    c->plan_tile_off[D + 1] = off;
    int32_t* cnt = c->plan_count.reserve(off + 1);
    int32_t* ub = c->plan_ubase.reserve(size_t(off) * kTileEntries + 1);
    int16_t* rw = c->plan_rowid.reserve(size_t(off) * kTileEntries + 1);
    // largest counts (kMaxDepth + 1 int32), then the sums (kMaxDepth + 1 int64, 8-aligned)
    constexpr int kSumAt = (kMaxDepth + 2) & ~1;
    int32_t* mx = c->plan_mx.reserve(kSumAt + 2 * (kMaxDepth + 1));
    long long* sm = reinterpret_cast<long long*>(mx + kSumAt);
    if (s != c->stream) {
        PRG_HIP(hipEventRecord(c->plan_ev, c->stream));
        PRG_HIP(hipStreamWaitEvent(s, c->plan_ev, 0));
    }
    // the leaf depth first: the vector field and divergence wait for its plan
    for (int k = 0; k <= D; k++) {
        const int d = k == 0 ? D : k - 1;
        const int64_t nt = c->plan_tile_off[d + 1] - c->plan_tile_off[d];
        if (c->counts[d] <= kSmallRows || nt == 0) continue;
        const int64_t t0 = c->plan_tile_off[d];
        TimedScope ts(c->kt, "spmv_tile_plan", 13.5 * double(c->counts[d]));
        k_tile_plan<<<unsigned(nt), kSpmvThreads, 0, s>>>(c->pnt.get() + 27 * c->pnt_off[d], c->counts[d] / 8,
                                                           cnt + t0, ub + t0 * kTileEntries, rw + t0 * kTileEntries);
        PRG_CHECK_LAUNCH();
    }
    {
        PlanOffsets P{};
        for (int d = 0; d <= D + 1; d++) P.off[d] = c->plan_tile_off[d];
        k_plan_max<<<unsigned(D + 1), 1024, 0, s>>>(cnt, P, mx, sm);
        PRG_CHECK_LAUNCH();
    }
    PRG_HIP(hipMemcpyAsync(plan_maxcount_host(c), mx, size_t(D + 1) * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipMemcpyAsync(plan_sumcount_host(c), sm, size_t(D + 1) * 8, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipEventRecord(c->plan_ev, s));
    c->plan_pending = true;
    c->plan_gen = c->tree_gen;
}

// plans started after the octree build, finished here (host wait)
void ensure_plans(prg_ctx* c) {
    if (c->plan_gen != c->tree_gen) start_plans(c);
    if (c->plan_pending) {
        PRG_HIP(hipEventSynchronize(c->plan_ev));
        for (int d = 0; d <= c->D; d++) {
            c->plan_maxcount[d] = plan_maxcount_host(c)[d];
            c->plan_cells[d] = plan_sumcount_host(c)[d];
        }
        c->plan_pending = false;
    }
}

bool tile_plan_on(prg_ctx* c, int d, TilePlan* P, hipStream_t s) {
    if (d < 0 || d > c->D || c->counts[d] <= kSmallRows) return false;
    if (c->plan_gen != c->tree_gen) start_plans(c);
    const int64_t t0 = c->plan_tile_off[d];
    if (c->plan_tile_off[d + 1] == t0) return false;
    if (c->plan_pending) PRG_HIP(hipStreamWaitEvent(s, c->plan_ev, 0));
    P->cnt = c->plan_count.get() + t0;
    P->ub = c->plan_ubase.get() + t0 * kTileEntries;
    P->rw = c->plan_rowid.get() + t0 * kTileEntries;
    return true;
}

prg_status solve_all(prg_ctx* c, prg_solve_stats* stats) {
    ensure_tables(c);
    StageTimer stage(c, PRG_STAGE_SOLVE);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M;
    ensure_plans(c);
    // every depth's solution is written by its solve (the large depths'
    // copies, the small depths' kernel); only a warm-started depth reads it first
    float* X = c->x.reserve(M);
    // the learned warm start predicts the depth of its network's grid; the
    // other depths start from zero
    int warm_depth = -1;
    if (c->cfg.warm_start == PRG_WARM_LEARNED) {
        if (!c->warm_net) throw Error{PRG_E_STATE, "PRG_WARM_LEARNED needs prg_set_warm_start_net"};
        int g = 0;
        while ((int64_t(1) << g) < prg_unet_grid(c->warm_net)) g++;
        if (g <= D && c->counts[g] > 0) warm_depth = g;
    }
    double* outv = c->cg_out.reserve(size_t(4) * (D + 1));
    PRG_HIP(hipMemsetAsync(outv, 0, size_t(4) * (D + 1) * 8, s));
    const float tol = c->cfg.cg_tol;
    const int32_t fixed = c->cfg.cg_fixed_iters, max_iters = c->cfg.cg_max_iters;
    const int32_t refresh = c->cfg.cg_refresh > 0 ? c->cfg.cg_refresh : (1 << 30);
    // With per-kernel timing on, everything runs on the context stream, one
    // chain after the other with the finest depth last, so each launch's
    // events time that launch alone.
    const bool serial = c->kt.enabled;

    // small depths: one launch, one workgroup each, on a stream of their own
    SmallJobs jobs{};
    for (int d = 0; d <= D; d++)
        if (c->counts[d] <= kSmallRows) {
            jobs.depth[jobs.count] = d;
            jobs.base[jobs.count] = c->bases[d];
            jobs.n[jobs.count] = c->counts[d];
            jobs.pnt_row[jobs.count] = d == 0 ? -1 : c->pnt_off[d];
            jobs.count++;
        }
    // chain 0: the finest depth; chain 1: every other large depth, together;
    // chain 2: a warm-started depth (other than the finest), after the
    // network's inference on its own stream
    constexpr int kChains = 3;
    Chain ch[kChains];
    const bool warm_large = warm_depth >= 0 && c->counts[warm_depth] > kSmallRows;
    for (int d = 0; d <= D; d++)
        if (c->counts[d] > kSmallRows) ch[d == D ? 0 : (warm_large && d == warm_depth ? 2 : 1)].depths.push_back(d);
    ch[0].s = s;
    ch[1].s = serial ? s : c->stream2;
    ch[2].s = serial ? s : c->stream4;
    hipStream_t small_s = serial ? s : c->stream3;

    // events come from the context's pool (created once, reused by every solve)
    size_t ev_used = 0;
    auto new_event = [&]() {
        if (ev_used == c->ev_pool.size()) {
            hipEvent_t e;
            PRG_HIP(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
        return c->ev_pool[ev_used++];
    };
    std::vector<hipEvent_t> dbeg(D + 1, nullptr), dend(D + 1, nullptr);
    std::vector<double> ms(D + 1, 0.0);
    std::vector<double> ov(size_t(4) * (D + 1), 0.0);
    prg_status result = PRG_OK;
    auto stencil_of = [&](int d) {
        Stencil S;
        for (int j = 0; j < 27; j++) S.L[j] = c->fs->stencil[size_t(d) * 27 + j];
        return S;
    };
    auto vec_grid = [&](int64_t n) {
        return int32_t(std::min<int64_t>(kMaxWG, (n + 4 * kVecThreads - 1) / (4 * kVecThreads)));
    };

    // segment table: chain 0's depth, then chain 1's depths
    int nsegs = 0;
    for (auto& h : ch) nsegs += int(h.depths.size());
    if (nsegs > kMaxSegs) throw Error{PRG_E_ARG, "too many large depths"};
    static_assert(kMaxSegs * sizeof(CgState) <= kHostStates, "chain states exceed their pinned area");
    static_assert(kHostStates + kMaxSegs * sizeof(Seg) <= kHostPlanMax, "segment table exceeds the pinned staging area");
    static_assert(kHostPlanSum + (kMaxDepth + 1) * 8 <= 4092, "plan counts exceed the pinned area (its last 4 bytes hold the owned-corner count)");
    CgState* hst = static_cast<CgState*>(c->cg_host);
    Seg* hseg = reinterpret_cast<Seg*>(static_cast<uint8_t*>(c->cg_host) + kHostStates);
    Seg* dseg = reinterpret_cast<Seg*>(c->cg_seg.reserve(std::max<size_t>(1, size_t(nsegs)) * sizeof(Seg)));
    {
        int si = 0;
        for (int k = 0; k < kChains; k++) {
            Chain& h = ch[k];
            h.seg0 = si;
            const int cnt = int(h.depths.size());
            if (!cnt) continue;
            // vectors: one 64-float-aligned segment per depth
            std::vector<int64_t> voff(cnt + 1, 0);
            int64_t tiles_all = 0;
            for (int i = 0; i < cnt; i++) {
                voff[i + 1] = voff[i] + ((c->counts[h.depths[i]] + 63) / 64) * 64;
                tiles_all += tiles_of(c->counts[h.depths[i]]);
            }
            float* V[5];
            for (int v = 0; v < 5; v++) V[v] = c->cg_vec[k][v].reserve(size_t(voff[cnt]));
            double* rr = c->cg_part[k][0].reserve(size_t(cnt) * 2 * kMaxWG);
            double* pq = c->cg_part[k][1].reserve(size_t(cnt) * kMaxSpmvWG);
            double* bb = c->cg_part[k][2].reserve(size_t(cnt) * std::max(kMaxWG, kMaxSpmvWG));
            h.st = reinterpret_cast<CgState*>(c->cg_state[k].reserve(size_t(cnt) * sizeof(CgState)));
            h.hst = hst + si;
            h.ev = new_event();
            // the chain's SpMV instance (cells per thread for its largest tile)
            // and workgroups: one resident round, shared by tile count
            int32_t kc = 1;
            for (int i = 0; i < cnt; i++)
                kc = std::max(kc, (8 * std::max(1, c->plan_maxcount[h.depths[i]]) + kSpmvThreads - 1) / kSpmvThreads);
            h.inst = spmv_instance(kc);
            size_t lds = 0;  // cell values of the largest tile among the chain's depths
            for (int i = 0; i < cnt; i++) lds = std::max(lds, size_t(8) * 4 * std::max(1, c->plan_maxcount[h.depths[i]]));
            const int64_t budget = int64_t(spmv_wg_per_cu(h.inst, lds)) * std::max(1, c->num_cu);
            int32_t wg = 0, vwg = 0;
            for (int i = 0; i < cnt; i++) {
                const int d = h.depths[i];
                const int64_t n = c->counts[d], nt = tiles_of(n);
                Seg& g = hseg[si + i];
                g.n = n;
                g.nwg = int32_t(std::min<int64_t>(
                    {int64_t(kMaxSpmvWG), nt, std::max<int64_t>(1, (budget * nt + tiles_all - 1) / tiles_all)}));
                g.wg0 = wg;
                wg += g.nwg;
                g.nvwg = vec_grid(n);
                g.vwg0 = vwg;
                vwg += g.nvwg;
                const int64_t t0 = c->plan_tile_off[d];
                g.cnt = c->plan_count.get() + t0;
                g.ub = c->plan_ubase.get() + t0 * kTileEntries;
                g.rw = c->plan_rowid.get() + t0 * kTileEntries;
                g.R = V[0] + voff[i];
                g.P0 = V[1] + voff[i];
                g.P1 = V[2] + voff[i];
                g.Q = V[3] + voff[i];
                g.XS = V[4] + voff[i];
                g.rr = rr + size_t(i) * 2 * kMaxWG;
                g.pq = pq + size_t(i) * kMaxSpmvWG;
                g.bb = bb + size_t(i) * std::max(kMaxWG, kMaxSpmvWG);
                g.st = h.st + i;
                const Stencil S = stencil_of(d);
                for (int j = 0; j < 27; j++) g.L[j] = S.L[j];
            }
            h.nwg = wg;
            h.nvwg = vwg;
            h.lds = lds;
            si += cnt;
        }
    }
    try {
        // the segment table and zeroed states, then the other streams fork
        // from the context stream
        // (k_cg_init resets each depth's state)
        if (nsegs) PRG_HIP(hipMemcpyAsync(dseg, hseg, size_t(nsegs) * sizeof(Seg), hipMemcpyHostToDevice, s));
        hipEvent_t es0 = new_event(), es1 = new_event();
        PRG_HIP(hipEventRecord(es0, s));
        for (hipStream_t x : {c->stream2, c->stream3, c->stream4}) PRG_HIP(hipStreamWaitEvent(x, es0, 0));
        // The network runs ahead of the solve of its depth: on that depth's
        // chain stream when the depth is large (the finest depth's chain or
        // chain 2), else on the small depths' stream.  A chain on the same
        // stream simply starts after it.
        hipEvent_t ew = nullptr;
        if (warm_depth >= 0) {
            hipStream_t ws = !warm_large ? small_s : (warm_depth == D ? ch[0].s : ch[2].s);
            unet_warm_start(c, warm_depth, ws);
            ew = new_event();
            PRG_HIP(hipEventRecord(ew, ws));
        }
        if (jobs.count) {
            es0 = new_event();
            PRG_HIP(hipEventRecord(es0, small_s));
            k_cg_small<<<jobs.count, kSmallBlock, 0, small_s>>>(jobs, c->d_stencil.get(), c->pnt.get(),
                                                                c->div.get(), X, warm_depth, tol, fixed, max_iters,
                                                                refresh, outv);
            PRG_CHECK_LAUNCH();
            PRG_HIP(hipEventRecord(es1, small_s));
        }

        auto begin_chain = [&](Chain& h) {
            if (ew && std::find(h.depths.begin(), h.depths.end(), warm_depth) != h.depths.end())
                PRG_HIP(hipStreamWaitEvent(h.s, ew, 0));
            for (size_t i = 0; i < h.depths.size(); i++) {
                const int d = h.depths[i];
                const Seg& g = hseg[h.seg0 + i];
                const int64_t n = c->counts[d], base = c->bases[d];
                dbeg[d] = new_event();
                dend[d] = new_event();
                PRG_HIP(hipEventRecord(dbeg[d], h.s));
                // b read, xs, r, p written; x0 read and A x0 formed unless x0 = 0
                TimedScope ts(c->kt, "cg_init", (d == warm_depth ? kPntRowBytes + 20.0 : 16.0) * double(n));
                k_cg_init<<<g.nvwg, kVecThreads, 0, h.s>>>(c->pnt.get() + 27 * c->pnt_off[d], n, stencil_of(d),
                                                          c->div.get() + base, X + base, d == warm_depth ? 0 : 1,
                                                          g.XS, g.R, g.P0, g.rr, g.bb, g.st);
                PRG_CHECK_LAUNCH();
                if (d == warm_depth) {
                    k_warm_reset_decide<<<1, kVecThreads, 0, h.s>>>(g.rr, g.bb, g.nvwg, g.st);
                    PRG_CHECK_LAUNCH();
                    k_warm_reset_apply<<<g.nvwg, kVecThreads, 0, h.s>>>(n, c->div.get() + base, g.XS, g.R, g.P0,
                                                                        g.rr, g.bb, g.st);
                    PRG_CHECK_LAUNCH();
                }
            }
            h.issued = 0;
            h.batch = kFirstBatch;
        };
        // iteration `it` (1-based) of a chain: SpMV, update, and the residual
        // refresh when due
        auto launch_iter = [&](Chain& h, int it) {
            const int32_t cnt = int32_t(h.depths.size());
            int64_t rows = 0;
            double spmv_bytes = 0.0;
            for (int d : h.depths) {
                rows += c->counts[d];
                spmv_bytes += spmv_plan_bytes(c, d);
            }
            const char* spmv_name = h.depths.back() == D ? "stencil_spmv_dot_finest" : "stencil_spmv_dot";
            {
                TimedScope ts(c->kt, spmv_name, spmv_bytes);
                h.launched.emplace_back(ts.idx, it);
                // the instance (cells per thread) that ran, counted as its own record
                static const char* kInstName[5] = {"spmv_cells_2", "spmv_cells_4", "spmv_cells_8", "spmv_cells_27",
                                                   "spmv_cells_5"};
                c->kt.count(kInstName[h.inst < 0 || h.inst > 4 ? 3 : h.inst]);
                launch_spmv(h.inst, h.nwg, h.lds, h.s, dseg + h.seg0, cnt, it, tol, fixed, max_iters);
            }
            {
                TimedScope ts(c->kt, "cg_update", 24.0 * double(rows));
                h.launched.emplace_back(ts.idx, it);
                k_cg_update<<<h.nvwg, kVecThreads, 0, h.s>>>(dseg + h.seg0, cnt, it, refresh);
                PRG_CHECK_LAUNCH();
            }
            if ((it % refresh) == 0) {
                for (int i = 0; i < cnt; i++) {
                    const int d = h.depths[i];
                    const Seg& g = hseg[h.seg0 + i];
                    TimedScope ts(c->kt, "cg_refresh", kResidualRowBytes * double(g.n));
                    h.launched.emplace_back(ts.idx, it);
                    k_cg_refresh<<<g.nvwg, kVecThreads, 0, h.s>>>(c->pnt.get() + 27 * c->pnt_off[d], g.n,
                                                                 stencil_of(d), c->div.get() + c->bases[d], g.XS,
                                                                 g.R, g.rr + (it & 1) * kMaxWG, g.st);
                    PRG_CHECK_LAUNCH();
                }
            }
        };
        // the chain's states read back after its queued iterations
        auto end_batch = [&](Chain& h) {
            h.issued += h.batch;
            PRG_HIP(hipMemcpyAsync(h.hst, h.st, h.depths.size() * sizeof(CgState), hipMemcpyDeviceToHost, h.s));
            PRG_HIP(hipEventRecord(h.ev, h.s));
            h.waiting = true;
        };
        auto issue_batch = [&](Chain& h) {
            h.launched.clear();
            for (int k = 0; k < h.batch; k++) launch_iter(h, h.issued + k + 1);
            end_batch(h);
        };
        auto end_chain = [&](Chain& h) {
            for (size_t i = 0; i < h.depths.size(); i++) {
                const int d = h.depths[i];
                const Seg& g = hseg[h.seg0 + i];
                const int64_t n = c->counts[d], base = c->bases[d];
                {
                    TimedScope ts(c->kt, "cg_true_residual", kResidualRowBytes * double(n));
                    k_true_residual<<<g.nvwg, kVecThreads, 0, h.s>>>(c->pnt.get() + 27 * c->pnt_off[d], n,
                                                                    stencil_of(d), c->div.get() + base, g.XS, g.bb);
                    PRG_CHECK_LAUNCH();
                }
                k_true_residual_finish<<<1, kVecThreads, 0, h.s>>>(g.bb, g.nvwg, g.rr, g.nvwg, kMaxWG, g.st,
                                                                  outv + 4 * d);
                PRG_CHECK_LAUNCH();
                PRG_HIP(hipMemcpyAsync(X + base, g.XS, size_t(n) * 4, hipMemcpyDeviceToDevice, h.s));
                PRG_HIP(hipEventRecord(dend[d], h.s));
            }
        };

        // Each chain: init, batches of iterations with a state read-back after
        // each, true residuals; the host polls the chains so none waits for
        // another.  Serial mode runs chain 2, then 1, the finest depth last.
        std::vector<int> serial_order;
        for (int k = kChains - 1; k >= 0; k--)
            if (!ch[k].depths.empty()) serial_order.push_back(k);
        size_t serial_next = 0;
        auto start_chain = [&](int k) {
            begin_chain(ch[k]);
            issue_batch(ch[k]);
        };
        if (serial) {
            if (!serial_order.empty()) start_chain(serial_order[serial_next++]);
        } else {
            // every chain's init, then the first batches one iteration at a
            // time across the chains, so no chain waits for another's launches
            for (int k = 0; k < kChains; k++)
                if (!ch[k].depths.empty()) {
                    begin_chain(ch[k]);
                    ch[k].launched.clear();
                }
            for (int i = 0; i < kFirstBatch; i++)
                for (int k = 0; k < kChains; k++)
                    if (!ch[k].depths.empty()) launch_iter(ch[k], i + 1);
            // The first batch usually covers the whole solve, so each chain's
            // true residuals and solution copy go right behind it (they are
            // queued again, and supersede these, if more iterations follow);
            // the end of a chain then waits for no host poll.
            for (int k = 0; k < kChains; k++)
                if (!ch[k].depths.empty()) {
                    end_batch(ch[k]);
                    end_chain(ch[k]);
                    ch[k].ended = true;
                }
        }
        while (true) {
            bool active = false;
            for (int k = 0; k < kChains; k++) {
                Chain& h = ch[k];
                if (!h.waiting) continue;
                active = true;
                const hipError_t q = hipEventQuery(h.ev);
                if (q == hipErrorNotReady) continue;
                PRG_HIP(q);
                h.waiting = false;
                bool all_done = true;
                int32_t last = 0;
                for (size_t i = 0; i < h.depths.size(); i++) {
                    all_done = all_done && (h.hst[i].done || h.issued >= max_iters);
                    last = std::max(last, h.hst[i].iters);
                }
                for (const auto& l : h.launched)
                    if (l.second > last && all_done) c->kt.relabel(l.first, "cg_idle");
                if (all_done) {
                    if (!h.ended) end_chain(h);
                    if (serial && serial_next < serial_order.size()) start_chain(serial_order[serial_next++]);
                } else {
                    h.batch = 6;
                    issue_batch(h);
                    h.ended = !serial;
                    if (h.ended) end_chain(h);
                }
            }
            if (!active) break;
        }
        if (jobs.count) {
            PRG_HIP(hipEventSynchronize(es1));
            float t = 0.0f;
            PRG_HIP(hipEventElapsedTime(&t, es0, es1));
            for (int k = 0; k < jobs.count; k++) ms[jobs.depth[k]] = t;  // solved together in one launch
            c->kt.recs["cg_small"].launches += 1;
            c->kt.recs["cg_small"].total_ms += t;
        }
        for (int d = 0; d <= D; d++)
            if (dbeg[d]) {
                PRG_HIP(hipEventSynchronize(dend[d]));
                float t = 0.0f;
                PRG_HIP(hipEventElapsedTime(&t, dbeg[d], dend[d]));
                ms[d] = t;
            }
        // the context stream continues after every chain and the small solves
        for (hipStream_t x : {c->stream2, c->stream3, c->stream4}) {
            hipEvent_t ej = new_event();
            PRG_HIP(hipEventRecord(ej, x));
            PRG_HIP(hipStreamWaitEvent(s, ej, 0));
        }
        PRG_HIP(hipMemcpyAsync(ov.data(), outv, ov.size() * 8, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
    } catch (...) {
        for (hipStream_t x : {c->stream2, c->stream3, c->stream4}) (void)hipStreamSynchronize(x);
        (void)hipStreamSynchronize(s);
        throw;
    }
    c->warm_depth_used = warm_depth;
    c->warm_rejected = warm_depth >= 0 && ov[4 * warm_depth + 3] > 0.5;
    for (int d = 0; d <= D; d++) {
        const int32_t iters = int32_t(ov[4 * d + 2]);
        if (stats) {
            stats[d].depth = d;
            stats[d].nodes = c->counts[d];
            stats[d].iters = iters;
            stats[d].rel_residual = ov[4 * d + 1];
            stats[d].true_rel_residual = ov[4 * d + 0];
            stats[d].ms = ms[d];
        }
        if (fixed <= 0 && iters >= max_iters && ov[4 * d + 1] > double(tol)) result = PRG_E_NOT_CONVERGED;
    }
    return result;
}

}  // namespace prg
