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
//                p.q; the last workgroup (atomic ticket) reduces the partials
//                in a fixed order and sets alpha;
//   k_cg_update  x += alpha p, r -= alpha q, FP64 partials of r.r; the last
//                workgroup sets beta, rr and the stop flag.
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

#include "ctx.h"

namespace prg {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kSmallBlock = 1024;
constexpr int64_t kSmallRows = 4096;
// Algorithmic bytes per SpMV row under the reference's CSR-free representation
// (int32 neighbour indices, FP32 values): 108 B of neighbour indices, p_i read,
// (Ap)_i written.  The kernels here read 27 parent-neighbour entries per sibling
// block instead (13.5 B per row), so their compulsory traffic is lower.
constexpr double kSpmvRowBytes = 116.0;

struct Stencil {
    float L[27];
};

struct CgState {
    double bb, rr, pq;
    float alpha, beta;
    int32_t done, iters;
    uint32_t ticket;
    int32_t pad;
};

// neighbour j (compile-time after unrolling) of slot s in the block whose
// parent-neighbour row is `row`; -1 if absent.
__device__ __forceinline__ int32_t pnt_neighbor(const int32_t* __restrict__ row, int s, int j) {
    const int vx = ((s >> 2) & 1) + (j / 9 - 1);
    const int vy = ((s >> 1) & 1) + ((j / 3) % 3 - 1);
    const int vz = (s & 1) + (j % 3 - 1);
    const int pj = 9 * ((vx >> 1) + 1) + 3 * ((vy >> 1) + 1) + ((vz >> 1) + 1);
    const int cs = ((vx & 1) << 2) | ((vy & 1) << 1) | (vz & 1);
    const int32_t cb = row[pj];
    return cb < 0 ? -1 : cb + cs;
}

// y_t = sum_j L[j] * v[neighbour_j(t)]  (gpuSpMV order, float)
__device__ __forceinline__ float spmv_pnt(const int32_t* __restrict__ pnt, int64_t t, const float* __restrict__ v,
                                          const Stencil& S) {
    const int32_t* row = pnt + 27 * (t >> 3);
    const int s = int(t & 7);
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int32_t nb = pnt_neighbor(row, s, j);
        if (nb < 0 || S.L[j] == 0.0f) continue;
        const float prod = S.L[j] * v[nb];
        out = out + prod;
    }
    return out;
}

// p_k at row t: r + beta * p_{k-1}
__device__ __forceinline__ float pdir(const float* __restrict__ r, const float* __restrict__ pold, float bf, int64_t t) {
    const float tp = bf * pold[t];
    return r[t] + tp;
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

// Publishes this workgroup's partial; returns true in every thread of the
// last workgroup to finish, after which `partial` is complete.
__device__ __forceinline__ bool publish(double blk, double* partial, CgState* st, unsigned* flag) {
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = blk;
        __threadfence();
        const unsigned t = atomicAdd(&st->ticket, 1u);
        *flag = (t == gridDim.x - 1) ? 1u : 0u;
    }
    __syncthreads();
    const bool last = *flag != 0u;
    if (last) __threadfence();
    return last;
}

// Sum of the partials by the last workgroup: fixed per-thread strides, fixed tree.
__device__ __forceinline__ double gather_partials(const double* partial, int nb, double* lds) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += ((volatile const double*)partial)[b];
    return block_allsum<kBlock>(s, lds);
}

__global__ void __launch_bounds__(kBlock) k_cg_init(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                    const float* x, float* r, float* p, double* partial,
                                                    double* partial2, CgState* st, float tol, int32_t fixed) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned flag;
    double bb = 0.0, rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_pnt(pnt, t, x, S);
        const float rv = b[t] - ax;
        r[t] = rv;
        p[t] = rv;
        bb += double(b[t]) * double(b[t]);
        rr += double(rv) * double(rv);
    }
    const double sbb = block_allsum<kBlock>(bb, lds);
    const double srr = block_allsum<kBlock>(rr, lds);
    if (threadIdx.x == 0) partial2[blockIdx.x] = sbb;
    if (publish(srr, partial, st, &flag)) {
        const double RR = gather_partials(partial, gridDim.x, lds);
        const double BB = gather_partials(partial2, gridDim.x, lds);
        if (threadIdx.x == 0) {
            st->bb = BB;
            st->rr = RR;
            st->iters = 0;
            st->beta = 0.0f;
            const double tol2 = double(tol) * double(tol) * BB;
            st->done = (BB <= 0.0) || (RR == 0.0) || (fixed > 0 ? 0 : (RR <= tol2));
            st->ticket = 0u;
        }
    }
}

// q = A p_k, LDS-staged per sibling block.  Eight lanes own the eight rows of
// a sibling block; together they load the block's 4x4x4 cell neighbourhood
// (64 cells reached through the block's 27 parent-neighbour entries), form
// p_k = r + beta p_{k-1} for each cell once (first != 0: p_1 = r is already
// in pnew), and then every lane applies its 27-point stencil from LDS in the
// order j = 0..26.  About 20 global loads per row instead of 81.
constexpr int kBlkPerWG = kBlock / 8;
__global__ void __launch_bounds__(kBlock) k_cg_spmv(const int32_t* pnt, int64_t n, Stencil S, const float* r,
                                                    const float* pold, float* pnew, float* q, int32_t first,
                                                    double* partial, CgState* st) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned flag;
    __shared__ int32_t prow[kBlkPerWG][28];
    __shared__ float cell[kBlkPerWG][64];
    if (st->done) return;
    const float bf = st->beta;
    const int g = threadIdx.x >> 3, s = threadIdx.x & 7;
    const int64_t nblk = n >> 3;
    double pq = 0.0;
    for (int64_t b0 = int64_t(blockIdx.x) * kBlkPerWG; b0 < nblk; b0 += int64_t(gridDim.x) * kBlkPerWG) {
        const int64_t blk = b0 + g;
        const bool valid = blk < nblk;
        if (valid) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int e = s + 8 * m;
                if (e < 27) prow[g][e] = pnt[27 * blk + e];
            }
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const int c = s + 8 * m;  // local cell (lx, ly, lz) = (c >> 4, (c >> 2) & 3, c & 3)
                const int ax = (c >> 4) + 1, ay = ((c >> 2) & 3) + 1, az = (c & 3) + 1;
                const int32_t cb = prow[g][9 * (ax >> 1) + 3 * (ay >> 1) + (az >> 1)];
                float v = 0.0f;
                if (cb >= 0) {
                    const int64_t idx = cb + (((ax & 1) << 2) | ((ay & 1) << 1) | (az & 1));
                    v = first ? pnew[idx] : pdir(r, pold, bf, idx);
                }
                cell[g][c] = v;
            }
        }
        __syncthreads();
        if (valid) {
            const int sx = (s >> 2) & 1, sy = (s >> 1) & 1, sz = s & 1;
            float out = 0.0f;
#pragma unroll
            for (int j = 0; j < 27; j++) {
                const int ax = 2 + sx + (j / 9 - 1), ay = 2 + sy + ((j / 3) % 3 - 1), az = 2 + sz + (j % 3 - 1);
                if (prow[g][9 * (ax >> 1) + 3 * (ay >> 1) + (az >> 1)] < 0 || S.L[j] == 0.0f) continue;
                const float prod = S.L[j] * cell[g][(ax - 1) * 16 + (ay - 1) * 4 + (az - 1)];
                out = out + prod;
            }
            const int64_t t = 8 * blk + s;
            const float pt = cell[g][(1 + sx) * 16 + (1 + sy) * 4 + (1 + sz)];
            if (!first) pnew[t] = pt;
            q[t] = out;
            pq += double(pt) * double(out);
        }
        __syncthreads();
    }
    const double sum = block_allsum<kBlock>(pq, lds);
    if (publish(sum, partial, st, &flag)) {
        const double PQ = gather_partials(partial, gridDim.x, lds);
        if (threadIdx.x == 0) {
            st->pq = PQ;
            st->alpha = float(st->rr / PQ);
            if (!(PQ > 0.0)) st->done = 1;  // breakdown once the residual has vanished
            st->ticket = 0u;
        }
    }
}

__device__ __forceinline__ void finish_iteration(CgState* st, double rr_new, float tol, int32_t fixed,
                                                 int32_t max_iters, bool count) {
    if (count) st->iters += 1;
    st->beta = float(rr_new / st->rr);
    st->rr = rr_new;
    const double tol2 = double(tol) * double(tol) * st->bb;
    const bool stop =
        (st->iters >= max_iters) || (rr_new == 0.0) || (fixed > 0 ? (st->iters >= fixed) : (rr_new <= tol2));
    st->done = stop ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_cg_update(int64_t n, const float* p, const float* q, float* x, float* r,
                                                      double* partial, CgState* st, float tol, int32_t fixed,
                                                      int32_t max_iters, int32_t refresh_next) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned flag;
    if (st->done) return;
    const float a = st->alpha;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float t1 = a * p[t];
        x[t] = x[t] + t1;
        const float t2 = a * q[t];
        const float rv = r[t] - t2;
        r[t] = rv;
        rr += double(rv) * double(rv);
    }
    const double sum = refresh_next ? 0.0 : block_allsum<kBlock>(rr, lds);
    if (publish(sum, partial, st, &flag)) {
        if (refresh_next) {
            // the residual is recomputed next; only count the iteration
            if (threadIdx.x == 0) {
                st->iters += 1;
                st->ticket = 0u;
            }
            return;
        }
        const double RR = gather_partials(partial, gridDim.x, lds);
        if (threadIdx.x == 0) {
            finish_iteration(st, RR, tol, fixed, max_iters, true);
            st->ticket = 0u;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_cg_refresh(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                       const float* x, float* r, double* partial, CgState* st,
                                                       float tol, int32_t fixed, int32_t max_iters) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned flag;
    if (st->done) return;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_pnt(pnt, t, x, S);
        const float rv = b[t] - ax;
        r[t] = rv;
        rr += double(rv) * double(rv);
    }
    const double sum = block_allsum<kBlock>(rr, lds);
    if (publish(sum, partial, st, &flag)) {
        const double RR = gather_partials(partial, gridDim.x, lds);
        if (threadIdx.x == 0) {
            finish_iteration(st, RR, tol, fixed, max_iters, false);
            st->ticket = 0u;
        }
    }
}

// ||b - A x|| / ||b|| of the final iterate (FP32 SpMV, FP64 norm)
__global__ void __launch_bounds__(kBlock) k_true_residual(const int32_t* pnt, int64_t n, Stencil S, const float* b,
                                                          const float* x, double* partial, CgState* st,
                                                          double* out) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned flag;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_pnt(pnt, t, x, S);
        const double rv = double(b[t]) - double(ax);
        rr += rv * rv;
    }
    const double sum = block_allsum<kBlock>(rr, lds);
    if (publish(sum, partial, st, &flag)) {
        const double RR = gather_partials(partial, gridDim.x, lds);
        if (threadIdx.x == 0) {
            out[0] = (st->bb > 0.0) ? sqrt(RR / st->bb) : 0.0;
            out[1] = (st->bb > 0.0) ? sqrt(st->rr / st->bb) : 0.0;
            out[2] = double(st->iters);
            st->ticket = 0u;
        }
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

// row t of a small depth: neighbours from the parent-neighbour table (depth
// >= 1); the root's only neighbour is itself.
__device__ __forceinline__ float spmv_small(const int32_t* __restrict__ pnt, int64_t t, const float* __restrict__ v,
                                            const float* __restrict__ L) {
    if (!pnt) return (L[13] == 0.0f) ? 0.0f : 0.0f + L[13] * v[0];
    const int32_t* row = pnt + 27 * (t >> 3);
    const int s = int(t & 7);
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int32_t nb = pnt_neighbor(row, s, j);
        if (nb < 0 || L[j] == 0.0f) continue;
        const float prod = L[j] * v[nb];
        out = out + prod;
    }
    return out;
}

__global__ void __launch_bounds__(kSmallBlock) k_cg_small(SmallJobs jobs, const float* stencils, const int32_t* pnt_all,
                                                          const float* div, float* xall, float* rall,
                                                          float* pall, float* qall, float tol, int32_t fixed,
                                                          int32_t max_iters, int32_t refresh, double* out) {
    __shared__ double lds[kSmallBlock / 64];
    __shared__ float L[27];
    const int job = blockIdx.x;
    const int d = jobs.depth[job];
    const int64_t base = jobs.base[job], n = jobs.n[job];
    const int32_t* pnt = jobs.pnt_row[job] < 0 ? nullptr : pnt_all + 27 * jobs.pnt_row[job];
    if (threadIdx.x < 27) L[threadIdx.x] = stencils[27 * d + threadIdx.x];
    __syncthreads();
    const float* b = div + base;
    float* x = xall + base;
    float* r = rall + base;
    float* p = pall + base;
    float* q = qall + base;
    double bb = 0.0, rr = 0.0;
    for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) {
        const float rv = b[t] - spmv_small(pnt, t, x, L);
        r[t] = rv;
        p[t] = rv;
        bb += double(b[t]) * double(b[t]);
        rr += double(rv) * double(rv);
    }
    bb = block_allsum<kSmallBlock>(bb, lds);
    rr = block_allsum<kSmallBlock>(rr, lds);
    const double tol2 = double(tol) * double(tol) * bb;
    int it = 0;
    while (bb > 0.0 && rr != 0.0 && it < max_iters && (fixed > 0 ? it < fixed : rr > tol2)) {
        __syncthreads();  // p complete
        double pq = 0.0;
        for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) {
            const float qv = spmv_small(pnt, t, p, L);
            q[t] = qv;
            pq += double(p[t]) * double(qv);
        }
        pq = block_allsum<kSmallBlock>(pq, lds);
        if (!(pq > 0.0)) break;
        const float af = float(rr / pq);
        for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) {
            const float t1 = af * p[t];
            x[t] = x[t] + t1;
            const float t2 = af * q[t];
            r[t] = r[t] - t2;
        }
        it++;
        if (refresh > 0 && it % refresh == 0) {
            __syncthreads();  // x complete
            for (int64_t t = threadIdx.x; t < n; t += kSmallBlock)
                r[t] = b[t] - spmv_small(pnt, t, x, L);
        }
        double rr2 = 0.0;
        for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) rr2 += double(r[t]) * double(r[t]);
        rr2 = block_allsum<kSmallBlock>(rr2, lds);
        const float bf = float(rr2 / rr);
        rr = rr2;
        for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) {
            const float tp = bf * p[t];
            p[t] = r[t] + tp;
        }
    }
    __syncthreads();
    double tr = 0.0;
    for (int64_t t = threadIdx.x; t < n; t += kSmallBlock) {
        const double rv = double(b[t]) - double(spmv_small(pnt, t, x, L));
        tr += rv * rv;
    }
    tr = block_allsum<kSmallBlock>(tr, lds);
    if (threadIdx.x == 0) {
        out[4 * d + 0] = (bb > 0.0) ? sqrt(tr / bb) : 0.0;
        out[4 * d + 1] = (bb > 0.0) ? sqrt(rr / bb) : 0.0;
        out[4 * d + 2] = double(it);
    }
}

}  // namespace

prg_status solve_all(prg_ctx* c, prg_solve_stats* stats) {
    ensure_tables(c);
    StageTimer stage(c, PRG_STAGE_SOLVE);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M;
    float* X = c->x.reserve(M);
    if (c->cfg.warm_start == PRG_WARM_NONE) PRG_HIP(hipMemsetAsync(X, 0, size_t(M) * 4, s));
    DBuf<float> R, P0, P1, Q, Rs, Ps, Qs;
    R.reserve(M);
    P0.reserve(M);
    P1.reserve(M);
    Q.reserve(M);
    DBuf<double> part, part2, outv;
    part.reserve(kMaxBlocks);
    part2.reserve(kMaxBlocks);
    outv.reserve(size_t(4) * (D + 1));
    DBuf<CgState> st;
    st.reserve(1);
    PRG_HIP(hipMemsetAsync(st.get(), 0, sizeof(CgState), s));
    PRG_HIP(hipMemsetAsync(outv.get(), 0, size_t(4) * (D + 1) * 8, s));
    const float tol = c->cfg.cg_tol;
    const int32_t fixed = c->cfg.cg_fixed_iters, max_iters = c->cfg.cg_max_iters;
    const int32_t refresh = c->cfg.cg_refresh > 0 ? c->cfg.cg_refresh : (1 << 30);
    std::vector<double> ms(D + 1, 0.0);
    hipEvent_t e0, e1, es0, es1;
    PRG_HIP(hipEventCreate(&e0));
    PRG_HIP(hipEventCreate(&e1));
    PRG_HIP(hipEventCreate(&es0));
    PRG_HIP(hipEventCreate(&es1));
    CgState* hst = nullptr;
    PRG_HIP(hipHostMalloc(&hst, sizeof(CgState), hipHostMallocDefault));
    prg_status result = PRG_OK;
    std::vector<double> ov(size_t(4) * (D + 1), 0.0);
    try {
        // small depths: one launch, one workgroup each
        SmallJobs jobs{};
        for (int d = 0; d <= D; d++)
            if (c->counts[d] <= kSmallRows) {
                jobs.depth[jobs.count] = d;
                jobs.base[jobs.count] = c->bases[d];
                jobs.n[jobs.count] = c->counts[d];
                jobs.pnt_row[jobs.count] = d == 0 ? -1 : c->pnt_off[d];
                jobs.count++;
            }
        // The small depths run on the second stream, overlapping the large
        // ones; they get their own scratch vectors (indexed by global node).
        if (jobs.count) {
            const int64_t small_end = jobs.base[jobs.count - 1] + jobs.n[jobs.count - 1];
            Rs.reserve(small_end);
            Ps.reserve(small_end);
            Qs.reserve(small_end);
            hipStream_t s2 = c->stream2;
            PRG_HIP(hipEventRecord(es0, s));
            PRG_HIP(hipStreamWaitEvent(s2, es0, 0));
            k_cg_small<<<jobs.count, kSmallBlock, 0, s2>>>(jobs, c->d_stencil.get(), c->pnt.get(), c->div.get(), X,
                                                           Rs.get(), Ps.get(), Qs.get(), tol, fixed, max_iters, refresh,
                                                           outv.get());
            PRG_CHECK_LAUNCH();
            PRG_HIP(hipEventRecord(es1, s2));
        }
        for (int d = 0; d <= D; d++) {
            const int64_t n = c->counts[d], base = c->bases[d];
            if (n <= kSmallRows) continue;
            Stencil S;
            for (int j = 0; j < 27; j++) S.L[j] = c->fs->stencil[size_t(d) * 27 + j];
            const unsigned nbk = unsigned(std::min<int64_t>(kMaxBlocks, (n + kBlock - 1) / kBlock));
            // the LDS-staged SpMV walks sibling blocks: 32 per workgroup
            const unsigned nbk_spmv = unsigned(std::min<int64_t>(kMaxBlocks, (n / 8 + kBlkPerWG - 1) / kBlkPerWG));
            const int32_t* pnt = c->pnt.get() + 27 * c->pnt_off[d];
            const float* B = c->div.get() + base;
            float* Xd = X + base;
            float* pb[2] = {P0.get(), P1.get()};
            PRG_HIP(hipEventRecord(e0, s));
            {
                TimedScope ts(c->kt, "cg_init", kSpmvRowBytes * double(n));
                k_cg_init<<<nbk, kBlock, 0, s>>>(pnt, n, S, B, Xd, R.get(), pb[0], part.get(), part2.get(), st.get(),
                                                 tol, fixed);
                PRG_CHECK_LAUNCH();
            }
            int issued = 0;
            int batch = 24;
            // (timer record, iteration) of this batch's launches: those past the
            // final iteration return at once and are filed as "cg_idle"
            std::vector<std::pair<int, int>> launched;
            const char* spmv_name = d == D ? "stencil_spmv_dot_finest" : "stencil_spmv_dot";
            while (true) {
                launched.clear();
                for (int k = 0; k < batch; k++) {
                    const int it = issued + k + 1;  // iteration number, 1-based
                    const bool first = (it == 1);
                    float* pnew = pb[(it - 1) & 1];
                    const float* pold = pb[it & 1];
                    {
                        // algorithmic bytes of one SpMV row in the reference's representation:
                        // 27 int32 neighbour indices, p_i read, (Ap)_i written (116 B)
                        TimedScope ts(c->kt, spmv_name, kSpmvRowBytes * double(n));
                        launched.emplace_back(ts.idx, it);
                        k_cg_spmv<<<nbk_spmv, kBlock, 0, s>>>(pnt, n, S, R.get(), pold, pnew, Q.get(),
                                                              first ? 1 : 0, part.get(), st.get());
                        PRG_CHECK_LAUNCH();
                    }
                    const bool do_refresh = (it % refresh) == 0;
                    {
                        TimedScope ts(c->kt, "cg_update", 24.0 * double(n));
                        launched.emplace_back(ts.idx, it);
                        k_cg_update<<<nbk, kBlock, 0, s>>>(n, pnew, Q.get(), Xd, R.get(), part.get(), st.get(), tol,
                                                           fixed, max_iters, do_refresh ? 1 : 0);
                        PRG_CHECK_LAUNCH();
                    }
                    if (do_refresh) {
                        TimedScope ts(c->kt, "cg_refresh", kSpmvRowBytes * double(n));
                        launched.emplace_back(ts.idx, it);
                        k_cg_refresh<<<nbk, kBlock, 0, s>>>(pnt, n, S, B, Xd, R.get(), part.get(), st.get(), tol,
                                                            fixed, max_iters);
                        PRG_CHECK_LAUNCH();
                    }
                }
                issued += batch;
                PRG_HIP(hipMemcpyAsync(hst, st.get(), sizeof(CgState), hipMemcpyDeviceToHost, s));
                PRG_HIP(hipStreamSynchronize(s));
                for (const auto& l : launched)
                    if (l.second > hst->iters) c->kt.relabel(l.first, "cg_idle");
                if (hst->done || issued >= max_iters) break;
                batch = 6;
            }
            {
                TimedScope ts(c->kt, "cg_true_residual", kSpmvRowBytes * double(n));
                k_true_residual<<<nbk, kBlock, 0, s>>>(pnt, n, S, B, Xd, part.get(), st.get(), outv.get() + 4 * d);
                PRG_CHECK_LAUNCH();
            }
            PRG_HIP(hipEventRecord(e1, s));
            PRG_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            PRG_HIP(hipEventElapsedTime(&t, e0, e1));
            ms[d] = t;
        }
        if (jobs.count) {
            PRG_HIP(hipStreamWaitEvent(s, es1, 0));
            PRG_HIP(hipEventSynchronize(es1));
            float t = 0.0f;
            PRG_HIP(hipEventElapsedTime(&t, es0, es1));
            for (int k = 0; k < jobs.count; k++) ms[jobs.depth[k]] = t;  // solved together in one launch
            c->kt.recs["cg_small"].launches += 1;
            c->kt.recs["cg_small"].total_ms += t;
        }
        PRG_HIP(hipMemcpyAsync(ov.data(), outv.get(), ov.size() * 8, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipStreamSynchronize(c->stream2);
        (void)hipHostFree(hst);
        for (hipEvent_t e : {e0, e1, es0, es1}) (void)hipEventDestroy(e);
        throw;
    }
    (void)hipHostFree(hst);
    for (hipEvent_t e : {e0, e1, es0, es1}) (void)hipEventDestroy(e);
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
