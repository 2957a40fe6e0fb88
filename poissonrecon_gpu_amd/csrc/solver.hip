// Per-depth Laplacian solves (reference: LaplacianIteration main.cu:1223-1332,
// gpuConjugateGradient CG_CUDA.cuh:256-324).
//
// The reference assembles a CSR matrix per depth in managed memory and runs
// a cooperative-groups CG with six grid syncs per iteration.  Its matrix is a
// constant 27-point stencil per depth (the dot tables depend on relative
// position only), so here the SpMV is matrix-free: y_i = sum_j L_d[j] *
// x[neigh_j(i)] over existing neighbours, in the CSR row order j = 0..26 and
// float arithmetic of gpuSpMV.  Per iteration:
//   k_cg_spmv    q = A p and the FP64 partial sums of p.q; the last workgroup
//                (atomic ticket) adds the partials in workgroup order and
//                sets alpha = rr / pq;
//   k_cg_update  x += alpha p, r -= alpha q, FP64 partial sums of r.r; the
//                last workgroup sets beta, rr and the done flag;
//   k_cg_refresh every `refresh` iterations r = b - A x recomputed (FP32
//                SpMV, FP64 norm) instead of the recurrence value;
//   k_cg_direction p = r + beta p.
// Vectors and SpMV are FP32, every dot product and scalar is FP64.  Kernels
// return immediately once the device `done` flag is set, so the host can
// queue iterations in batches and synchronise once per batch.
#include <cmath>

#include "ctx.h"

namespace prg {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;

struct Stencil {
    float L[27];
};

struct CgState {
    double bb, rr, pq, rr_new;
    float alpha, beta;
    int32_t done, iters;
    uint32_t ticket;
    int32_t pad;
};

__device__ __forceinline__ float spmv_row(const int32_t* neigh, int64_t M, int64_t i, int64_t base,
                                          const float* v, const Stencil& S) {
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 27; j++) {
        const int32_t nb = neigh[int64_t(j) * M + i];
        if (nb == -1 || S.L[j] == 0.0f) continue;
        const float prod = S.L[j] * v[nb - base];
        out = out + prod;
    }
    return out;
}

// Block sum of one double; returns the total in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; w++) t += lds[w];
    __syncthreads();
    return t;
}

// Publishes the block partial; true in exactly one thread of the last block.
__device__ __forceinline__ bool publish_partial(double blk, double* partial, CgState* st, unsigned* flag_lds) {
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = blk;
        __threadfence();
        const unsigned t = atomicAdd(&st->ticket, 1u);
        *flag_lds = (t == gridDim.x - 1) ? 1u : 0u;
    }
    __syncthreads();
    return *flag_lds != 0u;
}

__device__ __forceinline__ double sum_partials(const double* partial, int nb) {
    __threadfence();
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += ((volatile const double*)partial)[b];
    return s;
}

__global__ void __launch_bounds__(kBlock) k_cg_init(const int32_t* neigh, int64_t M, int64_t base, int64_t n,
                                                    Stencil S, const float* b, float* x, float* r, float* p,
                                                    double* partial, double* partial2, CgState* st, float tol,
                                                    int32_t fixed) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned last;
    double bb = 0.0, rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_row(neigh, M, base + t, base, x, S);
        const float rv = b[t] - ax;
        r[t] = rv;
        p[t] = rv;
        bb += double(b[t]) * double(b[t]);
        rr += double(rv) * double(rv);
    }
    const double sbb = block_sum(bb, lds);
    const double srr = block_sum(rr, lds);
    if (threadIdx.x == 0) partial2[blockIdx.x] = sbb;
    if (publish_partial(srr, partial, st, &last) && threadIdx.x == 0) {
        const double RR = sum_partials(partial, gridDim.x);
        const double BB = sum_partials(partial2, gridDim.x);
        st->bb = BB;
        st->rr = RR;
        st->iters = 0;
        const double tol2 = double(tol) * double(tol) * BB;
        st->done = (BB <= 0.0) || (RR == 0.0) || (fixed > 0 ? 0 : (RR <= tol2));
        st->ticket = 0u;
    }
}

__global__ void __launch_bounds__(kBlock) k_cg_spmv(const int32_t* neigh, int64_t M, int64_t base, int64_t n,
                                                    Stencil S, const float* p, float* q, double* partial,
                                                    CgState* st) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned last;
    if (st->done) return;
    double pq = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float qv = spmv_row(neigh, M, base + t, base, p, S);
        q[t] = qv;
        pq += double(p[t]) * double(qv);
    }
    const double s = block_sum(pq, lds);
    if (publish_partial(s, partial, st, &last) && threadIdx.x == 0) {
        const double PQ = sum_partials(partial, gridDim.x);
        st->pq = PQ;
        st->alpha = float(st->rr / PQ);
        // breakdown (p.Ap <= 0 once the residual has vanished): stop before the update
        if (!(PQ > 0.0)) st->done = 1;
        st->ticket = 0u;
    }
}

// Ends an iteration: beta, rr, iteration count and the stop test.
__device__ __forceinline__ void finish_iteration(CgState* st, double rr_new, float tol, int32_t fixed,
                                                 int32_t max_iters, bool count) {
    if (count) st->iters += 1;
    st->beta = float(rr_new / st->rr);
    st->rr = rr_new;
    const double tol2 = double(tol) * double(tol) * st->bb;
    // an exactly zero residual ends the solve in either mode (the next step would divide 0 by 0)
    const bool stop = (st->iters >= max_iters) || (rr_new == 0.0) || (fixed > 0 ? (st->iters >= fixed) : (rr_new <= tol2));
    st->done = stop ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_cg_update(int64_t n, const float* p, const float* q, float* x, float* r,
                                                      double* partial, CgState* st, float tol, int32_t fixed,
                                                      int32_t max_iters, int32_t refresh_next) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned last;
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
    if (refresh_next) {
        // the residual is recomputed next; only count the iteration
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const unsigned t = atomicAdd(&st->ticket, 1u);
            if (t == gridDim.x - 1) {
                st->iters += 1;
                st->ticket = 0u;
            }
        }
        return;
    }
    const double s = block_sum(rr, lds);
    if (publish_partial(s, partial, st, &last) && threadIdx.x == 0) {
        finish_iteration(st, sum_partials(partial, gridDim.x), tol, fixed, max_iters, true);
        st->ticket = 0u;
    }
}

__global__ void __launch_bounds__(kBlock) k_cg_refresh(const int32_t* neigh, int64_t M, int64_t base, int64_t n,
                                                       Stencil S, const float* b, const float* x, float* r,
                                                       double* partial, CgState* st, float tol, int32_t fixed,
                                                       int32_t max_iters) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned last;
    if (st->done) return;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_row(neigh, M, base + t, base, x, S);
        const float rv = b[t] - ax;
        r[t] = rv;
        rr += double(rv) * double(rv);
    }
    const double s = block_sum(rr, lds);
    if (publish_partial(s, partial, st, &last) && threadIdx.x == 0) {
        finish_iteration(st, sum_partials(partial, gridDim.x), tol, fixed, max_iters, false);
        st->ticket = 0u;
    }
}

__global__ void __launch_bounds__(kBlock) k_cg_direction(int64_t n, const float* r, float* p, const CgState* st) {
    if (st->done) return;
    const float bt = st->beta;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float tp = bt * p[t];
        p[t] = r[t] + tp;
    }
}

// true relative residual ||b - A x|| / ||b|| of the final iterate
__global__ void __launch_bounds__(kBlock) k_true_residual(const int32_t* neigh, int64_t M, int64_t base, int64_t n,
                                                          Stencil S, const float* b, const float* x,
                                                          double* partial, CgState* st, double* out) {
    __shared__ double lds[kBlock / 64];
    __shared__ unsigned last;
    double rr = 0.0;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < n; t += int64_t(gridDim.x) * kBlock) {
        const float ax = spmv_row(neigh, M, base + t, base, x, S);
        const double rv = double(b[t]) - double(ax);
        rr += rv * rv;
    }
    const double s = block_sum(rr, lds);
    if (publish_partial(s, partial, st, &last) && threadIdx.x == 0) {
        const double RR = sum_partials(partial, gridDim.x);
        out[0] = (st->bb > 0.0) ? sqrt(RR / st->bb) : 0.0;
        out[1] = (st->bb > 0.0) ? sqrt(st->rr / st->bb) : 0.0;
        out[2] = double(st->iters);
        st->ticket = 0u;
    }
}

}  // namespace

prg_status solve_all(prg_ctx* c, prg_solve_stats* stats) {
    ensure_tables(c);
    StageTimer stage(c, PRG_STAGE_SOLVE);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M;
    int64_t maxn = 0;
    for (int d = 0; d <= D; d++) maxn = std::max(maxn, c->counts[d]);
    float* X = c->x.reserve(M);
    if (c->cfg.warm_start == PRG_WARM_NONE) PRG_HIP(hipMemsetAsync(X, 0, size_t(M) * 4, s));
    DBuf<float> R, P, Q;
    R.reserve(maxn);
    P.reserve(maxn);
    Q.reserve(maxn);
    DBuf<double> part, part2, outv;
    part.reserve(kMaxBlocks);
    part2.reserve(kMaxBlocks);
    outv.reserve(4);
    DBuf<CgState> st;
    st.reserve(1);
    PRG_HIP(hipMemsetAsync(st.get(), 0, sizeof(CgState), s));
    CgState* hst = nullptr;
    PRG_HIP(hipHostMalloc(&hst, sizeof(CgState), hipHostMallocDefault));
    const float tol = c->cfg.cg_tol;
    const int32_t fixed = c->cfg.cg_fixed_iters, max_iters = c->cfg.cg_max_iters;
    const int32_t refresh = c->cfg.cg_refresh > 0 ? c->cfg.cg_refresh : (1 << 30);
    const int batch = 20;
    prg_status result = PRG_OK;
    hipEvent_t e0, e1;
    PRG_HIP(hipEventCreate(&e0));
    PRG_HIP(hipEventCreate(&e1));
    try {
        for (int d = 0; d <= D; d++) {
            const int64_t n = c->counts[d], base = c->bases[d];
            Stencil S;
            for (int j = 0; j < 27; j++) S.L[j] = c->fs->stencil[size_t(d) * 27 + j];
            const unsigned nbk = unsigned(std::min<int64_t>(kMaxBlocks, (n + kBlock - 1) / kBlock));
            const float* B = c->div.get() + base;
            float* Xd = X + base;
            PRG_HIP(hipEventRecord(e0, s));
            {
                TimedScope ts(c->kt, "cg_init", 128.0 * double(n));
                k_cg_init<<<nbk, kBlock, 0, s>>>(c->neigh.get(), M, base, n, S, B, Xd, R.get(), P.get(),
                                                 part.get(), part2.get(), st.get(), tol, fixed);
                PRG_CHECK_LAUNCH();
            }
            int issued = 0;
            while (true) {
                for (int k = 0; k < batch; k++) {
                    const int it = issued + k + 1;
                    const bool do_refresh = (it % refresh) == 0;
                    {
                        TimedScope ts(c->kt, "stencil_spmv_dot", 120.0 * double(n));
                        k_cg_spmv<<<nbk, kBlock, 0, s>>>(c->neigh.get(), M, base, n, S, P.get(), Q.get(),
                                                         part.get(), st.get());
                        PRG_CHECK_LAUNCH();
                    }
                    {
                        TimedScope ts(c->kt, "cg_update", 24.0 * double(n));
                        k_cg_update<<<nbk, kBlock, 0, s>>>(n, P.get(), Q.get(), Xd, R.get(), part.get(), st.get(),
                                                           tol, fixed, max_iters, do_refresh ? 1 : 0);
                        PRG_CHECK_LAUNCH();
                    }
                    if (do_refresh) {
                        TimedScope ts(c->kt, "cg_refresh", 120.0 * double(n));
                        k_cg_refresh<<<nbk, kBlock, 0, s>>>(c->neigh.get(), M, base, n, S, B, Xd, R.get(),
                                                            part.get(), st.get(), tol, fixed, max_iters);
                        PRG_CHECK_LAUNCH();
                    }
                    {
                        TimedScope ts(c->kt, "cg_direction", 12.0 * double(n));
                        k_cg_direction<<<nbk, kBlock, 0, s>>>(n, R.get(), P.get(), st.get());
                        PRG_CHECK_LAUNCH();
                    }
                }
                issued += batch;
                PRG_HIP(hipMemcpyAsync(hst, st.get(), sizeof(CgState), hipMemcpyDeviceToHost, s));
                PRG_HIP(hipStreamSynchronize(s));
                if (hst->done || issued >= max_iters) break;
            }
            {
                TimedScope ts(c->kt, "cg_true_residual", 120.0 * double(n));
                k_true_residual<<<nbk, kBlock, 0, s>>>(c->neigh.get(), M, base, n, S, B, Xd, part.get(), st.get(),
                                                       outv.get());
                PRG_CHECK_LAUNCH();
            }
            PRG_HIP(hipEventRecord(e1, s));
            double ov[3];
            PRG_HIP(hipMemcpyAsync(ov, outv.get(), sizeof(ov), hipMemcpyDeviceToHost, s));
            PRG_HIP(hipStreamSynchronize(s));
            float ms = 0.0f;
            PRG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (stats) {
                stats[d].depth = d;
                stats[d].nodes = n;
                stats[d].iters = int32_t(ov[2]);
                stats[d].rel_residual = ov[1];
                stats[d].true_rel_residual = ov[0];
                stats[d].ms = ms;
            }
            if (fixed <= 0 && int32_t(ov[2]) >= max_iters && ov[1] > double(tol)) result = PRG_E_NOT_CONVERGED;
        }
    } catch (...) {
        (void)hipHostFree(hst);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipHostFree(hst);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return result;
}

}  // namespace prg
