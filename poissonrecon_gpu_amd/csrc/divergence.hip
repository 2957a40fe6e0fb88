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
//   - coarser depths: the term list of each node is cut into chunks of
//     kChunk terms, one wave per chunk accumulates in double and reduces with
//     cross-lane shuffles, and a finishing pass adds each node's chunk partials
//     in chunk order.  No global atomics; the result is deterministic.
#include <cmath>

#include "ctx.h"

namespace prg {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 2048;  // terms per wave-chunk

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

__global__ void __launch_bounds__(kBlock) k_vector_field(const uint32_t* key, const int32_t* neigh,
                                                         const int32_t* pidx, const int32_t* pnum, const float* px,
                                                         const float* py, const float* pz, const float* nx,
                                                         const float* ny, const float* nz, int64_t M, int64_t baseD,
                                                         int64_t MD, int D, BaseMD bm, float* vfx, float* vfy,
                                                         float* vfz) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int64_t i = baseD + t;
    const uint32_t k = key[i];
    const float width = float(1.0 / double(1 << D));
    const float ocx = float((0.5 + double(compact3(k >> 2))) * double(width));
    const float ocy = float((0.5 + double(compact3(k >> 1))) * double(width));
    const float ocz = float((0.5 + double(compact3(k))) * double(width));
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    for (int j = 0; j < 27; j++) {
        const int32_t nb = neigh[int64_t(j) * M + i];
        if (nb == -1) continue;
        const int32_t p0 = pidx[nb], pn = pnum[nb];
        for (int32_t q = p0; q < p0 + pn; q++) {
            const float fx = shifted_value(bm, px[q], ocx);
            const float fy = shifted_value(bm, py[q], ocy);
            const float fz = shifted_value(bm, pz[q], ocz);
            float w = fx * fy;
            w = w * fz;
            vx = vx + w * nx[q];
            vy = vy + w * ny[q];
            vz = vz + w * nz[q];
        }
    }
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

// leaf level: 27 terms per node
__global__ void __launch_bounds__(kBlock) k_div_leaf(const int32_t* neigh, const float* vfx, const float* vfy,
                                                     const float* vfz, const float* T, int64_t M, int64_t baseD,
                                                     int64_t MD, float* div) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= MD) return;
    const int64_t i = baseD + t;
    double val = 0.0;
    for (int j = 0; j < 27; j++) {
        const int32_t nb = neigh[int64_t(j) * M + i];
        if (nb == -1) continue;
        const int64_t l = nb - baseD;
        const float r = fdot(vfx[l], vfy[l], vfz[l], T[j / 9], T[(j / 3) % 3], T[j % 3]);
        val += double(r);
    }
    div[i] = float(val);
}

__device__ __forceinline__ int depth_of(const Bases& B, int D, int64_t i) {
    int d = 0;
    while (d < D && i >= B.b[d + 1]) d++;
    return d;
}

// number of terms of every node below the leaf level
__global__ void __launch_bounds__(kBlock) k_div_terms(const int32_t* neigh, const int32_t* dnum, int64_t M,
                                                      int64_t n, int32_t* chunks, int64_t* terms) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    int64_t L = 0;
    for (int j = 0; j < 27; j++) {
        const int32_t nb = neigh[int64_t(j) * M + i];
        if (nb != -1) L += dnum[nb];
    }
    terms[i] = L;
    chunks[i] = int32_t((L + kChunk - 1) / kChunk);
}

__global__ void __launch_bounds__(kBlock) k_div_chunk_map(const int32_t* chunks, const int32_t* chunk_off, int64_t n,
                                                          int32_t* chunk_node) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t o = chunk_off[i];
    for (int32_t c = 0; c < chunks[i]; c++) chunk_node[o + c] = int32_t(i);
}

__global__ void __launch_bounds__(kBlock) k_div_chunks(const uint32_t* key, const int32_t* neigh,
                                                       const int32_t* didx, const int32_t* dnum, const float* vfx,
                                                       const float* vfy, const float* vfz, const float* table,
                                                       const int64_t* table_off, const int32_t* chunk_node,
                                                       const int32_t* chunk_off, const int64_t* terms,
                                                       const int32_t* d_nchunks, int64_t M, Bases B, int D,
                                                       double* partial) {
    const int lane = threadIdx.x & 63;
    const int64_t nchunks = *d_nchunks;
    const int64_t wave0 = (int64_t(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (int64_t(gridDim.x) * kBlock) >> 6;
    const int64_t baseD = B.b[D];
    for (int64_t c = wave0; c < nchunks; c += nwaves) {
        const int64_t i = chunk_node[c];
        const int64_t k0 = int64_t(c - chunk_off[i]) * kChunk;
        const int64_t k1 = min(k0 + kChunk, terms[i]);
        const int d = depth_of(B, D, i);
        const int64_t s = int64_t(1) << (D - d);
        const float* T = table + table_off[d];
        const uint32_t ki = key[i];
        const int64_t oix = compact3(ki >> 2) >> (D - d), oiy = compact3(ki >> 1) >> (D - d),
                      oiz = compact3(ki) >> (D - d);
        double acc = 0.0;
        int64_t prefix = 0;
        for (int j = 0; j < 27; j++) {
            const int32_t nb = neigh[int64_t(j) * M + i];
            if (nb == -1) continue;
            const int64_t len = dnum[nb];
            const int64_t lo = max(prefix, k0), hi = min(prefix + len, k1);
            const int64_t start = didx[nb];
            for (int64_t t = lo + lane; t < hi; t += 64) {
                const int64_t leaf = start + (t - prefix);
                const uint32_t kl = key[baseD + leaf];
                const int64_t ux = int64_t(compact3(kl >> 2)) - (oix - 1) * s;
                const int64_t uy = int64_t(compact3(kl >> 1)) - (oiy - 1) * s;
                const int64_t uz = int64_t(compact3(kl)) - (oiz - 1) * s;
                const float r = fdot(vfx[leaf], vfy[leaf], vfz[leaf], T[ux], T[uy], T[uz]);
                acc += double(r);
            }
            prefix += len;
            if (prefix >= k1) break;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (lane == 0) partial[c] = acc;
    }
}

__global__ void __launch_bounds__(kBlock) k_div_finish(const int32_t* chunks, const int32_t* chunk_off,
                                                       const double* partial, int64_t n, float* div) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    const int32_t o = chunk_off[i];
    for (int32_t c = 0; c < chunks[i]; c++) v += partial[o + c];
    div[i] = float(v);
}

}  // namespace

void ensure_tables(prg_ctx* c) {
    if (c->fs && c->fs->D == c->D) return;
    std::unique_ptr<FsTables> t(new FsTables());
    build_fs_tables(c->D, t.get());
    hipStream_t s = c->stream;
    PRG_HIP(hipMemcpyAsync(c->d_base_fn.reserve(t->base_fn.size()), t->base_fn.data(), t->base_fn.size() * 4,
                           hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(c->d_stencil.reserve(t->stencil.size()), t->stencil.data(), t->stencil.size() * 4,
                           hipMemcpyHostToDevice, s));
    PRG_HIP(hipMemcpyAsync(c->d_div_table.reserve(t->div_table.size()), t->div_table.data(),
                           t->div_table.size() * 4, hipMemcpyHostToDevice, s));
    PRG_HIP(hipStreamSynchronize(s));
    c->fs = std::move(t);
}

void compute_divergence(prg_ctx* c) {
    ensure_tables(c);
    hipStream_t s = c->stream;
    const int D = c->D;
    const int64_t M = c->M, MD = c->MD(), baseD = c->bases[D];
    BaseMD bm;
    for (int k = 0; k < 16; k++) bm.v[k] = c->fs->base_maxdepth[k];
    float* VX = c->vfx.reserve(MD);
    float* VY = c->vfy.reserve(MD);
    float* VZ = c->vfz.reserve(MD);
    {
        StageTimer st(c, PRG_STAGE_VECTOR_FIELD);
        TimedScope ts(c->kt, "vector_field", 12.0 * double(MD) + 116.0 * double(MD) + 24.0 * double(c->N));
        k_vector_field<<<grid_for(MD, kBlock), kBlock, 0, s>>>(
            c->key.get(), c->neigh.get(), c->pidx.get(), c->pnum.get(), c->px.get(), c->py.get(), c->pz.get(),
            c->nx.get(), c->ny.get(), c->nz.get(), M, baseD, MD, D, bm, VX, VY, VZ);
        PRG_CHECK_LAUNCH();
    }
    StageTimer st(c, PRG_STAGE_DIVERGENCE);
    float* DV = c->div.reserve(M);
    DBuf<int64_t> toff;
    PRG_HIP(hipMemcpyAsync(toff.reserve(c->fs->div_off.size()), c->fs->div_off.data(), c->fs->div_off.size() * 8,
                           hipMemcpyHostToDevice, s));
    {
        TimedScope ts(c->kt, "div_leaf", 124.0 * double(MD));
        k_div_leaf<<<grid_for(MD, kBlock), kBlock, 0, s>>>(c->neigh.get(), VX, VY, VZ,
                                                           c->d_div_table.get() + c->fs->div_off[D], M, baseD, MD, DV);
        PRG_CHECK_LAUNCH();
    }
    const int64_t n = baseD;  // nodes of depths 0 .. D-1
    if (n > 0) {
        DBuf<int32_t> chunks, coff, cnode, ctot;
        DBuf<int64_t> terms;
        DBuf<double> partial;
        chunks.reserve(n);
        coff.reserve(n);
        terms.reserve(n);
        ctot.reserve(1);
        k_div_terms<<<grid_for(n, kBlock), kBlock, 0, s>>>(c->neigh.get(), c->dnum.get(), M, n, chunks.get(),
                                                           terms.get());
        PRG_CHECK_LAUNCH();
        scan_exclusive(chunks.get(), coff.get(), n, ctot.get(), s, c->scratch, c->kt);
        int32_t nchunks = 0;
        PRG_HIP(hipMemcpyAsync(&nchunks, ctot.get(), 4, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
        cnode.reserve(nchunks);
        partial.reserve(nchunks);
        k_div_chunk_map<<<grid_for(n, kBlock), kBlock, 0, s>>>(chunks.get(), coff.get(), n, cnode.get());
        PRG_CHECK_LAUNCH();
        Bases B{};
        for (int d = 0; d <= D; d++) B.b[d] = c->bases[d];
        B.b[D + 1] = M;
        const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(int64_t(nchunks), 8192));
        {
            TimedScope ts(c->kt, "div_chunks", 16.0 * double(MD) * double(D));
            k_div_chunks<<<grid_for(waves * 64, kBlock), kBlock, 0, s>>>(
                c->key.get(), c->neigh.get(), c->didx.get(), c->dnum.get(), VX, VY, VZ, c->d_div_table.get(),
                toff.get(), cnode.get(), coff.get(), terms.get(), ctot.get(), M, B, D, partial.get());
            PRG_CHECK_LAUNCH();
        }
        k_div_finish<<<grid_for(n, kBlock), kBlock, 0, s>>>(chunks.get(), coff.get(), partial.get(), n, DV);
        PRG_CHECK_LAUNCH();
        PRG_HIP(hipStreamSynchronize(s));
    }
}

}  // namespace prg
