// Learned warm start: a 3-D U-Net on the dense grid of one coarse depth
// (G^3 cells, G = 2^depth) mapping the depth's divergence (normalised by its
// root-mean-square) and node mask to the solution on the same scale.  It has
// no counterpart in the reference, whose per-depth solves always start from
// zero (main.cu:1237-1329); CG still runs to the same tolerance, so the
// network can only change the iteration count.
//
// Architecture (4 levels, 16-32-64-64 channels, every 3x3x3 convolution
// followed by ReLU, a 1x1 head):
//   e0: conv 16->16, 16->16 (G)      e1: pool, 16->32, 32->32 (G/2)
//   e2: pool, 32->64, 64->64 (G/4)   b:  pool, 64->64, 64->64 (G/8)
//   d2: up, concat e2, 128->64, 64->64;  d1: up, concat e1, 96->32, 32->32
//   d0: up, concat e0, 48->16, 16->16;   head: 16->1
// Input channels: normalised divergence, node mask, zeros (16 in all).
//
// Convolutions are implicit GEMMs on the BF16 matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate): M = voxels, N = output
// channels, K = 27 taps x input channels.  Activations are FP32 in memory
// (channels last) and rounded to BF16 as the fragments load; weights keep an
// FP32 master copy (Adam) and a BF16 copy.  The data gradient is the same
// kernel with the weights flipped and transposed; the weight gradient is a
// GEMM over voxels in fixed chunks whose partials are summed in a fixed
// order, so training is deterministic.
#include <cmath>
#include <vector>

#include "../../include/prg_unet.h"
#include "ctx.h"

namespace prg {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kT = 256;          // threads per workgroup (4 waves)
constexpr int kWChunk = 2048;    // voxels per weight-gradient partial

__device__ __forceinline__ short bf(float f) {
    const __bf16 h = (__bf16)f;
    return __builtin_bit_cast(short, h);
}

__device__ __forceinline__ int64_t nb_voxel(int64_t v, int t, int G) {
    const int x = int(v / (int64_t(G) * G)), y = int((v / G) % G), z = int(v % G);
    const int nx = x + t / 9 - 1, ny = y + (t / 3) % 3 - 1, nz = z + t % 3 - 1;
    if (nx < 0 || ny < 0 || nz < 0 || nx >= G || ny >= G || nz >= G) return -1;
    return (int64_t(nx) * G + ny) * G + nz;
}

// Input of a convolution: channels [0, clo) come from `lo` (grid G / 2,
// nearest upsampling), channels [clo, clo + cskip) from `skip` (grid G); clo
// is 0 for a plain input.  Channel counts are 16, 32 or 64 per source.
struct ConvIn {
    const float* lo;
    int clo;
    const float* skip;
    int cskip;
};

__device__ __forceinline__ int log2i(int v) { return 31 - __builtin_clz(v); }

// a / d for 0 <= a < 2^20 and 1 <= d < 2^10 (inv = 1.0f / d): the float
// estimate is within one of the quotient and is corrected
__device__ __forceinline__ int small_div(int a, int d, float inv) {
    int q = int(float(a) * inv);
    if (q * d > a) q--;
    if ((q + 1) * d <= a) q++;
    return q;
}

constexpr int kLoadBatch = 8;  // halo loads in flight per thread

// One source of the halo: HX x HY x HZ voxels x c channels into LDS channels
// [choff, choff + c); HALF reads the source at G / 2 (nearest upsampling).
// Each thread issues kLoadBatch independent loads before it converts and
// stores any of them.
template <bool HALF>
__device__ __forceinline__ void load_halo(s4* halo4, int H, int HY, int HZ, const float* __restrict__ src, int c,
                                          int choff, int ox, int oy, int oz, int G) {
    const int shift = log2i(c >> 2);
    const int total = H << shift;
    const int Gs = HALF ? G >> 1 : G;
    const float invz = 1.0f / float(HZ), invy = 1.0f / float(HY);
    for (int base = threadIdx.x; base < total; base += kT * kLoadBatch) {
        float4 v[kLoadBatch];
        int dst[kLoadBatch];
#pragma unroll
        for (int k = 0; k < kLoadBatch; k++) {
            const int it = base + k * kT;
            v[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            dst[k] = -1;
            if (it < total) {
                const int h = it >> shift, c4 = it & ((1 << shift) - 1);
                const int hxy = small_div(h, HZ, invz), hz = h - hxy * HZ;
                const int hx = small_div(hxy, HY, invy), hy = hxy - hx * HY;
                const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
                if (gx >= 0 && gy >= 0 && gz >= 0 && gx < G && gy < G && gz < G) {
                    const int64_t vox = HALF ? (int64_t(gx >> 1) * Gs + (gy >> 1)) * Gs + (gz >> 1)
                                             : (int64_t(gx) * G + gy) * G + gz;
                    v[k] = *reinterpret_cast<const float4*>(src + vox * c + 4 * c4);
                }
                const int ch = choff + 4 * c4;
                dst[k] = ((ch >> 4) * H + h) * 4 + ((ch >> 2) & 3);
            }
        }
#pragma unroll
        for (int k = 0; k < kLoadBatch; k++)
            if (dst[k] >= 0) halo4[dst[k]] = s4{bf(v[k].x), bf(v[k].y), bf(v[k].z), bf(v[k].w)};
    }
}

// out[v, co] = act(bias[co] + sum_t sum_ci bf16(in[v + off_t, ci]) * w[t][co][ci])
//
// One workgroup per brick of BX x BY x BZ output voxels (x, y, z; z fastest).
// The brick's input halo, (BX + 2)(BY + 2)(BZ + 2) voxels, is staged in LDS
// as BF16, 16-channel slice after slice: lds[s][h][16]; each wave loads whole
// halo rows (contiguous in memory).  The brick's voxels in (x, y, z) order
// form M-tiles of 16; wave w computes M-tiles [MA (w % MG), MA (w % MG) + MA)
// for N-tiles [NA (w / MG), NA (w / MG) + NA) (16 output channels each).
// MFMA lane layout (16x16x16): lane l holds A[l % 16][4 (l / 16) + e],
// B[4 (l / 16) + e][l % 16] and D[4 (l / 16) + e][l % 16], e = 0..3.  K runs
// over slices, then taps; a slice's 27 weight fragments are loaded together.
template <int MA, int NA>
__global__ void __launch_bounds__(kT) k_conv3(ConvIn src, const short* __restrict__ w, const float* __restrict__ bias,
                                             int G, int BX, int BY, int BZ, int MG, int relu, float* __restrict__ out,
                                             int cout) {
    extern __shared__ s4 halo4[];
    const short* halo = reinterpret_cast<const short*>(halo4);
    const int cin = src.clo + src.cskip;
    const int HY = BY + 2, HZ = BZ + 2, HYZ = HY * HZ, H = (BX + 2) * HYZ;
    const int nbz = G / BZ, nby = G / BY;
    int b = blockIdx.x;
    const int oz = (b % nbz) * BZ;
    b /= nbz;
    const int oy = (b % nby) * BY;
    const int ox = (b / nby) * BX;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int i = l & 15, kq = l >> 4;
    const int mg = wave % MG, ng = wave / MG;
    const int NT = cout >> 4, S = cin >> 4, nvox = BX * BY * BZ;
    const bool active = (ng + 1) * NA <= NT;
    // the first slice's weight fragments are in flight while the halo loads
    const short* wl = w + int64_t((active ? ng : 0) * NA * 16 + i) * cin + 4 * kq;
    s4 wr[27][NA];
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int n = 0; n < NA; n++) wr[t][n] = *reinterpret_cast<const s4*>(wl + (int64_t(t) * cout + n * 16) * cin);
    if (src.clo) load_halo<true>(halo4, H, HY, HZ, src.lo, src.clo, 0, ox, oy, oz, G);
    load_halo<false>(halo4, H, HY, HZ, src.skip, src.cskip, src.clo, ox, oy, oz, G);
    __syncthreads();
    if (!active) return;  // idle wave; no barrier follows
    // lanes past the brick's voxels (grids under 16^3) read the brick's first
    // voxel, whose 27 neighbours are all in the halo, and use zeros
    int hb[MA];
    bool ok[MA];
#pragma unroll
    for (int m = 0; m < MA; m++) {
        const int q = (mg * MA + m) * 16 + i;
        ok[m] = q < nvox;
        const int qx = q / (BY * BZ), qy = (q / BZ) % BY, qz = q % BZ;
        hb[m] = ok[m] ? (qx + 1) * HYZ + (qy + 1) * HZ + qz + 1 : HYZ + HZ + 1;
    }
    // three partial sums over alternating taps keep independent MFMAs in flight
    f4 acc[3][MA][NA];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int m = 0; m < MA; m++)
#pragma unroll
            for (int n = 0; n < NA; n++) acc[x][m][n] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const s4 zero4 = s4{0, 0, 0, 0};
    for (int sl = 0; sl < S; sl++) {
        const short* hs = halo + sl * H * 16 + 4 * kq;
        // next slice's weights (the last slice re-reads its own: in bounds, unused)
        const int nsl = sl + 1 < S ? sl + 1 : sl;
#pragma unroll
        for (int g = 0; g < 3; g++) {
            // the nine taps with dx = g - 1: their A fragments first, then the MFMAs
            s4 a[9][MA];
#pragma unroll
            for (int j = 0; j < 9; j++) {
                const int toff = (g - 1) * HYZ + (j / 3 - 1) * HZ + (j % 3 - 1);
#pragma unroll
                for (int m = 0; m < MA; m++) a[j][m] = *reinterpret_cast<const s4*>(hs + (hb[m] + toff) * 16);
            }
#pragma unroll
            for (int j = 0; j < 9; j++)
#pragma unroll
                for (int m = 0; m < MA; m++) a[j][m] = ok[m] ? a[j][m] : zero4;
#pragma unroll
            for (int j = 0; j < 9; j++) {
                const int t = 9 * g + j;
#pragma unroll
                for (int n = 0; n < NA; n++) {
#pragma unroll
                    for (int m = 0; m < MA; m++)
                        acc[j % 3][m][n] =
                            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[j][m], wr[t][n], acc[j % 3][m][n], 0, 0, 0);
                    wr[t][n] = *reinterpret_cast<const s4*>(wl + (int64_t(t) * cout + n * 16) * cin + nsl * 16);
                }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < MA; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int q = (mg * MA + m) * 16 + 4 * kq + e;
            if (q >= nvox) continue;
            const int gx = ox + q / (BY * BZ), gy = oy + (q / BZ) % BY, gz = oz + q % BZ;
            float* o = out + ((int64_t(gx) * G + gy) * G + gz) * cout + ng * NA * 16 + i;
#pragma unroll
            for (int n = 0; n < NA; n++) {
                float val = (acc[0][m][n][e] + acc[1][m][n][e]) + acc[2][m][n][e];
                if (bias) val = val + bias[(ng * NA + n) * 16 + i];
                if (relu) val = val > 0.0f ? val : 0.0f;
                o[n * 16] = val;
            }
        }
}

// dW partials: part[ch][t][co][ci] = sum over the chunk's voxels v of
// bf16(g[v, co]) * bf16(in[v + off_t, ci]); one wave per (chunk, t, co tile, ci tile).
__global__ void __launch_bounds__(kT) k_conv3_wgrad(const float* __restrict__ in, int cin,
                                                   const float* __restrict__ g, int cout, int G, int nchunk,
                                                   float* __restrict__ part) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int ntc = cout / 16, nti = cin / 16;
    const int64_t job = int64_t(blockIdx.x) * 4 + wave;
    const int64_t njobs = int64_t(nchunk) * 27 * ntc * nti;
    if (job >= njobs) return;
    const int cit = int(job % nti);
    const int cot = int((job / nti) % ntc);
    const int t = int((job / (int64_t(nti) * ntc)) % 27);
    const int ch = int(job / (int64_t(nti) * ntc * 27));
    const int64_t V = int64_t(G) * G * G;
    const int i = l & 15, kq = l >> 4;
    const int co = cot * 16 + i, ci = cit * 16 + i;
    f4 acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t vb0 = int64_t(ch) * kWChunk;
    const int64_t vb1 = vb0 + kWChunk < V ? vb0 + kWChunk : V;
    for (int64_t vb = vb0; vb < vb1; vb += 16) {
        s4 a, b;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t v = vb + 4 * kq + e;
            float ga = 0.0f, xb = 0.0f;
            if (v < vb1) {
                ga = g[v * cout + co];
                const int64_t nv = nb_voxel(v, t, G);
                if (nv >= 0) xb = in[nv * cin + ci];
            }
            a[e] = bf(ga);
            b[e] = bf(xb);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float* dst = part + ((int64_t(ch) * 27 + t) * cout) * cin;
#pragma unroll
    for (int e = 0; e < 4; e++) dst[int64_t(cot * 16 + 4 * kq + e) * cin + ci] = acc[e];
}

// grad[k] += sum over chunks (in order) of part[ch][k]
__global__ void k_sum_chunks(const float* part, int nchunk, int64_t n, float* grad) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= n) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; c++) s += part[int64_t(c) * n + k];
    grad[k] += s;
}

// bf16 copies: wb[t][co][ci] = bf16(w[t][co][ci]); wf[t][ci][co] = bf16(w[26 - t][co][ci])
__global__ void k_pack(const float* w, int cin, int cout, short* wb, short* wf) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    const int64_t n = int64_t(27) * cout * cin;
    if (k >= n) return;
    const int ci = int(k % cin), co = int((k / cin) % cout), t = int(k / (int64_t(cin) * cout));
    const short h = bf(w[k]);
    wb[k] = h;
    wf[(int64_t(26 - t) * cin + ci) * cout + co] = h;
}

// g[v, c] *= (y[v, c] > 0); the bias gradient's per-workgroup partials by channel
__global__ void k_relu_back(float* g, const float* y, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k < n && !(y[k] > 0.0f)) g[k] = 0.0f;
}

// db[c] += sum_v g[v, c]: one workgroup per channel, fixed order
__global__ void k_bias_grad(const float* g, int64_t V, int C, float* db) {
    __shared__ float red[kT];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int64_t v = threadIdx.x; v < V; v += kT) s += g[v * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] += red[0];
}

// average pooling 2x2x2 (G -> G/2) and its gradient
__global__ void k_pool(const float* in, int G, int C, float* out) {
    const int H = G / 2;
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= int64_t(H) * H * H * C) return;
    const int c = int(k % C);
    const int64_t v = k / C;
    const int x = int(v / (int64_t(H) * H)), y = int((v / H) % H), z = int(v % H);
    float s = 0.0f;
    for (int q = 0; q < 8; q++) {
        const int64_t u = (int64_t(2 * x + (q >> 2)) * G + 2 * y + ((q >> 1) & 1)) * G + 2 * z + (q & 1);
        s += in[u * C + c];
    }
    out[k] = s * 0.125f;
}

__global__ void k_pool_back(const float* g, int G, int C, float* gin, int accumulate) {
    const int H = G / 2;
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= int64_t(G) * G * G * C) return;
    const int c = int(k % C);
    const int64_t u = k / C;
    const int x = int(u / (int64_t(G) * G)), y = int((u / G) % G), z = int(u % G);
    const float val = g[((int64_t(x / 2) * H + y / 2) * H + z / 2) * C + c] * 0.125f;
    gin[k] = accumulate ? gin[k] + val : val;
}

// nearest upsampling (G/2 -> G) into channels [0, C) of a concat tensor of
// width W, and the skip tensor into channels [C, C + S)
__global__ void k_up_concat(const float* lo, int C, const float* skip, int S, int G, float* out) {
    const int W = C + S, H = G / 2;
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= int64_t(G) * G * G * W) return;
    const int c = int(k % W);
    const int64_t u = k / W;
    if (c < C) {
        const int x = int(u / (int64_t(G) * G)), y = int((u / G) % G), z = int(u % G);
        out[k] = lo[((int64_t(x / 2) * H + y / 2) * H + z / 2) * C + c];
    } else {
        out[k] = skip[u * S + (c - C)];
    }
}

// gradient of k_up_concat: g_lo (sum of 8 children, fixed order) and the skip part
__global__ void k_up_concat_back_lo(const float* g, int C, int S, int G, float* glo) {
    const int W = C + S, H = G / 2;
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= int64_t(H) * H * H * C) return;
    const int c = int(k % C);
    const int64_t v = k / C;
    const int x = int(v / (int64_t(H) * H)), y = int((v / H) % H), z = int(v % H);
    float s = 0.0f;
    for (int q = 0; q < 8; q++) {
        const int64_t u = (int64_t(2 * x + (q >> 2)) * G + 2 * y + ((q >> 1) & 1)) * G + 2 * z + (q & 1);
        s += g[u * W + c];
    }
    glo[k] = s;
}

__global__ void k_up_concat_back_skip(const float* g, int C, int S, int G, float* gskip) {
    const int W = C + S;
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= int64_t(G) * G * G * S) return;
    const int64_t u = k / S;
    gskip[k] = g[u * W + C + int(k % S)];
}

// head: pred[v] = b + sum_c w[c] * in[v, c] (float, c = 0..15 in order)
__global__ void k_head(const float* in, const float* w, int64_t V, float* pred) {
    const int64_t v = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (v >= V) return;
    float s = 0.0f;
    for (int c = 0; c < 16; c++) s += w[c] * in[v * 16 + c];
    pred[v] = s + w[16];
}

// masked MSE: sum_v m (p - t)^2 / sum_v m and gpred = 2 m (p - t) / sum m;
// one workgroup, fixed order (double sums)
__global__ void k_loss(const float* pred, const float* tgt, const float* mask, int64_t V, float* gpred, double* loss) {
    __shared__ double r1[kT], r2[kT];
    double sm = 0.0, sl = 0.0;
    for (int64_t v = threadIdx.x; v < V; v += kT) {
        const double m = mask[v], d = double(pred[v]) - double(tgt[v]);
        sm += m;
        sl += m * d * d;
    }
    r1[threadIdx.x] = sm;
    r2[threadIdx.x] = sl;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
        }
        __syncthreads();
    }
    const double msum = r1[0];
    if (threadIdx.x == 0) loss[0] = msum > 0.0 ? r2[0] / msum : 0.0;
    for (int64_t v = threadIdx.x; v < V; v += kT)
        gpred[v] = msum > 0.0 ? float(2.0 * mask[v] * (double(pred[v]) - double(tgt[v])) / msum) : 0.0f;
}

// head gradients: g_in[v, c] = gpred[v] * w[c]; dw[c], db: one workgroup each, fixed order
__global__ void k_head_back_in(const float* gpred, const float* w, int64_t V, float* gin) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= V * 16) return;
    gin[k] = gpred[k / 16] * w[k % 16];
}

__global__ void k_head_back_w(const float* gpred, const float* in, int64_t V, float* dw) {
    __shared__ float red[kT];
    const int c = blockIdx.x;  // 0..15 weights, 16 the bias
    float s = 0.0f;
    for (int64_t v = threadIdx.x; v < V; v += kT) s += gpred[v] * (c < 16 ? in[v * 16 + c] : 1.0f);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dw[c] += red[0];
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters
__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float c1, float c2,
                       float gscale) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k] * gscale;
    m[k] = 0.9f * m[k] + 0.1f * gk;
    v[k] = 0.999f * v[k] + 0.001f * gk * gk;
    p[k] = p[k] - lr * (m[k] / c1) / (sqrtf(v[k] / c2) + 1e-8f);
}

// input tensor: channel 0 = divergence / scale, channel 1 = mask, others 0
__global__ void k_input(const float* b, const float* mask, float inv_scale, int64_t V, float* x) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= V * 16) return;
    const int c = int(k % 16);
    const int64_t v = k / 16;
    x[k] = c == 0 ? b[v] * inv_scale : (c == 1 ? mask[v] : 0.0f);
}

__global__ void k_scale(float* a, int64_t n, float f) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k < n) a[k] = a[k] * f;
}

// ---- dense grids of one depth of a context ----
// cell of node i at depth d: its key's per-axis offsets
__device__ __forceinline__ uint32_t compact3u(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

__device__ __forceinline__ int64_t node_cell(uint32_t k, int d, int D) {
    const int64_t G = int64_t(1) << d;
    const int64_t x = compact3u(k >> 2) >> (D - d), y = compact3u(k >> 1) >> (D - d), z = compact3u(k) >> (D - d);
    return (x * G + y) * G + z;
}

// grid[cell(i)] = val[i] (and mask = 1) for the n nodes of one depth
__global__ void k_scatter_nodes(const uint32_t* key, const float* val, int64_t n, int d, int D, float* grid,
                                float* mask) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    const int64_t v = node_cell(key[i], d, D);
    if (grid) grid[v] = val[i];
    if (mask) mask[v] = 1.0f;
}

// out[0] = sum of val^2 (double, one workgroup, fixed order); out[1] = n
__global__ void k_sumsq(const float* val, int64_t n, double* out) {
    __shared__ double red[kT];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kT) s += double(val[i]) * double(val[i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = double(n);
    }
}

__device__ __forceinline__ float rms_of(const double* ss) {
    const double r = ss[1] > 0.0 ? sqrt(ss[0] / ss[1]) : 0.0;
    return r > 0.0 ? float(r) : 1.0f;
}

__global__ void k_input_dev(const float* b, const float* mask, const double* ss, int64_t V, float* x) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= V * 16) return;
    const int c = int(k % 16);
    const int64_t v = k / 16;
    x[k] = c == 0 ? b[v] * (1.0f / rms_of(ss)) : (c == 1 ? mask[v] : 0.0f);
}

// x[i] = pred[cell(i)] * scale for the n nodes of one depth (times `extra`,
// a test setting that is 1 otherwise)
__global__ void k_gather_nodes(const uint32_t* key, const float* pred, const double* ss, int64_t n, int d, int D,
                               float extra, float* x) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    x[i] = pred[node_cell(key[i], d, D)] * rms_of(ss) * extra;
}

unsigned blocks(int64_t n) { return unsigned((n + kT - 1) / kT); }

}  // namespace

// ---- the network object -------------------------------------------------------------
struct UNetConv {
    int cin, cout, G;
    int64_t woff, boff;
    const float* in;  // the input tensor (cin channels) the weight gradient reads
    const float* lo;  // decoder convolutions: the upsampled part of the input (clo channels, grid G / 2)
    int clo;          // 0 for a plain input
    const float* skip;
    float* cat;  // decoder convolutions: `in`, written only for the backward pass
    float* out;
    DBuf<short> wb, wf;
};

}  // namespace prg

struct prg_unet {
    int device = 0;
    int G = 64;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nparams = 0;
    prg::DBuf<float> params, grads, adam_m, adam_v;
    int64_t adam_t = 0;
    prg::UNetConv conv[14];  // the convolutions, encoder to decoder
    // activations
    prg::DBuf<float> x0, p1, p2, p3, c2, c1, c0, pred, target, mask, bgrid;
    prg::DBuf<float> act[14];
    // gradients
    prg::DBuf<float> ga, gb, ge0, ge1, ge2, gpred, part;
    prg::DBuf<double> loss;
    int64_t part_cap = 0;
};

namespace prg {
namespace {

const int kConvDef[14][3] = {{16, 16, 0}, {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2}, {64, 64, 2}, {64, 64, 3},
                             {64, 64, 3}, {128, 64, 2}, {64, 64, 2}, {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};

// Brick and wave split of a convolution with output grid G and NT = cout / 16
// N-tiles: few voxels and many channels per workgroup on the small grids, so
// every grid gives the chip enough workgroups.
struct ConvPlan {
    int BX, BY, BZ, MA, NA, MG;
};

ConvPlan conv_plan(int G, int NT) {
    ConvPlan p{};
    p.BZ = G < 16 ? G : 16;
    int bxy = 1;  // BX = BY = bxy before small-grid widening
    switch (NT) {
        case 1: bxy = int64_t(G) * G * G >= int64_t(256) * 1024 ? 4 : 2; p.MA = bxy == 4 ? 4 : 1; p.NA = 1; break;
        case 2: p.MA = 1; p.NA = 1; break;
        case 3: bxy = 2; p.MA = 1; p.NA = 3; break;
        case 4: p.MA = 1; p.NA = 1; break;
        case 6: p.MA = 1; p.NA = 3; break;
        case 8: p.MA = 1; p.NA = 2; break;
        default: throw Error{PRG_E_ARG, "unsupported channel count"};
    }
    p.BX = bxy < G ? bxy : G;
    p.BY = bxy < G ? bxy : G;
    if (NT == 2 || NT == 6) p.BY = 2 < G ? 2 : G;
    // at least one full M-tile of 16 voxels where the grid has them
    while (p.BX * p.BY * p.BZ < 16 && p.BY < G) p.BY *= 2;
    while (p.BX * p.BY * p.BZ < 16 && p.BX < G) p.BX *= 2;
    const int MT = (p.BX * p.BY * p.BZ + 15) / 16;
    if (MT % p.MA || NT % p.NA) throw Error{PRG_E_ARG, "convolution plan does not tile"};
    p.MG = MT / p.MA;
    if (p.MG * (NT / p.NA) > kT / 64) throw Error{PRG_E_ARG, "convolution plan needs more waves"};
    if (G % p.BX || G % p.BY || G % p.BZ) throw Error{PRG_E_ARG, "convolution brick does not divide the grid"};
    return p;
}

bool pow2_channels(int c) { return c == 16 || c == 32 || c == 64; }

void launch_conv(hipStream_t s, ConvIn src, const short* w, const float* bias, int G, int relu, float* out, int cout) {
    const int cin = src.clo + src.cskip;
    if (!pow2_channels(src.cskip) || (src.clo && !pow2_channels(src.clo)) || cout % 16 || cout <= 0 ||
        (src.clo && G < 2))
        throw Error{PRG_E_ARG, "unsupported convolution channels"};
    const ConvPlan p = conv_plan(G, cout / 16);
    const size_t lds = size_t(p.BX + 2) * (p.BY + 2) * (p.BZ + 2) * cin * 2;
    if (lds > 64 * 1024) throw Error{PRG_E_ARG, "convolution halo exceeds 64 KiB of LDS"};
    const unsigned nb = unsigned((G / p.BX) * (G / p.BY) * (G / p.BZ));
    const int MA = p.MA, NA = p.NA;
#define PRG_CONV(ma, na)                                                                                         \
    if (MA == ma && NA == na) {                                                                                   \
        k_conv3<ma, na><<<nb, kT, lds, s>>>(src, w, bias, G, p.BX, p.BY, p.BZ, p.MG, relu, out, cout);             \
        PRG_CHECK_LAUNCH();                                                                                       \
        return;                                                                                                   \
    }
    PRG_CONV(4, 1)
    PRG_CONV(1, 1)
    PRG_CONV(1, 2)
    PRG_CONV(1, 3)
#undef PRG_CONV
    throw Error{PRG_E_ARG, "no convolution kernel for the plan"};
}

void pack_all(prg_unet* n) {
    for (auto& L : n->conv) {
        const int64_t cnt = int64_t(27) * L.cout * L.cin;
        k_pack<<<blocks(cnt), kT, 0, n->stream>>>(n->params.get() + L.woff, L.cin, L.cout, L.wb.get(), L.wf.get());
        PRG_CHECK_LAUNCH();
    }
}

void forward(prg_unet* n) {
    hipStream_t s = n->stream;
    const int G = n->G;
    auto& C = n->conv;
    auto run = [&](int k) {
        auto& L = C[k];
        launch_conv(s, ConvIn{L.lo, L.clo, L.skip, L.cin - L.clo}, L.wb.get(), n->params.get() + L.boff, L.G, 1,
                    L.out, L.cout);
    };
    auto pool = [&](const float* in, int g, int c, float* out) {
        k_pool<<<blocks(int64_t(g / 2) * (g / 2) * (g / 2) * c), kT, 0, s>>>(in, g, c, out);
        PRG_CHECK_LAUNCH();
    };
    run(0);
    run(1);
    pool(C[1].out, G, 16, n->p1.get());
    run(2);
    run(3);
    pool(C[3].out, G / 2, 32, n->p2.get());
    run(4);
    run(5);
    pool(C[5].out, G / 4, 64, n->p3.get());
    run(6);
    run(7);
    run(8);
    run(9);
    run(10);
    run(11);
    run(12);
    run(13);
    const int64_t V = int64_t(G) * G * G;
    k_head<<<blocks(V), kT, 0, s>>>(C[13].out, n->params.get() + (n->nparams - 17), V, n->pred.get());
    PRG_CHECK_LAUNCH();
}

// gradient of conv k for incoming gradient g (in place: ReLU mask); writes the
// input gradient to gin unless gin is null; accumulates dW and db.
void conv_back(prg_unet* n, int k, float* g, float* gin) {
    hipStream_t s = n->stream;
    auto& L = n->conv[k];
    const int64_t V = int64_t(L.G) * L.G * L.G;
    k_relu_back<<<blocks(V * L.cout), kT, 0, s>>>(g, L.out, V * L.cout);
    PRG_CHECK_LAUNCH();
    k_bias_grad<<<L.cout, kT, 0, s>>>(g, V, L.cout, n->grads.get() + L.boff);
    PRG_CHECK_LAUNCH();
    if (L.clo) {
        // the decoder input the forward pass read through its fused loader
        k_up_concat<<<blocks(V * L.cin), kT, 0, s>>>(L.lo, L.clo, L.skip, L.cin - L.clo, L.G, L.cat);
        PRG_CHECK_LAUNCH();
    }
    const int nchunk = int((V + kWChunk - 1) / kWChunk);
    const int64_t nw = int64_t(27) * L.cout * L.cin;
    const int64_t jobs = int64_t(nchunk) * 27 * (L.cout / 16) * (L.cin / 16);
    k_conv3_wgrad<<<unsigned((jobs + 3) / 4), kT, 0, s>>>(L.in, L.cin, g, L.cout, L.G, nchunk, n->part.get());
    PRG_CHECK_LAUNCH();
    k_sum_chunks<<<blocks(nw), kT, 0, s>>>(n->part.get(), nchunk, nw, n->grads.get() + L.woff);
    PRG_CHECK_LAUNCH();
    if (gin) launch_conv(s, ConvIn{nullptr, 0, g, L.cout}, L.wf.get(), nullptr, L.G, 0, gin, L.cin);
}

void backward(prg_unet* n) {
    hipStream_t s = n->stream;
    const int G = n->G;
    const int64_t V0 = int64_t(G) * G * G, V1 = V0 / 8, V2 = V1 / 8;
    auto& C = n->conv;
    float* ga = n->ga.get();
    float* gb = n->gb.get();
    // head
    const float* hw = n->params.get() + (n->nparams - 17);
    k_head_back_in<<<blocks(V0 * 16), kT, 0, s>>>(n->gpred.get(), hw, V0, ga);
    PRG_CHECK_LAUNCH();
    k_head_back_w<<<17, kT, 0, s>>>(n->gpred.get(), C[13].out, V0, n->grads.get() + (n->nparams - 17));
    PRG_CHECK_LAUNCH();
    // d0
    conv_back(n, 13, ga, gb);
    conv_back(n, 12, gb, ga);  // ga: gradient of c0 (48 channels)
    k_up_concat_back_skip<<<blocks(V0 * 16), kT, 0, s>>>(ga, 32, 16, G, n->ge0.get());
    PRG_CHECK_LAUNCH();
    k_up_concat_back_lo<<<blocks(V1 * 32), kT, 0, s>>>(ga, 32, 16, G, gb);
    PRG_CHECK_LAUNCH();
    // d1 (gb: gradient of d1 output, G/2, 32 channels)
    conv_back(n, 11, gb, ga);
    conv_back(n, 10, ga, gb);  // gb: c1 gradient (96)
    k_up_concat_back_skip<<<blocks(V1 * 32), kT, 0, s>>>(gb, 64, 32, G / 2, n->ge1.get());
    PRG_CHECK_LAUNCH();
    k_up_concat_back_lo<<<blocks(V2 * 64), kT, 0, s>>>(gb, 64, 32, G / 2, ga);
    PRG_CHECK_LAUNCH();
    // d2 (ga: gradient of d2 output, G/4, 64)
    conv_back(n, 9, ga, gb);
    conv_back(n, 8, gb, ga);  // ga: c2 gradient (128)
    k_up_concat_back_skip<<<blocks(V2 * 64), kT, 0, s>>>(ga, 64, 64, G / 4, n->ge2.get());
    PRG_CHECK_LAUNCH();
    k_up_concat_back_lo<<<blocks(V2 / 8 * 64), kT, 0, s>>>(ga, 64, 64, G / 4, gb);
    PRG_CHECK_LAUNCH();
    // bottleneck (gb: G/8, 64)
    conv_back(n, 7, gb, ga);
    conv_back(n, 6, ga, gb);  // gb: gradient of p3 (G/8, 64)
    k_pool_back<<<blocks(V2 * 64), kT, 0, s>>>(gb, G / 4, 64, n->ge2.get(), 1);
    PRG_CHECK_LAUNCH();
    // e2
    conv_back(n, 5, n->ge2.get(), ga);
    conv_back(n, 4, ga, gb);  // gb: gradient of p2 (G/4, 32)
    k_pool_back<<<blocks(V1 * 32), kT, 0, s>>>(gb, G / 2, 32, n->ge1.get(), 1);
    PRG_CHECK_LAUNCH();
    // e1
    conv_back(n, 3, n->ge1.get(), ga);
    conv_back(n, 2, ga, gb);  // gb: gradient of p1 (G/2, 16)
    k_pool_back<<<blocks(V0 * 16), kT, 0, s>>>(gb, G, 16, n->ge0.get(), 1);
    PRG_CHECK_LAUNCH();
    // e0
    conv_back(n, 1, n->ge0.get(), ga);
    conv_back(n, 0, ga, nullptr);
}

void upload(prg_unet* n, float* dst, const float* src, int64_t count) {
    PRG_HIP(hipMemcpyAsync(dst, src, size_t(count) * 4, hipMemcpyHostToDevice, n->stream));
}

// one sample: input (divergence grid, mask grid) normalised by `scale`
void load_sample(prg_unet* n, const float* bgrid, const float* mask, const float* target, float scale) {
    const int64_t V = int64_t(n->G) * n->G * n->G;
    upload(n, n->bgrid.get(), bgrid, V);
    upload(n, n->mask.get(), mask, V);
    if (target) {
        upload(n, n->target.get(), target, V);
        k_scale<<<blocks(V), kT, 0, n->stream>>>(n->target.get(), V, 1.0f / scale);
        PRG_CHECK_LAUNCH();
    }
    k_input<<<blocks(V * 16), kT, 0, n->stream>>>(n->bgrid.get(), n->mask.get(), 1.0f / scale, V, n->x0.get());
    PRG_CHECK_LAUNCH();
}

template <class F>
prg_status unet_guard(prg_unet* n, F&& f) {
    if (!n) return PRG_E_ARG;
    try {
        PRG_HIP(hipSetDevice(n->device));
        return f();
    } catch (const Error& e) {
        n->err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        n->err = "host allocation failed";
        return PRG_E_NOMEM;
    }
}

}  // namespace

void unet_warm_start(prg_ctx* c, int depth, hipStream_t s) {
    prg_unet* n = c->warm_net;
    if (!n || (int64_t(1) << depth) != n->G) throw Error{PRG_E_STATE, "warm-start network does not match the depth"};
    if (n->device != c->cfg.device) throw Error{PRG_E_ARG, "warm-start network is on another device"};
    hipStream_t keep = n->stream;
    n->stream = s;  // the network's kernels run in the order of the chain that solves `depth`
    try {
        const int64_t V = int64_t(n->G) * n->G * n->G, cnt = c->counts[depth], base = c->bases[depth];
        PRG_HIP(hipMemsetAsync(n->bgrid.get(), 0, size_t(V) * 4, s));
        PRG_HIP(hipMemsetAsync(n->mask.get(), 0, size_t(V) * 4, s));
        PoolBuf<double> ss{&c->pool[96]};
        ss.reserve(2);
        k_scatter_nodes<<<blocks(cnt), kT, 0, s>>>(c->key.get() + base, c->div.get() + base, cnt, depth, c->D,
                                                   n->bgrid.get(), n->mask.get());
        PRG_CHECK_LAUNCH();
        k_sumsq<<<1, kT, 0, s>>>(c->div.get() + base, cnt, ss.get());
        PRG_CHECK_LAUNCH();
        k_input_dev<<<blocks(V * 16), kT, 0, s>>>(n->bgrid.get(), n->mask.get(), ss.get(), V, n->x0.get());
        PRG_CHECK_LAUNCH();
        {
            TimedScope ts(c->kt, "unet_forward", 0.0);
            forward(n);
        }
        k_gather_nodes<<<blocks(cnt), kT, 0, s>>>(c->key.get() + base, n->pred.get(), ss.get(), cnt, depth, c->D,
                                                  c->warm_scale, c->x.get() + base);
        PRG_CHECK_LAUNCH();
    } catch (...) {
        n->stream = keep;
        throw;
    }
    n->stream = keep;
}

}  // namespace prg

extern "C" {

prg_status prg_unet_create(prg_unet** out, int32_t device, int32_t grid, uint64_t seed) {
    if (!out) return PRG_E_ARG;
    *out = nullptr;
    if (grid < 16 || grid > 128 || (grid & (grid - 1))) return PRG_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return PRG_E_NODEVICE;
    prg_unet* n = new prg_unet();
    n->device = device;
    n->G = grid;
    const prg_status st = prg::unet_guard(n, [&]() {
        PRG_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        const int G = grid;
        const int64_t V0 = int64_t(G) * G * G;
        int64_t off = 0;

        for (int k = 0; k < 14; k++) {
            auto& L = n->conv[k];
            L.cin = prg::kConvDef[k][0];
            L.cout = prg::kConvDef[k][1];
            L.G = G >> prg::kConvDef[k][2];
            L.woff = off;
            off += int64_t(27) * L.cin * L.cout;
            L.boff = off;
            off += L.cout;
            L.wb.reserve(size_t(27) * L.cin * L.cout);
            L.wf.reserve(size_t(27) * L.cin * L.cout);
            const int64_t V = int64_t(L.G) * L.G * L.G;
            L.out = n->act[k].reserve(V * L.cout);
        }
        off += 17;  // head: 16 weights, bias
        n->nparams = off;
        n->x0.reserve(V0 * 16);
        n->p1.reserve(V0 / 8 * 16);
        n->p2.reserve(V0 / 64 * 32);
        n->p3.reserve(V0 / 512 * 64);
        n->c2.reserve(V0 / 64 * 128);
        n->c1.reserve(V0 / 8 * 96);
        n->c0.reserve(V0 * 48);
        n->pred.reserve(V0);
        n->target.reserve(V0);
        n->mask.reserve(V0);
        n->bgrid.reserve(V0);
        n->ga.reserve(V0 * 48);
        n->gb.reserve(V0 * 48);
        n->ge0.reserve(V0 * 16);
        n->ge1.reserve(V0 / 8 * 32);
        n->ge2.reserve(V0 / 64 * 64);
        n->gpred.reserve(V0);
        n->loss.reserve(1);
        int64_t pc = 0;
        for (auto& L : n->conv) {
            const int64_t V = int64_t(L.G) * L.G * L.G;
            pc = std::max<int64_t>(pc, ((V + prg::kWChunk - 1) / prg::kWChunk) * 27 * L.cin * L.cout);
        }
        n->part.reserve(pc);
        auto& C = n->conv;
        C[0].in = n->x0.get();
        C[1].in = C[0].out;
        C[2].in = n->p1.get();
        C[3].in = C[2].out;
        C[4].in = n->p2.get();
        C[5].in = C[4].out;
        C[6].in = n->p3.get();
        C[7].in = C[6].out;
        C[8].in = n->c2.get();
        C[9].in = C[8].out;
        C[10].in = n->c1.get();
        C[11].in = C[10].out;
        C[12].in = n->c0.get();
        C[13].in = C[12].out;
        for (int k = 0; k < 14; k++) {
            C[k].lo = nullptr;
            C[k].clo = 0;
            C[k].skip = C[k].in;
            C[k].cat = nullptr;
        }
        // decoder inputs: concat(upsampled coarser output, encoder skip)
        const int dec[3][3] = {{8, 7, 5}, {10, 9, 3}, {12, 11, 1}};
        float* cats[3] = {n->c2.get(), n->c1.get(), n->c0.get()};
        for (int j = 0; j < 3; j++) {
            auto& L = C[dec[j][0]];
            L.lo = C[dec[j][1]].out;
            L.clo = C[dec[j][1]].cout;
            L.skip = C[dec[j][2]].out;
            L.cat = cats[j];
        }
        // He-normal initial weights (splitmix64 + Box-Muller), zero biases
        std::vector<float> h(size_t(n->nparams), 0.0f);
        uint64_t st = seed;
        auto next = [&]() {
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto gauss = [&]() {
            const double u1 = (double((next() >> 11) + 1)) * (1.0 / 9007199254740993.0);
            const double u2 = double(next() >> 11) * (1.0 / 9007199254740992.0);
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        };
        for (auto& L : C) {
            const double sd = std::sqrt(2.0 / (27.0 * L.cin));
            for (int64_t k = 0; k < int64_t(27) * L.cin * L.cout; k++) h[size_t(L.woff + k)] = float(gauss() * sd);
        }
        for (int c = 0; c < 16; c++) h[size_t(n->nparams - 17 + c)] = float(gauss() * 0.25);
        n->params.reserve(n->nparams);
        n->grads.reserve(n->nparams);
        n->adam_m.reserve(n->nparams);
        n->adam_v.reserve(n->nparams);
        PRG_HIP(hipMemcpyAsync(n->params.get(), h.data(), size_t(n->nparams) * 4, hipMemcpyHostToDevice, n->stream));
        PRG_HIP(hipMemsetAsync(n->adam_m.get(), 0, size_t(n->nparams) * 4, n->stream));
        PRG_HIP(hipMemsetAsync(n->adam_v.get(), 0, size_t(n->nparams) * 4, n->stream));
        prg::pack_all(n);
        PRG_HIP(hipStreamSynchronize(n->stream));
        return PRG_OK;
    });
    if (st != PRG_OK) {
        prg_unet_destroy(n);
        return st;
    }
    *out = n;
    return PRG_OK;
}

void prg_unet_destroy(prg_unet* n) {
    if (!n) return;
    (void)hipSetDevice(n->device);
    if (n->stream) {
        (void)hipStreamSynchronize(n->stream);
        (void)hipStreamDestroy(n->stream);
    }
    delete n;
}

const char* prg_unet_last_error(prg_unet* n) { return n ? n->err.c_str() : "null network"; }

int64_t prg_unet_num_params(prg_unet* n) { return n ? n->nparams : -1; }

int32_t prg_unet_grid(prg_unet* n) { return n ? n->G : -1; }

prg_status prg_unet_get_params(prg_unet* n, float* out, int64_t count) {
    return prg::unet_guard(n, [&]() {
        if (!out || count != n->nparams) return PRG_E_ARG;
        PRG_HIP(hipMemcpyAsync(out, n->params.get(), size_t(count) * 4, hipMemcpyDeviceToHost, n->stream));
        PRG_HIP(hipStreamSynchronize(n->stream));
        return PRG_OK;
    });
}

prg_status prg_unet_set_params(prg_unet* n, const float* in, int64_t count) {
    return prg::unet_guard(n, [&]() {
        if (!in || count != n->nparams) return PRG_E_ARG;
        PRG_HIP(hipMemcpyAsync(n->params.get(), in, size_t(count) * 4, hipMemcpyHostToDevice, n->stream));
        prg::pack_all(n);
        PRG_HIP(hipStreamSynchronize(n->stream));
        return PRG_OK;
    });
}

prg_status prg_unet_predict(prg_unet* n, const float* bgrid, const float* mask, float scale, float* out) {
    return prg::unet_guard(n, [&]() {
        if (!bgrid || !mask || !out || !(scale > 0.0f)) return PRG_E_ARG;
        const int64_t V = int64_t(n->G) * n->G * n->G;
        prg::load_sample(n, bgrid, mask, nullptr, scale);
        prg::forward(n);
        prg::k_scale<<<prg::blocks(V), prg::kT, 0, n->stream>>>(n->pred.get(), V, scale);
        PRG_CHECK_LAUNCH();
        PRG_HIP(hipMemcpyAsync(out, n->pred.get(), size_t(V) * 4, hipMemcpyDeviceToHost, n->stream));
        PRG_HIP(hipStreamSynchronize(n->stream));
        return PRG_OK;
    });
}

prg_status prg_unet_gradients(prg_unet* n, const float* bgrid, const float* mask, const float* target, float scale,
                              float* grads, double* loss) {
    return prg::unet_guard(n, [&]() {
        if (!bgrid || !mask || !target || !grads || !loss || !(scale > 0.0f)) return PRG_E_ARG;
        PRG_HIP(hipMemsetAsync(n->grads.get(), 0, size_t(n->nparams) * 4, n->stream));
        prg::load_sample(n, bgrid, mask, target, scale);
        prg::forward(n);
        const int64_t V = int64_t(n->G) * n->G * n->G;
        prg::k_loss<<<1, prg::kT, 0, n->stream>>>(n->pred.get(), n->target.get(), n->mask.get(), V, n->gpred.get(),
                                                  n->loss.get());
        PRG_CHECK_LAUNCH();
        prg::backward(n);
        PRG_HIP(hipMemcpyAsync(grads, n->grads.get(), size_t(n->nparams) * 4, hipMemcpyDeviceToHost, n->stream));
        PRG_HIP(hipMemcpyAsync(loss, n->loss.get(), 8, hipMemcpyDeviceToHost, n->stream));
        PRG_HIP(hipStreamSynchronize(n->stream));
        return PRG_OK;
    });
}

prg_status prg_unet_train_step(prg_unet* n, const float* bgrids, const float* masks, const float* targets,
                               const float* scales, int32_t batch, float lr, double* mean_loss) {
    return prg::unet_guard(n, [&]() {
        if (!bgrids || !masks || !targets || !scales || batch < 1 || !(lr > 0.0f)) return PRG_E_ARG;
        const int64_t V = int64_t(n->G) * n->G * n->G;
        PRG_HIP(hipMemsetAsync(n->grads.get(), 0, size_t(n->nparams) * 4, n->stream));
        double total = 0.0;
        for (int b = 0; b < batch; b++) {
            prg::load_sample(n, bgrids + b * V, masks + b * V, targets + b * V, scales[b]);
            prg::forward(n);
            prg::k_loss<<<1, prg::kT, 0, n->stream>>>(n->pred.get(), n->target.get(), n->mask.get(), V,
                                                      n->gpred.get(), n->loss.get());
            PRG_CHECK_LAUNCH();
            prg::backward(n);
            double l = 0.0;
            PRG_HIP(hipMemcpyAsync(&l, n->loss.get(), 8, hipMemcpyDeviceToHost, n->stream));
            PRG_HIP(hipStreamSynchronize(n->stream));
            total += l;
        }
        n->adam_t++;
        const float c1 = float(1.0 - std::pow(0.9, double(n->adam_t)));
        const float c2 = float(1.0 - std::pow(0.999, double(n->adam_t)));
        prg::k_adam<<<prg::blocks(n->nparams), prg::kT, 0, n->stream>>>(n->params.get(), n->grads.get(),
                                                                        n->adam_m.get(), n->adam_v.get(),
                                                                        n->nparams, lr, c1, c2, 1.0f / batch);
        PRG_CHECK_LAUNCH();
        prg::pack_all(n);
        PRG_HIP(hipStreamSynchronize(n->stream));
        if (mean_loss) *mean_loss = total / batch;
        return PRG_OK;
    });
}

prg_status prg_dense_grid(prg_ctx* c, int32_t depth, int32_t which, float* grid, float* div_rms) {
    if (!c) return PRG_E_ARG;
    try {
        if (!c->have_tree) return PRG_E_STATE;
        if (depth < 1 || depth > c->D || depth > 7 || which < 0 || which > 2 || !grid) return PRG_E_ARG;
        if (which == 0 && !c->have_div) return PRG_E_STATE;
        if (which == 1 && !c->have_x) return PRG_E_STATE;
        PRG_HIP(hipSetDevice(c->cfg.device));
        hipStream_t s = c->stream;
        const int64_t G = int64_t(1) << depth, V = G * G * G, cnt = c->counts[depth], base = c->bases[depth];
        prg::DBuf<float> g;
        prg::DBuf<double> ss;
        g.reserve(V);
        ss.reserve(2);
        PRG_HIP(hipMemsetAsync(g.get(), 0, size_t(V) * 4, s));
        const float* src = which == 0 ? c->div.get() : c->x.get();
        prg::k_scatter_nodes<<<prg::blocks(cnt), prg::kT, 0, s>>>(c->key.get() + base,
                                                                  which == 2 ? nullptr : src + base, cnt, depth,
                                                                  c->D, which == 2 ? nullptr : g.get(),
                                                                  which == 2 ? g.get() : nullptr);
        PRG_CHECK_LAUNCH();
        double h[2] = {0.0, 0.0};
        if (div_rms && c->have_div) {
            prg::k_sumsq<<<1, prg::kT, 0, s>>>(c->div.get() + base, cnt, ss.get());
            PRG_CHECK_LAUNCH();
            PRG_HIP(hipMemcpyAsync(h, ss.get(), 16, hipMemcpyDeviceToHost, s));
        }
        PRG_HIP(hipMemcpyAsync(grid, g.get(), size_t(V) * 4, hipMemcpyDeviceToHost, s));
        PRG_HIP(hipStreamSynchronize(s));
        if (div_rms) *div_rms = (h[1] > 0.0 && h[0] > 0.0) ? float(std::sqrt(h[0] / h[1])) : 1.0f;
        return PRG_OK;
    } catch (const prg::Error& e) {
        return e.status;
    }
}

prg_status prg_set_warm_start_net(prg_ctx* c, prg_unet* n) {
    if (!c) return PRG_E_ARG;
    c->warm_net = n;
    return PRG_OK;
}

prg_status prg_warm_start_info(prg_ctx* c, int32_t* depth, int32_t* rejected) {
    if (!c || !depth || !rejected) return PRG_E_ARG;
    *depth = c->warm_depth_used;
    *rejected = c->warm_rejected;
    return PRG_OK;
}

}  // extern "C"
