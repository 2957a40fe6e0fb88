// Optional cross-batch vertex welding.  The reference concatenates the
// leaf-level batch and the subdivision batches without merging the vertices
// they share at the seams (main.cu:3793,4183,4543); this pass merges vertices
// whose three coordinates are bit-identical.  It is not part of the
// reference's output and runs only when asked for.
//
// Result: the first occurrence of each position keeps its place, in the
// original vertex order, so vertices that have no duplicate earlier in the
// list keep their relative order; every triangle entry is renumbered to its
// position's first occurrence.  Triangles are kept as they are (a triangle
// whose corners merge stays, degenerate).
//
// On the GPU: the vertex indices are sorted by (x, y, z) bit patterns with
// three stable 32-bit radix sorts (z, then y, then x), so equal positions form
// runs that start with their smallest index; a scan over run heads numbers the
// runs, and a scan over the first occurrences numbers the output vertices.
#include <mutex>
#include <string>

#include "ctx.h"

namespace prg {
namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_weld_init(const float* __restrict__ v, int64_t n, uint32_t* keys,
                                                      uint32_t* perm) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    keys[i] = __float_as_uint(v[3 * i + 2]);
    perm[i] = uint32_t(i);
}

// keys[i] = bits of coordinate `comp` of vertex perm[i]
__global__ void __launch_bounds__(kBlock) k_weld_gather(const float* __restrict__ v, int comp,
                                                        const uint32_t* __restrict__ perm, int64_t n, uint32_t* keys) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    keys[i] = __float_as_uint(v[3 * int64_t(perm[i]) + comp]);
}

__device__ __forceinline__ bool same_position(const float* __restrict__ v, uint32_t a, uint32_t b) {
    const int64_t pa = 3 * int64_t(a), pb = 3 * int64_t(b);
    return __float_as_uint(v[pa]) == __float_as_uint(v[pb]) && __float_as_uint(v[pa + 1]) == __float_as_uint(v[pb + 1]) &&
           __float_as_uint(v[pa + 2]) == __float_as_uint(v[pb + 2]);
}

// head[i] = 1 where sorted position i starts a run of equal positions
__global__ void __launch_bounds__(kBlock) k_weld_heads(const float* __restrict__ v, const uint32_t* __restrict__ perm,
                                                       int64_t n, int32_t* head) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || !same_position(v, perm[i], perm[i - 1])) ? 1 : 0;
}

// first[r] = the vertex that starts run r (its smallest index: the sorts are stable)
__global__ void __launch_bounds__(kBlock) k_weld_firsts(const uint32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ head,
                                                        const int32_t* __restrict__ run_pos, int64_t n,
                                                        int32_t* first) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n || !head[i]) return;
    first[run_pos[i]] = int32_t(perm[i]);
}

// owner[j] = the first occurrence of vertex j's position; keep[j] = (owner[j] == j)
__global__ void __launch_bounds__(kBlock) k_weld_owner(const uint32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ head,
                                                       const int32_t* __restrict__ run_pos,
                                                       const int32_t* __restrict__ first, int64_t n, int32_t* owner,
                                                       int32_t* keep) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t o = first[run_pos[i] + head[i] - 1];
    const uint32_t j = perm[i];
    owner[j] = o;
    keep[j] = (o == int32_t(j)) ? 1 : 0;
}

// remap[j] = output number of j's first occurrence; kept vertices are compacted
__global__ void __launch_bounds__(kBlock) k_weld_compact(const float* __restrict__ v, const int32_t* __restrict__ owner,
                                                         const int32_t* __restrict__ keep,
                                                         const int32_t* __restrict__ newid, int64_t n, int32_t* remap,
                                                         float* out) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    remap[j] = newid[owner[j]];
    if (keep[j]) {
        const int64_t o = 3 * int64_t(newid[j]);
        out[o] = v[3 * j];
        out[o + 1] = v[3 * j + 1];
        out[o + 2] = v[3 * j + 2];
    }
}

// triangle entries (checked on the host to lie in [0, n)) renumbered
__global__ void __launch_bounds__(kBlock) k_weld_triangles(int32_t* t, int64_t m3, const int32_t* __restrict__ remap) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < m3) t[i] = remap[t[i]];
}


// Device arrays of one weld (n vertices, m triangles).
struct WeldBufs {
    float *V, *out;
    uint32_t *keys, *perm, *keys_alt, *perm_alt;
    int32_t *head, *run_pos, *first, *owner, *keep, *newid, *remap, *T, *totals;
};

// The weld proper on device arrays `b`, with the scan and sort temporaries in `scr`.
void weld_core(prg_ctx* c, const WeldBufs& b, Scratch& scr, float* vertices, int64_t n, int32_t* triangles,
               int64_t m, int64_t n_leaf, int64_t* n_out, int64_t* n_leaf_out) {
    hipStream_t s = c->stream;
    const unsigned g = grid_for(n, kBlock);
    PRG_HIP(hipMemcpyAsync(b.V, vertices, size_t(3 * n) * 4, hipMemcpyHostToDevice, s));
    if (m > 0) PRG_HIP(hipMemcpyAsync(b.T, triangles, size_t(3 * m) * 4, hipMemcpyHostToDevice, s));
    k_weld_init<<<g, kBlock, 0, s>>>(b.V, n, b.keys, b.perm);
    PRG_CHECK_LAUNCH();
    radix_sort_pairs(b.keys, b.perm, b.keys_alt, b.perm_alt, n, 32, s, scr, c->kt);
    for (int comp = 1; comp >= 0; comp--) {
        k_weld_gather<<<g, kBlock, 0, s>>>(b.V, comp, b.perm, n, b.keys);
        PRG_CHECK_LAUNCH();
        radix_sort_pairs(b.keys, b.perm, b.keys_alt, b.perm_alt, n, 32, s, scr, c->kt);
    }
    k_weld_heads<<<g, kBlock, 0, s>>>(b.V, b.perm, n, b.head);
    PRG_CHECK_LAUNCH();
    scan_exclusive(b.head, b.run_pos, n, b.totals, s, scr, c->kt);
    k_weld_firsts<<<g, kBlock, 0, s>>>(b.perm, b.head, b.run_pos, n, b.first);
    PRG_CHECK_LAUNCH();
    k_weld_owner<<<g, kBlock, 0, s>>>(b.perm, b.head, b.run_pos, b.first, n, b.owner, b.keep);
    PRG_CHECK_LAUNCH();
    scan_exclusive(b.keep, b.newid, n, b.totals + 1, s, scr, c->kt);
    k_weld_compact<<<g, kBlock, 0, s>>>(b.V, b.owner, b.keep, b.newid, n, b.remap, b.out);
    PRG_CHECK_LAUNCH();
    if (m > 0) {
        k_weld_triangles<<<grid_for(3 * m, kBlock), kBlock, 0, s>>>(b.T, 3 * m, b.remap);
        PRG_CHECK_LAUNCH();
    }
    int32_t counts[2] = {0, 0};  // kept vertices; kept among the first n_leaf
    PRG_HIP(hipMemcpyAsync(&counts[0], b.totals + 1, 4, hipMemcpyDeviceToHost, s));
    if (n_leaf > 0 && n_leaf < n) PRG_HIP(hipMemcpyAsync(&counts[1], b.newid + n_leaf, 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    const int64_t kept = counts[0];
    PRG_HIP(hipMemcpyAsync(vertices, b.out, size_t(3 * kept) * 4, hipMemcpyDeviceToHost, s));
    if (m > 0) PRG_HIP(hipMemcpyAsync(triangles, b.T, size_t(3 * m) * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));  // the device arrays are reused or freed after return
    *n_out = kept;
    *n_leaf_out = n_leaf <= 0 ? 0 : (n_leaf < n ? int64_t(counts[1]) : kept);
}

// ---- the weld without allocation ------------------------------------------------------
// Fixed per-device array holding the device copy of the mesh, the weld's
// arrays and its scan and sort temporaries, for up to kNoAllocVertices
// vertices and kNoAllocTriangles triangles.
constexpr int64_t kNoAllocVertices = PRG_WELD_NOALLOC_MAX_VERTICES;
constexpr int64_t kNoAllocTriangles = PRG_WELD_NOALLOC_MAX_TRIANGLES;
constexpr int64_t kAlign = 256;
constexpr int64_t align_up(int64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr int64_t kSortTiles = (kNoAllocVertices + 4095) / 4096 + 1;  // radix sort tiles (4096 keys each)
constexpr int64_t kScanTemp = 256 * kSortTiles + 1024;               // histogram scan; >= n / 2048 scan tiles
constexpr int64_t kNoAllocBytes = 2 * align_up(12 * kNoAllocVertices) + 4 * align_up(4 * kNoAllocVertices) +
                                  7 * align_up(4 * kNoAllocVertices) + align_up(12 * kNoAllocTriangles) +
                                  align_up(16) + 3 * align_up(4 * kScanTemp);
static_assert(kNoAllocBytes == PRG_WELD_NOALLOC_BYTES, "prg.h states the size of the weld's fixed array");

__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// The device address of `scratch` on `device` (current), looked up once per
// device.  In host code the array's name is only a stand-in for the symbol.
uint8_t* weld_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error{PRG_E_ARG, "device ordinal out of range"};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        PRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error{PRG_E_HIP, "hipGetSymbolAddress(scratch) returned null"};
        g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    }
    return g_scratch_ptr[device];
}

void weld_mesh(prg_ctx* c, float* vertices, int64_t n, int32_t* triangles, int64_t m, int64_t n_leaf,
               int64_t* n_out, int64_t* n_leaf_out) {
    *n_out = n;
    *n_leaf_out = n_leaf < n ? n_leaf : n;
    if (n <= 1) return;
    DBuf<float> V, out;
    DBuf<uint32_t> keys, perm, keys_alt, perm_alt;
    DBuf<int32_t> head, run_pos, first, owner, keep, newid, remap, T, totals;
    WeldBufs b{V.reserve(size_t(3 * n)),  out.reserve(size_t(3 * n)),  keys.reserve(size_t(n)),
               perm.reserve(size_t(n)),   keys_alt.reserve(size_t(n)), perm_alt.reserve(size_t(n)),
               head.reserve(size_t(n)),   run_pos.reserve(size_t(n)),  first.reserve(size_t(n)),
               owner.reserve(size_t(n)),  keep.reserve(size_t(n)),     newid.reserve(size_t(n)),
               remap.reserve(size_t(n)),  m > 0 ? T.reserve(size_t(3 * m)) : nullptr, totals.reserve(4)};
    weld_core(c, b, c->scratch, vertices, n, triangles, m, n_leaf, n_out, n_leaf_out);
}

void weld_mesh_noalloc(prg_ctx* c, float* vertices, int64_t n, int32_t* triangles, int64_t m, int64_t n_leaf,
                       int64_t* n_out, int64_t* n_leaf_out) {
    if (n > kNoAllocVertices || m > kNoAllocTriangles)
        throw Error{PRG_E_ARG, "mesh exceeds the fixed weld array (" + std::to_string(kNoAllocVertices) +
                                   " vertices, " + std::to_string(kNoAllocTriangles) + " triangles)"};
    *n_out = n;
    *n_leaf_out = n_leaf < n ? n_leaf : n;
    if (n <= 1) return;
    uint8_t* base = weld_noalloc_scratch(c->cfg.device);
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        uint8_t* p = base + off;
        off += align_up(bytes);
        return p;
    };
    const int64_t N = kNoAllocVertices;
    WeldBufs b{};
    b.V = reinterpret_cast<float*>(take(12 * N));
    b.out = reinterpret_cast<float*>(take(12 * N));
    b.keys = reinterpret_cast<uint32_t*>(take(4 * N));
    b.perm = reinterpret_cast<uint32_t*>(take(4 * N));
    b.keys_alt = reinterpret_cast<uint32_t*>(take(4 * N));
    b.perm_alt = reinterpret_cast<uint32_t*>(take(4 * N));
    int32_t** ints[7] = {&b.head, &b.run_pos, &b.first, &b.owner, &b.keep, &b.newid, &b.remap};
    for (int32_t** q : ints) *q = reinterpret_cast<int32_t*>(take(4 * N));
    b.T = reinterpret_cast<int32_t*>(take(12 * kNoAllocTriangles));
    b.totals = reinterpret_cast<int32_t*>(take(16));
    // the scan and sort temporaries, lent to a Scratch whose buffers never grow
    // past these capacities (so they are never reallocated or freed)
    Scratch scr;
    struct Lend {
        Scratch& s;
        ~Lend() {
            s.a.p = nullptr;
            s.b.p = nullptr;
            s.c.p = nullptr;
            s.a.n = s.b.n = s.c.n = 0;
        }
    } lend{scr};
    scr.a.p = reinterpret_cast<int32_t*>(take(4 * kScanTemp));
    scr.b.p = reinterpret_cast<int32_t*>(take(4 * kScanTemp));
    scr.c.p = reinterpret_cast<int32_t*>(take(4 * kScanTemp));
    scr.a.n = scr.b.n = scr.c.n = size_t(kScanTemp);
    if (off > kNoAllocBytes) throw Error{PRG_E_HIP, "weld array layout exceeds its size"};
    weld_core(c, b, scr, vertices, n, triangles, m, n_leaf, n_out, n_leaf_out);
}

}  // namespace prg
