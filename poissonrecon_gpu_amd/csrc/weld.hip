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

}  // namespace

void weld_mesh(prg_ctx* c, float* vertices, int64_t n, int32_t* triangles, int64_t m, int64_t n_leaf,
               int64_t* n_out, int64_t* n_leaf_out) {
    *n_out = n;
    *n_leaf_out = n_leaf < n ? n_leaf : n;
    if (n <= 1) return;
    hipStream_t s = c->stream;
    const unsigned g = grid_for(n, kBlock);
    DBuf<float> V, out;
    DBuf<uint32_t> keys, perm, keys_alt, perm_alt;
    DBuf<int32_t> head, run_pos, first, owner, keep, newid, remap, T, totals;
    V.reserve(size_t(3 * n));
    out.reserve(size_t(3 * n));
    keys.reserve(size_t(n));
    perm.reserve(size_t(n));
    keys_alt.reserve(size_t(n));
    perm_alt.reserve(size_t(n));
    head.reserve(size_t(n));
    run_pos.reserve(size_t(n));
    first.reserve(size_t(n));
    owner.reserve(size_t(n));
    keep.reserve(size_t(n));
    newid.reserve(size_t(n));
    remap.reserve(size_t(n));
    totals.reserve(4);
    PRG_HIP(hipMemcpyAsync(V.get(), vertices, size_t(3 * n) * 4, hipMemcpyHostToDevice, s));
    if (m > 0) {
        T.reserve(size_t(3 * m));
        PRG_HIP(hipMemcpyAsync(T.get(), triangles, size_t(3 * m) * 4, hipMemcpyHostToDevice, s));
    }
    k_weld_init<<<g, kBlock, 0, s>>>(V.get(), n, keys.get(), perm.get());
    PRG_CHECK_LAUNCH();
    radix_sort_pairs(keys.get(), perm.get(), keys_alt.get(), perm_alt.get(), n, 32, s, c->scratch, c->kt);
    for (int comp = 1; comp >= 0; comp--) {
        k_weld_gather<<<g, kBlock, 0, s>>>(V.get(), comp, perm.get(), n, keys.get());
        PRG_CHECK_LAUNCH();
        radix_sort_pairs(keys.get(), perm.get(), keys_alt.get(), perm_alt.get(), n, 32, s, c->scratch, c->kt);
    }
    k_weld_heads<<<g, kBlock, 0, s>>>(V.get(), perm.get(), n, head.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(head.get(), run_pos.get(), n, totals.get(), s, c->scratch, c->kt);
    k_weld_firsts<<<g, kBlock, 0, s>>>(perm.get(), head.get(), run_pos.get(), n, first.get());
    PRG_CHECK_LAUNCH();
    k_weld_owner<<<g, kBlock, 0, s>>>(perm.get(), head.get(), run_pos.get(), first.get(), n, owner.get(), keep.get());
    PRG_CHECK_LAUNCH();
    scan_exclusive(keep.get(), newid.get(), n, totals.get() + 1, s, c->scratch, c->kt);
    k_weld_compact<<<g, kBlock, 0, s>>>(V.get(), owner.get(), keep.get(), newid.get(), n, remap.get(), out.get());
    PRG_CHECK_LAUNCH();
    if (m > 0) {
        k_weld_triangles<<<grid_for(3 * m, kBlock), kBlock, 0, s>>>(T.get(), 3 * m, remap.get());
        PRG_CHECK_LAUNCH();
    }
    int32_t counts[2] = {0, 0};  // kept vertices; kept among the first n_leaf
    PRG_HIP(hipMemcpyAsync(&counts[0], totals.get() + 1, 4, hipMemcpyDeviceToHost, s));
    if (n_leaf > 0 && n_leaf < n)
        PRG_HIP(hipMemcpyAsync(&counts[1], newid.get() + n_leaf, 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));
    const int64_t kept = counts[0];
    PRG_HIP(hipMemcpyAsync(vertices, out.get(), size_t(3 * kept) * 4, hipMemcpyDeviceToHost, s));
    if (m > 0) PRG_HIP(hipMemcpyAsync(triangles, T.get(), size_t(3 * m) * 4, hipMemcpyDeviceToHost, s));
    PRG_HIP(hipStreamSynchronize(s));  // the device buffers are freed on return
    *n_out = kept;
    *n_leaf_out = n_leaf <= 0 ? 0 : (n_leaf < n ? int64_t(counts[1]) : kept);
}

}  // namespace prg
