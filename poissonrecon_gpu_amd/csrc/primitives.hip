// Device-wide primitives: exclusive scan of int32 and a stable LSD radix sort
// of (uint32 key, uint32 value) pairs.
//
// Radix sort, one 8-bit digit per pass (3D key bits -> ceil(3D/8) passes):
//   upsweep   - each workgroup counts the digits of its 4096-key tile in LDS
//               and writes a digit-major histogram [256][tiles];
//   scan      - exclusive scan of that histogram gives every (digit, tile)
//               its global output offset;
//   downsweep - each wave ranks its 1024 keys stably with wave64 ballots
//               (peers with the same digit via 8 ballots, rank = popcount of
//               lower peers + running per-wave count), the tile is reordered
//               by digit in LDS, and written out in digit runs so the global
//               stores are contiguous.
// Stability holds because waves own consecutive sub-tiles, items are visited
// in input order and tiles are laid out digit-major in the scan.
#include "common.h"

namespace prg {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t* lds_waves, int32_t& total) {
    // wave-level inclusive scan
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    int32_t wsum = 0, tot = 0;
    for (int w = 0; w < nw; w++) {
        int32_t t = lds_waves[w];
        if (w < wave) wsum += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wsum + x - v;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const int32_t* in, int64_t n, int32_t* sums) {
    __shared__ int32_t w[kScanBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k * kScanBlock + threadIdx.x;
        if (i < n) acc += in[i];
    }
    int32_t tot;
    (void)block_exclusive_scan(acc, w, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the tile sums, in chunks with a carry.
__global__ void __launch_bounds__(1024) k_scan_sums(int32_t* sums, int64_t n, int32_t* d_total) {
    __shared__ int32_t w[16];
    int32_t carry = 0;
    for (int64_t base = 0; base < n; base += 1024) {
        int64_t i = base + threadIdx.x;
        int32_t v = (i < n) ? sums[i] : 0;
        int32_t tot;
        int32_t ex = block_exclusive_scan(v, w, tot);
        if (i < n) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const int32_t* in, int32_t* out, int64_t n,
                                                             const int32_t* sums) {
    __shared__ int32_t w[kScanBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    // each thread owns kScanItems consecutive elements
    int32_t v[kScanItems];
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + int64_t(threadIdx.x) * kScanItems + k;
        v[k] = (i < n) ? in[i] : 0;
        acc += v[k];
    }
    int32_t tot;
    int32_t ex = block_exclusive_scan(acc, w, tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + int64_t(threadIdx.x) * kScanItems + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

// Two arrays of the same length scanned together: one reduce, one
// two-workgroup sums pass and one apply (sums_a, sums_b: the tile sums).
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce2(const int32_t* in_a, const int32_t* in_b, int64_t n,
                                                               int32_t* sums_a, int32_t* sums_b) {
    __shared__ int32_t w[kScanBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    int32_t acc_a = 0, acc_b = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k * kScanBlock + threadIdx.x;
        if (i < n) {
            acc_a += in_a[i];
            acc_b += in_b[i];
        }
    }
    int32_t tot_a, tot_b;
    (void)block_exclusive_scan(acc_a, w, tot_a);
    (void)block_exclusive_scan(acc_b, w, tot_b);
    if (threadIdx.x == 0) {
        sums_a[blockIdx.x] = tot_a;
        sums_b[blockIdx.x] = tot_b;
    }
}

__global__ void __launch_bounds__(1024) k_scan_sums2(int32_t* sums_a, int32_t* sums_b, int64_t n, int32_t* total_a,
                                                     int32_t* total_b) {
    __shared__ int32_t w[16];
    int32_t* sums = blockIdx.x == 0 ? sums_a : sums_b;
    int32_t* d_total = blockIdx.x == 0 ? total_a : total_b;
    int32_t carry = 0;
    for (int64_t base = 0; base < n; base += 1024) {
        int64_t i = base + threadIdx.x;
        int32_t v = (i < n) ? sums[i] : 0;
        int32_t tot;
        int32_t ex = block_exclusive_scan(v, w, tot);
        if (i < n) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply2(const int32_t* in_a, int32_t* out_a, const int32_t* in_b,
                                                              int32_t* out_b, int64_t n, const int32_t* sums_a,
                                                              const int32_t* sums_b) {
    __shared__ int32_t w[kScanBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    int32_t va[kScanItems], vb[kScanItems];
    int32_t acc_a = 0, acc_b = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + int64_t(threadIdx.x) * kScanItems + k;
        va[k] = (i < n) ? in_a[i] : 0;
        vb[k] = (i < n) ? in_b[i] : 0;
        acc_a += va[k];
        acc_b += vb[k];
    }
    int32_t tot;
    int32_t ea = block_exclusive_scan(acc_a, w, tot) + sums_a[blockIdx.x];
    int32_t eb = block_exclusive_scan(acc_b, w, tot) + sums_b[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + int64_t(threadIdx.x) * kScanItems + k;
        if (i < n) {
            out_a[i] = ea;
            out_b[i] = eb;
        }
        ea += va[k];
        eb += vb[k];
    }
}

// ---- radix sort ------------------------------------------------------------------
constexpr int kSortBlock = 256;           // 4 waves
constexpr int kSortItems = 16;            // keys per thread
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kWaveTile = 64 * kSortItems;  // 1024 keys per wave

__global__ void __launch_bounds__(kSortBlock) k_radix_upsweep(const uint32_t* keys, int64_t n, int shift,
                                                                int32_t* hist, int tiles) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        int64_t i = base + k * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kSortBlock) k_radix_downsweep(const uint32_t* keys_in, const uint32_t* vals_in,
                                                                  uint32_t* keys_out, uint32_t* vals_out, int64_t n,
                                                                  int shift, const int32_t* offsets, int tiles) {
    __shared__ int32_t wcnt[4][256];     // per-wave digit counts, then per-wave digit offsets
    __shared__ int32_t tile_off[256];    // digit start inside the tile
    __shared__ int32_t glob_off[256];    // digit start in the output
    __shared__ uint32_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    __shared__ int32_t wsum[4];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 4 * 256; i += kSortBlock) (&wcnt[0][0])[i] = 0;
    glob_off[threadIdx.x] = offsets[int64_t(threadIdx.x) * tiles + blockIdx.x];
    __syncthreads();

    const int64_t wbase = int64_t(blockIdx.x) * kSortTile + int64_t(wave) * kWaveTile;
    uint32_t k_reg[kSortItems], v_reg[kSortItems];
    int32_t r_reg[kSortItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int64_t i = wbase + k * 64 + lane;
        const bool valid = i < n;
        k_reg[k] = valid ? keys_in[i] : 0u;
        v_reg[k] = valid ? vals_in[i] : 0u;
        const uint32_t digit = (k_reg[k] >> shift) & 255u;
        // lanes with the same digit: intersect 8 ballots
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bb : ~bb;
        }
        int32_t before = 0;
        if (valid) before = wcnt[wave][digit];
        const int32_t rank = before + __popcll(peers & lt_mask);
        // the lowest lane of each peer group advances the running count
        if (valid && (peers & lt_mask) == 0ull) wcnt[wave][digit] = before + __popcll(peers);
        r_reg[k] = valid ? rank : -1;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // thread t owns digit t: exclusive offsets across waves, tile totals
    {
        const int d = threadIdx.x;
        int32_t run = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            int32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
        int32_t tot;
        int32_t ex = block_exclusive_scan(run, wsum, tot);
        tile_off[d] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        if (r_reg[k] < 0) continue;
        const uint32_t digit = (k_reg[k] >> shift) & 255u;
        const int32_t pos = tile_off[digit] + wcnt[wave][digit] + r_reg[k];
        skeys[pos] = k_reg[k];
        svals[pos] = v_reg[k];
    }
    __syncthreads();
    const int64_t tbase = int64_t(blockIdx.x) * kSortTile;
    const int valid_in_tile = int((n - tbase) < kSortTile ? (n - tbase) : kSortTile);
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int t = k * kSortBlock + threadIdx.x;
        if (t >= valid_in_tile) continue;
        const uint32_t key = skeys[t];
        const uint32_t digit = (key >> shift) & 255u;
        const int64_t dst = int64_t(glob_off[digit]) + (t - tile_off[digit]);
        keys_out[dst] = key;
        vals_out[dst] = svals[t];
    }
}

}  // namespace

void scan_exclusive(const int32_t* d_in, int32_t* d_out, int64_t n, int32_t* d_total, hipStream_t s,
                    Scratch& scratch, KernelTimer& kt) {
    if (n <= 0) {
        if (d_total) PRG_HIP(hipMemsetAsync(d_total, 0, sizeof(int32_t), s));
        return;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int32_t* sums = scratch.a.reserve(size_t(tiles));
    TimedScope ts(kt, "scan_exclusive", 8.0 * double(n));
    k_scan_reduce<<<unsigned(tiles), kScanBlock, 0, s>>>(d_in, n, sums);
    PRG_CHECK_LAUNCH();
    k_scan_sums<<<1, 1024, 0, s>>>(sums, tiles, d_total);
    PRG_CHECK_LAUNCH();
    k_scan_apply<<<unsigned(tiles), kScanBlock, 0, s>>>(d_in, d_out, n, sums);
    PRG_CHECK_LAUNCH();
}

void scan_exclusive2(const int32_t* a_in, int32_t* a_out, int32_t* a_total, const int32_t* b_in, int32_t* b_out,
                     int32_t* b_total, int64_t n, hipStream_t s, Scratch& scratch, KernelTimer& kt) {
    if (n <= 0) {
        scan_exclusive(a_in, a_out, n, a_total, s, scratch, kt);
        scan_exclusive(b_in, b_out, n, b_total, s, scratch, kt);
        return;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int32_t* sums = scratch.a.reserve(size_t(2 * tiles));
    TimedScope ts(kt, "scan_exclusive", 16.0 * double(n));
    k_scan_reduce2<<<unsigned(tiles), kScanBlock, 0, s>>>(a_in, b_in, n, sums, sums + tiles);
    PRG_CHECK_LAUNCH();
    k_scan_sums2<<<2, 1024, 0, s>>>(sums, sums + tiles, tiles, a_total, b_total);
    PRG_CHECK_LAUNCH();
    k_scan_apply2<<<unsigned(tiles), kScanBlock, 0, s>>>(a_in, a_out, b_in, b_out, n, sums, sums + tiles);
    PRG_CHECK_LAUNCH();
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                      hipStream_t s, Scratch& scratch, KernelTimer& kt) {
    if (n <= 1 || bits <= 0) return;
    const int64_t tiles64 = (n + kSortTile - 1) / kSortTile;
    const int tiles = int(tiles64);
    const int64_t hn = 256 * tiles64;
    int32_t* hist = scratch.b.reserve(size_t(hn));
    int32_t* offs = scratch.c.reserve(size_t(hn));
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        {
            TimedScope ts(kt, "radix_upsweep", 4.0 * double(n));
            k_radix_upsweep<<<unsigned(tiles), kSortBlock, 0, s>>>(ki, n, shift, hist, tiles);
            PRG_CHECK_LAUNCH();
        }
        scan_exclusive(hist, offs, hn, nullptr, s, scratch, kt);
        {
            TimedScope ts(kt, "radix_downsweep", 16.0 * double(n));
            k_radix_downsweep<<<unsigned(tiles), kSortBlock, 0, s>>>(ki, vi, ko, vo, n, shift, offs, tiles);
            PRG_CHECK_LAUNCH();
        }
        std::swap(ki, ko);
        std::swap(vi, vo);
        passes++;
    }
    if (passes & 1) {
        PRG_HIP(hipMemcpyAsync(keys, ki, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
        PRG_HIP(hipMemcpyAsync(vals, vi, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace prg
