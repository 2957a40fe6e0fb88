// Internal shared declarations of the MI355X Poisson reconstruction library.
// Everything here is private to libprg.so; the public boundary is include/prg.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/prg.h"

namespace prg {

constexpr int kMaxDepth = 10;  // 3D key bits must fit in uint32

// ---- errors ------------------------------------------------------------------
struct Error {
    prg_status status;
    std::string msg;
};

#define PRG_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::prg::Error{e_ == hipErrorOutOfMemory ? PRG_E_NOMEM : PRG_E_HIP,            \
                               std::string(#call) + ": " + hipGetErrorString(e_)};             \
    } while (0)

#define PRG_CHECK_LAUNCH() PRG_HIP(hipGetLastError())

// ---- device buffers ------------------------------------------------------------
// Owning device allocation that only grows (re-used across calls of a context).
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;  // capacity in elements
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* reserve(size_t count) {
        if (count <= n && p) return p;
        release();
        size_t c = count ? count : 1;
        PRG_HIP(hipMalloc(&p, c * sizeof(T)));
        n = c;
        return p;
    }
    T* get() const { return p; }
};

// A typed view of one slot of a context's buffer pool: grow-only device
// memory kept for the context's lifetime, so a repeated stage allocates
// nothing.  Each function uses its own slot numbers.
template <class T>
struct PoolBuf {
    DBuf<uint8_t>* b;
    T* reserve(size_t count) { return reinterpret_cast<T*>(b->reserve((count ? count : 1) * sizeof(T))); }
    T* get() const { return reinterpret_cast<T*>(b->get()); }
};

// ---- launch helpers --------------------------------------------------------------
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return unsigned(g);
}

// Per-kernel hipEvent timing (enabled by prg_config.timing).
struct KernelTimer {
    bool enabled = false;
    hipStream_t stream = nullptr;
    struct Rec {
        int64_t launches = 0;
        double total_ms = 0.0;
        double bytes = 0.0;  // algorithmic bytes summed over launches
    };
    std::map<std::string, Rec> recs;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        PRG_HIP(hipEventCreate(&e));
        return e;
    }
    // Returns an index to pass to end(); -1 if disabled.
    int begin(const char* name, double bytes) {
        if (!enabled) return -1;
        Pending p{name, get_event(), get_event(), bytes};
        PRG_HIP(hipEventRecord(p.a, stream));
        pending.push_back(p);
        return int(pending.size()) - 1;
    }
    void end(int idx) {
        if (idx < 0) return;
        PRG_HIP(hipEventRecord(pending[idx].b, stream));
    }
    // Counts a launch under `name` without timing it (no events: a record
    // that only says which variant ran, beside a timed one).
    void count(const char* name) {
        if (enabled) recs[name].launches++;
    }
    // Files a recorded launch under another name with other algorithmic bytes
    // (e.g. one that found its work already finished and moved nothing), before flush().
    void relabel(int idx, const char* name, double bytes = 0.0) {
        if (idx >= 0 && idx < int(pending.size())) {
            pending[idx].name = name;
            pending[idx].bytes = bytes;
        }
    }
    // Resolves recorded events (synchronises on them).
    void flush() {
        for (auto& p : pending) {
            PRG_HIP(hipEventSynchronize(p.b));
            float ms = 0.0f;
            PRG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            Rec& r = recs[p.name];
            r.launches++;
            r.total_ms += ms;
            r.bytes += p.bytes;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct TimedScope {
    KernelTimer& t;
    int idx;
    TimedScope(KernelTimer& tm, const char* name, double bytes) : t(tm), idx(tm.begin(name, bytes)) {}
    ~TimedScope() { t.end(idx); }
};

// ---- function-space tables (host-built, funcspace.cpp) -----------------------------
struct FsTables {
    int D = 0;
    int64_t res = 0;
    // baseFunctions[i] as ConfirmedPPolynomial<3,4>: 4 pieces x (start, c0..c3)
    std::vector<float> base_fn;
    // depth-D base function as ConfirmedPPolynomial<2,4>: 4 pieces x (start, c0..c2)
    float base_maxdepth[16] = {};
    // Laplacian stencil per depth, 27 floats each (0 where the reference drops the entry)
    std::vector<float> stencil;
    // divergence weights per node depth d: 3 * 2^(D-d) floats starting at div_off[d]
    std::vector<float> div_table;
    std::vector<int64_t> div_off;
};
void build_fs_tables(int D, FsTables* out);

// ---- primitives (primitives.hip) ------------------------------------------------------
struct Scratch;  // grows on demand
// Exclusive scan of n int32 values; writes the total to *d_total (device) if non-null.
// two arrays of the same length (exclusive scans and totals), in one pass each
void scan_exclusive2(const int32_t* a_in, int32_t* a_out, int32_t* a_total, const int32_t* b_in, int32_t* b_out,
                     int32_t* b_total, int64_t n, hipStream_t s, Scratch& scratch, KernelTimer& kt);
void scan_exclusive(const int32_t* d_in, int32_t* d_out, int64_t n, int32_t* d_total, hipStream_t s,
                    Scratch& scratch, KernelTimer& kt);
// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key.
// Result is in (keys, vals); alt buffers are scratch of the same size.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n,
                      int bits, hipStream_t s, Scratch& scratch, KernelTimer& kt);

struct Scratch {
    DBuf<int32_t> a, b, c;
    DBuf<uint8_t> bytes;
};

}  // namespace prg
