// Many independent clouds in one pass (prg_reconstruct_batch, include/prg.h).
//
// Worker threads each own one context (two HIP streams) and pull clouds from
// a shared queue sorted by decreasing N x depth (longest first), so the GPU
// overlaps several clouds' kernels, copies and host round-trips.  Contexts
// are reused across clouds (their device buffers only grow), and a context
// whose configured depth differs from the next cloud's is rebuilt.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/prg.h"

extern "C" prg_status prg_reconstruct_batch(prg_batch_item* items, int64_t count, const prg_config* config,
                                            int32_t concurrency, int32_t subdivide) {
    if (!items || count < 0 || concurrency < 1 || concurrency > 64) return PRG_E_ARG;
    prg_config base;
    prg_config_default(&base);
    if (config) base = *config;
    std::vector<int64_t> order(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; i++) {
        order[size_t(i)] = i;
        items[i].status = PRG_OK;
        std::memset(&items[i].mesh, 0, sizeof(prg_mesh));
        items[i].ms = 0.0;
        items[i].max_true_rel_residual = 0.0;
        items[i].total_iters = 0;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return double(items[a].n) * items[a].max_depth > double(items[b].n) * items[b].max_depth;
    });
    std::atomic<int64_t> next{0};
    std::atomic<int32_t> first_error{PRG_OK};
    auto worker = [&]() {
        prg_ctx* ctx = nullptr;
        int32_t ctx_depth = -1;
        while (true) {
            const int64_t k = next.fetch_add(1);
            if (k >= count) break;
            prg_batch_item& it = items[order[size_t(k)]];
            const auto t0 = std::chrono::steady_clock::now();
            if (!ctx || ctx_depth != it.max_depth) {
                if (ctx) prg_destroy(ctx);
                ctx = nullptr;
                prg_config cfg = base;
                cfg.max_depth = it.max_depth;
                const prg_status st = prg_create(&ctx, &cfg);
                if (st != PRG_OK) {
                    it.status = st;
                    ctx = nullptr;
                    ctx_depth = -1;
                    int32_t expect = PRG_OK;
                    first_error.compare_exchange_strong(expect, st);
                    continue;
                }
                ctx_depth = it.max_depth;
            }
            prg_solve_stats stats[16];
            std::memset(stats, 0, sizeof(stats));
            prg_status st = prg_set_points(ctx, it.xyz, it.nxyz, it.n, 0);
            if (st == PRG_OK) st = prg_reconstruct(ctx, subdivide, &it.mesh, stats);
            it.status = st;
            if (st == PRG_OK || st == PRG_E_NOT_CONVERGED) {
                for (int d = 0; d <= it.max_depth; d++) {
                    it.max_true_rel_residual = std::max(it.max_true_rel_residual, stats[d].true_rel_residual);
                    it.total_iters += stats[d].iters;
                }
            } else {
                int32_t expect = PRG_OK;
                first_error.compare_exchange_strong(expect, st);
            }
            it.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        if (ctx) prg_destroy(ctx);
    };
    std::vector<std::thread> pool;
    const int nthreads = int(std::min<int64_t>(concurrency, std::max<int64_t>(1, count)));
    for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    return prg_status(first_error.load());
}
