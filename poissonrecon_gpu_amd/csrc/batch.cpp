// Many independent clouds in one pass (prg_reconstruct_batch, include/prg.h).
//
// Worker threads each own one context and pull clouds from a shared queue
// sorted by decreasing N x depth (longest first), so the GPU overlaps several
// clouds' kernels, copies and host round-trips.  A worker keeps its context
// for every cloud (its device buffers only grow; prg_set_max_depth switches
// the depth).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/prg.h"

namespace {

// Grow-only pinned host buffer of one worker: clouds are staged in it for the
// upload and meshes land in it before they are copied to their own arrays, so
// every transfer is a direct DMA.
struct Pinned {
    void* p = nullptr;
    size_t cap = 0;
    bool reserve(size_t bytes) {
        if (bytes <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4;
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

// One cloud through the stages on a context, the mesh through pinned staging
// into malloc'd arrays (released by prg_mesh_release).
prg_status reconstruct_staged(prg_ctx* ctx, prg_batch_item& it, int32_t subdivide, Pinned& in, Pinned& vbuf,
                              Pinned& tbuf, prg_solve_stats* stats) {
    const size_t pb = size_t(3) * size_t(it.n) * sizeof(float);
    prg_status st;
    if (in.reserve(2 * pb)) {
        float* px = static_cast<float*>(in.p);
        std::memcpy(px, it.xyz, pb);
        std::memcpy(px + 3 * it.n, it.nxyz, pb);
        st = prg_set_points(ctx, px, px + 3 * it.n, it.n, 0);
    } else {
        st = prg_set_points(ctx, it.xyz, it.nxyz, it.n, 0);
    }
    if (st != PRG_OK) return st;
    if ((st = prg_build_octree(ctx)) != PRG_OK) return st;
    if ((st = prg_compute_divergence(ctx)) != PRG_OK) return st;
    const prg_status solved = prg_solve(ctx, stats);
    if (solved != PRG_OK && solved != PRG_E_NOT_CONVERGED) return solved;
    if ((st = prg_isovalue(ctx, nullptr)) != PRG_OK) return st;
    prg_mesh m;
    std::memset(&m, 0, sizeof(m));
    // surface meshes of sampled clouds have about 0.8 vertices and 1.6
    // triangles per point; staging sized for more rarely needs the retry below
    (void)vbuf.reserve(size_t(it.n) * 3 * 4);
    (void)tbuf.reserve(size_t(it.n) * 6 * 4);
    for (int attempt = 0; attempt < 2; attempt++) {
        st = prg_extract_mesh_into(ctx, subdivide, static_cast<float*>(vbuf.p), int64_t(vbuf.cap / 4),
                                   static_cast<int32_t*>(tbuf.p), int64_t(tbuf.cap / 4), &m);
        if (st != PRG_E_ARG || attempt == 1) break;
        // the counts are in m: grow the staging (the extraction runs again)
        if (!vbuf.reserve(size_t(3 * m.n_vertices + 1) * 4) || !tbuf.reserve(size_t(3 * m.n_triangles + 1) * 4))
            return PRG_E_NOMEM;
    }
    if (st != PRG_OK) return st;
    it.mesh = m;
    it.mesh.vertices = static_cast<float*>(std::malloc(size_t(3 * m.n_vertices + 1) * 4));
    it.mesh.triangles = static_cast<int32_t*>(std::malloc(size_t(3 * m.n_triangles + 1) * 4));
    if (!it.mesh.vertices || !it.mesh.triangles) {
        prg_mesh_release(&it.mesh);
        return PRG_E_NOMEM;
    }
    std::memcpy(it.mesh.vertices, m.vertices, size_t(3 * m.n_vertices) * 4);
    std::memcpy(it.mesh.triangles, m.triangles, size_t(3 * m.n_triangles) * 4);
    return solved;
}

}  // namespace

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
        Pinned in, vbuf, tbuf;
        prg_ctx* ctx = nullptr;
        int32_t ctx_depth = -1;
        while (true) {
            const int64_t k = next.fetch_add(1);
            if (k >= count) break;
            prg_batch_item& it = items[order[size_t(k)]];
            const auto t0 = std::chrono::steady_clock::now();
            if (ctx && ctx_depth != it.max_depth) {
                // keep the context (and its grown buffers) across depths
                const prg_status st = prg_set_max_depth(ctx, it.max_depth);
                if (st == PRG_OK) {
                    ctx_depth = it.max_depth;
                } else {
                    prg_destroy(ctx);
                    ctx = nullptr;
                }
            }
            if (!ctx) {
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
            const prg_status st = reconstruct_staged(ctx, it, subdivide, in, vbuf, tbuf, stats);
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
