// C ABI of the library (include/prg.h): argument checks, state machine,
// error capture and host <-> device buffer access.  Every entry point
// catches prg::Error and returns its status; nothing aborts the process.
#include <cstdlib>
#include <cstring>
#include <new>

#include "ctx.h"

using prg::Error;

namespace {

prg_status fail(prg_ctx* c, prg_status st, const std::string& msg) {
    if (c) c->err = msg;
    return st;
}

template <class F>
prg_status guarded(prg_ctx* c, F&& f) {
    if (!c) return PRG_E_ARG;
    try {
        return f();
    } catch (const Error& e) {
        c->err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return PRG_E_NOMEM;
    } catch (const std::exception& e) {
        c->err = e.what();
        return PRG_E_HIP;
    }
}

struct BufInfo {
    int64_t bytes;
};

int64_t buffer_bytes(const prg_ctx* c, int which) {
    const int64_t N = c->N, M = c->M, MD = c->MD();
    switch (which) {
        case PRG_BUF_POINTS:
        case PRG_BUF_NORMALS: return c->have_points ? 12 * N : -1;
        case PRG_BUF_POINT_LEAF: return c->have_tree ? 4 * N : -1;
        case PRG_BUF_KEY:
        case PRG_BUF_PIDX:
        case PRG_BUF_PNUM:
        case PRG_BUF_PARENT:
        case PRG_BUF_DIDX:
        case PRG_BUF_DNUM: return c->have_tree ? 4 * M : -1;
        case PRG_BUF_CHILDREN: return c->have_tree ? 32 * M : -1;
        case PRG_BUF_NEIGHS: return c->have_tree ? 108 * M : -1;
        case PRG_BUF_VECTOR_FIELD: return c->have_div ? 12 * MD : -1;
        case PRG_BUF_DIVERGENCE:
        case PRG_BUF_SOLUTION: return c->have_tree ? 4 * M : -1;
        case PRG_BUF_POINT_VALUES: return c->have_iso ? 4 * N : -1;
        case PRG_BUF_LEAF_CORNER_VALUES: return (c->mesh_nv > 0 || c->mesh_nt > 0 || c->corner_vals.get()) ? 32 * MD : -1;
        default: return -1;
    }
}

void d2h(prg_ctx* c, void* dst, const void* src, size_t bytes) {
    PRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
}

// Interleaves three device arrays of n floats into host AoS.
void read_xyz(prg_ctx* c, const float* a, const float* b, const float* d, int64_t n, float* host) {
    std::vector<float> t(static_cast<size_t>(3 * n));
    d2h(c, t.data(), a, size_t(n) * 4);
    d2h(c, t.data() + n, b, size_t(n) * 4);
    d2h(c, t.data() + 2 * n, d, size_t(n) * 4);
    PRG_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; i++) {
        host[3 * i] = t[i];
        host[3 * i + 1] = t[n + i];
        host[3 * i + 2] = t[2 * n + i];
    }
}

}  // namespace

extern "C" {

prg_status prg_config_default(prg_config* cfg) {
    if (!cfg) return PRG_E_ARG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->max_depth = 8;
    cfg->device = 0;
    cfg->cg_tol = 1e-6f;
    cfg->cg_max_iters = 2000;
    cfg->cg_refresh = 20;
    cfg->cg_fixed_iters = 0;
    cfg->warm_start = PRG_WARM_NONE;
    cfg->timing = 0;
    return PRG_OK;
}

prg_status prg_create(prg_ctx** out, const prg_config* cfg) {
    if (!out) return PRG_E_ARG;
    *out = nullptr;
    prg_config c0;
    prg_config_default(&c0);
    const prg_config& cf = cfg ? *cfg : c0;
    if (cf.max_depth < 2 || cf.max_depth > prg::kMaxDepth) return PRG_E_ARG;
    if (!(cf.cg_tol >= 0.0f) || cf.cg_max_iters < 1) return PRG_E_ARG;
    if (cf.warm_start != PRG_WARM_NONE && cf.warm_start != PRG_WARM_LEARNED) return PRG_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PRG_E_NODEVICE;
    if (cf.device < 0 || cf.device >= ndev) return PRG_E_ARG;
    prg_ctx* c = new (std::nothrow) prg_ctx();
    if (!c) return PRG_E_NOMEM;
    c->cfg = cf;
    c->D = cf.max_depth;
    prg_status st = guarded(c, [&]() {
        PRG_HIP(hipSetDevice(cf.device));
        PRG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        PRG_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        PRG_HIP(hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking));
        PRG_HIP(hipStreamCreateWithFlags(&c->stream4, hipStreamNonBlocking));
        PRG_HIP(hipEventCreateWithFlags(&c->plan_ev, hipEventDisableTiming));
        PRG_HIP(hipEventCreate(&c->topo_ev));
        PRG_HIP(hipEventCreate(&c->mesh_ev));
        PRG_HIP(hipEventCreate(&c->mesh_ev2));
        for (auto& e : c->stage_ev) {
            PRG_HIP(hipEventCreate(&e[0]));
            PRG_HIP(hipEventCreate(&e[1]));
        }
        PRG_HIP(hipHostMalloc(&c->transform_host, 16, hipHostMallocDefault));
        PRG_HIP(hipHostMalloc(&c->cg_host, 4096, hipHostMallocDefault));
        PRG_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->sub_host), 128, hipHostMallocDefault));
        PRG_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, cf.device));
        prg::weld_noalloc_scratch(cf.device);  // the fixed weld array's address on this device, once
        if (const char* lim = std::getenv("PRG_TILE_LDS_LIMIT")) c->tile_limit = std::max(0, std::atoi(lim));
        if (const char* ev = std::getenv("PRG_SUBDIVIDE_EVALUATE_ALL")) c->sub_certify = std::atoi(ev) == 0;
        if (const char* ws = std::getenv("PRG_WARM_START_SCALE")) c->warm_scale = float(std::atof(ws));
        c->kt.enabled = cf.timing != 0;
        c->kt.stream = c->stream;
        return PRG_OK;
    });
    if (st != PRG_OK) {
        prg_destroy(c);
        return st;
    }
    *out = c;
    return PRG_OK;
}

void prg_destroy(prg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t x : {c->stream2, c->stream3, c->stream4})
        if (x) (void)hipStreamSynchronize(x);
    if (c->plan_ev) (void)hipEventDestroy(c->plan_ev);
    if (c->topo_ev) (void)hipEventDestroy(c->topo_ev);
    if (c->mesh_ev) (void)hipEventDestroy(c->mesh_ev);
    if (c->mesh_ev2) (void)hipEventDestroy(c->mesh_ev2);
    for (auto& e : c->stage_ev)
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    if (c->transform_host) (void)hipHostFree(c->transform_host);
    if (c->cg_host) (void)hipHostFree(c->cg_host);
    if (c->sub_host) (void)hipHostFree(c->sub_host);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    hipStream_t ss[4] = {c->stream, c->stream2, c->stream3, c->stream4};
    delete c;  // releases device buffers
    for (hipStream_t x : ss)
        if (x) (void)hipStreamDestroy(x);
}

const char* prg_last_error(const prg_ctx* c) { return c ? c->err.c_str() : "null context"; }

prg_status prg_set_max_depth(prg_ctx* c, int32_t max_depth) {
    return guarded(c, [&]() {
        if (max_depth < 2 || max_depth > prg::kMaxDepth) return fail(c, PRG_E_ARG, "max_depth must be in [2, 10]");
        PRG_HIP(hipSetDevice(c->cfg.device));
        // work of the previous octree still queued on the other streams ends first
        for (hipStream_t x : {c->stream, c->stream2, c->stream3, c->stream4}) PRG_HIP(hipStreamSynchronize(x));
        c->plan_pending = c->topo_pending = false;
        c->cfg.max_depth = max_depth;
        c->D = max_depth;
        c->have_points = c->have_tree = c->have_div = c->have_x = c->have_iso = c->iso_ready = false;
        c->have_sub = false;
        c->mesh_nv = c->mesh_nt = 0;
        c->tree_gen++;
        return PRG_OK;
    });
}

prg_status prg_set_points(prg_ctx* c, const float* xyz, const float* nxyz, int64_t n, int32_t on_device) {
    return guarded(c, [&]() {
        if (!xyz || !nxyz) return fail(c, PRG_E_ARG, "null point pointer");
        if (n < 1 || n >= (int64_t(1) << 31)) return fail(c, PRG_E_ARG, "point count must be in [1, 2^31)");
        PRG_HIP(hipSetDevice(c->cfg.device));
        c->have_points = c->have_tree = c->have_div = c->have_x = c->have_iso = c->iso_ready = false;
        c->mesh_nv = c->mesh_nt = 0;
        c->stage_ms.fill(0.0);
        for (bool& p : c->stage_pending) p = false;
        prg::StageTimer st(c, PRG_STAGE_NORMALIZE);
        prg::normalize_points(c, xyz, nxyz, n, on_device != 0);
        // the caller's arrays are consumed when this call returns
        if (on_device) PRG_HIP(hipStreamSynchronize(c->stream));
        c->have_points = true;
        return PRG_OK;
    });
}

prg_status prg_build_octree(prg_ctx* c) {
    return guarded(c, [&]() {
        if (!c->have_points) return fail(c, PRG_E_STATE, "prg_set_points first");
        PRG_HIP(hipSetDevice(c->cfg.device));
        // plans or topology of the previous octree may still read its tables
        if (c->plan_pending) PRG_HIP(hipStreamWaitEvent(c->stream, c->plan_ev, 0));
        if (c->topo_pending) PRG_HIP(hipStreamWaitEvent(c->stream, c->topo_ev, 0));
        {
            prg::StageTimer st(c, PRG_STAGE_OCTREE);
            prg::build_octree(c);
        }
        {
            prg::StageTimer st(c, PRG_STAGE_TABLES);
            prg::ensure_tables(c);
        }
        c->have_tree = true;
        c->tree_gen++;
        prg::start_plans(c);
        prg::start_topology(c);
        // divergence and solution read as zeros until computed (zeroed when
        // read; the divergence stage writes every node and the solve starts
        // from zero itself)
        c->have_div = c->have_x = c->have_iso = false;
        return PRG_OK;  // the later stages are ordered on the context's stream
    });
}

prg_status prg_level_info(prg_ctx* c, int64_t* counts, int64_t* bases) {
    return guarded(c, [&]() {
        if (!c->have_tree) return fail(c, PRG_E_STATE, "prg_build_octree first");
        if (!counts || !bases) return fail(c, PRG_E_ARG, "null output");
        for (int d = 0; d <= c->D; d++) {
            counts[d] = c->counts[d];
            bases[d] = c->bases[d];
        }
        return PRG_OK;
    });
}

prg_status prg_level_view(prg_ctx* c, int32_t depth, prg_level_desc* v) {
    return guarded(c, [&]() {
        if (!c->have_tree) return fail(c, PRG_E_STATE, "prg_build_octree first");
        if (!v || depth < 0 || depth > c->D) return fail(c, PRG_E_ARG, "bad depth");
        prg::ensure_neighbours(c);
        PRG_HIP(hipStreamSynchronize(c->stream));  // the viewed arrays are complete
        const int64_t b = c->bases[depth];
        v->depth = depth;
        v->count = c->counts[depth];
        v->base = b;
        v->key = c->key.get() + b;
        v->pidx = c->pidx.get() + b;
        v->pnum = c->pnum.get() + b;
        v->parent = c->parent.get() + b;
        v->child_base = c->child_base.get() + b;
        v->didx = c->didx.get() + b;
        v->dnum = c->dnum.get() + b;
        v->neighs = c->neigh.get() + b;
        v->neigh_stride = c->M;
        return PRG_OK;
    });
}

prg_status prg_compute_divergence(prg_ctx* c) {
    return guarded(c, [&]() {
        if (!c->have_tree) return fail(c, PRG_E_STATE, "prg_build_octree first");
        PRG_HIP(hipSetDevice(c->cfg.device));
        prg::compute_divergence(c);
        c->have_div = true;
        c->have_x = c->have_iso = false;
        return PRG_OK;
    });
}

prg_status prg_solve(prg_ctx* c, prg_solve_stats* per_depth) {
    return guarded(c, [&]() {
        if (!c->have_div) return fail(c, PRG_E_STATE, "prg_compute_divergence first");
        PRG_HIP(hipSetDevice(c->cfg.device));
        prg_status st = prg::solve_all(c, per_depth);
        c->have_x = true;
        c->x_gen++;
        c->have_iso = c->iso_ready = false;
        if (st == PRG_E_NOT_CONVERGED) c->err = "CG reached its iteration cap";
        return st;
    });
}

prg_status prg_isovalue(prg_ctx* c, float* out) {
    return guarded(c, [&]() {
        if (!c->have_x) return fail(c, PRG_E_STATE, "prg_solve first");
        PRG_HIP(hipSetDevice(c->cfg.device));
        prg::compute_isovalue(c);
        c->have_iso = true;
        c->iso_ready = true;
        if (out) {
            PRG_HIP(hipMemcpyAsync(&c->iso, c->iso_dev.get(), 4, hipMemcpyDeviceToHost, c->stream));
            PRG_HIP(hipStreamSynchronize(c->stream));
            *out = c->iso;
        }
        return PRG_OK;
    });
}

prg_status prg_set_isovalue(prg_ctx* c, float iso) {
    return guarded(c, [&]() {
        PRG_HIP(hipSetDevice(c->cfg.device));
        c->iso = iso;
        PRG_HIP(hipMemcpyAsync(c->iso_dev.reserve(1), &c->iso, 4, hipMemcpyHostToDevice, c->stream));
        c->iso_ready = true;
        return PRG_OK;
    });
}

namespace {

// The host copy of the normalisation transform, once its copy has landed.
static void resolve_transform(prg_ctx* c) {
    if (!c->transform_pending) return;
    PRG_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; a++) c->center[a] = c->transform_host[a];
    c->scale = c->transform_host[3];
    c->transform_pending = false;
}

// Leaf-level marching cubes, then (subdivide != 0) the coarse-leaf
// subdivision batches; sets the counts of `out`.
static void build_mesh(prg_ctx* c, int32_t subdivide, prg_mesh* out) {
    PRG_HIP(hipSetDevice(c->cfg.device));
    prg::extract_leaf_mesh(c);
    c->sub_nv = c->sub_nt = 0;
    if (subdivide) prg::extract_subdivision(c);
    std::memset(out, 0, sizeof(*out));
    resolve_transform(c);
    out->n_leaf_vertices = c->mesh_nv;
    out->n_leaf_triangles = c->mesh_nt;
    out->n_vertices = c->mesh_nv + c->sub_nv;
    out->n_triangles = c->mesh_nt + c->sub_nt;
    for (int a = 0; a < 3; a++) out->center[a] = c->center[a];
    out->scale = c->scale;
}

// Copies the leaf-level batch and the subdivision batches, in that order
// (the leaf-level batch may be on its way already: out_copied).
static void copy_mesh(prg_ctx* c, float* vertices, int32_t* triangles) {
    prg::StageTimer st(c, PRG_STAGE_MESH_COPY);
    if (!c->out_copied) {
        if (c->mesh_nv) d2h(c, vertices, c->mesh_v.get(), size_t(3 * c->mesh_nv) * 4);
        if (c->mesh_nt) d2h(c, triangles, c->mesh_t.get(), size_t(3 * c->mesh_nt) * 4);
    }
    if (c->sub_nv) d2h(c, vertices + 3 * c->mesh_nv, c->sub_v.get(), size_t(3 * c->sub_nv) * 4);
    if (c->sub_nt) d2h(c, triangles + 3 * c->mesh_nt, c->sub_t.get(), size_t(3 * c->sub_nt) * 4);
    if (c->out_copied && c->out_async) PRG_HIP(hipStreamWaitEvent(c->stream, c->mesh_ev, 0));
    PRG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

prg_status prg_extract_mesh(prg_ctx* c, int32_t subdivide, prg_mesh* out) {
    return guarded(c, [&]() {
        if (!c->have_x) return fail(c, PRG_E_STATE, "prg_solve first");
        if (!c->iso_ready) return fail(c, PRG_E_STATE, "prg_isovalue or prg_set_isovalue first");
        if (!out) return fail(c, PRG_E_ARG, "null mesh");
        build_mesh(c, subdivide, out);
        out->vertices = static_cast<float*>(std::malloc(size_t(3 * out->n_vertices + 1) * 4));
        out->triangles = static_cast<int32_t*>(std::malloc(size_t(3 * out->n_triangles + 1) * 4));
        if (!out->vertices || !out->triangles) {
            prg_mesh_release(out);
            return fail(c, PRG_E_NOMEM, "mesh host allocation");
        }
        copy_mesh(c, out->vertices, out->triangles);
        return PRG_OK;
    });
}

prg_status prg_extract_mesh_into(prg_ctx* c, int32_t subdivide, float* vertices, int64_t vcap, int32_t* triangles,
                                 int64_t tcap, prg_mesh* out) {
    return guarded(c, [&]() {
        if (!c->have_x) return fail(c, PRG_E_STATE, "prg_solve first");
        if (!c->iso_ready) return fail(c, PRG_E_STATE, "prg_isovalue or prg_set_isovalue first");
        if (!out) return fail(c, PRG_E_ARG, "null mesh");
        // the leaf-level batch goes straight into the caller's buffers while
        // it is produced; with the subdivision, its copy overlaps the
        // subdivision's kernels
        c->out_v = vertices;
        c->out_t = triangles;
        c->out_vcap = vcap;
        c->out_tcap = tcap;
        c->out_async = subdivide != 0;
        struct Reset {
            prg_ctx* c;
            ~Reset() {
                // an exception may leave copies in flight into the caller's buffers
                if (c->out_copied && c->out_async) (void)hipStreamSynchronize(c->stream2);
                c->out_v = nullptr;
                c->out_t = nullptr;
                c->out_vcap = c->out_tcap = 0;
                c->out_async = false;
            }
        } reset{c};
        build_mesh(c, subdivide, out);
        const int64_t nv = out->n_vertices, nt = out->n_triangles;
        if (3 * nv > vcap || 3 * nt > tcap || (!vertices && nv) || (!triangles && nt))
            return fail(c, PRG_E_ARG, "mesh buffers too small");
        copy_mesh(c, vertices, triangles);
        out->vertices = vertices;
        out->triangles = triangles;
        return PRG_OK;
    });
}

prg_status prg_decimate_mesh(prg_mesh* mesh, int32_t mode, float edge_ratio) {
    if (!mesh || (mode != 0 && mode != 1) || !(edge_ratio > 0.0f)) return PRG_E_ARG;
    const int64_t n = mesh->n_vertices, m = mesh->n_triangles;
    if (n < 0 || m < 0 || (n > 0 && !mesh->vertices) || (m > 0 && !mesh->triangles)) return PRG_E_ARG;
    if (n >= (int64_t(1) << 31)) return PRG_E_ARG;
    for (int64_t i = 0; i < 3 * m; i++)
        if (mesh->triangles[i] < 0 || mesh->triangles[i] >= n) return PRG_E_ARG;
    try {
        int64_t nv = n, nt = m;
        prg::decimate(mode, edge_ratio, mesh->vertices, &nv, mesh->triangles, &nt);
        mesh->n_vertices = nv;
        mesh->n_triangles = nt;
        mesh->n_leaf_vertices = mesh->n_leaf_triangles = 0;
    } catch (const std::bad_alloc&) {
        return PRG_E_NOMEM;
    }
    return PRG_OK;
}

prg_status prg_subdivision_info(prg_ctx* c, int64_t* n_roots, int64_t* n_virtual_leaves) {
    return guarded(c, [&]() {
        if (!n_roots || !n_virtual_leaves) return fail(c, PRG_E_ARG, "null output");
        *n_roots = c->sub_roots;
        *n_virtual_leaves = c->sub_leaves;
        return PRG_OK;
    });
}

prg_status prg_subdivision_bricks(prg_ctx* c, int64_t* n_bricks, int64_t* n_exact, int64_t* n_evaluated) {
    return guarded(c, [&]() {
        if (!n_bricks || !n_exact || !n_evaluated) return fail(c, PRG_E_ARG, "null output");
        *n_bricks = c->sub_bricks;
        *n_exact = c->sub_exact;
        *n_evaluated = c->sub_evaluated;
        return PRG_OK;
    });
}

namespace {
static prg_status weld_checked(prg_ctx* c, prg_mesh* mesh, int64_t* n_merged, bool noalloc) {
    return guarded(c, [&]() {
        if (!mesh) return fail(c, PRG_E_ARG, "null mesh");
        const int64_t n = mesh->n_vertices, m = mesh->n_triangles;
        if (n < 0 || m < 0 || (n > 0 && !mesh->vertices) || (m > 0 && !mesh->triangles))
            return fail(c, PRG_E_ARG, "mesh arrays missing");
        if (n >= (int64_t(1) << 31)) return fail(c, PRG_E_ARG, "more than 2^31 - 1 vertices");
        if (noalloc && (n > PRG_WELD_NOALLOC_MAX_VERTICES || m > PRG_WELD_NOALLOC_MAX_TRIANGLES))
            return fail(c, PRG_E_ARG, "mesh larger than the fixed weld array (PRG_WELD_NOALLOC_MAX_*)");
        for (int64_t i = 0; i < 3 * m; i++)
            if (mesh->triangles[i] < 0 || mesh->triangles[i] >= n)
                return fail(c, PRG_E_ARG, "triangle index out of range");
        PRG_HIP(hipSetDevice(c->cfg.device));
        int64_t nv = n, nleaf = mesh->n_leaf_vertices;
        if (noalloc)
            prg::weld_mesh_noalloc(c, mesh->vertices, n, mesh->triangles, m, mesh->n_leaf_vertices, &nv, &nleaf);
        else
            prg::weld_mesh(c, mesh->vertices, n, mesh->triangles, m, mesh->n_leaf_vertices, &nv, &nleaf);
        if (n_merged) *n_merged = n - nv;
        mesh->n_vertices = nv;
        mesh->n_leaf_vertices = nleaf;
        return PRG_OK;
    });
}
}  // namespace

prg_status prg_weld_mesh(prg_ctx* c, prg_mesh* mesh, int64_t* n_merged) { return weld_checked(c, mesh, n_merged, false); }

prg_status prg_weld_mesh_noalloc(prg_ctx* c, prg_mesh* mesh, int64_t* n_merged) {
    return weld_checked(c, mesh, n_merged, true);
}

void prg_mesh_release(prg_mesh* m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->triangles);
    m->vertices = nullptr;
    m->triangles = nullptr;
    m->n_vertices = m->n_triangles = 0;
}

prg_status prg_reconstruct(prg_ctx* c, int32_t subdivide, prg_mesh* out, prg_solve_stats* per_depth) {
    prg_status st = prg_build_octree(c);
    if (st != PRG_OK) return st;
    if ((st = prg_compute_divergence(c)) != PRG_OK) return st;
    st = prg_solve(c, per_depth);
    if (st != PRG_OK && st != PRG_E_NOT_CONVERGED) return st;
    prg_status st2 = prg_isovalue(c, nullptr);
    if (st2 != PRG_OK) return st2;
    st2 = prg_extract_mesh(c, subdivide, out);
    if (st2 != PRG_OK) return st2;
    return st;
}

prg_status prg_stage_times(prg_ctx* c, prg_timings* out) {
    return guarded(c, [&]() {
        if (!out) return fail(c, PRG_E_ARG, "null output");
        for (int i = 0; i < PRG_STAGE_COUNT; i++) {
            prg::resolve_stage(c, i);
            out->ms[i] = c->stage_ms[i];
        }
        return PRG_OK;
    });
}

prg_status prg_buffer_size(prg_ctx* c, int32_t which, int64_t* bytes) {
    return guarded(c, [&]() {
        if (!bytes) return fail(c, PRG_E_ARG, "null output");
        int64_t b = buffer_bytes(c, which);
        if (b < 0) return fail(c, PRG_E_STATE, "buffer not available");
        *bytes = b;
        return PRG_OK;
    });
}

prg_status prg_read_buffer(prg_ctx* c, int32_t which, void* host, int64_t bytes) {
    return guarded(c, [&]() {
        const int64_t b = buffer_bytes(c, which);
        if (b < 0) return fail(c, PRG_E_STATE, "buffer not available");
        if (!host || bytes != b) return fail(c, PRG_E_ARG, "size mismatch");
        PRG_HIP(hipSetDevice(c->cfg.device));
        const int64_t M = c->M;
        switch (which) {
            case PRG_BUF_POINTS:
                prg::unpack_points(c);
                read_xyz(c, c->px.get(), c->py.get(), c->pz.get(), c->N, (float*)host);
                break;
            case PRG_BUF_NORMALS:
                prg::unpack_points(c);
                read_xyz(c, c->nx.get(), c->ny.get(), c->nz.get(), c->N, (float*)host);
                break;
            case PRG_BUF_POINT_LEAF: d2h(c, host, c->point_leaf.get(), size_t(b)); break;
            case PRG_BUF_KEY: d2h(c, host, c->key.get(), size_t(b)); break;
            case PRG_BUF_PIDX: d2h(c, host, c->pidx.get(), size_t(b)); break;
            case PRG_BUF_PNUM: d2h(c, host, c->pnum.get(), size_t(b)); break;
            case PRG_BUF_PARENT: d2h(c, host, c->parent.get(), size_t(b)); break;
            case PRG_BUF_DIDX: d2h(c, host, c->didx.get(), size_t(b)); break;
            case PRG_BUF_DNUM: d2h(c, host, c->dnum.get(), size_t(b)); break;
            case PRG_BUF_CHILDREN: {
                std::vector<int32_t> cb(static_cast<size_t>(M));
                d2h(c, cb.data(), c->child_base.get(), size_t(M) * 4);
                PRG_HIP(hipStreamSynchronize(c->stream));
                int32_t* o = (int32_t*)host;
                for (int64_t i = 0; i < M; i++)
                    for (int k = 0; k < 8; k++) o[8 * i + k] = cb[i] < 0 ? -1 : cb[i] + k;
                break;
            }
            case PRG_BUF_NEIGHS: {
                prg::ensure_neighbours(c);
                std::vector<int32_t> t(static_cast<size_t>(27 * M));
                d2h(c, t.data(), c->neigh.get(), size_t(27 * M) * 4);
                PRG_HIP(hipStreamSynchronize(c->stream));
                int32_t* o = (int32_t*)host;
                for (int64_t i = 0; i < M; i++)
                    for (int j = 0; j < 27; j++) o[27 * i + j] = t[size_t(j) * M + i];
                break;
            }
            case PRG_BUF_VECTOR_FIELD: read_xyz(c, c->vfx.get(), c->vfy.get(), c->vfz.get(), c->MD(), (float*)host); break;
            case PRG_BUF_DIVERGENCE:
                if (!c->have_div) PRG_HIP(hipMemsetAsync(c->div.reserve(M), 0, size_t(b), c->stream));
                d2h(c, host, c->div.get(), size_t(b));
                break;
            case PRG_BUF_SOLUTION:
                if (!c->have_x) PRG_HIP(hipMemsetAsync(c->x.reserve(M), 0, size_t(b), c->stream));
                d2h(c, host, c->x.get(), size_t(b));
                break;
            case PRG_BUF_POINT_VALUES: d2h(c, host, c->pvals.get(), size_t(b)); break;
            case PRG_BUF_LEAF_CORNER_VALUES: d2h(c, host, c->corner_vals.get(), size_t(b)); break;
            default: return fail(c, PRG_E_ARG, "unknown buffer");
        }
        PRG_HIP(hipStreamSynchronize(c->stream));
        return PRG_OK;
    });
}

prg_status prg_write_buffer(prg_ctx* c, int32_t which, const void* host, int64_t bytes) {
    return guarded(c, [&]() {
        if (which != PRG_BUF_DIVERGENCE && which != PRG_BUF_SOLUTION)
            return fail(c, PRG_E_ARG, "buffer is not writable");
        const int64_t b = buffer_bytes(c, which);
        if (b < 0) return fail(c, PRG_E_STATE, "prg_build_octree first");
        if (!host || bytes != b) return fail(c, PRG_E_ARG, "size mismatch");
        PRG_HIP(hipSetDevice(c->cfg.device));
        float* dst = (which == PRG_BUF_DIVERGENCE) ? c->div.reserve(c->M) : c->x.reserve(c->M);
        PRG_HIP(hipMemcpyAsync(dst, host, size_t(b), hipMemcpyHostToDevice, c->stream));
        PRG_HIP(hipStreamSynchronize(c->stream));
        // results derived from the overwritten buffer are stale: a new
        // divergence invalidates the solution, a new solution the isovalue
        // and the values at the samples (prg_isovalue or prg_set_isovalue
        // must run again before a mesh is extracted)
        if (which == PRG_BUF_DIVERGENCE) {
            prg::ensure_tables(c);
            c->have_div = true;
            c->have_x = false;
        } else {
            c->have_x = true;
            c->x_gen++;
        }
        c->have_iso = c->iso_ready = false;
        return PRG_OK;
    });
}

prg_status prg_kernel_stats(prg_ctx* c, const char* name, int64_t* launches, double* total_ms,
                            double* bytes_per_launch) {
    return guarded(c, [&]() {
        if (!name) return fail(c, PRG_E_ARG, "null name");
        c->kt.flush();
        auto it = c->kt.recs.find(name);
        if (it == c->kt.recs.end()) return fail(c, PRG_E_ARG, std::string("no launches of ") + name);
        if (launches) *launches = it->second.launches;
        if (total_ms) *total_ms = it->second.total_ms;
        if (bytes_per_launch)
            *bytes_per_launch = it->second.launches ? it->second.bytes / double(it->second.launches) : 0.0;
        return PRG_OK;
    });
}

prg_status prg_kernel_stats_reset(prg_ctx* c) {
    return guarded(c, [&]() {
        c->kt.flush();
        c->kt.recs.clear();
        return PRG_OK;
    });
}

prg_status prg_get_transform(prg_ctx* c, float* center3, float* scale) {
    return guarded(c, [&]() {
        if (!c->have_points) return fail(c, PRG_E_STATE, "prg_set_points first");
        resolve_transform(c);
        if (center3)
            for (int a = 0; a < 3; a++) center3[a] = c->center[a];
        if (scale) *scale = c->scale;
        return PRG_OK;
    });
}

}  // extern "C"
