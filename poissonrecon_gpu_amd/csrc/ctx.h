// Internal context of one reconstruction (one HIP device, one stream).
// Device layout (all structure-of-arrays, node arrays indexed by the global
// node index; depth d occupies [base[d], base[d] + count[d]) as in the
// reference's concatenated NodeArray, main.cu:772-794):
//   points     px,py,pz,nx,ny,nz      float[N]   normalised, then sorted
//   nodes      key                    uint32[M]
//              pidx,pnum,parent,child_base,didx,dnum   int32[M]
//              neigh                  int32[27][M]  (plane j holds neighbour j of every node)
//              pnt                    int32[M/8][27] per depth >= 1: for each sibling block, the
//                                     level-local child base of each of its parent's 27
//                                     neighbours (-1 if absent); neighbour j of slot s is
//                                     pnt[b][LUTparent[s][j]] + LUTchild[s][j]
//   fields     vfield_{x,y,z}         float[M_D]
//              div, x                 float[M]
#pragma once

#include <array>
#include <memory>

#include "common.h"

struct prg_unet;

struct prg_ctx {
    prg_config cfg{};
    int D = 8;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // independent work overlapping `stream` (CG chains, small-depth solves)
    hipStream_t stream3 = nullptr;
    hipStream_t stream4 = nullptr;  // the warm-started depth (after the network's inference)
    std::string err;
    prg::KernelTimer kt;
    prg::Scratch scratch;
    prg::Scratch scratch2;  // scans of the work started on the second stream

    // points
    int64_t N = 0;
    float center[3] = {0, 0, 0};
    float scale = 1.0f;
    prg::DBuf<float> px, py, pz, nx, ny, nz;       // normalised input, sorted in place by the octree build
    // normalised points in input order, 8 floats each (x, y, z, nx, ny, nz, 0, 0);
    // px .. nz hold them in key order once the octree is built
    prg::DBuf<float> pstage;
    bool points_sorted = false;
    prg::DBuf<int32_t> point_leaf;                 // local leaf-level slot of each sorted point

    // octree
    bool have_points = false, have_tree = false, have_div = false, have_x = false, have_iso = false;
    std::vector<int64_t> counts, bases;
    int64_t M = 0;
    prg::DBuf<uint32_t> key;
    prg::DBuf<int32_t> pidx, pnum, parent, child_base, didx, dnum, neigh;
    prg::DBuf<int32_t> pnt;            // concatenated per depth 1..D, block rows of 27
    std::vector<int64_t> pnt_off;      // first row of depth d (rows = blocks)

    // function space
    std::shared_ptr<const prg::FsTables> fs;  // shared by every context of the same depth
    prg::DBuf<float> d_base_fn, d_stencil, d_div_table;

    // fields
    prg::DBuf<float> vfx, vfy, vfz, div, x, pvals;
    prg::DBuf<double> dscalars;  // reductions and CG scalars
    // CG scratch, kept across solves: three concurrent chains of large depths
    // (r, p0, p1, q, aligned x; FP64 partials; state), the small-depth vectors
    prg::DBuf<float> cg_vec[3][5];
    prg::DBuf<double> cg_part[3][4];  // r.r (two slots), p.q, b.b / true-residual partials
    prg::DBuf<uint8_t> cg_state[3];
    prg::DBuf<uint8_t> cg_seg;  // the solver's per-depth segment table
    prg::DBuf<double> cg_out;
    // SpMV tile plans of the large depths, valid for octree `plan_gen`
    prg::DBuf<int32_t> plan_count, plan_ubase;
    prg::DBuf<int16_t> plan_rowid;
    int64_t plan_tile_off[prg::kMaxDepth + 2] = {};
    int32_t plan_maxcount[prg::kMaxDepth + 1] = {};
    int64_t plan_cells[prg::kMaxDepth + 1] = {};  // per depth: distinct blocks listed over all tiles
    uint64_t tree_gen = 0, plan_gen = 0;
    // per-node rows of the coarse depths (27 neighbour solution values and the
    // parent), for the solution `x_gen` of octree `rows_tree`
    prg::DBuf<float> coarse_rows;
    int32_t rows_depth = 0;
    uint64_t x_gen = 0, rows_x = ~uint64_t(0), rows_tree = ~uint64_t(0);
    prg::DBuf<int32_t> plan_mx;       // per-depth largest tile count and sum of tile counts (device)
    hipEvent_t plan_ev = nullptr;     // completion of the plans started after an octree build
    bool plan_pending = false;
    // marching-cubes topology of the octree `topo_gen` (corner owners and
    // their numbering, the corner basis table), started after the build
    uint64_t topo_gen = 0;
    hipEvent_t topo_ev = nullptr;
    // caller buffers a leaf-level mesh is copied into while it is produced
    // (prg_extract_mesh_into), and whether the last extraction did so
    float* out_v = nullptr;
    int32_t* out_t = nullptr;
    int64_t out_vcap = 0, out_tcap = 0;
    bool out_copied = false;
    // with the subdivision to follow: the leaf-level copies run on the second
    // stream behind mesh_ev while the subdivision runs (mesh_ev2 orders them
    // after the triangle emission)
    bool out_async = false;
    hipEvent_t mesh_ev = nullptr;
    hipEvent_t mesh_ev2 = nullptr;
    bool topo_pending = false;
    int32_t topo_corners = 0;  // owned corners
    bool neigh_ready = false;  // the 27 neighbour planes (depths 1..D) are built on demand
    prg_unet* warm_net = nullptr;  // learned warm start (not owned)
    // normalisation transform (centre x, y, z, scale) computed on the device;
    // the host copy is read at the next synchronisation
    prg::DBuf<float> transform_dev;
    float* transform_host = nullptr;  // pinned, 4 floats
    bool transform_pending = false;
    // per-stage temporaries (slots: octree 0-59, divergence 60-75, marching cubes 76-95, warm start 96)
    prg::DBuf<uint8_t> pool[144];
    float iso = 0.0f;
    bool iso_ready = false;  // iso computed or set by the caller
    // the isovalue in device memory (read by the marching-cubes kernels);
    // `iso` holds the last value read back or set by the caller
    prg::DBuf<float> iso_dev;

    // leaf-level marching cubes results (device)
    prg::DBuf<float> corner_vals;  // 8 * M_D, square corner numbering
    prg::DBuf<float> mesh_v;
    prg::DBuf<int32_t> mesh_t;
    int64_t mesh_nv = 0, mesh_nt = 0;
    // coarse-leaf subdivision batches (vertex indices continue the leaf-level batch's)
    prg::DBuf<float> sub_v;
    prg::DBuf<int32_t> sub_t;
    int64_t sub_nv = 0, sub_nt = 0, sub_roots = 0, sub_leaves = 0;
    int64_t sub_bricks = 0, sub_exact = 0, sub_evaluated = 0;  // bricks; left by the interval; evaluated
    bool have_sub = false;

    // stage timing: start/end events per stage, resolved lazily (when the
    // stage runs again or the times are read), so no stage waits for its own
    // measurement
    std::array<double, PRG_STAGE_COUNT> stage_ms{};
    hipEvent_t stage_ev[PRG_STAGE_COUNT][2] = {};
    bool stage_pending[PRG_STAGE_COUNT] = {};
    std::vector<hipEvent_t> ev_pool;  // the solver's events, reused across solves
    int num_cu = 0;                   // compute units of the device
    void* cg_host = nullptr;          // pinned, 4 KiB: CG state read-backs and the segment table's staging
    int32_t* sub_host = nullptr;      // pinned, 32 ints: the subdivision's counters
    // test setting (environment PRG_TILE_LDS_LIMIT at creation): caps the
    // listed blocks and points the leaf-level tile kernels stage in LDS, so
    // their per-row and global-memory paths run on ordinary inputs
    int32_t tile_limit = 0;
    // test knob (PRG_SUBDIVIDE_EVALUATE_ALL=1): evaluate every subdivision brick, no certificates
    bool sub_certify = true;
    // test knob (PRG_WARM_START_SCALE): multiplies the learned warm start's prediction
    float warm_scale = 1.0f;
    // last solve: the warm-started depth (-1: none) and whether its start was
    // rejected (||b - A x0|| > ||b||: the depth started from zero instead)
    int32_t warm_depth_used = -1, warm_rejected = 0;

    int64_t MD() const { return counts.empty() ? 0 : counts[D]; }
    int64_t base_of(int d) const { return bases[d]; }
};

namespace prg {

// Stage entry points implemented in the .hip files.
void normalize_points(prg_ctx* c, const float* xyz, const float* nxyz, int64_t n, bool on_device);
void build_octree(prg_ctx* c);
// px .. nz in input order (before an octree is built)
void unpack_points(prg_ctx* c);
// merges bit-identical vertices of a host mesh in place (triangle entries
// already checked to lie in [0, n))
void weld_mesh(prg_ctx* c, float* vertices, int64_t n, int32_t* triangles, int64_t m, int64_t n_leaf,
               int64_t* n_out, int64_t* n_leaf_out);
// the same without allocation: every device array in the library's fixed
// per-device weld array (limits in prg.h)
void weld_mesh_noalloc(prg_ctx* c, float* vertices, int64_t n, int32_t* triangles, int64_t m, int64_t n_leaf,
                       int64_t* n_out, int64_t* n_leaf_out);
// the device address of the fixed weld array on `device` (current), looked up once per device
uint8_t* weld_noalloc_scratch(int device);
// EdgeCollapse (mode 0) / TriangleCollapse (mode 1) of a host mesh in place (decimate.cpp)
void decimate(int mode, float edge_ratio, float* vertices, int64_t* n_io, int32_t* triangles, int64_t* m_io);
// the 27-neighbour planes of every depth (kernels use the parent-neighbour
// rows; these serve level views, buffer reads and the subdivision)
void ensure_neighbours(prg_ctx* c);
// SpMV tile plans of the new octree, started on the second stream right after
// its build (they overlap the vector field and divergence); the solve waits
// for them
void start_plans(prg_ctx* c);
// waits (on the host) for the plans of the current octree
void ensure_plans(prg_ctx* c);
// the plan of depth d for work on stream s (which waits for the plans on the
// device); false where the depth has none (kSmallRows rows or fewer)
struct TilePlan;
bool tile_plan_on(prg_ctx* c, int d, TilePlan* P, hipStream_t s);
// marching-cubes topology of the new octree, on the second stream after the
// plans (it depends on the tree only)
void start_topology(prg_ctx* c);
void ensure_tables(prg_ctx* c);
void compute_divergence(prg_ctx* c);
prg_status solve_all(prg_ctx* c, prg_solve_stats* stats);
void compute_isovalue(prg_ctx* c);
void extract_leaf_mesh(prg_ctx* c);
void extract_subdivision(prg_ctx* c);
// learned warm start: enqueues on `s` the network's prediction for depth
// `depth`, written into the solution vector x (device)
void unet_warm_start(prg_ctx* c, int depth, hipStream_t s);

// adds a recorded measurement of stage st to stage_ms
inline void resolve_stage(prg_ctx* c, int st) {
    if (!c->stage_pending[st]) return;
    PRG_HIP(hipEventSynchronize(c->stage_ev[st][1]));
    float ms = 0.0f;
    PRG_HIP(hipEventElapsedTime(&ms, c->stage_ev[st][0], c->stage_ev[st][1]));
    c->stage_ms[st] += ms;
    c->stage_pending[st] = false;
}

// Measures one stage with hipEvents on the context stream.
struct StageTimer {
    prg_ctx* c;
    int stage;
    StageTimer(prg_ctx* ctx, int st) : c(ctx), stage(st) {
        resolve_stage(c, st);
        PRG_HIP(hipEventRecord(c->stage_ev[st][0], c->stream));
    }
    ~StageTimer() noexcept(false) {
        PRG_HIP(hipEventRecord(c->stage_ev[stage][1], c->stream));
        c->stage_pending[stage] = true;
    }
};

}  // namespace prg
