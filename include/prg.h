/*
 * prg.h - C ABI of the MI355X Poisson surface-reconstruction pipeline.
 *
 * The reference (DavidXu-JJ/PoissonRecon_GPU) has no library API: its stage
 * boundaries are host functions inside main.cu that pass raw device pointers.
 * This header keeps those stage boundaries as plain C entry points (plain
 * pointers and sizes, no C++ or torch types).  Each declaration names the
 * reference code it replaces (file:line in the reference tree).
 *
 * Conventions
 *   - every call returns prg_status; nothing calls exit();
 *   - the library owns device memory; host outputs are caller-provided or
 *     released with the matching *_release call;
 *   - a context is single-threaded and owns one HIP stream;
 *   - sizes are int64_t.
 */
#ifndef PRG_H
#define PRG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum prg_status {
    PRG_OK = 0,
    PRG_E_ARG = 1,           /* bad argument (depth out of range, n >= 2^31, ...) */
    PRG_E_NOMEM = 2,         /* device or host allocation failed                  */
    PRG_E_HIP = 3,           /* a HIP runtime call or kernel launch failed        */
    PRG_E_IO = 4,            /* unreadable / malformed point or mesh file         */
    PRG_E_NOT_CONVERGED = 5, /* CG hit its iteration cap (stats still returned)   */
    PRG_E_STATE = 6,         /* stage called before the stage it depends on       */
    PRG_E_NODEVICE = 7       /* no HIP device                                     */
} prg_status;

typedef enum prg_warm_start {
    PRG_WARM_NONE = 0,   /* x0 = 0 at every depth (reference behaviour, main.cu:1237-1329)        */
    PRG_WARM_LEARNED = 2 /* the attached U-Net's prediction at its grid's depth (prg_unet.h);
                            every other depth from zero                                          */
} prg_warm_start;

/* Runtime configuration; replaces the compile-time #defines of main.cu:29-78. */
typedef struct prg_config {
    int32_t max_depth;      /* octree depth D, 2..10 (reference: #define maxDepth, main.cu:69) */
    int32_t device;         /* HIP device ordinal                                              */
    float cg_tol;           /* relative residual target ||b-Ax||/||b|| (default 1e-6)          */
    int32_t cg_max_iters;   /* iteration cap per depth (reference: 10000, CG_CUDA.cuh:263)     */
    int32_t cg_refresh;     /* recompute r = b - A x every this many iterations (default 20)  */
    int32_t cg_fixed_iters; /* > 0: run exactly this many iterations per depth, ignore cg_tol  */
    int32_t warm_start;     /* prg_warm_start                                                  */
    int32_t timing;         /* 1: record per-kernel hipEvent timings (prg_kernel_stats)        */
} prg_config;

typedef struct prg_ctx prg_ctx;

/* Per-depth solve report (reference prints these, CG_CUDA.cuh:469-470,506). */
typedef struct prg_solve_stats {
    int32_t depth;
    int32_t iters;
    int64_t nodes;
    double rel_residual;      /* recurrence residual at exit, relative to ||b||          */
    double true_rel_residual; /* ||b - A x|| / ||b|| recomputed from x at exit           */
    double ms;                /* device time of this depth's solve (hipEvents)           */
} prg_solve_stats;

/* Output triangle mesh.  Vertices are in the normalised unit cube exactly as
 * the reference's VertexBuffer (main.cu:2599-2627); `center` and `scale`
 * map them back to input coordinates as PlyWriteTriangles does
 * (plyfile.cu:2800-2806): out = v * scale + center. */
typedef struct prg_mesh {
    int64_t n_vertices;
    int64_t n_triangles;
    float* vertices;    /* 3 * n_vertices  */
    int32_t* triangles; /* 3 * n_triangles */
    float center[3];
    float scale;
    int64_t n_leaf_vertices;  /* vertices / triangles of the leaf-level batch (main.cu:3793) */
    int64_t n_leaf_triangles;
} prg_mesh;

/* Stage timer slots; names mirror the reference's printed stages (main.cu:575-4571). */
typedef enum prg_stage {
    PRG_STAGE_NORMALIZE = 0,  /* bounding box + normalisation (main.cu:530-571)          */
    PRG_STAGE_OCTREE = 1,     /* "GPU NodeArray build" (main.cu:839)                     */
    PRG_STAGE_TABLES = 2,     /* "CPU generate precomputed inner product table" (:3351)  */
    PRG_STAGE_VECTOR_FIELD = 3, /* "Compute Vector Field" (:3379)                        */
    PRG_STAGE_DIVERGENCE = 4, /* finer + coarser divergence (:3413,3462)                 */
    PRG_STAGE_SOLVE = 5,      /* "GPU Laplacian Iteration" (:3476)                       */
    PRG_STAGE_ISOVALUE = 6,   /* "GPU calculate isoValue" (:3499)                        */
    PRG_STAGE_MC_TOPOLOGY = 7, /* Vertex/Edge/FaceArray builds (:3575,3609,3642)         */
    PRG_STAGE_MC_EVAL = 8,    /* vertex values, addresses, interpolation, triangles (:3665-3788) */
    PRG_STAGE_SUBDIVIDE = 9,  /* coarse-leaf subdivision (:4209,4564)                    */
    PRG_STAGE_MESH_COPY = 10, /* device mesh -> host (insertTriangle, :3220-3245)        */
    PRG_STAGE_COUNT = 11
} prg_stage;

typedef struct prg_timings {
    double ms[PRG_STAGE_COUNT]; /* device time per stage (hipEvents on the context stream) */
} prg_timings;

/* Device views of one depth of the octree (SoA).  Valid until the next
 * mutating call.  Replaces the OctNode AoS records (OctNode.cuh:8-36). */
typedef struct prg_level_desc {
    int32_t depth;
    int64_t count;            /* M_d: nodes at this depth (sibling-complete)        */
    int64_t base;             /* global index of the first node of this depth       */
    const uint32_t* key;      /* Morton key, reference bit layout (main.cu:142-157) */
    const int32_t* pidx;      /* first sorted point of the node                     */
    const int32_t* pnum;      /* number of points                                   */
    const int32_t* parent;    /* global parent index, -1 at the root                */
    const int32_t* child_base;/* global index of children[0], -1 for leaves         */
    const int32_t* didx;      /* first leaf-level slot of the subtree               */
    const int32_t* dnum;      /* number of leaf-level slots of the subtree          */
    const int32_t* neighs;    /* 27 planes: neighs[j * neigh_stride + (i - base)]   */
    int64_t neigh_stride;
} prg_level_desc;

/* Host-readable buffers (tests, parity harnesses). */
typedef enum prg_buffer {
    PRG_BUF_POINTS = 0,       /* float[3N] normalised sample points in sorted order (main.cu:596-603) */
    PRG_BUF_NORMALS = 1,      /* float[3N] scaled normals in sorted order            */
    PRG_BUF_POINT_LEAF = 2,   /* int32[N] leaf-level slot of each sorted point (PointToNodeArrayD) */
    PRG_BUF_KEY = 3,          /* uint32[M]                                           */
    PRG_BUF_PIDX = 4,         /* int32[M]                                            */
    PRG_BUF_PNUM = 5,         /* int32[M]                                            */
    PRG_BUF_PARENT = 6,       /* int32[M]                                            */
    PRG_BUF_CHILDREN = 7,     /* int32[8M] children[k] as in OctNode (-1 for leaves) */
    PRG_BUF_NEIGHS = 8,       /* int32[27M] neighs[j] per node, node-major           */
    PRG_BUF_DIDX = 9,         /* int32[M]                                            */
    PRG_BUF_DNUM = 10,        /* int32[M]                                            */
    PRG_BUF_VECTOR_FIELD = 11,/* float[3 M_D] (main.cu:932-964)                      */
    PRG_BUF_DIVERGENCE = 12,  /* float[M] (main.cu:1007-1141); writable; 0 until computed */
    PRG_BUF_SOLUTION = 13,    /* float[M] (d_x, main.cu:1223-1332); writable; 0 until solved */
    PRG_BUF_POINT_VALUES = 14,/* float[N] implicit function at sorted samples (main.cu:1334-1381) */
    PRG_BUF_LEAF_CORNER_VALUES = 15, /* float[8 M_D] f - iso at each leaf-level node corner,
                                        corner j in square numbering (main.cu:1659-1668) */
    PRG_BUF_COUNT = 16
} prg_buffer;

/* ---- configuration and context ---------------------------------------- */

/* Fills the defaults (D=8, tol 1e-6, 2000 iterations, refresh 20). */
prg_status prg_config_default(prg_config* cfg);

/* Creates a context on cfg->device with its own stream.  Replaces the
 * implicit global CUDA state of main() (main.cu:3247). */
prg_status prg_create(prg_ctx** out, const prg_config* cfg);
void prg_destroy(prg_ctx* ctx);
/* Changes the context's maximum depth (2..10), keeping its device buffers for
 * reuse; points, octree and results are dropped (call prg_set_points next).
 * No counterpart in the reference, whose depth is a compile-time constant
 * (main.cu:69); it serves callers that reconstruct clouds of different depths
 * in one context, such as prg_reconstruct_batch's workers. */
prg_status prg_set_max_depth(prg_ctx* ctx, int32_t max_depth);
const char* prg_last_error(const prg_ctx* ctx);

/* ---- stages ------------------------------------------------------------ */

/* Uploads n oriented points and applies the reference normalisation A1
 * (pipelineBuildNodeArray, main.cu:530-571): bbox, scale = 1.25 * max extent,
 * p' = (p - center) / scale, n' = n * 2^(D+1) / |n|.  `on_device` = 1 means the
 * pointers are device memory.  n must be in [1, 2^31).  The arrays have been
 * read when the call returns. */
prg_status prg_set_points(prg_ctx* ctx, const float* xyz, const float* nxyz, int64_t n, int32_t on_device);

/* Morton keys, sort, run-length leaves, per-depth sibling-complete arrays and
 * the 27-neighbour tables.  Replaces pipelineBuildNodeArray after the read
 * (main.cu:577-841). */
prg_status prg_build_octree(prg_ctx* ctx);

/* Per-depth node counts and bases (NodeArrayCount_h / BaseAddressArray_h,
 * main.cu:772-776); arrays of max_depth + 1 entries. */
prg_status prg_level_info(prg_ctx* ctx, int64_t* counts, int64_t* bases);
prg_status prg_level_view(prg_ctx* ctx, int32_t depth, prg_level_desc* out);

/* Vector field + divergence for every node (main.cu:3361-3462). */
prg_status prg_compute_divergence(prg_ctx* ctx);

/* Per-depth Laplacian solves (LaplacianIteration, main.cu:1223-1332, and
 * solverCG_DeviceToDevice, CG_CUDA.cuh:344-509).  per_depth may be NULL or
 * hold max_depth + 1 entries.  Returns PRG_E_NOT_CONVERGED if any depth hit
 * the cap (stats are still filled). */
prg_status prg_solve(prg_ctx* ctx, prg_solve_stats* per_depth);

/* Mean implicit-function value at the samples (main.cu:3480-3499).  The
 * value stays in device memory for prg_extract_mesh; with out == NULL the
 * call does not wait for it. */
prg_status prg_isovalue(prg_ctx* ctx, float* out);
/* Overrides the isovalue used by prg_extract_mesh (stage-wise parity). */
prg_status prg_set_isovalue(prg_ctx* ctx, float iso);

/* Marching cubes on the leaf level and, when subdivide != 0, on the
 * virtually refined coarse leaves (main.cu:3502-4561); the mesh is copied to
 * host memory the library allocates (insertTriangle, main.cu:3220-3245). */
prg_status prg_extract_mesh(prg_ctx* ctx, int32_t subdivide, prg_mesh* out);
void prg_mesh_release(prg_mesh* mesh);

/* As prg_extract_mesh, but copies into caller-provided host buffers (for
 * example pinned memory reused across calls): vertices holds vcap floats,
 * triangles tcap ints.  out receives the counts and, on success, the given
 * pointers; when a capacity is too small the call returns PRG_E_ARG with the
 * required counts in out and copies nothing.  With subdivide == 0 and
 * buffers that fit, the vertices are copied while the triangles are still
 * being emitted.  prg_mesh_release must not be called on such a mesh. */
prg_status prg_extract_mesh_into(prg_ctx* ctx, int32_t subdivide, float* vertices, int64_t vcap, int32_t* triangles,
                                 int64_t tcap, prg_mesh* out);

/* Optional cross-batch vertex welding (no counterpart in the reference, which
 * leaves the vertices that the leaf-level and subdivision batches share at
 * their seams duplicated, main.cu:3793,4183,4543).  Vertices whose three
 * coordinates are bit-identical are merged into their first occurrence: the
 * first occurrences keep the original order, vertices are compacted in place
 * (mesh->vertices is not reallocated), every triangle entry is renumbered,
 * and n_vertices / n_leaf_vertices are reduced accordingly; triangles are
 * kept (one whose corners merge stays, degenerate).  Runs on ctx's GPU.
 * n_merged (may be NULL) receives the number of vertices removed.  Every
 * triangle entry must lie in [0, n_vertices) (PRG_E_ARG otherwise). */
prg_status prg_weld_mesh(prg_ctx* ctx, prg_mesh* mesh, int64_t* n_merged);

/* prg_weld_mesh without allocation.  No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The device copy of
 * the mesh and every temporary of the weld live in one fixed array of the
 * library per GPU (PRG_WELD_NOALLOC_BYTES, held until the process ends), so
 * the mesh may hold at most PRG_WELD_NOALLOC_MAX_VERTICES vertices and
 * PRG_WELD_NOALLOC_MAX_TRIANGLES triangles (PRG_E_ARG beyond; nothing is
 * changed then).  The call allocates nothing (no hipMalloc, hipFree or
 * allocating library routine) and returns after its work is done, with the
 * result of prg_weld_mesh.  Calls that overlap on one GPU (from several
 * threads or contexts) race on that array: serialise them. */
#define PRG_WELD_NOALLOC_MAX_VERTICES (1 << 19)
#define PRG_WELD_NOALLOC_MAX_TRIANGLES (1 << 20)
#define PRG_WELD_NOALLOC_BYTES 48643328
prg_status prg_weld_mesh_noalloc(prg_ctx* ctx, prg_mesh* mesh, int64_t* n_merged);

/* Mesh decimation on a host mesh, in place: EdgeCollapse (mode 0) or
 * TriangleCollapse (mode 1) of Geometry.inl:50-254 (declared
 * Geometry.cuh:135-137; the reference's main() never calls them).  Triangles
 * are visited from the last to the first; one whose sum of squared edge
 * lengths exceeds edge_ratio times that of an equilateral triangle of its
 * area has its shortest edge (or itself) collapsed into the smallest vertex
 * index, positions averaging.  Vertices are compacted in order and triangles
 * renumbered; n_vertices and n_triangles are updated and the leaf-batch
 * counts set to 0 (the batches no longer separate).  Host code: each
 * collapse depends on all earlier ones.  edge_ratio must be > 0. */
prg_status prg_decimate_mesh(prg_mesh* mesh, int32_t mode, float edge_ratio);

/* After a prg_extract_mesh* call with subdivide != 0: the number of coarse
 * leaves that were refined (SubdivideNum, main.cu:3826-3834) and of the
 * virtual depth-D leaves marching cubes ran over. */
prg_status prg_subdivision_info(prg_ctx* ctx, int64_t* n_roots, int64_t* n_virtual_leaves);

/* After a prg_extract_mesh* call with subdivide != 0: how the subdivision's
 * bricks (blocks of up to 8^3 virtual leaves) were settled: all of them, those
 * the one-interval certificate did not settle, and those whose corner values
 * were evaluated (the exact certificate did not settle them either).  No
 * counterpart in the reference, which evaluates every virtual leaf. */
prg_status prg_subdivision_bricks(prg_ctx* ctx, int64_t* n_bricks, int64_t* n_exact, int64_t* n_evaluated);

/* Whole pipeline: build, divergence, solve, isovalue, mesh. */
prg_status prg_reconstruct(prg_ctx* ctx, int32_t subdivide, prg_mesh* out, prg_solve_stats* per_depth);

/* ---- many clouds in one pass ---------------------------------------------
 * The reference reconstructs one cloud per process run (main.cu:3247-4573).
 * prg_reconstruct_batch runs a list of independent clouds: `concurrency`
 * worker threads each own a context on `device` and take the next cloud from
 * a queue ordered by decreasing points x depth, so several clouds are in
 * flight on the GPU at once.  Each cloud's mesh is written to its item
 * (release with prg_mesh_release); a failed cloud reports its status there
 * and the others continue. */
typedef struct prg_batch_item {
    const float* xyz;       /* 3 * n floats, host memory                         */
    const float* nxyz;      /* 3 * n floats, host memory                         */
    int64_t n;
    int32_t max_depth;      /* 2 .. 10                                           */
    int32_t status;         /* out: prg_status of this cloud                     */
    prg_mesh mesh;          /* out                                                */
    double ms;              /* out: wall time of this cloud on its worker        */
    double max_true_rel_residual; /* out: largest per-depth ||b - Ax|| / ||b||    */
    int32_t total_iters;    /* out: CG iterations summed over depths             */
} prg_batch_item;

prg_status prg_reconstruct_batch(prg_batch_item* items, int64_t count, const prg_config* config,
                                 int32_t concurrency, int32_t subdivide);

/* ---- introspection ----------------------------------------------------- */

prg_status prg_stage_times(prg_ctx* ctx, prg_timings* out);
prg_status prg_buffer_size(prg_ctx* ctx, int32_t which, int64_t* bytes);
prg_status prg_read_buffer(prg_ctx* ctx, int32_t which, void* host, int64_t bytes);
/* Writable: PRG_BUF_DIVERGENCE, PRG_BUF_SOLUTION (inject a stage input). */
prg_status prg_write_buffer(prg_ctx* ctx, int32_t which, const void* host, int64_t bytes);

/* Average device time of one kernel over the launches recorded since the
 * last reset (config.timing = 1).  `name` is the kernel's symbol name without
 * template arguments (e.g. "stencil_spmv_dot").  Returns PRG_E_ARG if unknown. */
prg_status prg_kernel_stats(prg_ctx* ctx, const char* name, int64_t* launches, double* total_ms,
                            double* bytes_per_launch);
prg_status prg_kernel_stats_reset(prg_ctx* ctx);

/* Normalisation transform (center, scale) used by prg_set_points. */
prg_status prg_get_transform(prg_ctx* ctx, float* center3, float* scale);

/* ---- point / mesh files (PointStream.inl:29-239, plyfile.cu:2769-2837) -- */

/* Reads oriented points: .ply (x y z nx ny nz float, ascii or binary),
 * .bnpts (6 float32 per point), anything else ASCII (6 floats per line);
 * dispatch by extension as main.cu:517-520.  Free with prg_points_release. */
prg_status prg_read_points(const char* path, float** xyz, float** nxyz, int64_t* n);
void prg_points_release(float* xyz, float* nxyz);
/* Writes an ASCII PLY triangle mesh, vertices as v * scale + center ("%g"). */
prg_status prg_write_mesh_ply(const char* path, const prg_mesh* mesh);

#ifdef __cplusplus
}
#endif

#endif /* PRG_H */
