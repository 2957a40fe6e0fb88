/* ORACLE (test infrastructure only): C interface of the CPU restatement of
 * the reference pipeline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load liboracle.so; the product never does.
 *
 * The state object mirrors the product context so tests can run the same
 * stage sequence on both and compare buffers.  Buffer ids are the product's
 * prg_buffer values (include/prg.h). */
#ifndef PRG_ORACLE_H
#define PRG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state* orc_create(int32_t max_depth);
void orc_destroy(orc_state* s);

/* A1 normalisation (main.cu:530-571) */
int32_t orc_set_points(orc_state* s, const float* xyz, const float* nxyz, int64_t n);
void orc_get_transform(orc_state* s, float* center3, float* scale);
/* A2-A8 (main.cu:126-841) */
int32_t orc_build_octree(orc_state* s);
void orc_level_info(orc_state* s, int64_t* counts, int64_t* bases);
/* A11-A13 (main.cu:913-1141, 3361-3462) */
int32_t orc_compute_divergence(orc_state* s);
/* A14-A15: per-depth solves.  mode 0 = the build's CG (FP32 vectors and SpMV,
 * FP64 dot products, relative tolerance, true residual every `refresh`
 * iterations); mode 1 = the reference's CG (CG_CUDA.cuh:256-324: float
 * scalars, absolute tolerance tol).  fixed_iters > 0 runs exactly that many.
 * iters/relres receive max_depth+1 entries (may be NULL). */
int32_t orc_solve(orc_state* s, int32_t mode, float tol, int32_t max_iters, int32_t refresh,
                  int32_t fixed_iters, int32_t* iters, double* relres);
/* FP64 solve of the same float systems (CG in double to rel tol) into x64. */
int32_t orc_solve_fp64(orc_state* s, double tol, int32_t max_iters, double* x64_out);
/* A16 (main.cu:1334-1381, 3480-3499) */
float orc_isovalue(orc_state* s);
void orc_set_isovalue(orc_state* s, float iso);
/* A17-A18 leaf-level marching cubes (main.cu:3502-3795). */
int32_t orc_extract_leaf_mesh(orc_state* s, int64_t* n_vertices, int64_t* n_triangles);
int32_t orc_copy_mesh(orc_state* s, float* vertices, int32_t* triangles);
/* A19 coarse-leaf subdivision (main.cu:2957-3217, 3799-4561): after
 * orc_extract_leaf_mesh, appends one batch per group of subdivided coarse
 * leaves; the counts returned are the whole mesh's. */
int32_t orc_extract_subdivision(orc_state* s, int64_t* n_vertices, int64_t* n_triangles);
int32_t orc_extract_subdivision_from(orc_state* s, int32_t min_depth, int64_t* n_vertices, int64_t* n_triangles);
void orc_subdivision_info(orc_state* s, int64_t* n_roots, int64_t* n_virtual_leaves);
int32_t orc_subdivision_roots(orc_state* s, int64_t* roots_per_depth);

/* Buffers: ids and layouts of prg_buffer (include/prg.h). */
int64_t orc_buffer_size(orc_state* s, int32_t which);
int32_t orc_read_buffer(orc_state* s, int32_t which, void* dst, int64_t bytes);
int32_t orc_write_buffer(orc_state* s, int32_t which, const void* src, int64_t bytes);

/* Function-space probes (FunctionData.inl:111-300, ConfirmedPPolynomial.cuh). */
double orc_fs_entry(int32_t max_depth, int32_t which, int32_t da, int64_t oa, int32_t db, int64_t ob);
/* Writes 4 pieces x (start, c0..c3) = 20 floats of baseFunctions[idx]. */
void orc_fs_base_function(int32_t max_depth, int64_t idx, float* out20);
/* Writes 4 pieces x (start, c0..c2) = 16 floats of the depth-D base function. */
void orc_fs_base_maxdepth(int32_t max_depth, float* out16);
/* Float evaluation (value(), ConfirmedPPolynomial.cuh:78-91) of baseFunctions[idx] at x. */
float orc_fs_value(int32_t max_depth, int64_t idx, float x);
/* The 27-entry Laplacian stencil of depth d as the reference assembles it (float). */
void orc_fs_stencil(int32_t max_depth, int32_t depth, float* out27);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int32_t orc_num_threads(void);
void orc_set_num_threads(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
