/* Learned warm start for the per-depth solves: a 3-D U-Net on the dense grid
 * of one coarse depth (G^3 cells, G = 2^depth), trained and run on the GPU.
 *
 * No counterpart in the reference: its per-depth CG always starts from zero
 * (LaplacianIteration, main.cu:1237-1329).  With the network attached and
 * prg_config.warm_start = PRG_WARM_LEARNED, prg_solve starts depth log2(G)
 * from the network's prediction; every depth still runs CG to the same
 * tolerance, so only the iteration count can change.
 *
 * Grids are G^3 floats, cell (x, y, z) at index (x G + y) G + z, in the
 * depth's own units; `scale` is the root-mean-square of the divergence over
 * the depth's nodes (the network sees divergence / scale and predicts
 * solution / scale).  Network parameters are one float32 vector in a fixed
 * layout (per 3x3x3 convolution: weights [27][Cout][Cin], bias [Cout]; then
 * the 1x1 head: 16 weights and a bias). */
#ifndef PRG_UNET_H
#define PRG_UNET_H
#include "prg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prg_unet prg_unet;

/* grid: 16, 32, 64 or 128; weights He-normal from `seed`. */
prg_status prg_unet_create(prg_unet** net, int32_t device, int32_t grid, uint64_t seed);
void prg_unet_destroy(prg_unet* net);
const char* prg_unet_last_error(prg_unet* net);
int64_t prg_unet_num_params(prg_unet* net);
int32_t prg_unet_grid(prg_unet* net);
prg_status prg_unet_get_params(prg_unet* net, float* out, int64_t count);
prg_status prg_unet_set_params(prg_unet* net, const float* in, int64_t count);

/* Predicted solution grid (host arrays of G^3 floats). */
prg_status prg_unet_predict(prg_unet* net, const float* div_grid, const float* mask_grid, float scale,
                            float* solution_grid);
/* Masked mean-squared error of one sample (in units of scale) and its gradient
 * for every parameter (count = prg_unet_num_params). */
prg_status prg_unet_gradients(prg_unet* net, const float* div_grid, const float* mask_grid,
                              const float* solution_grid, float scale, float* grads, double* loss);
/* One Adam step (beta 0.9/0.999) on the mean gradient of `batch` samples
 * stored back to back; mean_loss may be NULL. */
prg_status prg_unet_train_step(prg_unet* net, const float* div_grids, const float* mask_grids,
                               const float* solution_grids, const float* scales, int32_t batch, float lr,
                               double* mean_loss);

/* Dense grid of one depth of a context (which: 0 divergence, 1 solution,
 * 2 node mask) into host memory (G^3 floats, G = 2^depth), and the
 * root-mean-square of the divergence over the depth's nodes. */
prg_status prg_dense_grid(prg_ctx* ctx, int32_t depth, int32_t which, float* grid, float* div_rms);

/* Attaches a network (or NULL) used when warm_start = PRG_WARM_LEARNED; the
 * context does not own it. */
prg_status prg_set_warm_start_net(prg_ctx* ctx, prg_unet* net);

/* After prg_solve: the depth the network warm-started (-1: none) and whether
 * its start was rejected because ||b - A x0|| > ||b|| (that depth then
 * started from x0 = 0, as every depth of the reference does,
 * main.cu:1237-1329). */
prg_status prg_warm_start_info(prg_ctx* ctx, int32_t* depth, int32_t* rejected);

#ifdef __cplusplus
}
#endif
#endif
