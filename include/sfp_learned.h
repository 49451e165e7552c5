/*
 * sfp_learned.h -- the learned preconditioner. Inside the FP64 flexible CG
 * each application runs the multigrid V-cycle, restricts the residual the
 * V-cycle leaves to the first coarse level, and there applies a 3-level 3-D
 * U-Net (BF16 activations and weights, FP32 accumulation on MFMA) that maps
 * that residual and three operator features to an approximation of the
 * level's exact correction; the correction is scaled to be optimal in the
 * level's energy norm, prolonged and post-smoothed on the fine grid. The solve
 * falls back to the plain V-cycle when an iteration reduces the residual by
 * less than sfp_params.stall_ratio.
 * No counterpart in the reference (it solves with amgcl only).
 *
 * Volumes are channels-last float arrays [batch][x][y][z][c]; every axis must
 * be a multiple of 4 (for the solver: the first coarse level, (n+1)/2 per axis).
 * Network input channels: 0 residual (scaled to unit RMS over the active
 * cells), 1 log(diag/6), 2 (diag - sum of the six couplings)/diag,
 * 3 clamp(phi/h, -4, 4)/4 with h the level's cell size and phi the mean of
 * the fine level set over the children; inactive cells are all zero.
 */
#ifndef SFP_LEARNED_H
#define SFP_LEARNED_H
#include "sfp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of network parameters (FP32). */
long long sfp_unet_param_count(void);

/* Give a context the network weights; afterwards sfp_set_params may select
 * SFP_PRECOND_LEARNED. */
sfp_status sfp_set_network(sfp_context ctx, const float *params, long long count);

/* Operator features of the last solve of `which` (0 pressure, 1 magnetic) on
 * the first coarse level, as fed to the network: out has 3 floats per coarse
 * cell (channels 1..3; dims from sfp_test_level_dims(ctx, which, 1, dims));
 * phi is the fine level set (host, N^3) or NULL for zeros in channel 3. */
sfp_status sfp_learned_features(sfp_context ctx, int which, const double *phi, float *out);

/* ---- stand-alone network (training and tests) ---- */
typedef struct sfp_unet_s *sfp_unet;
sfp_status sfp_unet_create(sfp_unet *out, int device, int batch, int nx, int ny, int nz,
                           int training);
void sfp_unet_destroy(sfp_unet net);
sfp_status sfp_unet_set_params(sfp_unet net, const float *params, long long count);
sfp_status sfp_unet_get_params(sfp_unet net, float *params, long long count);
sfp_status sfp_unet_init_params(sfp_unet net, unsigned long long seed);
/* input: [B][X][Y][Z][4], output: [B][X][Y][Z] (host arrays) */
sfp_status sfp_unet_forward(sfp_unet net, const float *input, float *output);
/* forward + backward of the masked MSE loss; grads has param_count floats */
sfp_status sfp_unet_gradients(sfp_unet net, const float *input, const float *target,
                              const float *mask, double *loss, float *grads);
/* one Adam step (lr) on the masked MSE; device pointers; *loss before the step */
sfp_status sfp_unet_train_step_device(sfp_unet net, const float *input, const float *target,
                                      const float *mask, float lr, double *loss);
const char *sfp_unet_last_error(sfp_unet net);
/* Profiling: `reps` forward passes on the network's device-resident input
 * (that of the last sfp_unet_forward), timed with HIP events on its
 * stream: *forward_ms the mean of a whole pass, conv_ms[k] the mean of
 * convolution k (10 values, in the order of the parameters). */
sfp_status sfp_unet_profile_forward(sfp_unet net, int reps, double *forward_ms, double *conv_ms);

#ifdef __cplusplus
}
#endif
#endif /* SFP_LEARNED_H */
