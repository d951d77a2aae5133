/*
 * tea_warmstart.h -- learned warm start for the TeaLeaf solve (no counterpart
 * in the reference: the reference always starts each timestep's solve from
 * u = u0, tea_leaf_common.cuknl:60-129 via tea_leaf_kernel_cuda.cu:254-283).
 *
 * A six-layer 3x3 convolutional network (3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1
 * channels, ReLU between layers, BF16 weights and activations, FP32
 * accumulation on the gfx950 matrix cores) predicts the change of u over a
 * timestep from three per-cell features:
 *     f0 = (log10 density - 0.5) / 1.5
 *     f1 = (u^n - u^{n-1}) / s,   s   = RMS over the mesh of (u^n - u^{n-1})
 *     f2 = u^n / s_u,             s_u = RMS over the mesh of u^n
 * (trained to output (u^{n+1} - u^n - (u^n - u^{n-1})) / s).  The guess
 * handed to the solver is
 *     u = u^n + c0 (u^n - u^{n-1}) + c1 (u^{n-1} - u^{n-2}) + ... + cN s net,
 * over the differences of the last H steps (the deck key
 * tl_warm_start_history H, 1..8, default 8), with the coefficients that
 * minimise the 2-norm of the residual ||u0 - A u||_2 over that span (least
 * squares: G c = g with G_ab = A v_a . A v_b, g_a = A v_a . (u0 - A u^n)), or
 * with tl_warm_start_norm=energy the error's energy norm ||u* - u||_A (a
 * Galerkin step: G_ab = v_a . A v_b, g_a = v_a . (u0 - A u^n)).  Either way
 * the guess is never worse in that norm than u^n, the linear extrapolation
 * or the network's own guess (c = 1, 0, 1).  The solver still runs to the
 * same FP64 residual target, measured from the u = u^n guess, so the guess
 * can only change the iteration count.
 *
 * The parameter vector (tea_ws_param_count() floats) is, in order:
 *     W1[27][32] (k = tap*3 + feature, tap = (dy+1)*3 + (dx+1)), b1[32],
 *     for layers 2..5: W[9][32][32] (tap, in, out), b[32],
 *     W6[9][32], b6.
 * Weight files hold the 8-byte magic "TLWS0001", an int32 parameter count
 * and the float32 parameters, little-endian.
 *
 * Trainer: all pointers named `samples` are device pointers to batch x 6 x
 * height x width floats: planes f0, f1, f2, the target
 * (u^{n+1} - u^n - (u^n - u^{n-1})) / s, and the operator's face
 * coefficients Kx (left face) and Ky (bottom face) of each cell.  Functions
 * return 0 on success and -1 with tea_ws_last_error() set on failure.
 */
#ifndef TEA_WARMSTART_H
#define TEA_WARMSTART_H

#ifdef __cplusplus
extern "C" {
#endif

int tea_ws_param_count(void);
const char *tea_ws_last_error(void);

/* ---- engine side (the chunk of tea_leaf_engine.h) ------------------------- */
/* Enables the warm start with a parameter vector (n == tea_ws_param_count();
 * needs x_cells % 64 == 0 and y_cells % 4 == 0), or with params == NULL
 * without a network (directions from previous steps only); forgets the
 * previous steps.  0 on success. */
int tea_engine_warm_start_load(const float *params, int n);
/* Called after tea_leaf_common_init (u = u0 = u^n): once u^{n-1} is known,
 * sets u to the guess above (*applied = 1); in every case then records
 * u^{n-1} (and u^{n-2}) for the next timestep. */
void tea_engine_warm_start_apply(int *applied);
/* Forgets the previous steps (the next apply only records u^n). */
void tea_engine_warm_start_reset(void);
/* The last guess's coefficients of the first two history differences and of
 * the network's direction (c[0], c[1], c[2]; all 0 when it only recorded u^n). */
void tea_engine_warm_start_coefficients(double *c);
/* Copies the interior of field field_id (TEA_FIELD_* / TEA_XFIELD_*) into a
 * device array of ny x nx doubles (row-major, x fastest).  Synchronous. */
int tea_engine_copy_field(int field_id, double *device_dst);

/* Device buffers for batches of `batch` crops of height x width cells
 * (height % 4 == 0, width % 64 == 0).  The loss is taken over the cells M at
 * least `margin` from the crop edge, with e = output - target:
 *   loss 0: mean squared error, sum_M e^2 / |M|;
 *   loss 1: energy, sum_M e (A_M e) / |M|, A_M the timestep operator
 *           I - dt div K grad restricted to M (the norm in which CG's error
 *           contracts).
 * NULL on failure. */
void *tea_ws_trainer_create(int batch, int height, int width, int margin, int loss);
void tea_ws_trainer_destroy(void *trainer);
/* Sets the FP32 master parameters (and clears the Adam moments). */
int tea_ws_trainer_set_params(void *trainer, const float *host_params);
int tea_ws_trainer_get_params(void *trainer, float *host_params);
/* Network output for a batch (out_host: batch x height x width floats). */
int tea_ws_trainer_forward(void *trainer, const float *samples, float *out_host);
/* Loss and its gradient with respect to every parameter (host arrays). */
int tea_ws_trainer_gradient(void *trainer, const float *samples, float *grad_host, double *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr; the loss
 * before the step is returned through `loss` when it is not NULL (that
 * synchronises; with NULL the step is queued without waiting). */
int tea_ws_trainer_step(void *trainer, const float *samples, float lr, double *loss);
int tea_ws_trainer_sync(void *trainer);

/* Training samples from dense device fields (ny x nx doubles each: density,
 * u^{n-2}, u^{n-1}, u^n, u^{n+1}, Kx, Ky): crop k of crop x crop cells at
 * device origins [2k] = x, [2k+1] = y goes to out + k*6*crop*crop; cells
 * outside the mesh are zero.  The target plane is the part of the step that
 * the guess u^n + c0 (u^n - u^{n-1}) + c1 (u^{n-1} - u^{n-2}) misses, over s:
 * (u^{n+1} - u^n - c0 (u^n - u^{n-1}) - c1 (u^{n-1} - u^{n-2})) / s
 * (c0 = 1, c1 = 0: the linear extrapolation).  scales[0] = s, scales[1] = s_u
 * (host).  Synchronous. */
int tea_ws_make_samples(const double *rho, const double *u_prev2, const double *u_prev, const double *u0,
                        const double *u_next, const double *kx, const double *ky, int nx, int ny, double c0,
                        double c1, const int *origins, int n, int crop, float *out, double *scales);

#ifdef __cplusplus
}
#endif
#endif
