/*
 * aqflip_unet.h -- C ABI of the learned pressure warm start (libaqflip).
 *
 * No counterpart in the reference: AmesingFlank/Aquarius always runs a fixed
 * number of Jacobi sweeps from the previous pressure (Fluid/Fluid_3D_common.cu:48-55)
 * and never measures the residual. This interface adds
 *   - a 3-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions, ReLU, 2x average
 *     pooling, nearest upsampling, skip concatenation, 1x1x1 output head) whose
 *     convolutions run in BF16 on MFMA with FP32 accumulation, trained on the GPU
 *     with the library's own forward/backward kernels and Adam on FP32 master
 *     weights;
 *   - a device-resident training set captured from FLIP handles: input channels
 *     b/sigma and the FLUID mask, target p_conv / (sigma*scale) with p_conv the converged pressure, where
 *     b = -div*density
 *     is the Jacobi right-hand side and sigma its RMS over FLUID cells;
 *   - pressure-solve modes for aq_flip_step: fixed sweep count (reference), to a
 *     tolerance from zero, or to a tolerance from the U-Net guess with plain
 *     zero-guess Jacobi as the fallback when the warm start is non-finite or its
 *     refinement stalls.
 * Conventions as in aqflip.h (status codes, thread-local aq_last_error()).
 */
#ifndef AQFLIP_UNET_H
#define AQFLIP_UNET_H

#include "aqflip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aq_unet aq_unet;

typedef struct aq_unet_config {
  int32_t channels[3];   /* 16, 32, 64 (fixed by the kernels) */
  uint64_t seed;         /* weight initialisation */
  float lr, beta1, beta2, eps;  /* Adam */
  float target_scale;    /* target = p / (sigma * target_scale * (D/64)^2) */
  int32_t infer_extent;  /* aq_flip_nn_guess restricts grids whose largest extent
                            exceeds this by 2x per level, runs the network on the
                            coarsest level and prolongs its pressure back (64) */
} aq_unet_config;

aq_status aq_unet_config_default(aq_unet_config* cfg);
aq_status aq_unet_create(const aq_unet_config* cfg, aq_unet** out);
void aq_unet_destroy(aq_unet* net);

/* FP32 master parameters, layer by layer: for each 3x3x3 convolution
 * W[cout][27][cin] (tap = (dd+1)*9 + (dh+1)*3 + (dw+1)) then bias[cout]; then
 * the head w[16], b[1]. */
int64_t aq_unet_param_count(const aq_unet* net);
aq_status aq_unet_get_params(aq_unet* net, float* host);
aq_status aq_unet_set_params(aq_unet* net, const float* host);

/* Host-tensor entry points (small problems, tests). x is B*D*H*W*2 floats
 * (channels b/sigma and mask, channels-last), y is B*D*H*W floats. D, H, W must be
 * multiples of 4. */
aq_status aq_unet_forward_host(aq_unet* net, const float* x, int32_t B, int32_t D, int32_t H,
                               int32_t W, float* y);
/* Masked MSE loss (mean over mask) and its gradient with respect to every
 * parameter (param_count floats). */
aq_status aq_unet_gradients_host(aq_unet* net, const float* x, const float* target,
                                 const float* mask, int32_t B, int32_t D, int32_t H, int32_t W,
                                 double* loss, float* grads);

/* One Adam step (Kingma & Ba, bias-corrected through lr_t = lr*sqrt(1-b2^t)/(1-b1^t))
 * on the FP32 master parameters with the given gradients (param_count floats), the
 * update aq_unet_train applies after each batch's backward pass. */
aq_status aq_unet_adam_step_host(aq_unet* net, const float* grads);
/* The Adam step size for the following steps (a schedule is a sequence of calls). */
aq_status aq_unet_set_lr(aq_unet* net, float lr);
/* The Adam moments m, v (param_count floats each; either may be NULL) and the step
 * count t. */
aq_status aq_unet_get_adam_state(aq_unet* net, float* m, float* v, int64_t* t);

/* Training set on the device. */
aq_status aq_unet_dataset_reserve(aq_unet* net, int32_t capacity, int32_t D, int32_t H, int32_t W);
/* Append the current pressure system of a FLIP handle: its right-hand side and
 * FLUID mask as input, its current pressure (solved to convergence by the caller)
 * as target. */
aq_status aq_unet_dataset_add_flip(aq_unet* net, aq_flip* flip);
int32_t aq_unet_dataset_size(const aq_unet* net);
/* `steps` Adam steps on batches drawn (seeded) from samples [0, train_count);
 * loss_history (may be NULL) receives each step's batch loss. */
aq_status aq_unet_train(aq_unet* net, int32_t steps, int32_t batch, int32_t train_count,
                        uint64_t seed, float* loss_history);
/* Mean masked MSE over samples [first, first+count). */
aq_status aq_unet_eval(aq_unet* net, int32_t first, int32_t count, double* loss);

/* Write the U-Net pressure guess for the handle's current right-hand side into its
 * pressure field (AIR cells 0). Grids larger than cfg.infer_extent are restricted
 * 2x per level for the network and its output prolonged back (cell-centred
 * trilinear); the guess p is then scaled by alpha = <p,b>/<p,Ap>, which minimises
 * the energy-norm error ||p* - alpha p||_A (A p* = b), so in that norm it is never
 * further from the solution than a zero guess. */
aq_status aq_flip_nn_guess(aq_flip* flip, aq_unet* net);
/* One U-Net correction of the current pressure: d = alpha * net(b - A p) with the
 * same restriction, prolongation and energy-norm scaling as the guess (alpha
 * relative to the residual equation), then p += d on FLUID cells. */
aq_status aq_flip_nn_correct(aq_flip* flip, aq_unet* net);

/* The AQ_SOLVE_TOL_NN solve on the handle's current system, outside a step: the U-Net
 * guess, then Jacobi to tol (FP64 relative residual checked every check_every sweeps,
 * at most max_iters sweeps, stalls detected as the solver's stall_checks says) with a
 * U-Net correction every nn_refresh sweeps (0: none). Corrections are kept while each
 * chunk of nn_refresh sweeps ends at least 1 % below the previous one. After a chunk
 * that does not (Jacobi near its round-off floor), the next correction is judged by the
 * residual after the sweeps that follow it: if that is not below the residual before
 * the correction, the iterate from before it is restored bit for bit and Jacobi
 * continues without further corrections. Outputs may be NULL: sweeps, the final
 * residual, corrections applied and corrections undone. AQ_ESTALL when tol was not
 * reached. */
aq_status aq_flip_solve_pressure_nn(aq_flip* flip, aq_unet* net, double tol, int32_t max_iters,
                                    int32_t check_every, int32_t nn_refresh, int32_t* iters,
                                    double* residual, int32_t* corrections, int32_t* rejected);

/* FIXED: cfg.pressure_iters Jacobi sweeps (the reference). TOL_ZERO: Jacobi from zero
 * to tol. TOL_NN: Jacobi from the U-Net guess to tol with U-Net corrections.
 * MGPCG: multigrid-preconditioned CG in FP64 to tol (aq_flip_solve_pressure_mgpcg;
 * max_iters and check_every count CG iterations, the stats' sweeps field too). */
enum { AQ_SOLVE_FIXED = 0, AQ_SOLVE_TOL_ZERO = 1, AQ_SOLVE_TOL_NN = 2, AQ_SOLVE_MGPCG = 3 };
typedef struct aq_flip_solver {
  int32_t mode;          /* AQ_SOLVE_* */
  double tol;            /* target FP64 relative residual */
  int32_t max_iters;     /* sweep cap per solve */
  int32_t check_every;   /* sweeps between residual checks */
  int32_t stall_checks;  /* consecutive checks without a 1% improvement that end a
                            solve as stalled (FP32 Jacobi's round-off floor) */
  int32_t nn_refresh;    /* AQ_SOLVE_TOL_NN: sweeps between U-Net corrections of the
                            iterate (the network applied to the residual equation
                            A d = b - A p, p += alpha d); 0 = initial guess only */
} aq_flip_solver;
typedef struct aq_flip_solver_stats {
  int64_t solves;         /* pressure solves since the last reset */
  int64_t sweeps;         /* Jacobi sweeps since the last reset */
  int64_t fallbacks;      /* U-Net guesses rejected (non-finite) for zero-guess Jacobi */
  int64_t capped;         /* solves that ended above tol (max_iters or stall) */
  int32_t last_sweeps;
  double last_residual;   /* FP64 relative residual at the end of the last solve */
  double last_guess_residual; /* residual of the last initial guess */
  int32_t last_fallback;
  int64_t nn_corrections; /* U-Net residual corrections applied */
  int64_t nn_rejected;    /* of those, undone (see aq_flip_solve_pressure_nn: judged
                             near the round-off floor, the residual after the sweeps
                             that followed was not below the one before it) */
} aq_flip_solver_stats;
/* Select how aq_flip_step solves for pressure; net is required for
 * AQ_SOLVE_TOL_NN and must outlive its use. */
aq_status aq_flip_set_solver(aq_flip* flip, const aq_flip_solver* solver, aq_unet* net);
aq_status aq_flip_solver_stats_get(aq_flip* flip, aq_flip_solver_stats* out, int32_t reset);

/* Lane-layout probe of v_mfma_f32_16x16x32_bf16: D(16x16) = A(16x32) * B(32x16),
 * all row-major FP32 on the host, operands rounded to bf16. */
aq_status aq_unet_mfma_probe(const float* A, const float* B, float* D);

#ifdef __cplusplus
}
#endif

#endif /* AQFLIP_UNET_H */
