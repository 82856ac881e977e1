/*
 * ntflow.h — C ABI of libntflow: the conditional rational-quadratic-spline
 * flow surrogate (arrival-time density + expected count per PMT) on MI355X.
 *
 * Replaces the reference's Flux model and its training/evaluation calls:
 *   RQNormFlow (src/surrogate_models/extended_cascade_model.jl:53-125),
 *   eval_transformed_normal_logpdf (src/surrogate_models/rq_spline_flow.jl:396-428),
 *   log_likelihood_with_poisson (extended_cascade_model.jl:132-139),
 *   train_model! with Adam (extended_cascade_model.jl:237-295).
 *
 * Network: P = W3ᵀ·relu(W2ᵀ·relu(W1ᵀ·x + b1) + b2) + b3 with x padded to
 * NTF_N_IN features, NTF_HIDDEN units per hidden layer, NTF_P_PAD outputs of
 * which the first NTF_N_PARAMS are used:
 *   [spline_1 … spline_L (3K+1 each: K widths, K heights, K+1 slopes),
 *    shift, scale logit, log λ]
 * Density of t: t = S_1(S_2(…S_L(z))), z ~ N(shift, 100·σ(scale logit)),
 * each S_l a rational-quadratic spline on [NTF_RANGE_MIN, NTF_RANGE_MAX] with
 * linear tails (Durkan et al. 2019); L = 1 is the reference's model.
 * Matrix products run on the matrix cores in bfloat16 with FP32
 * accumulation; the flow and its gradient run in FP32.
 *
 * Weights: one FP32 vector of NTF_N_WEIGHTS values,
 *   W1 [NTF_N_IN][NTF_HIDDEN] (row = input), b1 [NTF_HIDDEN],
 *   W2 [NTF_HIDDEN][NTF_HIDDEN], b2 [NTF_HIDDEN],
 *   W3 [NTF_HIDDEN][NTF_P_PAD], b3 [NTF_P_PAD].
 * Rows (device arrays): x [n][NTF_N_IN] FP32 features, t [n] FP32 time,
 * count [n] FP32 count, flags [n] uint8 (NTF_ROW_TIME: the time term
 * −logpdf(t) is in the row's loss; NTF_ROW_COUNT: the Poisson term
 * −(count·log λ − λ − lgamma(count+1)) is).  Status codes are ntprop.h's.
 */
#ifndef NTFLOW_H
#define NTFLOW_H

#include <stdint.h>
#include "ntprop.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NTF_N_IN 16
#define NTF_HIDDEN 128
#define NTF_BINS 8
#define NTF_LAYERS 3
#define NTF_N_PARAMS (NTF_LAYERS * (3 * NTF_BINS + 1) + 3)   /* 78 */
#define NTF_P_PAD 80
#define NTF_N_WEIGHTS (NTF_N_IN * NTF_HIDDEN + NTF_HIDDEN + NTF_HIDDEN * NTF_HIDDEN + NTF_HIDDEN + \
                       NTF_HIDDEN * NTF_P_PAD + NTF_P_PAD)   /* 29008 */
#define NTF_RANGE_MIN (-20.0f)
#define NTF_RANGE_MAX 100.0f

enum { NTF_ROW_TIME = 1, NTF_ROW_COUNT = 2 };

typedef struct ntf_ctx ntf_ctx;

int32_t ntf_create(int32_t device, ntf_ctx** out);
void ntf_destroy(ntf_ctx* ctx);

/* Copy weights in/out (host FP32, NTF_N_WEIGHTS values).  Setting weights
   resets the Adam state. */
int32_t ntf_set_weights(ntf_ctx* ctx, const float* weights);
int32_t ntf_get_weights(ntf_ctx* ctx, float* weights);

/* Forward pass over n rows: the NTF_P_PAD outputs per row (d_params, may be
   NULL) and the per-row loss (d_row_loss, may be NULL).  The model call of
   extended_cascade_model.jl:110-125 (logpdf = −row loss for a time row). */
int32_t ntf_forward(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                    const uint8_t* d_flags, int64_t n, float* d_params, float* d_row_loss, void* stream);

/* Loss and gradient without an update: d_grad (device, NTF_N_WEIGHTS) =
   scale · Σ_rows ∂loss_row/∂weights, *d_loss (device double) = Σ_rows loss_row. */
int32_t ntf_loss_grad(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                      const uint8_t* d_flags, int64_t n, float scale, float* d_grad, double* d_loss,
                      void* stream);

/* One Adam step (Flux.Optimise.Adam: β = (0.9, 0.999), ε = 1e-8, bias
   corrected) on the mean row loss of the n rows; d_loss (device double, may be
   NULL) receives the mean loss before the update. */
int32_t ntf_train_step(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                       const uint8_t* d_flags, int64_t n, float lr, double* d_loss, void* stream);

#ifdef __cplusplus
}
#endif
#endif
