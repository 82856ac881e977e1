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
 * −(count·log λ − λ − lgamma(count+1)) is), and optionally aux [n][2] FP32
 * = (offset, weight): log λ = network output + offset and the row's loss is
 * multiplied by weight (NULL: offset 0, weight 1).  An offset that carries
 * the known scale of the expected count (energy, distance) and a weight
 * e^−offset make the Poisson gradient (λ − n)·e^−offset scale free while
 * keeping its zero at λ = E[n].  Status codes are ntprop.h's.
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
                    const uint8_t* d_flags, const float* d_aux, int64_t n, float* d_params, float* d_row_loss,
                    void* stream);

/* Loss and gradient without an update: d_grad (device, NTF_N_WEIGHTS) =
   scale · Σ_rows ∂loss_row/∂weights, *d_loss (device double) = Σ_rows loss_row. */
int32_t ntf_loss_grad(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                      const uint8_t* d_flags, const float* d_aux, int64_t n, float scale, float* d_grad,
                      double* d_loss, void* stream);

/* One Adam step (Flux.Optimise.Adam: β = (0.9, 0.999), ε = 1e-8, bias
   corrected) on the mean row loss of the n rows; d_loss (device double, may be
   NULL) receives the mean loss before the update. */
int32_t ntf_train_step(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                       const uint8_t* d_flags, const float* d_aux, int64_t n, float lr, double* d_loss,
                       void* stream);

/* The same step on a batch drawn from larger arrays: the batch is rows
   d_idx[0 … batch−1] (device int64, each in [0, n_rows); an index outside
   makes its row count as absent) of the n_rows-row arrays, read by the kernel
   in place of a gathered copy (the reference's DataLoader batches,
   extended_cascade_model.jl:179-195, 237-295).  The result equals
   ntf_train_step on the gathered rows. */
int32_t ntf_train_step_gather(ntf_ctx* ctx, const float* d_x, const float* d_t, const float* d_count,
                              const uint8_t* d_flags, const float* d_aux, int64_t n_rows, const int64_t* d_idx,
                              int64_t batch, float lr, double* d_loss, void* stream);

/* ---- event likelihood -------------------------------------------------
   Rows of every (source, module, PMT) pair of an event set, built on the
   device (the reference's calc_flow_inputs, extended_cascade_model.jl:
   367-437, for this build's five conditioning numbers):
     d_sources  [n_sources][8]: position xyz, unit direction xyz, log10 E, unused
     d_modules  [n_modules][3]: module centres
     d_pmt_axes [n_pmts][3]: PMT axes (world frame)
     norm       host [10]: feature means then standard deviations (z-scoring,
                the reference's Normalizer)
     scale_knots host [n_knots][2]: (ln d, ln g) of the count scale profile
   Pair q = (s·n_modules + m)·n_pmts + k gets x[q] (NTF_N_IN features) and
   aux[q] = (ln(yield(E)·g(d)), 1) with yield(E) = exp(log_yield_1gev)·E. */
#define NTF_MAX_SCALE_KNOTS 64
int32_t ntf_pair_rows(ntf_ctx* ctx, const float* d_sources, int32_t n_sources, const float* d_modules,
                      int32_t n_modules, const float* d_pmt_axes, int32_t n_pmts, const float* norm,
                      const float* scale_knots, int32_t n_knots, float log_yield_1gev, float* d_x, float* d_aux,
                      void* stream);

/* logpdf of hit times: hit i uses the NTF_P_PAD outputs of pair d_pair[i]
   (from ntf_forward) at time d_t[i]; out-of-range pair indices give −inf. */
int32_t ntf_time_logpdf(ntf_ctx* ctx, const float* d_params, int64_t n_pairs, const int32_t* d_pair,
                        const float* d_t, int64_t n, float* d_logpdf, void* stream);

/* Per-segment sums of row terms: d_out[s] (= or += when accumulate) sign ·
   Σ d_values[q] over the chunks [d_segment_chunks[s], d_segment_chunks[s+1])
   of segment s, chunk c covering q ∈ [d_chunk_offsets[c], d_chunk_offsets[c+1]),
   with d_skip[q] == 0 (d_skip may be NULL).  Summed in double in a fixed
   order; bounded chunk lengths keep segments of very different sizes
   balanced. */
int32_t ntf_segment_sum(ntf_ctx* ctx, const float* d_values, const uint8_t* d_skip, const int64_t* d_chunk_offsets,
                        int64_t n_chunks, const int64_t* d_segment_chunks, int64_t n_segments, float sign,
                        int32_t accumulate, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
