/*
 * ntpmt.h — C ABI of the PMT front end on MI355X (compiled into libntprop.so).
 *
 * Replaces make_reco_pulses (src/pmt_frontend/pipeline.jl:146-160) for every
 * (module, PMT) channel of an event at once, after resample_simulation
 * (pipeline.jl:86-116; on the device: nt_resample in ntprop.h).  Per
 * detected hit and per channel, in FP64 as the reference (its pulse series
 * promote the times to Float64):
 *   apply_tt! / subtract_mean_tt!  pipeline.jl:127-143
 *       t' = t + Gamma(tt_shape, tt_scale) − tt_subtract
 *   PulseSeries(df, spe, pulse)    pulse_templates.jl:147-156, spe_templates.jl:40-48
 *       charge ~ spe_expon_weight · Exponential(spe_expon_scale)
 *              + (1 − spe_expon_weight) · Normal(mu, sigma) truncated to [trunc_low, ∞)
 *   make_waveform                  waveforms.jl:35-53
 *       t_i = t0 + i/sampling_freq, t0 = min t' − lead_time, i = 0 … n_wave − 1
 *       (n_wave = time_range·sampling_freq + 1), v_i = Σ_j q_j · pulse(t_i − t'_j)
 *       + noise_amp · N(0, 1); pulse(x) = amplitude · Gumbel(mu, beta) density
 *       truncated to [lo, hi] (PDFPulseTemplate, pulse_templates.jl:51-54,79-86)
 *   digitize_waveform              waveforms.jl:56-86
 *       DSP.filt of the first-order low-pass y_i = b0 v_i + b1 v_{i−1} − a1 y_{i−1},
 *       then DSP.resample by adc_freq/sampling_freq = 1/R (R integer): digi_k =
 *       Σ_m h_m · y_{R·k + (n_taps − 1)/2 − m}, k = 0 … n_adc − 1 (y = 0 outside),
 *       n_adc = time_range·adc_freq + 1 = ceil(n_wave / R)
 *   unfold_waveform(…, min_charge, :fnnls)   waveforms.jl:128-147, 89-126
 *       charges x = argmin_{x ≥ 0} ‖A x − digi‖, A[k, i] = filtered pulse at
 *       t0 + k/adc_freq − (t0 + i·unf_pulse_res), i = 0 … n_pulse − 1 (n_pulse =
 *       time_range/unf_pulse_res + 1), by the active-set method of
 *       NonNegLeastSquares.fnnls (Bro & de Jong 1997) with tolerance nnls_tol;
 *       pulses = (t0 + i·unf_pulse_res, x_i) for x_i > min_charge, in time order.
 *
 * Random numbers: Philox4x32-10 with key = seed (lo, hi).  Hit draws are keyed
 * by the hit's photon (photon id lo, hi, source index), so a channel's result
 * does not depend on the order the hits are stored in:
 *   transit time, trial t:   counters (0x72000000 + 2t, …) → Box-Muller normal,
 *                            (0x72000000 + 2t + 1, …) → acceptance uniform
 *                            (Marsaglia & Tsang 2000)
 *   charge:                  (0x73000000, …) → component and exponential
 *                            uniforms; (0x73000001 + t, …) → two normals per trial
 *   noise, sample i:         (0x74000000, channel, i / 2, 0) → normal i mod 2
 * with channel = module_id · 17 + min(pmt_id, 16) and 53-bit uniforms (the
 * noise depends on the seed and the channel only: give separate events
 * separate seeds).  The reference draws from Julia's default RNG (not
 * reproducible).
 *
 * Tables (host arrays, copied at ntpmt_create):
 *   decim_taps   the resampling filter phase used at the ratio 1/R (odd
 *                length, centred): DSP.resample_filter's Kaiser-windowed sinc
 *   filt_pulse   make_filtered_pulse (pulse_templates.jl:115-127): the pulse
 *                sampled at filt_pulse_t0 + i·filt_pulse_dt and low-pass
 *                filtered; linear between samples, 0 outside (extrapolation_bc = 0)
 * Requirements: R = sampling_freq/adc_freq and (1/adc_freq)/unf_pulse_res
 * integers (within 1e-9), n_wave ≤ NTPMT_MAX_WAVE, n_adc ≤ NTPMT_MAX_ADC,
 * n_pulse ≤ NTPMT_MAX_PULSE.  Status codes are ntprop.h's.
 */
#ifndef NTPMT_H
#define NTPMT_H

#include <stdint.h>
#include "ntprop.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NTPMT_MAX_WAVE 2048
#define NTPMT_MAX_ADC 96
#define NTPMT_MAX_PULSE 8192
#define NTPMT_MAX_TAPS 1024
#define NTPMT_MAX_FILT_PULSE 1024

/* Per-channel status bits (ntpmt_pulses.status). */
enum {
    NTPMT_ST_TT_GIVEUP = 1,      /* a transit-time draw used the mean after 64 rejected trials */
    NTPMT_ST_SPE_GIVEUP = 2,     /* a charge draw used mu after 32 rejected normals */
    NTPMT_ST_DEGENERATE = 4,     /* a column numerically dependent on the active set was skipped */
    NTPMT_ST_ITER_CAP = 8,       /* the active-set iteration limit ended the unfolding */
    NTPMT_ST_RANK_CAP = 16       /* the active set reached n_adc columns */
};

/* PMTConfig (pipeline.jl:42-82) with its derived quantities. */
typedef struct ntpmt_config {
    double tt_shape, tt_scale;       /* transit-time Gamma(α, θ) (calc_gamma_shape_mean_fwhm) */
    double tt_subtract;              /* subtract_mean_tt!: mean(tt_dist) = α·θ */
    double spe_expon_weight, spe_expon_scale, spe_norm_mu, spe_norm_sigma, spe_trunc_low;
    double pulse_mu, pulse_beta, pulse_lo, pulse_hi, pulse_amplitude;
    double noise_amp;                /* mode(pulse) / 10^(snr_db/10) */
    double sampling_freq, adc_freq;  /* GHz */
    double unf_pulse_res;            /* ns */
    double time_range, lead_time;    /* make_waveform: 300 and 50 ns */
    double min_charge;               /* unfold_waveform: 0.2 */
    double iir_b0, iir_b1, iir_a1;   /* digital low-pass (digitalfilter(Lowpass, Butterworth(1))) */
    double nnls_tol;                 /* fnnls tol (1e-8) */
    int64_t nnls_max_iter;           /* fnnls max_iter (30 · n_pulse) */
    const double* decim_taps;        /* host, n_decim_taps (odd) */
    int32_t n_decim_taps;
    const double* filt_pulse;        /* host, n_filt_pulse */
    int32_t n_filt_pulse;
    double filt_pulse_t0, filt_pulse_dt;
    int32_t lds_factor_rows;         /* rows of the unfolding's Cholesky factor kept in LDS, the rest
                                        in HBM (0: the library's choice, the most that let twelve
                                        channels share a CU: 35 for STD_PMT_CONFIG); fewer rows fit
                                        more channels on a CU but send more rows to HBM; results do
                                        not depend on it */
} ntpmt_config;

typedef struct ntpmt_ctx ntpmt_ctx;

int32_t ntpmt_create(int32_t device, const ntpmt_config* cfg, ntpmt_ctx** out);
void ntpmt_destroy(ntpmt_ctx* ctx);

/* Sizes the configuration implies: n_wave, n_adc, n_pulse (waveform,
   digitised and unfolding-grid samples). */
int32_t ntpmt_sizes(const ntpmt_ctx* ctx, int32_t* n_wave, int32_t* n_adc, int32_t* n_pulse);

/* Output columns (DEVICE arrays).  Channel c (0 … *d_n_channels − 1, in
   increasing channel = module_id·17 + pmt_id) holds its pulses at
   [c·n_adc, c·n_adc + n_pulses[c]) of time / charge.  waveform (optional,
   [capacity][n_adc]) receives the digitised waveform, t0 its first sample
   time. */
typedef struct ntpmt_pulses {
    uint16_t* module_id;      /* [capacity] */
    uint8_t* pmt_id;          /* [capacity] */
    int32_t* n_hits;          /* [capacity] detected hits of the channel */
    int32_t* n_pulses;        /* [capacity] */
    int32_t* status;          /* [capacity] NTPMT_ST_* bits */
    double* t0;               /* [capacity] */
    double* time;             /* [capacity][n_adc] */
    double* charge;           /* [capacity][n_adc] */
    double* waveform;         /* [capacity][n_adc], may be NULL */
} ntpmt_pulses;

/* make_reco_pulses for the hits of dev_hits (DEVICE columns: time (time_column
   0) or tres (1), module_id, pmt_id, source_index, photon_id) with
   d_select[i] != 0 (d_select may be NULL: every hit; e.g. nt_resample's
   d_detected), grouped by channel.  At most `capacity` channels are written;
   *d_n_channels (device int64) receives the number of channels, which may
   exceed it.  Asynchronous on `stream` (a hipStream_t, NULL = the context's);
   calls on one context run one at a time (they share its workspace). */
int32_t ntpmt_make_reco_pulses(ntpmt_ctx* ctx, const nt_hits* dev_hits, const uint8_t* d_select, int64_t n_hits,
                               int32_t time_column, uint64_t seed, int64_t capacity, const ntpmt_pulses* out_dev,
                               int64_t* d_n_channels, void* stream);

/* Wait for the context's last call. */
int32_t ntpmt_synchronize(ntpmt_ctx* ctx);

/* Message of the last NT_ERR_* status an ntpmt_* call returned on this thread. */
const char* ntpmt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
