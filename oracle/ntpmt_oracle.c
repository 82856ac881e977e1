/*
 * ntpmt_oracle.c — CPU restatement of the reference's PMT front end
 * (src/pmt_frontend/), in FP64 as the reference (its pulse series promote
 * times and charges to Float64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libntprop's ntpmt_* entry points.
 * It may be built and called by tests/ only, never by the library or any
 * product path.
 *
 * Third-party algorithms the reference calls and which are not in the
 * reference repository (Project.toml lists them; there is no Manifest, so no
 * version is pinned):
 *   DSP.jl            filt (first-order section, zero state), resample
 *                     (polyphase FIR, Kaiser window design)
 *   NonNegLeastSquares.jl  fnnls (Bro & de Jong 1997, "A fast
 *                     non-negativity-constrained least squares algorithm",
 *                     J. Chemometrics 11:393-401)
 *   Distributions.jl  Gamma, Exponential, truncated Normal samplers (law
 *                     only: the reference's random streams are not reproducible)
 * Their published algorithms are restated below; tests pin them by their
 * mathematical properties (frequency response, KKT conditions, sampling laws).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ntoracle.h"
#include "ntpmt_oracle.h"

/* ---------------------------------------------------------------- RNG --- */
static void block(uint64_t seed, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t out[4]) {
    const uint32_t ctr[4] = {w0, w1, w2, w3};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ntor_philox4x32_10(ctr, key, out);
}

/* 53-bit uniform in [0, 1) from two words */
static double u53(uint32_t hi, uint32_t lo) { return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * 0x1p-53; }

/* Box-Muller: two standard normals from one Philox block */
void ntor_pmt_normals(const uint32_t r[4], double out[2]) {
    const double u1 = 1.0 - u53(r[0], r[1]);           /* (0, 1] */
    const double u2 = u53(r[2], r[3]);
    const double rad = sqrt(-2.0 * log(u1));
    out[0] = rad * cos(2.0 * M_PI * u2);
    out[1] = rad * sin(2.0 * M_PI * u2);
}

#define TT_WORD 0x72000000u
#define SPE_WORD 0x73000000u
#define NOISE_WORD 0x74000000u
#define TT_TRIALS 64
#define SPE_TRIALS 16

/* Gamma(α, θ) by Marsaglia & Tsang (2000), "A simple method for generating
   gamma variables", ACM TOMS 26:363-372, for α ≥ 1 (the sampler
   Distributions.jl uses for α ≥ 1), without the squeeze (same law). */
static double tt_draw(const ntpmt_config* c, uint64_t seed, uint32_t plo, uint32_t phi, uint32_t src, int* giveup) {
    const double d = c->tt_shape - 1.0 / 3.0, cc = 1.0 / sqrt(9.0 * d);
    for (uint32_t t = 0; t < TT_TRIALS; ++t) {
        uint32_t r[4], a[4];
        double nrm[2];
        block(seed, TT_WORD + 2u * t, plo, phi, src, r);
        ntor_pmt_normals(r, nrm);
        const double x = nrm[0];
        double v = 1.0 + cc * x;
        if (!(v > 0.0)) continue;
        v = v * v * v;
        block(seed, TT_WORD + 2u * t + 1u, plo, phi, src, a);
        const double u = 1.0 - u53(a[0], a[1]);
        if (log(u) < 0.5 * x * x + d - d * v + d * log(v)) return d * v * c->tt_scale;
    }
    *giveup = 1;
    return c->tt_shape * c->tt_scale;
}

/* rand(MixtureModel([Exponential(rate), truncated(Normal(mu, sigma), low, Inf)],
   [w, 1 − w])), spe_templates.jl:40-48 */
static double spe_draw(const ntpmt_config* c, uint64_t seed, uint32_t plo, uint32_t phi, uint32_t src, int* giveup) {
    uint32_t r[4];
    block(seed, SPE_WORD, plo, phi, src, r);
    if (u53(r[0], r[1]) < c->spe_expon_weight) return -log(1.0 - u53(r[2], r[3])) * c->spe_expon_scale;
    for (uint32_t t = 0; t < SPE_TRIALS; ++t) {
        double nrm[2];
        block(seed, SPE_WORD + 1u + t, plo, phi, src, r);
        ntor_pmt_normals(r, nrm);
        for (int k = 0; k < 2; ++k) {
            const double v = c->spe_norm_mu + c->spe_norm_sigma * nrm[k];
            if (v >= c->spe_trunc_low) return v;
        }
    }
    *giveup = 1;
    return c->spe_norm_mu;
}

void ntor_pmt_hit_draws(const ntpmt_config* c, int64_t n, const float* t, const uint32_t* src, const uint64_t* pid,
                        uint64_t seed, double* t_out, double* q_out, int32_t* flags) {
    for (int64_t i = 0; i < n; ++i) {
        int g1 = 0, g2 = 0;
        const uint32_t plo = (uint32_t)pid[i], phi = (uint32_t)(pid[i] >> 32);
        const double tt = tt_draw(c, seed, plo, phi, src[i], &g1);
        t_out[i] = ((double)t[i] + tt) - c->tt_subtract;                 /* apply_tt!, subtract_mean_tt! */
        q_out[i] = spe_draw(c, seed, plo, phi, src[i], &g2);
        if (flags) flags[i] = (g1 ? NTPMT_ST_TT_GIVEUP : 0) | (g2 ? NTPMT_ST_SPE_GIVEUP : 0);
    }
}

/* ------------------------------------------------------- pulse shapes --- */
/* PDFPulseTemplate (pulse_templates.jl:79-86): amplitude · pdf of
   truncated(Gumbel(mu, beta), lo, hi) = Gumbel pdf / (F(hi) − F(lo)) on [lo, hi] */
double ntor_pmt_pulse(const ntpmt_config* c, double x) {
    if (!(x >= c->pulse_lo && x <= c->pulse_hi)) return 0.0;
    const double F = exp(-exp(-(c->pulse_hi - c->pulse_mu) / c->pulse_beta)) -
                     exp(-exp(-(c->pulse_lo - c->pulse_mu) / c->pulse_beta));
    const double z = (x - c->pulse_mu) / c->pulse_beta;
    return exp(-(z + exp(-z))) / c->pulse_beta / F * c->pulse_amplitude;
}

/* InterpolatedPulse (pulse_templates.jl:90-97, 115-127): linear between the
   filt_pulse samples, 0 outside */
double ntor_pmt_filt_pulse(const ntpmt_config* c, double s) {
    const int n = c->n_filt_pulse;
    const double u = (s - c->filt_pulse_t0) / c->filt_pulse_dt;
    if (!(u >= 0.0 && u <= (double)(n - 1))) return 0.0;
    int i = (int)floor(u);
    if (i >= n - 1) return c->filt_pulse[n - 1];
    const double f = u - i;
    return c->filt_pulse[i] * (1.0 - f) + c->filt_pulse[i + 1] * f;
}

/* DSP.filt of the first-order section, zero initial state (transposed direct
   form: y = b0 x + z, z = b1 x − a1 y) */
void ntor_pmt_iir(const ntpmt_config* c, int32_t n, const double* x, double* y) {
    double z = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        const double yi = c->iir_b0 * x[i] + z;
        z = c->iir_b1 * x[i] - c->iir_a1 * yi;
        y[i] = yi;
    }
}

/* DSP.resample(y, 1/R) at an exact integer ratio: output k is the centred
   filter at input R·k, zeros outside the input */
void ntor_pmt_decimate(const ntpmt_config* c, int32_t n_in, const double* y, int32_t R, int32_t n_out, double* d) {
    const int32_t half = (c->n_decim_taps - 1) / 2;
    for (int32_t k = 0; k < n_out; ++k) {
        double acc = 0.0;
        for (int32_t m = 0; m < c->n_decim_taps; ++m) {
            const int64_t j = (int64_t)R * k + half - m;
            if (j >= 0 && j < n_in) acc += c->decim_taps[m] * y[j];
        }
        d[k] = acc;
    }
}

/* --------------------------------------------------------------- NNLS --- */
/* Solve G z = r (G p×p, row-major) by Gaussian elimination with partial
   pivoting (what Julia's `\` does for a dense square matrix); 1 if singular. */
static int gepp_solve(int p, double* G, double* r) {
    for (int k = 0; k < p; ++k) {
        int piv = k;
        for (int i = k + 1; i < p; ++i)
            if (fabs(G[i * p + k]) > fabs(G[piv * p + k])) piv = i;
        if (G[piv * p + k] == 0.0) return 1;
        if (piv != k) {
            for (int j = 0; j < p; ++j) {
                const double t = G[k * p + j];
                G[k * p + j] = G[piv * p + j];
                G[piv * p + j] = t;
            }
            const double t = r[k];
            r[k] = r[piv];
            r[piv] = t;
        }
        for (int i = k + 1; i < p; ++i) {
            const double l = G[i * p + k] / G[k * p + k];
            for (int j = k; j < p; ++j) G[i * p + j] -= l * G[k * p + j];
            r[i] -= l * r[k];
        }
    }
    for (int k = p - 1; k >= 0; --k) {
        double s = r[k];
        for (int j = k + 1; j < p; ++j) s -= G[k * p + j] * r[j];
        r[k] = s / G[k * p + k];
    }
    return 0;
}

/* s[P] = AtA[P,P] \ Atb[P], s[!P] = 0, with AtA[P,P] = A_Pᵀ A_P (A row-major
   m × n) and the indices of P in increasing order */
static int solve_passive(int m, int n, const double* A, const double* atb, const unsigned char* P, double* s,
                         int* idx, double* G, double* r) {
    int p = 0;
    for (int j = 0; j < n; ++j)
        if (P[j]) idx[p++] = j;
    for (int a = 0; a < p; ++a) {
        for (int b = 0; b < p; ++b) {
            double acc = 0.0;
            for (int k = 0; k < m; ++k) acc += A[(int64_t)k * n + idx[a]] * A[(int64_t)k * n + idx[b]];
            G[a * p + b] = acc;
        }
        r[a] = atb[idx[a]];
    }
    if (p && gepp_solve(p, G, r)) return 1;
    memset(s, 0, sizeof(double) * (size_t)n);
    for (int a = 0; a < p; ++a) s[idx[a]] = r[a];
    return 0;
}

/* w = Atb − AtA x = Aᵀ(b − A x) */
static void gradient(int m, int n, const double* A, const double* b, const double* x, double* w, double* res) {
    for (int k = 0; k < m; ++k) {
        double acc = b[k];
        for (int j = 0; j < n; ++j)
            if (x[j] != 0.0) acc -= A[(int64_t)k * n + j] * x[j];
        res[k] = acc;
    }
    for (int j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int k = 0; k < m; ++k) acc += A[(int64_t)k * n + j] * res[k];
        w[j] = acc;
    }
}

/* NonNegLeastSquares.fnnls(AtA, Atb; tol, max_iter) as its published
   algorithm: outer loop while some w_i > tol outside P (and iter < max_iter):
   move argmax_{i∉P} w_i (first maximum) into P, solve the unconstrained
   problem on P; inner loop while some s_j ≤ tol on P (iter += 1): step
   x += α(s − x) with α = min x_j/(x_j − s_j) over those j, drop every j with
   x_j ≤ tol from P, solve again; then x = s.  Returns 0, 1 (singular
   system, Julia's SingularException) or 2 (outer-loop guard of 100·n). */
int32_t ntor_fnnls(int32_t m, int32_t n, const double* A, const double* b, double tol, int64_t max_iter, double* x,
                   int64_t* iter_out, int32_t* outer_out) {
    double* atb = calloc((size_t)n, sizeof(double));
    double* w = calloc((size_t)n, sizeof(double));
    double* s = calloc((size_t)n, sizeof(double));
    double* res = calloc((size_t)m, sizeof(double));
    unsigned char* P = calloc((size_t)n, 1);
    int* idx = calloc((size_t)n, sizeof(int));
    const int pmax = n < m + 8 ? n : m + 8;
    double* G = calloc((size_t)pmax * pmax, sizeof(double));
    double* r = calloc((size_t)n, sizeof(double));
    int32_t st = 0;
    memset(x, 0, sizeof(double) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int k = 0; k < m; ++k) acc += A[(int64_t)k * n + j] * b[k];
        atb[j] = acc;
    }
    memcpy(w, atb, sizeof(double) * (size_t)n);
    int64_t iter = 0;
    int32_t outer = 0;
    for (;;) {
        int np = 0, any = 0, best = -1;
        double bw = 0.0;
        for (int j = 0; j < n; ++j) {
            if (P[j]) {
                ++np;
                continue;
            }
            if (w[j] > tol) any = 1;
            if (best < 0 || w[j] > bw) {
                bw = w[j];
                best = j;
            }
        }
        if (!(np < n && any && iter < max_iter)) break;
        if (++outer > 100 * n) {
            st = 2;
            break;
        }
        /* argmax(w .* !P): entries of P count as ±0, below the positive max */
        P[best] = 1;
        if (np + 1 > pmax) {            /* more columns than rows (+8): singular in exact arithmetic */
            st = 1;
            break;
        }
        if (solve_passive(m, n, A, atb, P, s, idx, G, r)) {
            st = 1;
            break;
        }
        for (;;) {
            int neg = 0;
            for (int j = 0; j < n; ++j)
                if (P[j] && s[j] <= tol) neg = 1;
            if (!neg) break;
            ++iter;
            double alpha = INFINITY;
            for (int j = 0; j < n; ++j)
                if (P[j] && s[j] <= tol) alpha = fmin(alpha, x[j] / (x[j] - s[j]));
            for (int j = 0; j < n; ++j) x[j] += alpha * (s[j] - x[j]);
            for (int j = 0; j < n; ++j)
                if (x[j] <= tol) P[j] = 0;
            if (solve_passive(m, n, A, atb, P, s, idx, G, r)) {
                st = 1;
                break;
            }
        }
        if (st) break;
        memcpy(x, s, sizeof(double) * (size_t)n);
        gradient(m, n, A, b, x, w, res);
    }
    if (iter_out) *iter_out = iter;
    if (outer_out) *outer_out = outer;
    free(atb);
    free(w);
    free(s);
    free(res);
    free(P);
    free(idx);
    free(G);
    free(r);
    return st;
}

/* ---------------------------------------------------- one channel chain --- */
int32_t ntor_pmt_sizes(const ntpmt_config* c, int32_t* n_wave, int32_t* n_adc, int32_t* n_pulse, int32_t* R) {
    *n_wave = (int32_t)floor(c->time_range * c->sampling_freq + 1e-9) + 1;
    *n_adc = (int32_t)floor(c->time_range * c->adc_freq + 1e-9) + 1;
    *n_pulse = (int32_t)floor(c->time_range / c->unf_pulse_res + 1e-9) + 1;
    *R = (int32_t)llround(c->sampling_freq / c->adc_freq);
    return 0;
}

typedef struct { float key; int32_t i; } order_t;
static int cmp_order(const void* a, const void* b) {
    const order_t *x = a, *y = b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* make_reco_pulses after the hit draws, for one channel's n hits (times t,
   charges q): make_waveform (waveforms.jl:35-53, evaluate_pulse_series
   pulse_templates.jl:175-190), digitize_waveform (waveforms.jl:56-86) and
   unfold_waveform (waveforms.jl:128-147, make_nnls_matrix :89-107, apply_nnls
   :110-126) with fnnls.  Outputs (any may be NULL): t0, wave[n_wave] (with
   noise), filt[n_wave], digi[n_adc], x[n_pulse] (the charges on the grid
   t0 + i·unf_pulse_res). */
int32_t ntor_pmt_channel(const ntpmt_config* c, int32_t n, const double* t, const double* q, uint32_t channel,
                         uint64_t seed, double* t0_out, double* wave_out, double* filt_out, double* digi_out,
                         double* x_out, int64_t* iter_out, int32_t* outer_out) {
    int32_t n_wave, n_adc, n_pulse, R;
    ntor_pmt_sizes(c, &n_wave, &n_adc, &n_pulse, &R);
    double tmin = INFINITY;
    for (int32_t j = 0; j < n; ++j) tmin = fmin(tmin, t[j]);
    const double t0 = tmin - c->lead_time;
    if (t0_out) *t0_out = t0;
    order_t* ord = malloc(sizeof(order_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t j = 0; j < n; ++j) {
        ord[j].key = (float)t[j];
        ord[j].i = j;
    }
    qsort(ord, (size_t)n, sizeof(order_t), cmp_order);
    double* v = malloc(sizeof(double) * (size_t)n_wave);
    double* y = malloc(sizeof(double) * (size_t)n_wave);
    const double dt = 1.0 / c->sampling_freq;
    for (int32_t i = 0; i < n_wave; ++i) {
        const double ts = t0 + (double)i * dt;
        double acc = 0.0;
        for (int32_t o = 0; o < n; ++o) {
            const int32_t j = ord[o].i;
            acc += ntor_pmt_pulse(c, ts - t[j]) * q[j];
        }
        if (c->noise_amp > 0.0) {
            uint32_t r[4];
            double nrm[2];
            block(seed, NOISE_WORD, channel, (uint32_t)(i >> 1), 0u, r);
            ntor_pmt_normals(r, nrm);
            acc += nrm[i & 1] * c->noise_amp;
        }
        v[i] = acc;
    }
    free(ord);
    ntor_pmt_iir(c, n_wave, v, y);
    double* d = malloc(sizeof(double) * (size_t)n_adc);
    ntor_pmt_decimate(c, n_wave, y, R, n_adc, d);
    if (wave_out) memcpy(wave_out, v, sizeof(double) * (size_t)n_wave);
    if (filt_out) memcpy(filt_out, y, sizeof(double) * (size_t)n_wave);
    if (digi_out) memcpy(digi_out, d, sizeof(double) * (size_t)n_adc);
    int32_t st = 0;
    if (x_out) {
        /* A[k, i] = filtered pulse at t_k − τ_i, t_k = t0 + k/adc_freq, τ_i = t0 + i·res */
        double* A = malloc(sizeof(double) * (size_t)n_adc * (size_t)n_pulse);
        const double dadc = 1.0 / c->adc_freq;
        for (int32_t k = 0; k < n_adc; ++k) {
            const double tk = t0 + (double)k * dadc;
            for (int32_t i = 0; i < n_pulse; ++i)
                A[(int64_t)k * n_pulse + i] = ntor_pmt_filt_pulse(c, tk - (t0 + (double)i * c->unf_pulse_res));
        }
        st = ntor_fnnls(n_adc, n_pulse, A, d, c->nnls_tol, c->nnls_max_iter, x_out, iter_out, outer_out);
        free(A);
    }
    free(v);
    free(y);
    free(d);
    return st;
}
