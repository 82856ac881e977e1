/*
 * ntoracle.c — CPU restatement of the reference photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): built by oracle/Makefile into
 * oracle/liboracle.so and loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
 *
 * The reference (Julia + CUDA.jl) has no CPU path and cannot run here; this
 * file restates its algorithm function by function.  Citations are
 * reference-relative file:line.  Build flags: -O3 -ffp-contract=off (no
 * implicit fma: Julia emits none), OpenMP over photons.
 */
#include "ntoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11.  Multipliers and Weyl constants from the paper.  */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void ntor_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform in [0,1) with 24 random bits */
static inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
/* uniform in (0,1] with 24 random bits (argument of a logarithm) */
static inline float u01_oc(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

static void draw_block(uint64_t seed, uint32_t src, uint64_t photon, uint32_t block,
                       uint32_t out[4]) {
    uint32_t ctr[4] = {block, (uint32_t)photon, (uint32_t)(photon >> 32), src};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ntor_philox4x32_10(ctr, key, out);
}

#define C_VAC_M_NS 0.299792458 /* PhysicalConstants SpeedOfLightInVacuum in m/ns */
#define TWO_PI_F 6.2831855f    /* Float32(2π), the T(2*pi) of photon_prop_cuda.jl:102,111 */
#define FINE_STRUCTURE 7.2973525693e-3 /* CODATA2018 */

/* ------------------------------------------------------------------------ */
/* Medium constructors                                                      */
/* ------------------------------------------------------------------------ */

/* DIPPR105 water density, medium.jl:26-48 (A,B,C,D from :34) */
double ntor_dippr105(double temperature_k) {
    const double A = 0.14395, B = 0.0112, C = 649.727, D = 0.05107;
    return A / pow(B, 1.0 + pow(1.0 - temperature_k / C, D));
}

/* _calc_quan_fry_params, medium.jl:174-205 */
void ntor_quan_fry_params(double s, double t, double p, double out[4]) {
    const double n0 = 1.31405, n1 = 1.45e-5, n2 = 1.779e-4, n3 = 1.05e-6, n4 = 1.6e-8;
    const double n5 = 2.02e-6, n6 = 15.868, n7 = 0.01155, n8 = 0.00423, n9 = 4382.0,
                 n10 = 1.1455e6;
    out[0] = n0 + (n2 - n3 * t + n4 * t * t) * s - n5 * t * t + n1 * p;
    out[1] = n6 + n7 * s - n8 * t;
    out[2] = -n9;
    out[3] = n10;
}

/* make_cascadia_medium_properties, medium.jl:131-138 through the WaterProperties
   constructor medium.jl:98-124 (fields converted to Float32). */
void ntor_make_cascadia_medium(float g, nt_medium* m, double* density, double* radiation_length) {
    const float salinity = 34.82f;
    const float pressure_atm = (float)(269.44088f * 1e5 / 101325.0); /* bar -> atm */
    const float temperature = 1.8f;
    double qf[4];
    ntor_quan_fry_params(salinity, temperature, pressure_atm, qf);
    memset(m, 0, sizeof(*m));
    for (int i = 0; i < 4; ++i) m->quan_fry[i] = (float)qf[i];
    m->vol_conc_small_part = 0.0075f;
    m->vol_conc_large_part = 0.0075f;
    m->mean_scattering_angle = g;
    m->sl_fraction = 0.0f;
    m->sca_len_scale = 1.0f;
    m->dispersion_mode = NT_DISPERSION_REF;
    m->max_steps = 15;
    if (density) *density = (float)ntor_dippr105((double)temperature + 273.15);
    if (radiation_length) *radiation_length = 36.08f;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre 10-point quadrature (utils.jl:17, 82-103)                  */
/* ------------------------------------------------------------------------ */
static const double GL10_X[10] = {
    -0.9739065285171717, -0.8650633666889845, -0.6794095682990244, -0.4333953941292472,
    -0.14887433898163122, 0.14887433898163122, 0.4333953941292472, 0.6794095682990244,
    0.8650633666889845, 0.9739065285171717};
static const double GL10_W[10] = {
    0.06667134430868807, 0.14945134915058036, 0.219086362515982, 0.2692667193099965,
    0.295524224714753,   0.295524224714753,   0.2692667193099965, 0.219086362515982,
    0.14945134915058036, 0.06667134430868807};

typedef double (*fn1)(double, const void*);

static double gauss_quad(fn1 f, const void* ctx, double a, double b) {
    double ba_half = (b - a) / 2.0, mid = (a + b) / 2.0, acc = 0.0;
    for (int i = 0; i < 10; ++i) acc += GL10_W[i] * (f(ba_half * GL10_X[i] + mid, ctx) * ba_half);
    return acc;
}

static double f_pow(double x, const void* ctx) { return pow(x, (double)*(const int*)ctx); }
static double f_cos(double x, const void* ctx) {
    (void)ctx;
    return cos(x);
}
double ntor_integrate_gauss_quad_pow(double a, double b, int power) {
    return gauss_quad(f_pow, &power, a, b);
}
double ntor_integrate_gauss_quad_cos(double a, double b) { return gauss_quad(f_cos, NULL, a, b); }

/* frank_tamm, spectrum.jl:32-34 */
double ntor_frank_tamm(double wl, double n) {
    return 2.0 * M_PI * FINE_STRUCTURE / (wl * wl) * (1.0 - 1.0 / (n * n));
}

static double f_frank_tamm(double wl, const void* ctx) {
    const float* qf = (const float*)ctx;
    double q[4] = {qf[0], qf[1], qf[2], qf[3]};
    return ntor_frank_tamm(wl, ntor_refractive_index_f64(wl, q));
}

/* frank_tamm_norm, spectrum.jl:42-45: photons per metre in [lo, hi] */
double ntor_frank_tamm_norm(double lo, double hi, const float qf[4]) {
    return gauss_quad(f_frank_tamm, qf, lo, hi) * 1e9;
}

/* CherenkovSpectrum knots, spectrum.jl:55-71 (inverse CDF at 1-nm steps) and
   :97-100 (LinearInterpolation(norms, wl_steps) evaluated at range(0,1,N)). */
int ntor_cherenkov_spectrum_knots(float wl_lo, float wl_hi, int n_knots, const float qf[4],
                                  float* knots) {
    if (n_knots < 2 || !(wl_hi > wl_lo)) return NT_ERR_ARG;
    int n_steps = (int)floor((double)wl_hi - (double)wl_lo) + 1;
    double* norms = (double*)malloc(sizeof(double) * n_steps);
    double* wls = (double*)malloc(sizeof(double) * n_steps);
    double full = ntor_frank_tamm_norm(wl_lo, wl_hi, qf);
    norms[0] = 0.0;
    wls[0] = wl_lo;
    for (int i = 1; i < n_steps; ++i) {
        wls[i] = (double)wl_lo + (double)i;
        norms[i] = ntor_frank_tamm_norm(wl_lo, wls[i], qf) / full;
    }
    /* norms increase monotonically (the integrand is positive), so the
       sortperm of spectrum.jl:68 is the identity. */
    for (int j = 0; j < n_knots; ++j) {
        double y = (double)j / (double)(n_knots - 1);
        if (y >= norms[n_steps - 1]) {
            knots[j] = (float)wls[n_steps - 1];
            continue;
        }
        int i = 0;
        while (i + 1 < n_steps - 1 && norms[i + 1] < y) ++i;
        if (norms[i + 1] < y) ++i;
        double t = (y - norms[i]) / (norms[i + 1] - norms[i]);
        knots[j] = (float)(wls[i] + t * (wls[i + 1] - wls[i]));
    }
    free(norms);
    free(wls);
    return NT_OK;
}

/* ------------------------------------------------------------------------ */
/* Device-function restatements, FP32                                        */
/* ------------------------------------------------------------------------ */

/* refractive_index_fry, medium.jl:234-244 */
float ntor_refractive_index(float wl, const float qf[4]) {
    float x = 1.0f / wl;
    float x2 = x * x;
    return fmaf(x, qf[1], qf[0]) + fmaf(x2, qf[2], x2 * x * qf[3]);
}

/* dispersion_fry, medium.jl:293-298.  REF mode keeps the destructuring of the
   first three coefficients as (a2, a3, a4); QUAN_FRY mode differentiates the
   Quan-Fry polynomial (dn/dλ = -(a2 + 2 a3 x + 3 a4 x²)/λ²). */
float ntor_dispersion(float wl, const float qf[4], int mode) {
    float a2, a3, a4;
    if (mode == NT_DISPERSION_REF) {
        a2 = qf[0];
        a3 = qf[1];
        a4 = qf[2];
    } else {
        a2 = qf[1];
        a3 = qf[2];
        a4 = qf[3];
    }
    float x = 1.0f / wl;
    float s = (a2 + 2.0f * x * a3) + 3.0f * (x * x) * a4;
    return (s * -1.0f) / (wl * wl);
}

/* group_velocity, medium.jl:313-318 (λ0 = n·λ in REF mode) */
float ntor_group_velocity(float wl, const nt_medium* m) {
    float n = ntor_refractive_index(wl, m->quan_fry);
    float disp = ntor_dispersion(wl, m->quan_fry, m->dispersion_mode);
    float lambda0 = (m->dispersion_mode == NT_DISPERSION_REF) ? n * wl : wl;
    return (float)C_VAC_M_NS / (n - lambda0 * disp);
}

/* _sca_len_part_conc, medium.jl:331-347 (× sca_len_scale, 1 in the reference) */
float ntor_scattering_length(float wl, const nt_medium* m) {
    float x = 550.0f / wl;
    float c = (0.0017f * powf(x, 4.3f) + 1.34f * m->vol_conc_small_part * powf(x, 1.7f)) +
              0.312f * m->vol_conc_large_part * powf(x, 0.3f);
    float l = 1.0f / c;
    return m->sca_len_scale == 1.0f ? l : l * m->sca_len_scale;
}

/* cherenkov_angle, medium.jl:309-311 */
float ntor_cherenkov_angle(float wl, const float qf[4]) {
    return acosf(1.0f / ntor_refractive_index(wl, qf));
}

static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* cuda_hg_scattering_func, photon_prop_cuda.jl:70-77 */
float ntor_hg_cos(float g, float eta) {
    float t = fmaf(-g, g, 1.0f) / fmaf(g, fmaf(2.0f, eta, -1.0f), 1.0f);
    float c = (1.0f / (2.0f * g)) * (fmaf(g, g, 1.0f) - t * t);
    return clampf(c, -1.0f, 1.0f);
}

/* Simplified-Liu phase function (no reference counterpart):
   cosθ = 2 ξ^{1/(1+α)} − 1 with α = 2g/(1−g), so that E[cosθ] = g. */
float ntor_sl_cos(float g, float xi) {
    float alpha = 2.0f * g / (1.0f - g);
    float c = 2.0f * powf(xi, 1.0f / (1.0f + alpha)) - 1.0f;
    return clampf(c, -1.0f, 1.0f);
}

/* sph_to_cart, utils.jl:105-115 */
void ntor_sph_to_cart(float theta, float phi, float out[3]) {
    float st = sinf(theta), ct = cosf(theta), sp = sinf(phi), cp = cosf(phi);
    out[0] = cp * st;
    out[1] = sp * st;
    out[2] = ct;
}

/* rot_from_ez_fast, utils.jl:228-243 (|a3| == 1 branch kept as written) */
void ntor_rot_from_ez(const float a[3], const float op[3], float out[3]) {
    if (fabsf(a[2]) == 1.0f) {
        out[0] = op[0];
        out[1] = copysignf(op[1], a[2]);
        out[2] = copysignf(op[2], a[2]);
        return;
    }
    float a1sq = a[0] * a[0], a2sq = a[1] * a[1];
    float fact = (1.0f - a[2]) / (a2sq + a1sq);
    float a1a2 = a[0] * a[1];
    float x = (fmaf(-a1sq, fact, 1.0f) * op[0] - a1a2 * fact * op[1]) + a[0] * op[2];
    float y = ((-a1a2) * fact * op[0] + fmaf(-a2sq, fact, 1.0f) * op[1]) + a[1] * op[2];
    float z = ((-a[0]) * op[0] - a[1] * op[1]) + a[2] * op[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

/* rot_to_ez_fast, utils.jl:200-225 */
void ntor_rot_to_ez(const float a[3], const float op[3], float out[3]) {
    if (fabsf(a[2]) == 1.0f) {
        out[0] = op[0];
        out[1] = copysignf(op[1], a[2]);
        out[2] = copysignf(op[2], a[2]);
        return;
    }
    float a1sq = a[0] * a[0], a2sq = a[1] * a[1];
    float fact = (1.0f - a[2]) / (a2sq + a1sq);
    float a1a2 = a[0] * a[1];
    float x = (fmaf(-a1sq, fact, 1.0f) * op[0] - a1a2 * fact * op[1]) - a[0] * op[2];
    float y = ((-a1a2) * fact * op[0] + fmaf(-a2sq, fact, 1.0f) * op[1]) - a[1] * op[2];
    float z = (a[0] * op[0] + a[1] * op[1]) + a[2] * op[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

/* initialize_direction_isotropic, photon_prop_cuda.jl:100-104 with
   uniform(min, max) = min + u·(max − min), photon_prop_cuda.jl:53-56 */
void ntor_iso_dir(float u_cos, float u_phi, float out[3]) {
    float theta = acosf(-1.0f + u_cos * 2.0f);
    float phi = 0.0f + u_phi * TWO_PI_F;
    ntor_sph_to_cart(theta, phi, out);
}

/* sample_cherenkov_track_direction, utils.jl:249-260 */
void ntor_cherenkov_track_dir(float u_cos, float u_phi, float out[3]) {
    const float a = 0.39f, b = 2.61f;
    float I = 1.0f - expf(-b * powf(2.0f, a));
    float t = -logf(1.0f - u_cos * I) / b;
    float c = fmaxf(1.0f - powf(t, 1.0f / a), -1.0f);
    float phi = TWO_PI_F * u_phi;
    ntor_sph_to_cart(acosf(c), phi, out);
}

/* Fixed Cherenkov emission angle (rad) of the configuration option
   NT_OPT_CHERENKOV_ANGLE_URAD; 0 = acos(1/n(λ)) as the reference. */
static double g_cherenkov_angle = 0.0;
void ntor_set_cherenkov_angle(double rad) { g_cherenkov_angle = rad; }

/* sample_cherenkov_direction, photon_prop_cuda.jl:107-125 */
void ntor_cherenkov_dir(float wl, float u_phi, float u_tcos, float u_tphi, const float src_dir[3],
                        const float qf[4], float out[3]) {
    float theta_c = g_cherenkov_angle > 0.0 ? acosf((float)cos(g_cherenkov_angle)) : ntor_cherenkov_angle(wl, qf);
    float phi = 0.0f + u_phi * TWO_PI_F;
    float ph[3], track[3], tmp[3];
    ntor_sph_to_cart(theta_c, phi, ph);
    ntor_cherenkov_track_dir(u_tcos, u_tphi, track);
    ntor_rot_from_ez(track, ph, tmp);
    ntor_rot_from_ez(src_dir, tmp, out);
}

/* update_direction, photon_prop_cuda.jl:204-222; with sl_fraction > 0 the
   scattering angle comes from Simplified-Liu when u_mix < sl_fraction. */
void ntor_update_dir(const float dir[3], float eta, float u_phi, float u_mix, const nt_medium* m,
                     float out[3]) {
    float g = m->mean_scattering_angle;
    float c = (u_mix < m->sl_fraction) ? ntor_sl_cos(g, eta) : ntor_hg_cos(g, eta);
    float s = sqrtf(fmaf(-c, c, 1.0f));
    float phi = 0.0f + u_phi * TWO_PI_F;
    float sp = sinf(phi), cp = cosf(phi);
    float nd[3] = {cp * s, sp * s, c};
    ntor_rot_from_ez(dir, nd, out);
}

/* check_intersection(::Spherical, ...), photon_prop_cuda.jl:233-271 */
int ntor_intersect(const float pos[3], const float dir[3], float step, const float tpos[3],
                   float radius, float* d_out) {
    float dp0 = pos[0] - tpos[0], dp1 = pos[1] - tpos[1], dp2 = pos[2] - tpos[2];
    float rsq = radius * radius;
    float a = (dir[0] * dp0 + dir[1] * dp1) + dir[2] * dp2;
    float pp = (dp0 * dp0 + dp1 * dp1) + dp2 * dp2;
    float b = fmaf(a, a, -pp + rsq);
    if (!(b >= 0.0f)) return 0;
    float d = -a - sqrtf(b);
    if ((d > 0.0f) & (d < step)) {
        *d_out = d;
        return 1;
    }
    return 0;
}

/* initialize_wavelength(::CherenkovSpectrum): spectrum.texture[u]
   (photon_prop_cuda.jl:130) — a 1-D texture with normalized coordinates,
   linear filtering and clamp addressing (spectrum.jl:102-103): sample point
   x = u·N − 0.5, i = floor(x), α = x − i, value (1−α)·T[i] + α·T[i+1] with
   indices clamped to [0, N−1].  Q8 rounds α to 1/256 (the texture unit's
   8-bit fractional weight). */
float ntor_spectrum_sample(const nt_spectrum* s, float u) {
    if (s->kind == NT_SPEC_MONOCHROMATIC) return s->wavelength;
    if (s->kind == NT_SPEC_INV_LAMBDA2) {
        float lo = s->wl_range[0], hi = s->wl_range[1];
        return 1.0f / ((1.0f - u) / lo + u / hi);
    }
    int n = s->n_knots;
    const float* k = s->knots;
    if (s->addressing == NT_TEX_ENDPOINT) {
        float x = u * (float)(n - 1);
        float fi = floorf(x);
        int i = (int)fi;
        if (i >= n - 1) return k[n - 1];
        float al = x - fi;
        return (1.0f - al) * k[i] + al * k[i + 1];
    }
    float x = u * (float)n - 0.5f;
    float fi = floorf(x);
    float al = x - fi;
    if (s->addressing == NT_TEX_TEXEL_CENTER_CLAMP_Q8) al = rintf(al * 256.0f) * (1.0f / 256.0f);
    int i0 = (int)fi, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > n - 1 ? n - 1 : i1);
    return (1.0f - al) * k[i0] + al * k[i1];
}

/* Box-Muller standard normal from two uniforms */
static inline float box_muller(uint32_t r0, uint32_t r1) {
    float u1 = u01_oc(r0), u2 = u01(r1);
    return sqrtf(-2.0f * logf(u1)) * cosf(TWO_PI_F * u2);
}

/* rand_gamma, utils.jl:267-294 (Marsaglia-Tsang); trial t uses Philox block
   0x80000000+t: normal from words 0,1 and the acceptance uniform from word 2.
   The reference loops until a trial is accepted; here (as in the kernel) at
   most NTOR_GAMMA_MAX_TRIALS trials run, after which d·scale is returned with
   *trials = NTOR_GAMMA_MAX_TRIALS + 1 (for a finite shape > 1 a trial is
   accepted with probability above 0.9, so the bound is never reached). */
#define NTOR_GAMMA_MAX_TRIALS 4096u
float ntor_rand_gamma(float shape, float scale, uint64_t seed, uint32_t src, uint64_t photon,
                      int* trials) {
    float d = (float)((double)shape - 1.0 / 3.0);
    float c = 1.0f / sqrtf(9.0f * d);
    for (uint32_t t = 0; t < NTOR_GAMMA_MAX_TRIALS; ++t) {
        uint32_t r[4];
        draw_block(seed, src, photon, 0x80000000u + t, r);
        float x = box_muller(r[0], r[1]);
        float v = fmaf(c, x, 1.0f);
        if (v <= 0.0f) continue;
        v = v * v * v;
        float u = u01_oc(r[2]);
        float xsq = x * x;
        /* squeeze of utils.jl:286 as written: 1 − fma(0.0331, x⁴, 1) is never
           positive, so this branch is never taken */
        if (u < 1.0f - fmaf(0.0331f, xsq * xsq, 1.0f)) {
            if (trials) *trials = (int)t + 1;
            return d * v * scale;
        }
        if (logf(u) < 0.5f * xsq + d * ((1.0f - v) + logf(v))) {
            if (trials) *trials = (int)t + 1;
            return d * v * scale;
        }
    }
    if (trials) *trials = (int)NTOR_GAMMA_MAX_TRIALS + 1;
    return d * scale;
}

/* ------------------------------------------------------------------------ */
/* FP64 evaluations                                                          */
/* ------------------------------------------------------------------------ */
double ntor_hg_cos_f64(double g, double eta) {
    double t = (1.0 - g * g) / (1.0 + g * (2.0 * eta - 1.0));
    double c = (1.0 / (2.0 * g)) * (1.0 + g * g - t * t);
    return c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
}
double ntor_sl_cos_f64(double g, double xi) {
    double alpha = 2.0 * g / (1.0 - g);
    double c = 2.0 * pow(xi, 1.0 / (1.0 + alpha)) - 1.0;
    return c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
}
void ntor_rot_from_ez_f64(const double a[3], const double op[3], double out[3]) {
    if (fabs(a[2]) == 1.0) {
        out[0] = op[0];
        out[1] = copysign(op[1], a[2]);
        out[2] = copysign(op[2], a[2]);
        return;
    }
    double a1sq = a[0] * a[0], a2sq = a[1] * a[1];
    double fact = (1.0 - a[2]) / (a2sq + a1sq);
    double a1a2 = a[0] * a[1];
    out[0] = (1.0 - a1sq * fact) * op[0] - a1a2 * fact * op[1] + a[0] * op[2];
    out[1] = -a1a2 * fact * op[0] + (1.0 - a2sq * fact) * op[1] + a[1] * op[2];
    out[2] = -a[0] * op[0] - a[1] * op[1] + a[2] * op[2];
}
/* sample_cherenkov_track_direction, utils.jl:249-260, with T = Float64 */
void ntor_cherenkov_track_dir_f64(double u_cos, double u_phi, double out[3]) {
    const double a = 0.39, b = 2.61;
    double I = 1.0 - exp(-b * pow(2.0, a));
    double c = fmax(1.0 - pow(-log(1.0 - u_cos * I) / b, 1.0 / a), -1.0);
    double phi = 2.0 * M_PI * u_phi, st = sqrt(fmax(1.0 - c * c, 0.0));
    out[0] = cos(phi) * st;
    out[1] = sin(phi) * st;
    out[2] = c;
}

/* sample_cherenkov_direction, photon_prop_cuda.jl:107-125, in FP64 from the
   FP32 wavelength, uniforms, Quan-Fry coefficients and source direction */
void ntor_cherenkov_dir_f64(float wl, float u_phi, float u_tcos, float u_tphi, const float src_dir[3],
                            const float qf[4], double out[3]) {
    double q[4] = {qf[0], qf[1], qf[2], qf[3]};
    double n = ntor_refractive_index_f64(wl, q);
    double ct = 1.0 / n, st = sqrt(1.0 - ct * ct), phi = 2.0 * M_PI * (double)u_phi;
    if (g_cherenkov_angle > 0.0) {
        ct = cos(g_cherenkov_angle);
        st = sin(g_cherenkov_angle);
    }
    double ph[3] = {cos(phi) * st, sin(phi) * st, ct}, track[3], tmp[3];
    double src[3] = {src_dir[0], src_dir[1], src_dir[2]};
    ntor_cherenkov_track_dir_f64(u_tcos, u_tphi, track);
    ntor_rot_from_ez_f64(track, ph, tmp);
    ntor_rot_from_ez_f64(src, tmp, out);
}

/* update_direction, photon_prop_cuda.jl:204-222, in FP64 (the Simplified-Liu
   extension keeps its FP32 cosine) */
void ntor_update_dir_f64(const float dir[3], float eta, float u_phi, float u_mix, const nt_medium* m,
                         double out[3]) {
    float g = m->mean_scattering_angle;
    double c = (u_mix < m->sl_fraction) ? (double)ntor_sl_cos(g, eta) : ntor_hg_cos_f64(g, eta);
    double s = sqrt(fmax(1.0 - c * c, 0.0)), phi = 2.0 * M_PI * (double)u_phi;
    double nd[3] = {cos(phi) * s, sin(phi) * s, c}, d[3] = {dir[0], dir[1], dir[2]};
    ntor_rot_from_ez_f64(d, nd, out);
}

static void round3(const double v[3], float out[3]) {
    out[0] = (float)v[0];
    out[1] = (float)v[1];
    out[2] = (float)v[2];
}

int ntor_intersect_f64(const double pos[3], const double dir[3], double step, const double tpos[3],
                       double radius, double* d_out) {
    double dp[3] = {pos[0] - tpos[0], pos[1] - tpos[1], pos[2] - tpos[2]};
    double a = dir[0] * dp[0] + dir[1] * dp[1] + dir[2] * dp[2];
    double pp = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    double b = a * a - pp + radius * radius;
    if (!(b >= 0.0)) return 0;
    double d = -a - sqrt(b);
    *d_out = d;
    return (d > 0.0) && (d < step);
}
double ntor_refractive_index_f64(double wl, const double qf[4]) {
    double x = 1.0 / wl;
    return qf[0] + x * qf[1] + x * x * qf[2] + x * x * x * qf[3];
}
double ntor_group_velocity_f64(double wl, const double qf[4], int mode) {
    double n = ntor_refractive_index_f64(wl, qf);
    double x = 1.0 / wl, disp, lambda0;
    if (mode == NT_DISPERSION_REF) {
        disp = -(qf[0] + 2.0 * x * qf[1] + 3.0 * x * x * qf[2]) / (wl * wl);
        lambda0 = n * wl;
    } else {
        disp = -(qf[1] + 2.0 * x * qf[2] + 3.0 * x * x * qf[3]) / (wl * wl);
        lambda0 = wl;
    }
    return C_VAC_M_NS / (n - lambda0 * disp);
}
double ntor_scattering_length_f64(double wl, double vcs, double vcl) {
    double x = 550.0 / wl;
    return 1.0 / (0.0017 * pow(x, 4.3) + 1.34 * vcs * pow(x, 1.7) + 0.312 * vcl * pow(x, 0.3));
}

/* ------------------------------------------------------------------------ */
/* Propagation: cuda_propagate_photons!, photon_prop_cuda.jl:402-506          */
/* ------------------------------------------------------------------------ */
typedef struct {
    float pos[3], dir[3], idir[3], wl, time, dist;
    uint16_t module_id;
    uint32_t src;
    uint64_t photon;
} hitrec;

typedef struct {
    hitrec* v;
    int64_t n, cap;
} hitbuf;

static void hb_push(hitbuf* b, const hitrec* h) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->v = (hitrec*)realloc(b->v, sizeof(hitrec) * b->cap);
    }
    b->v[b->n++] = *h;
}

/* the medium's direction arithmetic: FP32 as the reference, or FP64 rounded once */
static void cherenkov_dir_mode(float wl, float u_phi, float u_tcos, float u_tphi, const float src[3],
                               const nt_medium* m, float out[3]) {
    if (m->direction_arithmetic == NT_DIR_FP64) {
        double v[3];
        ntor_cherenkov_dir_f64(wl, u_phi, u_tcos, u_tphi, src, m->quan_fry, v);
        round3(v, out);
    } else {
        ntor_cherenkov_dir(wl, u_phi, u_tcos, u_tphi, src, m->quan_fry, out);
    }
}

static void update_dir_mode(const float dir[3], float eta, float u_phi, float u_mix, const nt_medium* m,
                            float out[3]) {
    if (m->direction_arithmetic == NT_DIR_FP64) {
        double v[3];
        ntor_update_dir_f64(dir, eta, u_phi, u_mix, m, v);
        round3(v, out);
    } else {
        ntor_update_dir(dir, eta, u_phi, u_mix, m, out);
    }
}

/* initialize_photon_state, photon_prop_cuda.jl:133-188 */
static void init_photon(const nt_source* s, const nt_medium* m, const nt_spectrum* sp,
                        uint64_t seed, uint32_t si, uint64_t p, float pos[3], float dir[3],
                        float* time, float* wl) {
    const int fp64 = m->direction_arithmetic == NT_DIR_FP64;
    uint32_t r[4];
    draw_block(seed, si, p, 0u, r);
    *wl = ntor_spectrum_sample(sp, u01(r[0]));
    pos[0] = s->position[0];
    pos[1] = s->position[1];
    pos[2] = s->position[2];
    *time = s->time;
    switch (s->kind) {
        case NT_SRC_POINTLIKE_ISOTROPIC:
            ntor_iso_dir(u01(r[1]), u01(r[2]), dir);
            break;
        case NT_SRC_POINTLIKE_TIME_RANGE:
            ntor_iso_dir(u01(r[1]), u01(r[2]), dir);
            *time = s->time + u01(r[3]) * (s->time_end - s->time);
            break;
        case NT_SRC_AXICONE: {
            float phi = 0.0f + u01(r[2]) * TWO_PI_F, v[3];
            ntor_sph_to_cart(s->angle, phi, v);
            if (fp64) {
                double a[3] = {s->direction[0], s->direction[1], s->direction[2]}, o[3] = {v[0], v[1], v[2]}, w[3];
                ntor_rot_from_ez_f64(a, o, w);
                round3(w, dir);
            } else {
                ntor_rot_from_ez(s->direction, v, dir);
            }
            break;
        }
        case NT_SRC_PENCIL:
            dir[0] = s->direction[0];
            dir[1] = s->direction[1];
            dir[2] = s->direction[2];
            break;
        case NT_SRC_EXTENDED_CHERENKOV: {
            /* photon_prop_cuda.jl:164-176 */
            float lp = ntor_rand_gamma(s->long_a, 1.0f / s->long_b, seed, si, p, NULL) * s->long_lrad;
            for (int k = 0; k < 3; ++k) pos[k] = s->position[k] + lp * s->direction[k];
            *time = s->time + lp / (float)C_VAC_M_NS;
            cherenkov_dir_mode(*wl, u01(r[1]), u01(r[2]), u01(r[3]), s->direction, m, dir);
            break;
        }
        case NT_SRC_POINTLIKE_CHERENKOV:
            cherenkov_dir_mode(*wl, u01(r[1]), u01(r[2]), u01(r[3]), s->direction, m, dir);
            break;
        default:
            dir[0] = dir[1] = 0.0f;
            dir[2] = 1.0f;
    }
}

static void propagate_one(const nt_source* s, uint32_t si, uint64_t p, const nt_target* tg,
                          int nt, const nt_medium* m, const nt_spectrum* sp, uint64_t seed,
                          hitbuf* hb) {
    float pos[3], dir[3], idir[3], time, wl;
    init_photon(s, m, sp, seed, si, p, pos, dir, &time, &wl);
    idir[0] = dir[0];
    idir[1] = dir[1];
    idir[2] = dir[2];
    float dist = 0.0f;
    float sca_len = ntor_scattering_length(wl, m);
    float c_grp = ntor_group_velocity(wl, m);
    /* Scattering steps read the Philox stream of blocks 1, 2, ... three words
       per step (four steps per three blocks): step length, HG/SL uniform and
       azimuth from the top 24 bits of the three words, the mixture uniform
       from their low bytes (bits 0-7 of words 1, 2, 3 as bits 8-15, 16-23,
       24-31). */
    uint32_t buf[4];
    uint32_t next_block = 1u;
    int used = 4;
    for (int k = 0; k < m->max_steps; ++k) {
        uint32_t r[4];
        for (int j = 0; j < 3; ++j) {
            if (used == 4) {
                draw_block(seed, si, p, next_block++, buf);
                used = 0;
            }
            r[j] = buf[used++];
        }
        r[3] = ((r[0] & 0xFFu) << 8) | ((r[1] & 0xFFu) << 16) | ((r[2] & 0xFFu) << 24);
        float step = -logf(u01_oc(r[0])) * sca_len;
        int hit = 0, hit_t = -1;
        float dt = 0.0f;
        for (int t = 0; t < nt; ++t) {
            float d;
            if (ntor_intersect(pos, dir, step, tg[t].position, tg[t].radius, &d)) {
                hit = 1;
                hit_t = t;
                dt = d;
                break;
            }
        }
        if (!hit) {
            for (int q = 0; q < 3; ++q) pos[q] = pos[q] + dir[q] * step;
            dist += step;
            time += step / c_grp;
            float nd[3];
            update_dir_mode(dir, u01(r[1]), u01(r[2]), u01(r[3]), m, nd);
            dir[0] = nd[0];
            dir[1] = nd[1];
            dir[2] = nd[2];
            continue;
        }
        for (int q = 0; q < 3; ++q) pos[q] = pos[q] + dir[q] * dt;
        dist += dt;
        time += dt / c_grp;
        hitrec h;
        memcpy(h.pos, pos, sizeof pos);
        memcpy(h.dir, dir, sizeof dir);
        memcpy(h.idir, idir, sizeof idir);
        h.wl = wl;
        h.time = time;
        h.dist = dist;
        h.module_id = tg[hit_t].module_id;
        h.src = si;
        h.photon = p;
        hb_push(hb, &h);
        break;
    }
}

int ntor_propagate(const nt_source* sources, int n_sources, const nt_target* targets,
                   int n_targets, const nt_medium* medium, const nt_spectrum* spectrum,
                   uint64_t seed, int n_threads, const nt_hits* out, int64_t* n_hits,
                   int64_t* n_simulated) {
    if (n_sources < 0 || n_targets < 0 || !medium || !spectrum || !out) return NT_ERR_ARG;
    int64_t* offs = (int64_t*)malloc(sizeof(int64_t) * (n_sources + 1));
    offs[0] = 0;
    for (int i = 0; i < n_sources; ++i) {
        if (sources[i].photons < 0) {
            free(offs);
            return NT_ERR_ARG;
        }
        offs[i + 1] = offs[i] + sources[i].photons;
    }
    int64_t total = offs[n_sources];
    if (n_threads < 1) n_threads = 1;
    hitbuf* bufs = (hitbuf*)calloc(n_threads, sizeof(hitbuf));
#pragma omp parallel num_threads(n_threads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int64_t lo = total * tid / nth, hi = total * (tid + 1) / nth;
        int s = 0;
        while (s < n_sources && offs[s + 1] <= lo) ++s;
        for (int64_t g = lo; g < hi; ++g) {
            while (offs[s + 1] <= g) ++s;
            propagate_one(&sources[s], (uint32_t)s, (uint64_t)(g - offs[s]), targets, n_targets,
                          medium, spectrum, seed, &bufs[tid]);
        }
    }
    int64_t nh = 0, stored = 0;
    for (int t = 0; t < n_threads; ++t) nh += bufs[t].n;
    for (int t = 0; t < n_threads; ++t) {
        for (int64_t i = 0; i < bufs[t].n && stored < out->capacity; ++i, ++stored) {
            const hitrec* h = &bufs[t].v[i];
            memcpy(out->position + 3 * stored, h->pos, 12);
            memcpy(out->direction + 3 * stored, h->dir, 12);
            memcpy(out->initial_direction + 3 * stored, h->idir, 12);
            out->wavelength[stored] = h->wl;
            out->time[stored] = h->time;
            out->dist_travelled[stored] = h->dist;
            out->module_id[stored] = h->module_id;
            if (out->source_index) out->source_index[stored] = h->src;
            if (out->photon_id) out->photon_id[stored] = h->photon;
        }
        free(bufs[t].v);
    }
    free(bufs);
    free(offs);
    *n_hits = stored;
    *n_simulated = total;
    return nh > out->capacity ? NT_ERR_HIT_OVERFLOW : NT_OK;
}

/* Photons [first, first + count) of one source with RNG source index
   `source_index` (the counter word 3 of its Philox stream), hits ordered by
   photon: lets a test check photons whose 64-bit index has a nonzero high
   word without propagating the 2^32 before them. */
int ntor_propagate_range(const nt_source* source, uint32_t source_index, uint64_t first, int64_t count,
                         const nt_target* targets, int n_targets, const nt_medium* medium,
                         const nt_spectrum* spectrum, uint64_t seed, int n_threads, const nt_hits* out,
                         int64_t* n_hits) {
    if (!source || count < 0 || n_targets < 0 || !medium || !spectrum || !out) return NT_ERR_ARG;
    if (n_threads < 1) n_threads = 1;
    hitbuf* bufs = (hitbuf*)calloc(n_threads, sizeof(hitbuf));
#pragma omp parallel num_threads(n_threads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int64_t lo = count * tid / nth, hi = count * (tid + 1) / nth;
        for (int64_t g = lo; g < hi; ++g)
            propagate_one(source, source_index, first + (uint64_t)g, targets, n_targets, medium, spectrum, seed,
                          &bufs[tid]);
    }
    int64_t nh = 0, stored = 0;
    for (int t = 0; t < n_threads; ++t) nh += bufs[t].n;
    for (int t = 0; t < n_threads; ++t) {
        for (int64_t i = 0; i < bufs[t].n && stored < out->capacity; ++i, ++stored) {
            const hitrec* h = &bufs[t].v[i];
            memcpy(out->position + 3 * stored, h->pos, 12);
            memcpy(out->direction + 3 * stored, h->dir, 12);
            memcpy(out->initial_direction + 3 * stored, h->idir, 12);
            out->wavelength[stored] = h->wl;
            out->time[stored] = h->time;
            out->dist_travelled[stored] = h->dist;
            out->module_id[stored] = h->module_id;
            if (out->source_index) out->source_index[stored] = h->src;
            if (out->photon_id) out->photon_id[stored] = h->photon;
        }
        free(bufs[t].v);
    }
    free(bufs);
    *n_hits = stored;
    return nh > out->capacity ? NT_ERR_HIT_OVERFLOW : NT_OK;
}

/* ------------------------------------------------------------------------ */
/* Post-processing                                                           */
/* ------------------------------------------------------------------------ */

/* fast_linear_interp(x, xs, ys), utils.jl:24-53 */
double ntor_fast_linear_interp(double x, const double* xs, const double* ys, int n) {
    double lower = xs[0], upper = xs[n - 1];
    x = x < lower ? lower : (x > upper ? upper : x);
    if (x == lower) return ys[0];
    if (x == upper) return ys[n - 1];
    int iu = 0;
    while (xs[iu] < x) ++iu; /* searchsortedfirst */
    int il = iu - 1;
    double along = (x - xs[il]) / (xs[iu] - xs[il]);
    return ys[il] + (ys[iu] - ys[il]) * along;
}

/* check_pmt_hit, detection.jl:72-84: first PMT (1-based) whose axis is within
   the opening angle of the hit direction, 0 if none */
int ntor_check_pmt_hit(const double rel[3], const double* pmt_dirs, int n_pmts, double opening) {
    for (int j = 0; j < n_pmts; ++j) {
        const double* q = pmt_dirs + 3 * j;
        double c = rel[0] * q[0] + rel[1] * q[1] + rel[2] * q[2];
        c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
        if (acos(c) < opening) return j + 1;
    }
    return 0;
}

/* cart_to_sph, utils.jl:118-133 */
static void cart_to_sph(double x, double y, double z, double* th, double* ph) {
    if (z == 1.0) {
        *th = 0.0;
        *ph = 0.0;
        return;
    }
    if (z == -1.0) {
        *th = M_PI;
        *ph = 0.0;
        return;
    }
    *th = acos(z);
    if (x == 0.0 && y > 0.0)
        *ph = M_PI / 2;
    else if (x == 0.0 && y < 0.0)
        *ph = -M_PI / 2;
    else
        *ph = atan2(y, x);
}

/* make_pom_pmt_coordinates, detection.jl:107-134: four rings of four PMTs,
   then rotated by calc_rot_matrix(e_z, e_x) (utils.jl:173-184). */
void ntor_pom_pmt_coordinates(double coords[32]) {
    const double deg = M_PI / 180.0;
    for (int j = 0; j < 16; ++j) {
        int ring = j / 4, k = j % 4;
        double th, ph;
        switch (ring) {
            case 0: th = (90 - 57.5) * deg; ph = M_PI / 4 + k * (M_PI / 2); break;
            case 1: th = (90 - 25) * deg; ph = k * (M_PI / 2); break;
            case 2: th = (90 + 25) * deg; ph = k * (M_PI / 2); break;
            default: th = (90 + 57.5) * deg; ph = M_PI / 4 + k * (M_PI / 2); break;
        }
        /* R = I + [v]x + [v]x^2 (1 - a.b)/|v|^2 with a = e_z, b = e_x: v = (0,1,0) */
        double c[3] = {cos(ph) * sin(th), sin(ph) * sin(th), cos(th)};
        /* [v]x = [[0,0,1],[0,0,0],[-1,0,0]]; [v]x^2 = diag(-1,0,-1); R = [[0,0,1],[0,1,0],[-1,0,0]] */
        double r[3] = {c[2], c[1], -c[0]};
        cart_to_sph(r[0], r[1], r[2], &th, &ph);
        coords[2 * j] = th;
        coords[2 * j + 1] = ph;
    }
}
