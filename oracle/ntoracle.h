/*
 * ntoracle.h — CPU restatement of the reference photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for libntprop: it may be
 * built and called by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the library or any product path.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  FP32 functions keep the reference's
 * operation order and its explicit fma calls and are compiled without
 * contraction (-ffp-contract=off), as Julia emits no implicit fma.  *_f64
 * variants evaluate the same formulas in double precision for ulp checks.
 */
#ifndef NTORACLE_H
#define NTORACLE_H

#include <stdint.h>
#include "../include/ntprop.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RNG: Philox4x32-10 */
void ntor_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* host-side constructors (FP64) */
double ntor_dippr105(double temperature_k);
void ntor_quan_fry_params(double salinity, double temperature, double pressure_atm, double out[4]);
void ntor_make_cascadia_medium(float g, nt_medium* out, double* density, double* radiation_length);
double ntor_integrate_gauss_quad_pow(double a, double b, int power);
double ntor_integrate_gauss_quad_cos(double a, double b);
double ntor_frank_tamm(double wl, double n);
double ntor_frank_tamm_norm(double wl_lo, double wl_hi, const float qf[4]);
int ntor_cherenkov_spectrum_knots(float wl_lo, float wl_hi, int n_knots, const float qf[4],
                                  float* knots_out);

/* device-function restatements (FP32) */
float ntor_refractive_index(float wl, const float qf[4]);
float ntor_dispersion(float wl, const float qf[4], int mode);
float ntor_group_velocity(float wl, const nt_medium* m);
float ntor_scattering_length(float wl, const nt_medium* m);
float ntor_cherenkov_angle(float wl, const float qf[4]);
float ntor_hg_cos(float g, float eta);
float ntor_sl_cos(float g, float xi);
void ntor_sph_to_cart(float theta, float phi, float out[3]);
void ntor_rot_from_ez(const float a[3], const float op[3], float out[3]);
void ntor_rot_to_ez(const float a[3], const float op[3], float out[3]);
void ntor_iso_dir(float u_cos, float u_phi, float out[3]);
void ntor_cherenkov_track_dir(float u_cos, float u_phi, float out[3]);
void ntor_cherenkov_dir(float wl, float u_phi, float u_tcos, float u_tphi, const float src_dir[3],
                        const float qf[4], float out[3]);
void ntor_update_dir(const float dir[3], float eta, float u_phi, float u_mix, const nt_medium* m,
                     float out[3]);
int ntor_intersect(const float pos[3], const float dir[3], float step, const float tpos[3],
                   float radius, float* d);
float ntor_spectrum_sample(const nt_spectrum* s, float u);
float ntor_rand_gamma(float shape, float scale, uint64_t seed, uint32_t src, uint64_t photon,
                      int* trials);

/* FP64 evaluations of the same formulas */
double ntor_hg_cos_f64(double g, double eta);
double ntor_sl_cos_f64(double g, double xi);
void ntor_rot_from_ez_f64(const double a[3], const double op[3], double out[3]);
void ntor_cherenkov_track_dir_f64(double u_cos, double u_phi, double out[3]);
void ntor_set_cherenkov_angle(double rad);
void ntor_cherenkov_dir_f64(float wl, float u_phi, float u_tcos, float u_tphi, const float src_dir[3],
                            const float qf[4], double out[3]);
void ntor_update_dir_f64(const float dir[3], float eta, float u_phi, float u_mix, const nt_medium* m,
                         double out[3]);
int ntor_intersect_f64(const double pos[3], const double dir[3], double step, const double tpos[3],
                       double radius, double* d);
double ntor_group_velocity_f64(double wl, const double qf[4], int mode);
double ntor_refractive_index_f64(double wl, const double qf[4]);
double ntor_scattering_length_f64(double wl, double vcs, double vcl);

/* Propagation (photon_prop_cuda.jl:402-506).  Hits are returned ordered by
   (source index, photon index).  Returns NT_OK or NT_ERR_HIT_OVERFLOW. */
int ntor_propagate(const nt_source* sources, int n_sources, const nt_target* targets,
                   int n_targets, const nt_medium* medium, const nt_spectrum* spectrum,
                   uint64_t seed, int n_threads, const nt_hits* out, int64_t* n_hits,
                   int64_t* n_simulated);
int ntor_propagate_range(const nt_source* source, uint32_t source_index, uint64_t first, int64_t count,
                         const nt_target* targets, int n_targets, const nt_medium* medium,
                         const nt_spectrum* spectrum, uint64_t seed, int n_threads, const nt_hits* out,
                         int64_t* n_hits);

/* Post-processing (photon_prop_cuda.jl:850-924, detection.jl:72-163) */
double ntor_fast_linear_interp(double x, const double* xs, const double* ys, int n);
int ntor_check_pmt_hit(const double rel_pos[3], const double* pmt_dirs, int n_pmts,
                       double opening_angle);
void ntor_pom_pmt_coordinates(double coords[2 * 16]);

#ifdef __cplusplus
}
#endif
#endif
