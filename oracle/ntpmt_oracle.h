/*
 * ntpmt_oracle.h — CPU restatement of the reference's PMT front end
 * (src/pmt_frontend/).  TEST INFRASTRUCTURE ONLY: the checker for the
 * ntpmt_* entry points of libntprop; only tests/ may build or call it.
 */
#ifndef NTPMT_ORACLE_H
#define NTPMT_ORACLE_H

#include <stdint.h>
#include "../include/ntpmt.h"

#ifdef __cplusplus
extern "C" {
#endif

void ntor_pmt_normals(const uint32_t r[4], double out[2]);
void ntor_pmt_hit_draws(const ntpmt_config* c, int64_t n, const float* t, const uint32_t* src, const uint64_t* pid,
                        uint64_t seed, double* t_out, double* q_out, int32_t* flags);
double ntor_pmt_pulse(const ntpmt_config* c, double x);
double ntor_pmt_filt_pulse(const ntpmt_config* c, double s);
void ntor_pmt_iir(const ntpmt_config* c, int32_t n, const double* x, double* y);
void ntor_pmt_decimate(const ntpmt_config* c, int32_t n_in, const double* y, int32_t R, int32_t n_out, double* d);
int32_t ntor_fnnls(int32_t m, int32_t n, const double* A, const double* b, double tol, int64_t max_iter, double* x,
                   int64_t* iter_out, int32_t* outer_out);
int32_t ntor_pmt_sizes(const ntpmt_config* c, int32_t* n_wave, int32_t* n_adc, int32_t* n_pulse, int32_t* R);
int32_t ntor_pmt_channel(const ntpmt_config* c, int32_t n, const double* t, const double* q, uint32_t channel,
                         uint64_t seed, double* t0_out, double* wave_out, double* filt_out, double* digi_out,
                         double* x_out, int64_t* iter_out, int32_t* outer_out);

#ifdef __cplusplus
}
#endif
#endif
