// ntpmt.hip — the PMT front end on the device (C ABI: include/ntpmt.h).
//
// make_reco_pulses (src/pmt_frontend/pipeline.jl:146-160) for every
// (module, PMT) channel of an event in one call:
//   1. hit_draws_kernel: per selected hit, the transit-time Gamma draw and the
//      SPE charge (FP64), and a sort key (channel, FP32 time);
//   2. one radix sort of the keys, a run-length encoding of the channels and
//      an exclusive scan of the run lengths (hipCUB): every channel's hits are
//      contiguous and in time order;
//   3. channel_kernel: one 64-lane workgroup per channel, taken from an atomic
//      work counter over the channels sorted by decreasing hit count (bright
//      channels cost more: they go first, so none is left alone at the end).
//      In LDS: the waveform (n_wave FP64 samples, each a sum over the hits
//      within the pulse's support, found by binary search in the sorted keys,
//      plus noise), the first-order low-pass as a parallel scan of affine maps
//      over the lanes, the decimation by R (one FIR phase, FP64), then the
//      unfolding.
// The unfolding is the fnnls active-set method (Bro & de Jong 1997) in FP64
// on the implicit matrix A[k, i] = filtered pulse at t_k − τ_i: with
// τ_i = t0 + i·res and Q = (1/adc_freq)/res an integer, column i = Q·q + p is
// pattern p shifted down by q rows, so A is J·Q numbers (14 × 40 for the
// standard configuration, rows zero for every pattern dropped) kept in LDS.
// The active set's Gram matrix A_Pᵀ A_P is kept as a packed Cholesky factor
// that grows by one row when a column enters and loses one by Givens
// rotations when a column leaves; its first K rows (53 by default) sit in LDS
// in the waveform's space, the rest in a per-workgroup HBM region, so that
// eight channels share a CU.  Solves are lane-parallel column sweeps (a
// readlane broadcast, a multiply, one fma per lane, the next row loaded a
// step ahead) with the reciprocal diagonal in registers.  Per outer iteration
// the gradient Aᵀ(b − A x) over all n_pulse columns runs with one lane per
// pattern: its taps in registers and a window of residual rows rotating
// through registers.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "ntpmt.h"
#include "ntprop_device.h"

struct ntpmt_ctx;

namespace ntpmt {
namespace {

using ntp::philox;
using ntp::u53;

constexpr uint32_t kTtWord = 0x72000000u;
constexpr uint32_t kSpeWord = 0x73000000u;
constexpr uint32_t kNoiseWord = 0x74000000u;
constexpr int kTtTrials = 64;
constexpr int kSpeTrials = 16;
constexpr uint32_t kNoChannel = (1u << 21) - 1;      // key of hits that are not selected
constexpr int kLanes = 64;
static_assert(NTPMT_MAX_ADC <= 2 * kLanes, "active-set positions live in two registers per lane");
constexpr double kDegenerate = 1e-14;                // a column with d² ≤ this · ‖a‖² is numerically dependent
constexpr int kTaps = 16;                            // rows a column reaches in the register-window gradient pass
constexpr int kPadLo = 16, kPadHi = 48;              // zero rows around the residual (padded LDS array)
constexpr size_t kLdsPerCU = 160 * 1024;             // LDS of a CU
constexpr int kChannelsPerCU = 12;                   // one-wave workgroups a CU runs at once (3 per SIMD)

struct DevCfg {
    double tt_d, tt_c, tt_scale, tt_subtract;
    double spe_w, spe_scale, spe_mu, spe_sigma, spe_low;
    double p_mu, p_inv_beta, p_lo, p_hi, p_norm;
    double noise_amp, dt_wave, dt_adc, res, lead, min_charge;
    double b0, b1, a1, tol;
    int64_t max_iter;
    int n_wave, n_adc, n_pulse, R, Q, J, m0, n_taps, chunk, max_outer;
    int lds_rows;            // factor rows kept in LDS
    const double* taps;      // device [n_taps]
    const double* table;     // device [J][Q]
};

struct HitIn {
    const float* time;
    const uint16_t* module_id;
    const uint8_t* pmt_id;
    const uint32_t* source_index;
    const uint64_t* photon_id;
    const uint8_t* select;
    int64_t n;
    uint64_t seed;
};

struct ChanArgs {
    double* spill;             // [gridDim.x][spill_stride] factor rows beyond the LDS ones
    int64_t spill_stride;
    const uint64_t* key;       // sorted keys
    const int32_t* idx;        // hit index of each sorted key
    const double* tq;          // [n][2] (t', charge) by hit index
    const uint8_t* hflag;      // per hit status bits
    const uint32_t* unique;    // channel of each run
    const int32_t* counts;     // run lengths
    const int32_t* offsets;    // run starts
    const int32_t* n_runs;
    const int32_t* order;      // runs by decreasing hit count (heaviest first)
    int32_t* next;             // work counter
    int64_t capacity;
    uint64_t seed;
    ntpmt_pulses out;
};

// ---------------------------------------------------------------- RNG ---
__device__ __forceinline__ void normals(uint4 r, double& n0, double& n1) {
    const double u1 = 1.0 - u53(r.x, r.y);          // (0, 1]
    const double u2 = u53(r.z, r.w);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    n0 = rad * c;
    n1 = rad * s;
}

// Gamma(α, θ): Marsaglia & Tsang (2000), no squeeze; d = α − 1/3, c = 1/√(9d)
__device__ double tt_draw(const DevCfg& c, uint32_t k0, uint32_t k1, uint32_t plo, uint32_t phi, uint32_t src,
                          uint8_t& flag) {
    for (int t = 0; t < kTtTrials; ++t) {
        double x, unused;
        normals(philox(make_uint4(kTtWord + 2u * t, plo, phi, src), k0, k1), x, unused);
        double v = 1.0 + c.tt_c * x;
        if (!(v > 0.0)) continue;
        v = v * v * v;
        const uint4 a = philox(make_uint4(kTtWord + 2u * t + 1u, plo, phi, src), k0, k1);
        const double u = 1.0 - u53(a.x, a.y);
        if (log(u) < 0.5 * x * x + c.tt_d - c.tt_d * v + c.tt_d * log(v)) return c.tt_d * v * c.tt_scale;
    }
    flag |= NTPMT_ST_TT_GIVEUP;
    return (c.tt_d + 1.0 / 3.0) * c.tt_scale;
}

// SPE charge: exponential with probability w, else Normal(mu, sigma) truncated to [low, ∞)
__device__ double spe_draw(const DevCfg& c, uint32_t k0, uint32_t k1, uint32_t plo, uint32_t phi, uint32_t src,
                           uint8_t& flag) {
    const uint4 r = philox(make_uint4(kSpeWord, plo, phi, src), k0, k1);
    if (u53(r.x, r.y) < c.spe_w) return -log(1.0 - u53(r.z, r.w)) * c.spe_scale;
    for (int t = 0; t < kSpeTrials; ++t) {
        double n0, n1;
        normals(philox(make_uint4(kSpeWord + 1u + t, plo, phi, src), k0, k1), n0, n1);
        const double v0 = c.spe_mu + c.spe_sigma * n0;
        if (v0 >= c.spe_low) return v0;
        const double v1 = c.spe_mu + c.spe_sigma * n1;
        if (v1 >= c.spe_low) return v1;
    }
    flag |= NTPMT_ST_SPE_GIVEUP;
    return c.spe_mu;
}

// ascending order of a float as ascending order of a uint32
__device__ __forceinline__ uint32_t ordered32(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b >> 31) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void hit_draws_kernel(HitIn h, DevCfg c, uint64_t* key, int32_t* idx, double* tq,
                                                        uint8_t* hflag) {
    const uint32_t k0 = (uint32_t)h.seed, k1 = (uint32_t)(h.seed >> 32);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < h.n; i += (int64_t)gridDim.x * 256) {
        uint64_t k = ((uint64_t)kNoChannel << 32) | 0xFFFFFFFFull;
        if (!h.select || h.select[i]) {
            const uint64_t pid = h.photon_id[i];
            const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32), src = h.source_index[i];
            uint8_t f = 0;
            const double tt = tt_draw(c, k0, k1, plo, phi, src, f);
            const double tp = ((double)h.time[i] + tt) - c.tt_subtract;     // apply_tt!, subtract_mean_tt!
            const double q = spe_draw(c, k0, k1, plo, phi, src, f);
            const uint32_t ch = (uint32_t)h.module_id[i] * 17u + (uint32_t)min((int)h.pmt_id[i], 16);
            k = ((uint64_t)ch << 32) | ordered32((float)tp);
            tq[2 * i] = tp;
            tq[2 * i + 1] = q;
            hflag[i] = f;
        }
        key[i] = k;
        idx[i] = (int32_t)i;
    }
}

__global__ void channel_ids_kernel(const uint64_t* __restrict__ key, int64_t n, uint32_t* ch) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        ch[i] = (uint32_t)(key[i] >> 32);
}

// count of channels (the run of unselected hits, if any, is last) and the work counter
__global__ void finish_count_kernel(const int32_t* n_runs, const uint32_t* unique, int64_t* n_channels,
                                    int32_t* next) {
    const int32_t r = *n_runs;
    const int64_t nc = (r > 0 && unique[r - 1] == kNoChannel) ? r - 1 : r;
    if (n_channels) *n_channels = nc;
    *next = 0;
}

// sort keys that put the channel runs in decreasing order of their hit count
// (the run of unselected hits and the unused entries last)
__global__ void order_keys_kernel(const int32_t* n_runs, const uint32_t* unique, const int32_t* counts, int64_t n,
                                  uint32_t* key, int32_t* run) {
    const int32_t r = *n_runs;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        key[i] = (i < r && unique[i] != kNoChannel) ? ~(uint32_t)counts[i] : 0xFFFFFFFFu;
        run[i] = (int32_t)i;
    }
}

__global__ void zero_count_kernel(int64_t* n_channels) {
    if (n_channels) *n_channels = 0;
}

// ------------------------------------------------------------- wave ops ---
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;                                 // a butterfly: every lane holds the same sum
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ bool wany(bool b) { return __ballot(b) != 0; }
// lane l's value (l uniform) by v_readlane: a scalar broadcast, no LDS round trip
__device__ __forceinline__ double readlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
// value of position j (uniform) from the two per-lane registers
__device__ __forceinline__ double getp(double a, double b, int j) {
    return j < kLanes ? readlane(a, j) : readlane(b, j - kLanes);
}
__device__ __forceinline__ int getpi(int a, int b, int j) {
    return j < kLanes ? __builtin_amdgcn_readlane(a, j) : __builtin_amdgcn_readlane(b, j - kLanes);
}
// set position j (uniform) of a register pair (selects on both registers:
// no choice between the two addresses, which would put them in memory)
__device__ __forceinline__ void setp(double& a, double& b, int j, double v, int lane) {
    const bool in_a = j < kLanes;
    a = (in_a && lane == j) ? v : a;
    b = (!in_a && lane == j - kLanes) ? v : b;
}
__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

struct Lds {
    double* T;          // [J][Q]
    double* buf;        // waveform [n_wave], then the Cholesky factor's rows 0 … K−1 (packed)
    double* glb;        // the factor's rows K … (packed, this workgroup's HBM region)
    int K, triK;        // rows in LDS, tri(K)
    double* b;          // [n_adc] digitised waveform
    double* r;          // [n_adc] residual b − A x, inside kPadLo / kPadHi zeros
    int2* pos_qp;       // [n_adc] (q, p) of the active positions (residual)
    double* pos_x;      // [n_adc] x of the active positions (residual)
    uint32_t* inP;      // [n_pulse / 32] active set
    uint32_t* excl;     // [n_pulse / 32] columns skipped as dependent
};

// a_colᵀ v over the rows the column reaches (k ascending)
__device__ __forceinline__ double adot(const Lds& s, const DevCfg& c, int q, int p, const double* v) {
    double acc = 0.0;
    for (int j = 0; j < c.J; ++j) {
        const int k = q + c.m0 + j;
        if (k >= 0 && k < c.n_adc) acc = fma(s.T[j * c.Q + p], v[k], acc);
    }
    return acc;
}

// a_iᵀ a_j for columns i = (qi, pi), j = (qj, pj)
__device__ __forceinline__ double gram(const Lds& s, const DevCfg& c, int qi, int pi, int qj, int pj) {
    double acc = 0.0;
    for (int ji = 0; ji < c.J; ++ji) {
        const int k = qi + c.m0 + ji, jj = ji + qi - qj;
        if (k >= 0 && k < c.n_adc && jj >= 0 && jj < c.J) acc = fma(s.T[ji * c.Q + pi], s.T[jj * c.Q + pj], acc);
    }
    return acc;
}

__device__ __forceinline__ bool bit(const uint32_t* m, int i) { return (m[i >> 5] >> (i & 31)) & 1u; }

// Factor entry (i, j), i ≥ j.  Rows below K sit in LDS; the rare rows beyond
// (more than K active columns) in the workgroup's HBM region.  kSpill = false
// is the all-LDS form the callers use while every row involved is below K.
template <bool kSpill>
__device__ __forceinline__ double fac_ld(const Lds& s, int i, int j) {
    if (!kSpill || i < s.K) return s.buf[tri(i) + j];
    return s.glb[tri(i) - s.triK + j];
}
template <bool kSpill>
__device__ __forceinline__ void fac_st(const Lds& s, int i, int j, double v) {
    if (!kSpill || i < s.K) s.buf[tri(i) + j] = v;
    else s.glb[tri(i) - s.triK + j] = v;
}

// Active set: positions 0 … P−1, position j in lane j mod 64, register A (j < 64) or B.
struct Active {
    int P = 0;
    int colA = 0, colB = 0;      // column q·Q + p
    int qA = 0, qB = 0, pA = 0, pB = 0;
    double xA = 0, xB = 0, sA = 0, sB = 0, zA = 0, zB = 0, atbA = 0, atbB = 0;
    double dA = 0, dB = 0;       // reciprocal diagonal of the factor's row j
};

// Append column `col` as position t: extends the Cholesky factor by row t
// (forward substitution with the Gram row of the active columns) and z by z_t.
// Returns false (nothing written) if the column is numerically dependent on
// positions 0 … t−1.
template <bool kSpill>
__device__ __forceinline__ bool chol_append(const Lds& s, const DevCfg& c, Active& a, int t, int col, double atb_col,
                                            int lane) {
    const int cq = col / c.Q, cp = col - cq * c.Q;
    double gA = (lane < t) ? gram(s, c, a.qA, a.pA, cq, cp) : 0.0;
    double gB = (lane + kLanes < t) ? gram(s, c, a.qB, a.pB, cq, cp) : 0.0;
    double lA = 0.0, lB = 0.0;
    // column-oriented forward substitution; each lane's factor entry of the
    // next step is loaded a step ahead (the LDS latency stays off the chain
    // readlane → multiply → fma)
    const int iA = lane, iB = lane + kLanes;
    double eA = (iA > 0 && iA < t) ? fac_ld<kSpill>(s, iA, 0) : 0.0;
    double eB = (iB < t) ? fac_ld<kSpill>(s, iB, 0) : 0.0;
    for (int j = 0; j < t; ++j) {
        const int jn = j + 1;
        const double nA = (iA > jn && iA < t) ? fac_ld<kSpill>(s, iA, jn) : 0.0;
        const double nB = (iB > jn && iB < t) ? fac_ld<kSpill>(s, iB, jn) : 0.0;
        const double lj = getp(gA, gB, j) * getp(a.dA, a.dB, j);
        setp(lA, lB, j, lj, lane);
        gA = fma(-eA, lj, gA);                     // eA = 0 on the lanes this step leaves alone
        gB = fma(-eB, lj, gB);
        eA = nA;
        eB = nB;
    }
    const double gcc = gram(s, c, cq, cp, cq, cp);
    const double d2 = gcc - wsum(lA * lA + lB * lB);
    if (!(d2 > kDegenerate * gcc)) return false;
    const double d = sqrt(d2), id = 1.0 / d;
    const double zt = (atb_col - wsum((lane < t ? lA * a.zA : 0.0) + (lane + kLanes < t ? lB * a.zB : 0.0))) * id;
    if (lane < t) fac_st<kSpill>(s, t, lane, lA);
    if (lane + kLanes < t) fac_st<kSpill>(s, t, lane + kLanes, lB);
    if (lane == 0) fac_st<kSpill>(s, t, t, d);
    const bool own_a = t < kLanes && lane == t, own_b = t >= kLanes && lane == t - kLanes;
    a.colA = own_a ? col : a.colA;
    a.qA = own_a ? cq : a.qA;
    a.pA = own_a ? cp : a.pA;
    a.qB = own_b ? cq : a.qB;
    a.pB = own_b ? cp : a.pB;
    a.zA = own_a ? zt : a.zA;
    a.atbA = own_a ? atb_col : a.atbA;
    a.dA = own_a ? id : a.dA;
    a.colB = own_b ? col : a.colB;
    a.zB = own_b ? zt : a.zB;
    a.atbB = own_b ? atb_col : a.atbB;
    a.dB = own_b ? id : a.dB;
    __syncthreads();
    return true;
}

// s = L⁻ᵀ z over positions 0 … P−1 (0 beyond); row j of the factor is
// loaded one step ahead
template <bool kSpill>
__device__ __forceinline__ void back_solve_t(const Lds& s, Active& a, int lane) {
    double zA = a.zA, zB = a.zB;
    a.sA = 0.0;
    a.sB = 0.0;
    int j = a.P - 1;
    double eA = (lane < j) ? fac_ld<kSpill>(s, j, lane) : 0.0;
    double eB = (lane + kLanes < j) ? fac_ld<kSpill>(s, j, lane + kLanes) : 0.0;
    for (; j >= 0; --j) {
        const int jn = j - 1;
        const double nA = (lane < jn) ? fac_ld<kSpill>(s, jn, lane) : 0.0;
        const double nB = (lane + kLanes < jn) ? fac_ld<kSpill>(s, jn, lane + kLanes) : 0.0;
        const double sj = getp(zA, zB, j) * getp(a.dA, a.dB, j);
        setp(a.sA, a.sB, j, sj, lane);
        zA = fma(-eA, sj, zA);
        zB = fma(-eB, sj, zB);
        eA = nA;
        eB = nB;
    }
}

// position j + 1 → j for j ≥ r (the A/B register pair of a per-position value)
template <typename T>
__device__ __forceinline__ void shift_down(T& A, T& B, int r, int lane) {
    const T a1 = __shfl_down(A, 1), b0 = __shfl(B, 0), b1 = __shfl_down(B, 1);
    const T na = lane < r ? A : (lane == kLanes - 1 ? b0 : a1);
    const T nb = lane + kLanes < r ? B : b1;
    A = na;
    B = nb;
}

// Remove position r: the factor without row r is lower triangular but for
// one entry right of the diagonal in each later row; Givens rotations of the
// column pairs (k, k+1), k = r … P−2, remove them (L Q Qᵀ Lᵀ = L Lᵀ) and the
// same rotations carry z (L z = Aᵀb stays true).  Then the later rows move up
// one slot and the per-position registers shift down one position.
template <bool kSpill>
__device__ __forceinline__ void chol_delete_t(const Lds& s, Active& a, int r, int lane) {
    const int P = a.P;
    for (int k = r; k < P - 1; ++k) {
        const double av = fac_ld<kSpill>(s, k + 1, k), bv = fac_ld<kSpill>(s, k + 1, k + 1);   // old row k + 1
        const double rr = sqrt(av * av + bv * bv), ir = 1.0 / rr;
        const double cs = av * ir, sn = bv * ir;
        for (int i = k + 2 + lane; i < P; i += kLanes) {
            const double x = fac_ld<kSpill>(s, i, k), y = fac_ld<kSpill>(s, i, k + 1);
            fac_st<kSpill>(s, i, k, cs * x + sn * y);
            fac_st<kSpill>(s, i, k + 1, cs * y - sn * x);
        }
        const double zk = getp(a.zA, a.zB, k), zk1 = getp(a.zA, a.zB, k + 1);
        setp(a.zA, a.zB, k, cs * zk + sn * zk1, lane);
        setp(a.zA, a.zB, k + 1, cs * zk1 - sn * zk, lane);
        setp(a.dA, a.dB, k, ir, lane);
        if (lane == 0) {
            fac_st<kSpill>(s, k + 1, k, rr);
            fac_st<kSpill>(s, k + 1, k + 1, 0.0);
        }
        __syncthreads();
    }
    // old row i (> r), columns 0 … i−1, to slot i − 1: each lane moves the
    // same columns of every row, so a row's source is read before the next
    // row's copy overwrites it, by the same lane
    for (int i = r + 1; i < P; ++i)
        for (int col = lane; col < i; col += kLanes) fac_st<kSpill>(s, i - 1, col, fac_ld<kSpill>(s, i, col));
    __syncthreads();
    shift_down(a.colA, a.colB, r, lane);
    shift_down(a.qA, a.qB, r, lane);
    shift_down(a.pA, a.pB, r, lane);
    shift_down(a.xA, a.xB, r, lane);
    shift_down(a.atbA, a.atbB, r, lane);
    a.P = P - 1;
    setp(a.zA, a.zB, a.P, 0.0, lane);
}

// the all-LDS forms while every row involved is below K
__device__ __forceinline__ void back_solve(const Lds& s, Active& a, int lane) {
    if (a.P > s.K) back_solve_t<true>(s, a, lane); else back_solve_t<false>(s, a, lane);
}
__device__ __forceinline__ void chol_delete(const Lds& s, Active& a, int r, int lane) {
    if (a.P > s.K) chol_delete_t<true>(s, a, r, lane); else chol_delete_t<false>(s, a, r, lane);
}
__device__ __forceinline__ bool chol_add(const Lds& s, const DevCfg& c, Active& a, int t, int col, double atb_col,
                                         int lane) {
    return t >= s.K ? chol_append<true>(s, c, a, t, col, atb_col, lane)
                    : chol_append<false>(s, c, a, t, col, atb_col, lane);
}

// r = b − A x: the positions' (q, p, x) are published in LDS, then each lane
// accumulates its rows lane and lane + 64 over them (loads unrolled ahead of
// the multiply-adds)
__device__ __forceinline__ void residual(const Lds& s, const DevCfg& c, const Active& a, int lane) {
    if (lane < a.P) {
        s.pos_qp[lane] = make_int2(a.qA, a.pA);
        s.pos_x[lane] = a.xA;
    }
    if (lane + kLanes < a.P) {
        s.pos_qp[lane + kLanes] = make_int2(a.qB, a.pB);
        s.pos_x[lane + kLanes] = a.xB;
    }
    __syncthreads();
    const int k0 = lane, k1 = lane + kLanes;
    double r0 = k0 < c.n_adc ? s.b[k0] : 0.0, r1 = k1 < c.n_adc ? s.b[k1] : 0.0;
#pragma unroll 4
    for (int j = 0; j < a.P; ++j) {
        const int2 qp = s.pos_qp[j];
        const double xj = s.pos_x[j];
        const int l0 = k0 - qp.x - c.m0, l1 = k1 - qp.x - c.m0;
        if (l0 >= 0 && l0 < c.J) r0 = fma(-s.T[l0 * c.Q + qp.y], xj, r0);
        if (l1 >= 0 && l1 < c.J) r1 = fma(-s.T[l1 * c.Q + qp.y], xj, r1);
    }
    if (k0 < c.n_adc) s.r[k0] = r0;
    if (k1 < c.n_adc) s.r[k1] = r1;
    __syncthreads();
}

// argmax over the columns outside the active set of w = Aᵀ r (first index
// on ties).  With J ≤ kTaps, −kPadLo ≤ m0 ≤ 0 and at most one pattern per
// lane (Q ≤ 64; the standard configuration): lane p < Q takes the columns
// q·Q + p, its pattern's taps in registers and a window of kTaps residual rows
// rotating through registers (one broadcast LDS read per column); otherwise
// one column per lane and step.
template <int kJ>
__device__ __forceinline__ void w_argmax(const Lds& s, const DevCfg& c, int lane, double& best, int& bi) {
    best = -INFINITY;
    bi = 0x7fffffff;
    if (c.J <= kJ && c.m0 >= -kPadLo && c.m0 <= 0 && c.Q <= kLanes) {
        if (lane < c.Q) {
            const int p = lane;
            double t[kTaps], win[kTaps];
#pragma unroll
            for (int j = 0; j < kTaps; ++j) t[j] = j < c.J ? s.T[j * c.Q + p] : 0.0;
            const int nq = (c.n_pulse - 1 - p) / c.Q + 1;      // columns of this phase
            const double* rb = s.r + c.m0;                      // rb[q + j] = r[q + m0 + j]
#pragma unroll
            for (int j = 0; j < kTaps; ++j) win[j] = rb[j];
            for (int q0 = 0; q0 < nq; q0 += kTaps) {
#pragma unroll
                for (int u = 0; u < kTaps; ++u) {
                    double w = 0.0;
#pragma unroll
                    for (int j = 0; j < kJ; ++j) w = fma(t[j], win[(u + j) & (kTaps - 1)], w);
                    win[u] = rb[q0 + u + kTaps];
                    const int q = q0 + u, col = q * c.Q + p;
                    if (q < nq && w > best && !bit(s.inP, col) && !bit(s.excl, col)) {
                        best = w;
                        bi = col;
                    }
                }
            }
        }
    } else {
        for (int col = lane; col < c.n_pulse; col += kLanes) {
            if (bit(s.inP, col) || bit(s.excl, col)) continue;
            const int q = col / c.Q;
            const double w = adot(s, c, q, col - q * c.Q, s.r);
            if (w > best) { best = w; bi = col; }
        }
    }
    for (int o = 32; o; o >>= 1) {                 // max, first index on ties
        const double bo = __shfl_xor(best, o);
        const int io = __shfl_xor(bi, o);
        if (bo > best || (bo == best && io < bi)) { best = bo; bi = io; }
    }
}

struct ChanKArgs {
    DevCfg c;
    ChanArgs g;
};
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) ChanKArgs* const_kargs_ptr;
#else
typedef const ChanKArgs* const_kargs_ptr;       // host pass: never called
#endif
// the kernel argument through a pointer the compiler cannot match with the
// loads at kernel entry: what is read through it per channel is live only
// while that channel is processed
__device__ __forceinline__ const_kargs_ptr chan_kargs() {
#if defined(__HIP_DEVICE_COMPILE__)
    const_kargs_ptr k = (const_kargs_ptr)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return k;
#else
    return nullptr;
#endif
}

// Three waves per SIMD (168 VGPRs, no spills; twelve channels per CU with the
// default LDS split): the solve is a chain of dependent steps, and what hides
// their latency is the number of channels a CU runs at once.
template <int kJ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void channel_kernel(ChanKArgs ka) {
    extern __shared__ double smem[];
    const int lane0 = threadIdx.x;
    {
        const const_kargs_ptr kp = chan_kargs();
        for (int i = lane0; i < kp->c.J * kp->c.Q; i += kLanes) smem[i] = kp->c.table[i];
    }
    __syncthreads();

    for (;;) {
        const const_kargs_ptr kp = chan_kargs();
        auto& g = kp->g;
        int wi = 0;
        if (lane0 == 0) wi = atomicAdd(g.next, 1);
        wi = __shfl(wi, 0);
        if (wi >= *g.n_runs) break;
        // per channel: the configuration, the lane index and the seed words
        // afresh, so that nothing derived from them stays live across channels
        const DevCfg c = kp->c;
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        uint32_t k0 = (uint32_t)g.seed, k1 = (uint32_t)(g.seed >> 32);
        asm volatile("" : "+s"(k0), "+s"(k1));
        const int nw32 = (c.n_pulse + 31) / 32;
        Lds s;
        s.T = smem;
        s.buf = s.T + c.J * c.Q;
        s.K = c.lds_rows;
        s.triK = tri(s.K);
        s.glb = g.spill + (int64_t)blockIdx.x * g.spill_stride;
        const int nbuf = max(c.n_wave, s.triK);
        s.b = s.buf + nbuf;
        s.r = s.b + c.n_adc + kPadLo;
        s.pos_x = s.r + c.n_adc + kPadHi;
        s.pos_qp = (int2*)(s.pos_x + c.n_adc);
        s.inP = (uint32_t*)(s.pos_qp + c.n_adc);
        s.excl = s.inP + nw32;
        const int ci = g.order[wi];
        const uint32_t ch = g.unique[ci];
        if (ch == kNoChannel || ci >= g.capacity) continue;
        const int beg = g.offsets[ci], n = g.counts[ci];
        int status = 0;

        // ---- first pulse time, status of the hit draws
        double tmin = INFINITY;
        for (int p = lane; p < n; p += kLanes) {
            const int h = g.idx[beg + p];
            tmin = fmin(tmin, g.tq[2 * h]);
            status |= g.hflag[h];
        }
        tmin = wmin(tmin);
        for (int o = 32; o; o >>= 1) status |= __shfl_xor(status, o);
        const double t0 = tmin - c.lead;

        // ---- make_waveform: Σ_j q_j · pulse(t_i − t'_j) + noise
        for (int g0 = 0; g0 < c.n_wave; g0 += kLanes) {
            const int i = g0 + lane;
            const double ts = t0 + (double)i * c.dt_wave;
            // hits with t' in [t_first − p_hi, t_last − p_lo] of this group's samples
            const double lo = (t0 + (double)g0 * c.dt_wave) - c.p_hi;
            const double hi = (t0 + (double)(g0 + kLanes - 1) * c.dt_wave) - c.p_lo;
            const uint32_t olo = ordered32((float)lo), ohi = ordered32((float)hi);
            int a = 0, bnd = n;                            // first key ≥ lo
            while (a < bnd) {
                const int m = (a + bnd) >> 1;
                if ((uint32_t)g.key[beg + m] < olo) a = m + 1; else bnd = m;
            }
            int e = a, e1 = n;                             // first key > hi
            while (e < e1) {
                const int m = (e + e1) >> 1;
                if ((uint32_t)g.key[beg + m] <= ohi) e = m + 1; else e1 = m;
            }
            double acc = 0.0;
            for (int p = a; p < e; ++p) {
                const int h = g.idx[beg + p];
                const double x = ts - g.tq[2 * h];
                if (x >= c.p_lo && x <= c.p_hi) {
                    const double z = (x - c.p_mu) * c.p_inv_beta;
                    acc += c.p_norm * exp(-(z + exp(-z))) * g.tq[2 * h + 1];
                }
            }
            if (c.noise_amp > 0.0) {
                double n0, n1;
                normals(philox(make_uint4(kNoiseWord, ch, (uint32_t)(i >> 1), 0u), k0, k1), n0, n1);
                acc += ((i & 1) ? n1 : n0) * c.noise_amp;
            }
            if (i < c.n_wave) s.buf[i] = acc;
        }
        __syncthreads();

        // ---- low-pass y_i = b0 v_i + (b1 v_{i−1} − a1 y_{i−1}): chunk per lane, affine-map scan
        {
            const int st = lane * c.chunk;
            const int len = max(0, min(c.chunk, c.n_wave - st));
            const double xprev = (st > 0 && st <= c.n_wave) ? s.buf[st - 1] : 0.0;
            double y = 0.0, xp = xprev, A = 1.0;
            for (int k = 0; k < len; ++k) {
                const double x = s.buf[st + k];
                y = c.b0 * x + (c.b1 * xp - c.a1 * y);
                xp = x;
                A *= -c.a1;
            }
            double B = y;
            for (int o = 1; o < kLanes; o <<= 1) {          // inclusive scan of (A, B) ∘ …
                const double Ao = __shfl_up(A, o), Bo = __shfl_up(B, o);
                if (lane >= o) {
                    B = A * Bo + B;
                    A = A * Ao;
                }
            }
            double Y = __shfl_up(B, 1);
            if (lane == 0) Y = 0.0;
            __syncthreads();                               // every lane has read its x_{st−1}
            y = Y;
            xp = xprev;
            for (int k = 0; k < len; ++k) {
                const double x = s.buf[st + k];
                y = c.b0 * x + (c.b1 * xp - c.a1 * y);
                xp = x;
                s.buf[st + k] = y;
            }
        }
        __syncthreads();

        // ---- resample by 1/R: digi_k = Σ_m h_m y_{R·k + half − m}
        {
            const int half = (c.n_taps - 1) / 2;
            for (int k = lane; k < c.n_adc; k += kLanes) {
                double acc = 0.0;
                for (int m = 0; m < c.n_taps; ++m) {
                    const int j = c.R * k + half - m;
                    if (j >= 0 && j < c.n_wave) acc += c.taps[m] * s.buf[j];
                }
                s.b[k] = acc;
                s.r[k] = acc;
                if (g.out.waveform) g.out.waveform[(int64_t)ci * c.n_adc + k] = acc;
            }
            for (int i = lane; i < 2 * nw32; i += kLanes) s.inP[i] = 0u;   // inP and excl
            for (int i = lane; i < kPadLo + kPadHi; i += kLanes)
                (i < kPadLo ? s.r[i - kPadLo] : s.r[c.n_adc + i - kPadLo]) = 0.0;
        }
        __syncthreads();

        // ---- fnnls
        Active a;
        int64_t iter = 0;
        int outer = 0;
        for (;;) {
            // lane-derived addresses are recomputed per iteration, not held
            // in registers across the solve
            asm volatile("" : "+v"(lane));
            double best;
            int bi;
            w_argmax<kJ>(s, c, lane, best, bi);
            if (!(best > c.tol) || iter >= c.max_iter || a.P >= c.n_pulse) break;
            if (++outer > c.max_outer) { status |= NTPMT_ST_ITER_CAP; break; }
            if (a.P >= c.n_adc) { status |= NTPMT_ST_RANK_CAP; break; }
            const double atb_col = adot(s, c, bi / c.Q, bi - (bi / c.Q) * c.Q, s.b);
            if (!chol_add(s, c, a, a.P, bi, atb_col, lane)) {
                if (lane == 0) s.excl[bi >> 5] |= 1u << (bi & 31);
                status |= NTPMT_ST_DEGENERATE;
                __syncthreads();
                continue;
            }
            setp(a.xA, a.xB, a.P, 0.0, lane);
            if (lane == 0) s.inP[bi >> 5] |= 1u << (bi & 31);
            a.P += 1;
            __syncthreads();
            back_solve(s, a, lane);
            for (;;) {
                const bool vA = lane < a.P, vB = lane + kLanes < a.P;
                if (!wany((vA && a.sA <= c.tol) || (vB && a.sB <= c.tol))) break;
                ++iter;
                double al = INFINITY;
                if (vA && a.sA <= c.tol) al = fmin(al, a.xA / (a.xA - a.sA));
                if (vB && a.sB <= c.tol) al = fmin(al, a.xB / (a.xB - a.sB));
                al = wmin(al);
                if (!(al < INFINITY)) al = 0.0;            // only 0/0 quotients: keep x, drop x ≤ tol
                if (vA) a.xA += al * (a.sA - a.xA);
                if (vB) a.xB += al * (a.sB - a.xB);
                // drop positions with x ≤ tol (highest first, so lower positions keep their index)
                const bool kA = vA && a.xA > c.tol, kB = vB && a.xB > c.tol;
                uint64_t dA = __ballot(vA && !kA), dB = __ballot(vB && !kB);
                if (!dA && !dB) { status |= NTPMT_ST_ITER_CAP; break; }    // no progress: cannot happen exactly
                while (dA | dB) {
                    const int r = dB ? kLanes + 63 - __builtin_clzll(dB) : 63 - __builtin_clzll(dA);
                    const int col = getpi(a.colA, a.colB, r);
                    if (lane == 0) s.inP[col >> 5] &= ~(1u << (col & 31));
                    chol_delete(s, a, r, lane);
                    if (r < kLanes) dA &= ~(1ull << r); else dB &= ~(1ull << (r - kLanes));
                }
                back_solve(s, a, lane);
            }
            // x = s
            a.xA = lane < a.P ? a.sA : 0.0;
            a.xB = lane + kLanes < a.P ? a.sB : 0.0;
            residual(s, c, a, lane);
        }

        // ---- pulses with charge > min_charge, in time (column) order
        {
            const bool oA = lane < a.P && a.xA > c.min_charge, oB = lane + kLanes < a.P && a.xB > c.min_charge;
            int rA = 0, rB = 0;
            for (int j = 0; j < a.P; ++j) {
                const int col = getpi(a.colA, a.colB, j);
                const double xj = getp(a.xA, a.xB, j);
                if (xj > c.min_charge) {
                    rA += col < a.colA;
                    rB += col < a.colB;
                }
            }
            const int64_t base = (int64_t)ci * c.n_adc;
            if (oA) {
                g.out.time[base + rA] = t0 + (double)a.colA * c.res;
                g.out.charge[base + rA] = a.xA;
            }
            if (oB) {
                g.out.time[base + rB] = t0 + (double)a.colB * c.res;
                g.out.charge[base + rB] = a.xB;
            }
            const int npulse = __popcll(__ballot(oA)) + __popcll(__ballot(oB));
            if (lane == 0) {
                g.out.module_id[ci] = (uint16_t)(ch / 17u);
                g.out.pmt_id[ci] = (uint8_t)(ch % 17u);
                g.out.n_hits[ci] = n;
                g.out.n_pulses[ci] = npulse;
                g.out.status[ci] = status;
                g.out.t0[ci] = t0;
            }
        }
        __syncthreads();
    }
}

thread_local char g_err[256];

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace ntpmt

using namespace ntpmt;

struct ntpmt_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevCfg cfg{};
    double* d_taps = nullptr;
    double* d_table = nullptr;
    size_t lds_bytes = 0;
    int blocks = 0;
    const void* kernel = nullptr;       // channel_kernel instance for the configuration's pattern length
    double* d_spill = nullptr;          // factor rows beyond the LDS ones, per workgroup
    int64_t spill_stride = 0;
    char* d_ws = nullptr;
    size_t ws_cap = 0;
    hipEvent_t last = nullptr;
};

namespace {

int32_t fail(hipError_t e, const char* what) {
    snprintf(ntpmt::g_err, sizeof ntpmt::g_err, "%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? NT_ERR_OOM : NT_ERR_DEVICE;
}

#define PMT_TRY(x)                                      \
    do {                                                \
        hipError_t _e = (x);                            \
        if (_e != hipSuccess) return fail(_e, #x);      \
    } while (0)

bool near_int(double v, int64_t* out) {
    const double r = std::nearbyint(v);
    if (!(std::fabs(v - r) <= 1e-9 * std::fmax(1.0, std::fabs(v))) || r < 1.0) return false;
    *out = (int64_t)r;
    return true;
}

int32_t arg_error(const char* msg) {
    snprintf(ntpmt::g_err, sizeof ntpmt::g_err, "ntpmt: %s", msg);
    return NT_ERR_ARG;
}

// linear interpolation of the filtered pulse, 0 outside (the oracle's formula)
double filt_pulse_at(const ntpmt_config* c, double s) {
    const int n = c->n_filt_pulse;
    const double u = (s - c->filt_pulse_t0) / c->filt_pulse_dt;
    if (!(u >= 0.0 && u <= (double)(n - 1))) return 0.0;
    const int i = (int)std::floor(u);
    if (i >= n - 1) return c->filt_pulse[n - 1];
    const double f = u - i;
    return c->filt_pulse[i] * (1.0 - f) + c->filt_pulse[i + 1] * f;
}

struct WsLayout {
    size_t key_a, key_b, idx_a, idx_b, tq, hflag, ch, unique, counts, offsets, n_runs, next, okey_a, okey_b, orun_a,
        orun_b, cub, cub_bytes, total;
};

WsLayout ws_layout(int64_t n) {
    WsLayout L{};
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o += align256(bytes); return at; };
    L.key_a = take(8 * nn);
    L.key_b = take(8 * nn);
    L.idx_a = take(4 * nn);
    L.idx_b = take(4 * nn);
    L.tq = take(16 * nn);
    L.hflag = take(nn);
    L.ch = take(4 * nn);
    L.unique = take(4 * nn);
    L.counts = take(4 * nn);
    L.offsets = take(4 * nn);
    L.n_runs = take(8);
    L.next = take(8);
    L.okey_a = take(4 * nn);
    L.okey_b = take(4 * nn);
    L.orun_a = take(4 * nn);
    L.orun_b = take(4 * nn);
    size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t1, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)nn, 0, 53);
    (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, t2, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (int32_t*)nullptr, (int32_t*)nullptr, (int)nn);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, t3, (const int32_t*)nullptr, (int32_t*)nullptr, (int)nn);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t4, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)nn, 0, 32);
    L.cub_bytes = std::max(std::max(t1, t2), std::max(t3, t4));
    L.cub = take(L.cub_bytes);
    L.total = o;
    return L;
}

}  // namespace

extern "C" {

int32_t ntpmt_create(int32_t device, const ntpmt_config* c, ntpmt_ctx** out) {
    if (!out) return arg_error("null output handle");
    *out = nullptr;
    if (!c) return arg_error("null configuration");
    const double vals[] = {c->tt_shape, c->tt_scale, c->tt_subtract, c->spe_expon_weight, c->spe_expon_scale,
                           c->spe_norm_mu, c->spe_norm_sigma, c->spe_trunc_low, c->pulse_mu, c->pulse_beta,
                           c->pulse_lo, c->pulse_hi, c->pulse_amplitude, c->noise_amp, c->sampling_freq,
                           c->adc_freq, c->unf_pulse_res, c->time_range, c->lead_time, c->min_charge, c->iir_b0,
                           c->iir_b1, c->iir_a1, c->nnls_tol, c->filt_pulse_t0, c->filt_pulse_dt};
    for (double v : vals)
        if (!std::isfinite(v)) return arg_error("configuration values must be finite");
    if (!(c->tt_shape >= 1.0) || !(c->tt_scale > 0.0)) return arg_error("transit-time Gamma needs shape >= 1, scale > 0");
    if (!(c->spe_expon_weight >= 0.0 && c->spe_expon_weight <= 1.0) || !(c->spe_expon_scale > 0.0) ||
        !(c->spe_norm_sigma > 0.0))
        return arg_error("SPE weight in [0, 1], positive exponential scale and sigma");
    if (!(c->pulse_beta > 0.0) || !(c->pulse_hi > c->pulse_lo)) return arg_error("pulse beta > 0 and lo < hi");
    const double F = std::exp(-std::exp(-(c->pulse_hi - c->pulse_mu) / c->pulse_beta)) -
                     std::exp(-std::exp(-(c->pulse_lo - c->pulse_mu) / c->pulse_beta));
    if (!(F > 0.0)) return arg_error("the pulse's truncation interval holds no probability");
    if (!(c->noise_amp >= 0.0) || !(c->sampling_freq > 0.0) || !(c->adc_freq > 0.0) || !(c->unf_pulse_res > 0.0) ||
        !(c->time_range > 0.0) || !(c->nnls_tol >= 0.0) || !(c->filt_pulse_dt > 0.0) || c->nnls_max_iter < 0)
        return arg_error("positive frequencies, resolution, time range, table step; tolerance >= 0");
    int64_t R = 0, Q = 0;
    if (!near_int(c->sampling_freq / c->adc_freq, &R)) return arg_error("sampling_freq/adc_freq must be an integer");
    if (!near_int((1.0 / c->adc_freq) / c->unf_pulse_res, &Q) || Q > 1024)
        return arg_error("(1/adc_freq)/unf_pulse_res must be an integer <= 1024");
    const int64_t n_wave = (int64_t)std::floor(c->time_range * c->sampling_freq + 1e-9) + 1;
    const int64_t n_adc = (int64_t)std::floor(c->time_range * c->adc_freq + 1e-9) + 1;
    const int64_t n_pulse = (int64_t)std::floor(c->time_range / c->unf_pulse_res + 1e-9) + 1;
    if (n_wave > NTPMT_MAX_WAVE || n_adc > NTPMT_MAX_ADC || n_pulse > NTPMT_MAX_PULSE)
        return arg_error("waveform, digitised or unfolding grid longer than the NTPMT_MAX_* limits");
    if (n_adc != (n_wave + R - 1) / R) return arg_error("the digitised grid and the resampled length differ");
    if (!c->decim_taps || c->n_decim_taps < 1 || c->n_decim_taps > NTPMT_MAX_TAPS || c->n_decim_taps % 2 == 0)
        return arg_error("decim_taps: an odd number of taps up to NTPMT_MAX_TAPS");
    if (!c->filt_pulse || c->n_filt_pulse < 2 || c->n_filt_pulse > NTPMT_MAX_FILT_PULSE)
        return arg_error("filt_pulse: 2 to NTPMT_MAX_FILT_PULSE samples");
    for (int i = 0; i < c->n_decim_taps; ++i)
        if (!std::isfinite(c->decim_taps[i])) return arg_error("decim_taps must be finite");
    for (int i = 0; i < c->n_filt_pulse; ++i)
        if (!std::isfinite(c->filt_pulse[i])) return arg_error("filt_pulse must be finite");

    // pattern table: T[j][p] = filtered pulse at dt_adc·(j + m0) − res·p
    const double dt_adc = 1.0 / c->adc_freq;
    const double fp_end = c->filt_pulse_t0 + (c->n_filt_pulse - 1) * c->filt_pulse_dt;
    const int m0 = (int)std::ceil(c->filt_pulse_t0 / dt_adc - 1e-9);
    const int m1 = (int)std::floor((fp_end + c->unf_pulse_res * (double)(Q - 1)) / dt_adc + 1e-9);
    const int J = m1 - m0 + 1;
    if (J < 1 || J > 256) return arg_error("the filtered pulse spans more than 256 digitiser samples");
    std::vector<double> full((size_t)J * Q);
    for (int j = 0; j < J; ++j)
        for (int p = 0; p < Q; ++p)
            full[(size_t)j * Q + p] = filt_pulse_at(c, dt_adc * (double)(j + m0) - c->unf_pulse_res * (double)p);
    // rows that are zero for every pattern contribute nothing: drop them
    auto zero_row = [&](int j) {
        for (int p = 0; p < Q; ++p)
            if (full[(size_t)j * Q + p] != 0.0) return false;
        return true;
    };
    int j_lo = 0, j_hi = J - 1;
    while (j_lo < j_hi && zero_row(j_lo)) ++j_lo;
    while (j_hi > j_lo && zero_row(j_hi)) --j_hi;
    const int m0_trim = m0 + j_lo, J_trim = j_hi - j_lo + 1;
    std::vector<double> table(full.begin() + (size_t)j_lo * Q, full.begin() + (size_t)(j_hi + 1) * Q);

    ntpmt_ctx* x = new ntpmt_ctx();
    x->device = device;
    auto cleanup = [&](int32_t st) {
        ntpmt_destroy(x);
        return st;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return cleanup(fail(e, "hipSetDevice"));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return cleanup(fail(e, "hipGetDeviceProperties"));
    x->num_cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup(fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&x->last, hipEventDisableTiming)) != hipSuccess)
        return cleanup(fail(e, "hipEventCreate"));
    if ((e = hipMalloc(&x->d_taps, sizeof(double) * c->n_decim_taps)) != hipSuccess) return cleanup(fail(e, "hipMalloc"));
    if ((e = hipMalloc(&x->d_table, sizeof(double) * table.size())) != hipSuccess) return cleanup(fail(e, "hipMalloc"));
    if ((e = hipMemcpy(x->d_taps, c->decim_taps, sizeof(double) * c->n_decim_taps, hipMemcpyHostToDevice)) !=
        hipSuccess)
        return cleanup(fail(e, "hipMemcpy"));
    if ((e = hipMemcpy(x->d_table, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(fail(e, "hipMemcpy"));

    DevCfg& d = x->cfg;
    d.tt_d = c->tt_shape - 1.0 / 3.0;
    d.tt_c = 1.0 / std::sqrt(9.0 * d.tt_d);
    d.tt_scale = c->tt_scale;
    d.tt_subtract = c->tt_subtract;
    d.spe_w = c->spe_expon_weight;
    d.spe_scale = c->spe_expon_scale;
    d.spe_mu = c->spe_norm_mu;
    d.spe_sigma = c->spe_norm_sigma;
    d.spe_low = c->spe_trunc_low;
    d.p_mu = c->pulse_mu;
    d.p_inv_beta = 1.0 / c->pulse_beta;
    d.p_lo = c->pulse_lo;
    d.p_hi = c->pulse_hi;
    d.p_norm = c->pulse_amplitude / c->pulse_beta / F;
    d.noise_amp = c->noise_amp;
    d.dt_wave = 1.0 / c->sampling_freq;
    d.dt_adc = dt_adc;
    d.res = c->unf_pulse_res;
    d.lead = c->lead_time;
    d.min_charge = c->min_charge;
    d.b0 = c->iir_b0;
    d.b1 = c->iir_b1;
    d.a1 = c->iir_a1;
    d.tol = c->nnls_tol;
    d.max_iter = c->nnls_max_iter;
    d.n_wave = (int)n_wave;
    d.n_adc = (int)n_adc;
    d.n_pulse = (int)n_pulse;
    d.R = (int)R;
    d.Q = (int)Q;
    d.J = J_trim;
    d.m0 = m0_trim;
    d.n_taps = c->n_decim_taps;
    d.chunk = (int)((n_wave + kLanes - 1) / kLanes);
    d.max_outer = 8 * (int)n_adc + 64;
    d.taps = x->d_taps;
    d.table = x->d_table;
    const size_t nw32 = (size_t)(n_pulse + 31) / 32;
    if (c->lds_factor_rows < 0) return cleanup(arg_error("lds_factor_rows >= 0"));
    // LDS per channel with the factor's first K rows in it; by default the
    // largest K that leaves room for kChannelsPerCU channels on a CU, but not
    // fewer rows than fit in the waveform's space (they cost no LDS)
    auto lds_for = [&](size_t K) {
        return sizeof(double) * ((size_t)J_trim * Q + std::max<size_t>(n_wave, K * (K + 1) / 2) + 2 * (size_t)n_adc +
                                 kPadLo + kPadHi + 2 * (size_t)n_adc) +
               sizeof(uint32_t) * 2 * nw32;
    };
    size_t K = (size_t)n_adc;
    if (c->lds_factor_rows > 0) {
        K = std::min<size_t>(K, (size_t)c->lds_factor_rows);
    } else {
        const size_t budget = (kLdsPerCU / kChannelsPerCU) & ~(size_t)511;
        while (K > 1 && lds_for(K) > budget && K * (K + 1) / 2 > (size_t)n_wave) --K;
    }
    d.lds_rows = (int)K;
    x->lds_bytes = lds_for(K);
    if (x->lds_bytes > 64 * 1024) return cleanup(arg_error("configuration needs more than 64 KB of LDS per channel"));
    x->kernel = J_trim <= 14 ? (const void*)channel_kernel<14> : (const void*)channel_kernel<16>;
    int per_cu = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, x->kernel, 64, x->lds_bytes)) != hipSuccess)
        return cleanup(fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor"));
    x->blocks = std::max(1, per_cu) * x->num_cus;
    x->spill_stride = (int64_t)(n_adc * (n_adc + 1) / 2) - (int64_t)(K * (K + 1) / 2);
    if (x->spill_stride > 0 &&
        (e = hipMalloc(&x->d_spill, sizeof(double) * (size_t)x->spill_stride * x->blocks)) != hipSuccess)
        return cleanup(fail(e, "hipMalloc"));
    *out = x;
    return NT_OK;
}

void ntpmt_destroy(ntpmt_ctx* x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    if (x->last) (void)hipEventSynchronize(x->last);
    if (x->d_ws) (void)hipFree(x->d_ws);
    if (x->d_taps) (void)hipFree(x->d_taps);
    if (x->d_table) (void)hipFree(x->d_table);
    if (x->d_spill) (void)hipFree(x->d_spill);
    if (x->last) (void)hipEventDestroy(x->last);
    if (x->stream) (void)hipStreamDestroy(x->stream);
    delete x;
}

int32_t ntpmt_sizes(const ntpmt_ctx* x, int32_t* n_wave, int32_t* n_adc, int32_t* n_pulse) {
    if (!x) return arg_error("null context");
    if (n_wave) *n_wave = x->cfg.n_wave;
    if (n_adc) *n_adc = x->cfg.n_adc;
    if (n_pulse) *n_pulse = x->cfg.n_pulse;
    return NT_OK;
}

int32_t ntpmt_make_reco_pulses(ntpmt_ctx* x, const nt_hits* h, const uint8_t* d_select, int64_t n, int32_t time_column,
                               uint64_t seed, int64_t capacity, const ntpmt_pulses* out, int64_t* d_n_channels,
                               void* stream) {
    if (!x || !h || !out) return arg_error("null context, hits or output");
    if (n < 0 || n > INT32_MAX) return arg_error("0 <= n_hits <= 2^31 - 1");
    if (capacity < 0) return arg_error("capacity >= 0");
    if (time_column != 0 && time_column != 1) return arg_error("time_column 0 (time) or 1 (tres)");
    const float* tcol = time_column == 0 ? h->time : h->tres;
    if (n > 0 && (!tcol || !h->module_id || !h->pmt_id || !h->source_index || !h->photon_id))
        return arg_error("the time column, module_id, pmt_id, source_index and photon_id are required");
    if (capacity > 0 && (!out->module_id || !out->pmt_id || !out->n_hits || !out->n_pulses || !out->status ||
                         !out->t0 || !out->time || !out->charge))
        return arg_error("every output column except waveform is required");
    std::lock_guard<std::mutex> lk(x->mu);
    PMT_TRY(hipSetDevice(x->device));
    hipStream_t s = stream ? (hipStream_t)stream : x->stream;
    // the workspace is the context's: the previous call's kernels finish first
    PMT_TRY(hipEventSynchronize(x->last));
    if (n == 0) {
        zero_count_kernel<<<1, 1, 0, s>>>(d_n_channels);
        PMT_TRY(hipGetLastError());
        PMT_TRY(hipEventRecord(x->last, s));
        return NT_OK;
    }
    const WsLayout L = ws_layout(n);
    if (L.total > x->ws_cap) {
        if (x->d_ws) (void)hipFree(x->d_ws);
        x->d_ws = nullptr;
        x->ws_cap = 0;
        PMT_TRY(hipMalloc(&x->d_ws, L.total));
        x->ws_cap = L.total;
    }
    char* b = x->d_ws;
    auto* key_a = (uint64_t*)(b + L.key_a);
    auto* key_b = (uint64_t*)(b + L.key_b);
    auto* idx_a = (int32_t*)(b + L.idx_a);
    auto* idx_b = (int32_t*)(b + L.idx_b);
    auto* tq = (double*)(b + L.tq);
    auto* hflag = (uint8_t*)(b + L.hflag);
    auto* ch = (uint32_t*)(b + L.ch);
    auto* uniq = (uint32_t*)(b + L.unique);
    auto* counts = (int32_t*)(b + L.counts);
    auto* offs = (int32_t*)(b + L.offsets);
    auto* n_runs = (int32_t*)(b + L.n_runs);
    auto* next = (int32_t*)(b + L.next);
    size_t cub = L.cub_bytes;

    HitIn hi{tcol, h->module_id, h->pmt_id, h->source_index, h->photon_id, d_select, n, seed};
    const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)x->num_cus * 16);
    hit_draws_kernel<<<grid, 256, 0, s>>>(hi, x->cfg, key_a, idx_a, tq, hflag);
    PMT_TRY(hipGetLastError());
    PMT_TRY(hipcub::DeviceRadixSort::SortPairs(b + L.cub, cub, key_a, key_b, idx_a, idx_b, (int)n, 0, 53, s));
    channel_ids_kernel<<<grid, 256, 0, s>>>(key_b, n, ch);
    PMT_TRY(hipGetLastError());
    cub = L.cub_bytes;
    PMT_TRY(hipcub::DeviceRunLengthEncode::Encode(b + L.cub, cub, ch, uniq, counts, n_runs, (int)n, s));
    cub = L.cub_bytes;
    PMT_TRY(hipcub::DeviceScan::ExclusiveSum(b + L.cub, cub, counts, offs, (int)n, s));
    finish_count_kernel<<<1, 1, 0, s>>>(n_runs, uniq, d_n_channels, next);
    PMT_TRY(hipGetLastError());
    auto* okey_a = (uint32_t*)(b + L.okey_a);
    auto* okey_b = (uint32_t*)(b + L.okey_b);
    auto* orun_a = (int32_t*)(b + L.orun_a);
    auto* orun_b = (int32_t*)(b + L.orun_b);
    order_keys_kernel<<<grid, 256, 0, s>>>(n_runs, uniq, counts, n, okey_a, orun_a);
    PMT_TRY(hipGetLastError());
    cub = L.cub_bytes;
    PMT_TRY(hipcub::DeviceRadixSort::SortPairs(b + L.cub, cub, okey_a, okey_b, orun_a, orun_b, (int)n, 0, 32, s));
    ChanArgs a{x->d_spill, x->spill_stride, key_b, idx_b, tq, hflag, uniq, counts, offs, n_runs, orun_b, next,
               capacity, seed, *out};
    DevCfg cfg = x->cfg;
    ChanKArgs ka{cfg, a};
    void* kargs[] = {&ka};
    PMT_TRY(hipLaunchKernel(x->kernel, dim3(x->blocks), dim3(64), kargs, x->lds_bytes, s));
    PMT_TRY(hipGetLastError());
    PMT_TRY(hipEventRecord(x->last, s));
    return NT_OK;
}

const char* ntpmt_last_error(void) { return ntpmt::g_err; }

int32_t ntpmt_synchronize(ntpmt_ctx* x) {
    if (!x) return arg_error("null context");
    PMT_TRY(hipSetDevice(x->device));
    PMT_TRY(hipEventSynchronize(x->last));
    return NT_OK;
}

}  // extern "C"
