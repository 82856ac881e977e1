// ntflow.hip — conditional rational-quadratic-spline flow surrogate for
// gfx950: fused forward + loss + backward training kernel on the matrix cores
// (bfloat16 MFMA, FP32 accumulation), deterministic gradient reduction fused
// with the Adam update.  C ABI in include/ntflow.h.
//
// Reference (paths relative to the reference repository root):
//   RQNormFlow MLP                 src/surrogate_models/extended_cascade_model.jl:53-125
//   constrain_spline_params        src/surrogate_models/rq_spline_flow.jl:127-216
//   safe_quadratic_root            rq_spline_flow.jl:215-254
//   inv_rqs_univariate             rq_spline_flow.jl:334-389
//   eval_transformed_normal_logpdf rq_spline_flow.jl:396-428
//   log_likelihood_with_poisson    extended_cascade_model.jl:132-139
//   train_model! / Adam            extended_cascade_model.jl:237-295
//
// Work decomposition: a workgroup (4 wavefronts) owns 64 rows.  Each
// wavefront runs the three matrix products of its 16 rows on 16×16×16 MFMA
// tiles with the weights resident in LDS, the per-row flow loss and its
// gradient with 16 lanes, and the two backward products; the workgroup then
// forms the weight gradients of its 64 rows (K = 64 MFMA reductions) and
// writes one partial gradient record.  The reduction kernel sums the records
// in a fixed order (bitwise reproducible) and applies Adam.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <new>

#include "../../include/ntflow.h"

namespace {

constexpr int kIn = NTF_N_IN, kH = NTF_HIDDEN, kP = NTF_P_PAD, kK = NTF_BINS, kL = NTF_LAYERS;
constexpr int kNP = NTF_N_PARAMS;
constexpr int kSpl = 3 * kK + 1;
constexpr int kRows = 64;                 // rows per workgroup
constexpr int kThreads = 256;             // 4 wavefronts
constexpr int kPStride = 81;              // FP32 row stride in LDS (odd: conflict-free per-lane rows)
constexpr float kRangeMin = NTF_RANGE_MIN, kRangeMax = NTF_RANGE_MAX;
constexpr float kMinBin = 1e-4f, kMinSlope = 1e-4f;
constexpr float kSlopeOffset = 0.5411666523385311f;   // log(exp(1 − 1e-4) − 1), rq_spline_flow.jl:151
constexpr float kLog2Pi = 1.8378770664093453f;

constexpr int oW1 = 0, ob1 = oW1 + kIn * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kP, kNW = ob3 + kP;
static_assert(kNW == NTF_N_WEIGHTS, "weight layout");
static_assert(kNP <= kP, "outputs");

// partial gradient record of one workgroup: dW1 | dW2 | dW3 | 4 × (db1, db2, db3)
constexpr int pW1 = 0, pW2 = pW1 + kIn * kH, pW3 = pW2 + kH * kH, pBias = pW3 + kH * kP;
constexpr int kBiasW = kH + kH + kP;
constexpr int kPartW = pBias + 4 * kBiasW;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 → bfloat16, round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct __align__(16) Smem {
    uint16_t w1t[kH][kIn];      // W1ᵀ [hidden][feature]
    uint16_t w2[kH][kH];        // W2 [in][out]
    uint16_t w3[kH][kP];        // W3 [in][out]
    uint16_t x[kRows][kIn];
    uint16_t h1[kRows][kH];
    uint16_t h2[kRows][kH];
    uint16_t dp[kRows][kP];
    uint16_t dh2[kRows][kH];
    uint16_t dh1[kRows][kH];
    float p[kRows][kPStride];   // network outputs, then the FP32 output gradient
    float b[kH + kH + kP];      // b1 | b2 | b3
    float loss[kRows];
};

// 16×16×16 bf16 MFMA operand layout (wavefront of 64): lane = 16·g + i holds
//   A: A[i][4g .. 4g+3]      B: B[4g .. 4g+3][i]      C: C[4g + j][i], j = 0..3
// A fragment, row-major source M[.][ld]: rows r0+i, columns k0+4g..
template <int LD>
__device__ __forceinline__ bf16x4 frag_rows(const uint16_t (*m)[LD], int r0, int k0, int i, int g) {
    return *reinterpret_cast<const bf16x4*>(&m[r0 + i][k0 + 4 * g]);
}
// fragment gathered down a column: M[k0+4g+j][c0+i], j = 0..3
template <int LD>
__device__ __forceinline__ bf16x4 frag_cols(const uint16_t (*m)[LD], int k0, int c0, int i, int g) {
    bf16x4 v;
    v[0] = (short)m[k0 + 4 * g + 0][c0 + i];
    v[1] = (short)m[k0 + 4 * g + 1][c0 + i];
    v[2] = (short)m[k0 + 4 * g + 2][c0 + i];
    v[3] = (short)m[k0 + 4 * g + 3][c0 + i];
    return v;
}

// ---------------------------------------------------------------------------
// Spline of one layer (FP32)
// ---------------------------------------------------------------------------
struct Knots {
    float xp[kK + 1], yp[kK + 1], d[kK + 1];
    float sw[kK], sh[kK];       // softmax of the width / height parameters
    float sg[kK + 1];           // σ(u + offset): derivative of the slope softplus
};

__device__ __forceinline__ void softmax8(const float* u, float* s) {
    float m = u[0];
#pragma unroll
    for (int i = 1; i < kK; ++i) m = fmaxf(m, u[i]);
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < kK; ++i) {
        s[i] = expf(u[i] - m);
        sum += s[i];
    }
    const float inv = 1.0f / sum;
#pragma unroll
    for (int i = 0; i < kK; ++i) s[i] *= inv;
}

// constrain_spline_params, rq_spline_flow.jl:170-216
__device__ __forceinline__ void make_knots(const float* u, Knots& k) {
    softmax8(u, k.sw);
    softmax8(u + kK, k.sh);
    const float span = (kRangeMax - kRangeMin) - kK * kMinBin;
    float cx = kRangeMin, cy = kRangeMin;
    k.xp[0] = kRangeMin;
    k.yp[0] = kRangeMin;
#pragma unroll
    for (int i = 1; i < kK; ++i) {
        cx += k.sw[i - 1] * span + kMinBin;
        cy += k.sh[i - 1] * span + kMinBin;
        k.xp[i] = cx;
        k.yp[i] = cy;
    }
    k.xp[kK] = kRangeMax;
    k.yp[kK] = kRangeMax;
#pragma unroll
    for (int i = 0; i <= kK; ++i) {
        const float a = u[2 * kK + i] + kSlopeOffset;
        k.d[i] = (a > 30.0f ? a : log1pf(expf(a))) + kMinSlope;
        k.sg[i] = 1.0f / (1.0f + expf(-a));
    }
}

// bin j with yp[j] <= v < yp[j+1]; the first bin when none holds v
__device__ __forceinline__ int find_bin(const Knots& k, float v) {
    int j = 0;
#pragma unroll
    for (int i = kK - 1; i >= 0; --i)
        if (v >= k.yp[i] && v < k.yp[i + 1]) j = i;
    return j;
}

__device__ __forceinline__ float safe_root(float a, float b, float c) {
    const float disc = b * b - 4.0f * a * c;
    const float s = disc > 0.0f ? sqrtf(disc) : 0.0f;
    return b >= 0.0f ? (2.0f * c) / (-b - s) : (-b + s) / (2.0f * a);
}

// inv_rqs_univariate (rq_spline_flow.jl:334-389) at v; returns x, ld = log|dx/dv|.
// With grad: accumulates d/du of (gx·x + gld·ld) into gu and returns d/dv.
template <bool kGrad>
__device__ float inv_layer(const float* u, float v, float& out, float gx, float gld, float* gu) {
    Knots k;
    make_knots(u, k);
    if (v <= k.yp[0] || v >= k.yp[kK]) {          // linear tails
        const int e = v <= k.yp[0] ? 0 : kK;
        const float de = k.d[e], edge = k.yp[e];
        if (!kGrad) {
            out = -logf(de);
            return (v - edge) / de + k.xp[e];
        }
        const float gde = -gx * (v - edge) / (de * de) - gld / de;
        gu[2 * kK + e] += gde * k.sg[e];
        return gx / de;
    }
    const int j = find_bin(k, v);
    const float x0 = k.xp[j], x1 = k.xp[j + 1], y0 = k.yp[j], y1 = k.yp[j + 1];
    const float d0 = k.d[j], d1 = k.d[j + 1];
    const float w = x1 - x0, h = y1 - y0, s = h / w;
    const float wv = fminf(fmaxf((v - y0) / h, 0.0f), 1.0f);
    const float st = d1 + d0 - 2.0f * s;
    const float qc = -s * wv;
    const float qb = d0 - st * wv;
    const float qa = s - qb;
    const float zr = safe_root(qa, qb, qc);
    const float z = fminf(fmaxf(zr, 0.0f), 1.0f);
    const float sq_z = z * z, z1mz = z - sq_z, omz = 1.0f - z, sq_1mz = omz * omz;
    const float Q = d1 * sq_z + 2.0f * s * z1mz + d0 * sq_1mz;
    const float D = s + st * z1mz;
    if (!kGrad) {
        out = -2.0f * logf(s) - logf(Q) + 2.0f * logf(D);
        return w * z + x0;
    }
    const float N = s * sq_z + d0 * z1mz;
    const float R = N / D;
    const float dDz = st * (1.0f - 2.0f * z);
    const float dQz = 2.0f * d1 * z + 2.0f * s * (1.0f - 2.0f * z) - 2.0f * d0 * omz;
    const float dNz = 2.0f * s * z + d0 * (1.0f - 2.0f * z);
    // ld = −2 log s − log Q + 2 log D
    const float ldz = -dQz / Q + 2.0f * dDz / D;
    const float lds = -2.0f / s - 2.0f * z1mz / Q + 2.0f * (1.0f - 2.0f * z1mz) / D;
    const float ldd0 = -sq_1mz / Q + 2.0f * z1mz / D;
    const float ldd1 = -sq_z / Q + 2.0f * z1mz / D;
    const float gz = gx * w + gld * ldz;
    float gw = gx * z, gs = gld * lds, gd0 = gld * ldd0, gd1 = gld * ldd1;
    float gv = 0.0f, gy0 = 0.0f, gh = 0.0f;
    if (zr == z) {
        // z solves h·R(z; s, d0, d1) = v − y0 (the forward spline in the bin)
        const float Rp = (dNz - R * dDz) / D;
        const float kappa = gz / (h * Rp);
        const float Rs = (sq_z - R * (1.0f - 2.0f * z1mz)) / D;
        const float Rd0 = (z1mz - R * z1mz) / D;
        const float Rd1 = (-R * z1mz) / D;
        gv = kappa;
        gy0 = -kappa;
        gh = -kappa * R;
        gs -= kappa * h * Rs;
        gd0 -= kappa * h * Rd0;
        gd1 -= kappa * h * Rd1;
    }
    gh += gs / w;                      // s = h / w
    gw -= gs * s / w;
    // knots: x0 = xp[j], w = xp[j+1] − xp[j] (likewise y); interior knots only
    float gxp[kK + 1], gyp[kK + 1];
#pragma unroll
    for (int i = 0; i <= kK; ++i) gxp[i] = gyp[i] = 0.0f;
    gxp[j] += gx - gw;
    gxp[j + 1] += gw;
    gyp[j] += gy0 - gh;
    gyp[j + 1] += gh;
    gu[2 * kK + j] += gd0 * k.sg[j];
    gu[2 * kK + j + 1] += gd1 * k.sg[j + 1];
    // xp[i] = lo + Σ_{m<i} (sw[m]·span + min): d xp[i] / d sw[m] = span for m < i < K
    const float span = (kRangeMax - kRangeMin) - kK * kMinBin;
    float gsw[kK], gsh[kK];
    float accx = 0.0f, accy = 0.0f;
#pragma unroll
    for (int m = kK - 1; m >= 0; --m) {
        if (m + 1 <= kK - 1) {
            accx += gxp[m + 1];
            accy += gyp[m + 1];
        }
        gsw[m] = accx * span;
        gsh[m] = accy * span;
    }
    float dotw = 0.0f, doth = 0.0f;
#pragma unroll
    for (int m = 0; m < kK; ++m) {
        dotw += k.sw[m] * gsw[m];
        doth += k.sh[m] * gsh[m];
    }
#pragma unroll
    for (int m = 0; m < kK; ++m) {
        gu[m] += k.sw[m] * (gsw[m] - dotw);
        gu[kK + m] += k.sh[m] * (gsh[m] - doth);
    }
    return gv;
}

// Row loss ω·(−[time·logpdf(t) + count·(n log λ − λ − lgamma(n+1))]) from the
// FP32 outputs p (LDS row), with log λ = p_λ + offset; with kGrad, writes
// scale·d loss / d p into g.
template <bool kGrad>
__device__ float row_loss(const float* p, float t, float cnt, int flags, float offset, float weight, float* g,
                          float scale) {
    const bool ft = (flags & NTF_ROW_TIME) != 0, fc = (flags & NTF_ROW_COUNT) != 0;
    const float* tail = p + kL * kSpl;
    const float loglam = tail[2] + offset;
    float loss = 0.0f;
    if (kGrad) {
#pragma unroll
        for (int c = 0; c < kP; ++c) g[c] = 0.0f;
    }
    if (ft) {
        float z[kL + 1];
        z[0] = t;
        float ldsum = 0.0f;
#pragma unroll
        for (int l = 0; l < kL; ++l) {
            float ld;
            z[l + 1] = inv_layer<false>(p + l * kSpl, z[l], ld, 0.0f, 0.0f, nullptr);
            ldsum += ld;
        }
        const float sig = 1.0f / (1.0f + expf(-tail[1]));
        const float S = 100.0f * sig;
        const float q = (z[kL] - tail[0]) / S;
        loss -= -logf(S) - 0.5f * (kLog2Pi + q * q) + ldsum;
        if (kGrad) {
            g[kL * kSpl] = -q / S;
            const float gS = 1.0f / S - q * q / S;
            g[kL * kSpl + 1] = gS * S * (1.0f - sig);
            float gx = q / S;
            for (int l = kL - 1; l >= 0; --l) {
                float unused;
                gx = inv_layer<true>(p + l * kSpl, z[l], unused, gx, -1.0f, g + l * kSpl);
            }
        }
    }
    if (fc) {
        const float lam = expf(loglam);
        loss -= cnt * loglam - lam - lgammaf(cnt + 1.0f);
        if (kGrad) g[kL * kSpl + 2] = lam - cnt;
    }
    if (kGrad) {
        const float ws = weight * scale;
#pragma unroll
        for (int c = 0; c < kP; ++c) g[c] *= ws;
    }
    return loss * weight;
}

// ---------------------------------------------------------------------------
// Fused kernel
// ---------------------------------------------------------------------------
struct Rows {
    const float* x;
    const float* t;
    const float* cnt;
    const uint8_t* flags;
    const float* aux;          // [n][2] (log λ offset, row weight) or NULL
    int64_t n;
};

template <bool kGrad>
__global__ __launch_bounds__(kThreads, 1) void flow_kernel(const float* __restrict__ w, Rows rows, float scale,
                                                          float* __restrict__ params_out,
                                                          float* __restrict__ loss_out,
                                                          float* __restrict__ partial,
                                                          double* __restrict__ partial_loss) {
    __shared__ Smem sm;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int r0 = wave * 16;
    const int64_t n_tiles = (rows.n + kRows - 1) / kRows;

    // weights → LDS (bfloat16) and biases (FP32), once per workgroup
    for (int e = tid; e < kIn * kH; e += kThreads) {
        const int f = e / kH, hcol = e % kH;
        sm.w1t[hcol][f] = f2bf(w[oW1 + e]);
    }
    for (int e = tid; e < kH * kH; e += kThreads) sm.w2[e / kH][e % kH] = f2bf(w[oW2 + e]);
    for (int e = tid; e < kH * kP; e += kThreads) sm.w3[e / kP][e % kP] = f2bf(w[oW3 + e]);
    for (int e = tid; e < kH; e += kThreads) {
        sm.b[e] = w[ob1 + e];
        sm.b[kH + e] = w[ob2 + e];
    }
    for (int e = tid; e < kP; e += kThreads) sm.b[2 * kH + e] = w[ob3 + e];

    // Row tiles: the gradient launch has one tile per workgroup (one partial
    // record each); the forward launch is persistent over the tiles.
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t row0 = tile * kRows;
    for (int e = tid; e < kRows * kIn; e += kThreads) {
        const int r = e / kIn, f = e % kIn;
        const int64_t gr = row0 + r;
        sm.x[r][f] = f2bf(gr < rows.n ? rows.x[gr * kIn + f] : 0.0f);
    }
    __syncthreads();

    // ---- forward: H1 = relu(X W1 + b1) -------------------------------------
    {
        const bf16x4 a = frag_rows(sm.x, r0, 0, i, g);
        for (int nb = 0; nb < kH / 16; ++nb) {
            const bf16x4 b = frag_rows(sm.w1t, nb * 16, 0, i, g);
            const f32x4 c = mfma16(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
            const int col = nb * 16 + i;
#pragma unroll
            for (int j = 0; j < 4; ++j) sm.h1[r0 + 4 * g + j][col] = f2bf(fmaxf(c[j] + sm.b[col], 0.0f));
        }
    }
    __syncthreads();
    // ---- H2 = relu(H1 W2 + b2) ----------------------------------------------
    {
        bf16x4 a[kH / 16];
#pragma unroll
        for (int kb = 0; kb < kH / 16; ++kb) a[kb] = frag_rows(sm.h1, r0, kb * 16, i, g);
        for (int nb = 0; nb < kH / 16; ++nb) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kH / 16; ++kb) c = mfma16(a[kb], frag_cols(sm.w2, kb * 16, nb * 16, i, g), c);
            const int col = nb * 16 + i;
#pragma unroll
            for (int j = 0; j < 4; ++j) sm.h2[r0 + 4 * g + j][col] = f2bf(fmaxf(c[j] + sm.b[kH + col], 0.0f));
        }
    }
    __syncthreads();
    // ---- P = H2 W3 + b3 ------------------------------------------------------
    {
        bf16x4 a[kH / 16];
#pragma unroll
        for (int kb = 0; kb < kH / 16; ++kb) a[kb] = frag_rows(sm.h2, r0, kb * 16, i, g);
        for (int nb = 0; nb < kP / 16; ++nb) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kH / 16; ++kb) c = mfma16(a[kb], frag_cols(sm.w3, kb * 16, nb * 16, i, g), c);
            const int col = nb * 16 + i;
#pragma unroll
            for (int j = 0; j < 4; ++j) sm.p[r0 + 4 * g + j][col] = c[j] + sm.b[2 * kH + col];
        }
    }
    __syncthreads();
    if (params_out) {
        for (int e = tid; e < kRows * kP; e += kThreads) {
            const int r = e / kP, c = e % kP;
            if (row0 + r < rows.n) params_out[(row0 + r) * kP + c] = sm.p[r][c];
        }
    }

    // ---- per-row loss (and d loss / d P): 16 lanes per wavefront -------------
    if (lane < 16) {
        const int r = r0 + lane;
        const int64_t gr = row0 + r;
        float loss = 0.0f;
        if (gr < rows.n) {
            float gp[kP];
            const float off = rows.aux ? rows.aux[2 * gr] : 0.0f;
            const float wgt = rows.aux ? rows.aux[2 * gr + 1] : 1.0f;
            loss = row_loss<kGrad>(sm.p[r], rows.t[gr], rows.cnt[gr], rows.flags[gr], off, wgt, gp, scale);
            if (kGrad) {
#pragma unroll
                for (int c = 0; c < kP; ++c) sm.p[r][c] = gp[c];
            }
        } else if (kGrad) {
#pragma unroll
            for (int c = 0; c < kP; ++c) sm.p[r][c] = 0.0f;
        }
        sm.loss[r] = loss;
        if (loss_out && gr < rows.n) loss_out[gr] = loss;
    }
    __syncthreads();
    if (tid == 0 && partial_loss) {
        double s = 0.0;
        for (int r = 0; r < kRows; ++r) s += (double)sm.loss[r];
        partial_loss[blockIdx.x] = s;
    }
    if (!kGrad) {
        __syncthreads();        // the next tile overwrites the LDS tiles
        continue;
    }

    float* part = partial + (int64_t)blockIdx.x * kPartW;
    float* pb = part + pBias + wave * kBiasW;
    // dP in bfloat16 for the products; db3 from the FP32 values of this wave's rows
    for (int e = lane; e < 16 * kP; e += 64) {
        const int r = r0 + e / kP, c = e % kP;
        sm.dp[r][c] = f2bf(sm.p[r][c]);
    }
    for (int c = lane; c < kP; c += 64) {
        float s = 0.0f;
        for (int r = 0; r < 16; ++r) s += sm.p[r0 + r][c];
        pb[2 * kH + c] = s;
    }
    __syncthreads();
    // ---- dH2 = (dP W3ᵀ) ⊙ [H2 > 0] --------------------------------------------
    {
        bf16x4 a[kP / 16];
#pragma unroll
        for (int kb = 0; kb < kP / 16; ++kb) a[kb] = frag_rows(sm.dp, r0, kb * 16, i, g);
        for (int nb = 0; nb < kH / 16; ++nb) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kP / 16; ++kb) c = mfma16(a[kb], frag_rows(sm.w3, nb * 16, kb * 16, i, g), c);
            const int col = nb * 16 + i;
            float colsum = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = r0 + 4 * g + j;
                const float v = sm.h2[r][col] != 0 ? c[j] : 0.0f;   // H2 ≥ 0: nonzero bits ⇔ > 0
                sm.dh2[r][col] = f2bf(v);
                colsum += v;
            }
            colsum += __shfl_xor(colsum, 16);
            colsum += __shfl_xor(colsum, 32);
            if (g == 0) pb[kH + col] = colsum;
        }
    }
    __syncthreads();
    // ---- dH1 = (dH2 W2ᵀ) ⊙ [H1 > 0] -------------------------------------------
    {
        bf16x4 a[kH / 16];
#pragma unroll
        for (int kb = 0; kb < kH / 16; ++kb) a[kb] = frag_rows(sm.dh2, r0, kb * 16, i, g);
        for (int nb = 0; nb < kH / 16; ++nb) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kH / 16; ++kb) c = mfma16(a[kb], frag_rows(sm.w2, nb * 16, kb * 16, i, g), c);
            const int col = nb * 16 + i;
            float colsum = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = r0 + 4 * g + j;
                const float v = sm.h1[r][col] != 0 ? c[j] : 0.0f;
                sm.dh1[r][col] = f2bf(v);
                colsum += v;
            }
            colsum += __shfl_xor(colsum, 16);
            colsum += __shfl_xor(colsum, 32);
            if (g == 0) pb[col] = colsum;
        }
    }
    __syncthreads();
    // ---- weight gradients over the workgroup's 64 rows (K = 64) ----------------
    // dW3 = H2ᵀ dP: 8 × 5 tiles
    for (int tile = wave; tile < (kH / 16) * (kP / 16); tile += 4) {
        const int mb = tile / (kP / 16), nb = tile % (kP / 16);
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kRows / 16; ++ks)
            c = mfma16(frag_cols(sm.h2, ks * 16, mb * 16, i, g), frag_cols(sm.dp, ks * 16, nb * 16, i, g), c);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[pW3 + (mb * 16 + 4 * g + j) * kP + nb * 16 + i] = c[j];
    }
    // dW2 = H1ᵀ dH2: 8 × 8 tiles
    for (int tile = wave; tile < (kH / 16) * (kH / 16); tile += 4) {
        const int mb = tile / (kH / 16), nb = tile % (kH / 16);
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kRows / 16; ++ks)
            c = mfma16(frag_cols(sm.h1, ks * 16, mb * 16, i, g), frag_cols(sm.dh2, ks * 16, nb * 16, i, g), c);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[pW2 + (mb * 16 + 4 * g + j) * kH + nb * 16 + i] = c[j];
    }
    // dW1 = Xᵀ dH1: 1 × 8 tiles
    for (int nb = wave; nb < kH / 16; nb += 4) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kRows / 16; ++ks)
            c = mfma16(frag_cols(sm.x, ks * 16, 0, i, g), frag_cols(sm.dh1, ks * 16, nb * 16, i, g), c);
#pragma unroll
        for (int j = 0; j < 4; ++j) part[pW1 + (4 * g + j) * kH + nb * 16 + i] = c[j];
    }
    __syncthreads();
    }  // tiles
}

constexpr int kSplits = 16;           // record ranges summed in parallel by the first reduction pass

// Source offset of weight q in a partial record and its number of sub-entries
// (biases: one per wavefront).
__device__ __forceinline__ int record_source(int q, int& nsub) {
    nsub = 1;
    if (q < ob1) return pW1 + (q - oW1);
    if (q < oW2) { nsub = 4; return pBias + (q - ob1); }
    if (q < ob2) return pW2 + (q - oW2);
    if (q < oW3) { nsub = 4; return pBias + kH + (q - ob2); }
    if (q < ob3) return pW3 + (q - oW3);
    nsub = 4;
    return pBias + 2 * kH + (q - ob3);
}

// First pass: tmp[split][q] = Σ over records [split·nb/S, (split+1)·nb/S) in order.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float* __restrict__ partial, int nblocks,
                                                             float* __restrict__ tmp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int split = blockIdx.y;
    if (q >= kNW) return;
    int nsub;
    const int src = record_source(q, nsub);
    const int b0 = (int)((int64_t)nblocks * split / kSplits), b1 = (int)((int64_t)nblocks * (split + 1) / kSplits);
    float acc = 0.0f;
    for (int b = b0; b < b1; ++b) {
        const float* rec = partial + (int64_t)b * kPartW + src;
        for (int s = 0; s < nsub; ++s) acc += rec[s * kBiasW];
    }
    tmp[(int64_t)split * kNW + q] = acc;
}

// Second pass: gradient = Σ_split tmp (fixed order), then either the gradient
// or the Adam update; thread 0 also sums the per-block losses.
__global__ __launch_bounds__(256) void reduce_kernel(const float* __restrict__ tmp, int nblocks,
                                                    const double* __restrict__ partial_loss, double loss_scale,
                                                    double* __restrict__ loss_out, float* __restrict__ grad_out,
                                                    float* __restrict__ w, float* __restrict__ m,
                                                    float* __restrict__ v, float lr, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0 && loss_out) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += partial_loss[b];
        *loss_out = s * loss_scale;
    }
    if (q >= kNW) return;
    float gsum = 0.0f;
#pragma unroll
    for (int sp = 0; sp < kSplits; ++sp) gsum += tmp[(int64_t)sp * kNW + q];
    if (grad_out) {
        grad_out[q] = gsum;
        return;
    }
    // Flux.Optimise.Adam: θ −= η·m̂/(√v̂ + ε)
    const float mq = 0.9f * m[q] + 0.1f * gsum;
    const float vq = 0.999f * v[q] + 0.001f * gsum * gsum;
    m[q] = mq;
    v[q] = vq;
    w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
}

__global__ void zero_kernel(float* p, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0.0f;
}

// ---------------------------------------------------------------------------
// Event likelihood helpers
// ---------------------------------------------------------------------------
struct PairArgs {
    float mean[5], inv_std[5];
    float knot_d[NTF_MAX_SCALE_KNOTS], knot_g[NTF_MAX_SCALE_KNOTS];
    int n_knots;
    float log_yield_1gev;
};

// Conditioning of one (source, module, PMT) pair: ln d, cos θ, cos β, cos φ,
// log10 E, z-scored; offset = ln(yield(E)·g(d)) with g interpolated in ln d
// (clamped to the end knots).  Pair index = (s·n_mod + m)·n_pmt + k.
__global__ __launch_bounds__(256) void pair_rows_kernel(const float* __restrict__ src, int n_src,
                                                       const float* __restrict__ mod, int n_mod,
                                                       const float* __restrict__ axes, int n_pmt, PairArgs a,
                                                       float* __restrict__ x, float* __restrict__ aux) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)n_src * n_mod * n_pmt;
    if (q >= n) return;
    const int k = (int)(q % n_pmt);
    const int64_t sm = q / n_pmt;
    const int m = (int)(sm % n_mod);
    const int si = (int)(sm / n_mod);
    const float* S = src + 8 * (int64_t)si;
    const float rx = mod[3 * m] - S[0], ry = mod[3 * m + 1] - S[1], rz = mod[3 * m + 2] - S[2];
    const float d = sqrtf(rx * rx + ry * ry + rz * rz);
    const float sx = -rx / d, sy = -ry / d, sz = -rz / d;             // module → source
    const float ux = S[3], uy = S[4], uz = S[5];
    const float cth = ux * sx + uy * sy + uz * sz;
    float px = ux - cth * sx, py = uy - cth * sy, pz = uz - cth * sz;
    float np_ = sqrtf(px * px + py * py + pz * pz);
    if (!(np_ > 1e-6f)) {                                          // any perpendicular
        const bool xs = fabsf(sx) < 0.9f;
        const float ax = xs ? 1.0f : 0.0f, ay = xs ? 0.0f : 1.0f;
        px = sy * 0.0f - sz * ay;
        py = sz * ax - sx * 0.0f;
        pz = sx * ay - sy * ax;
        np_ = sqrtf(px * px + py * py + pz * pz);
    }
    const float e1x = px / np_, e1y = py / np_, e1z = pz / np_;
    const float e2x = sy * e1z - sz * e1y, e2y = sz * e1x - sx * e1z, e2z = sx * e1y - sy * e1x;
    const float ax = axes[3 * k], ay = axes[3 * k + 1], az = axes[3 * k + 2];
    const float cb = ax * sx + ay * sy + az * sz;
    const float a1 = ax * e1x + ay * e1y + az * e1z;
    const float a2 = ax * e2x + ay * e2y + az * e2z;
    const float na = sqrtf(a1 * a1 + a2 * a2);
    const float cphi = na > 1e-12f ? a1 / na : 1.0f;
    const float lnd = logf(d);
    const float f[5] = {lnd, cth, cb, cphi, S[6]};
    float* xr = x + q * kIn;
#pragma unroll
    for (int j = 0; j < 5; ++j) xr[j] = (f[j] - a.mean[j]) * a.inv_std[j];
#pragma unroll
    for (int j = 5; j < kIn; ++j) xr[j] = 0.0f;
    float lg;
    if (lnd <= a.knot_d[0]) {
        lg = a.knot_g[0];
    } else if (lnd >= a.knot_d[a.n_knots - 1]) {
        lg = a.knot_g[a.n_knots - 1];
    } else {
        int j = 0;
        while (j + 2 < a.n_knots && lnd >= a.knot_d[j + 1]) ++j;
        const float u = (lnd - a.knot_d[j]) / (a.knot_d[j + 1] - a.knot_d[j]);
        lg = a.knot_g[j] + u * (a.knot_g[j + 1] - a.knot_g[j]);
    }
    aux[2 * q] = a.log_yield_1gev + 2.302585093f * S[6] + lg;
    aux[2 * q + 1] = 1.0f;
}

// Σ values over each chunk [chunk_off[c], chunk_off[c+1]) (elements with
// skip != 0 left out), in double, one workgroup per chunk; then Σ of each
// segment's chunks [seg_chunk[s], seg_chunk[s+1]) in order.  Chunks are
// bounded in length, so events of very different sizes cost alike, and the
// order of additions is fixed (bitwise reproducible sums).
__global__ __launch_bounds__(256) void chunk_sum_kernel(const float* __restrict__ v,
                                                       const uint8_t* __restrict__ skip,
                                                       const int64_t* __restrict__ chunk_off,
                                                       double* __restrict__ partial) {
    __shared__ double part[256];
    const int64_t c = blockIdx.x;
    const int64_t a = chunk_off[c], b = chunk_off[c + 1];
    double acc = 0.0;
    for (int64_t q = a + threadIdx.x; q < b; q += blockDim.x)
        if (!skip || !skip[q]) acc += (double)v[q];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[c] = part[0];
}

__global__ __launch_bounds__(256) void segment_sum_kernel(const double* __restrict__ partial,
                                                         const int64_t* __restrict__ seg_chunk, int64_t n_seg,
                                                         float sign, int accumulate, double* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    double acc = 0.0;
    for (int64_t c = seg_chunk[s]; c < seg_chunk[s + 1]; ++c) acc += partial[c];
    out[s] = (accumulate ? out[s] : 0.0) + (double)sign * acc;
}

// logpdf of each hit's time under the flow parameters of its pair.
__global__ __launch_bounds__(256) void time_logpdf_kernel(const float* __restrict__ params, int64_t n_pairs,
                                                         const int32_t* __restrict__ pair,
                                                         const float* __restrict__ t, int64_t n,
                                                         float* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t pq = pair[q];
    if (pq < 0 || pq >= n_pairs) {
        out[q] = -INFINITY;
        return;
    }
    float p[kNP];
    const float* src = params + pq * kP;
#pragma unroll
    for (int c = 0; c < kNP; ++c) p[c] = src[c];
    out[q] = -row_loss<false>(p, t[q], 0.0f, NTF_ROW_TIME, 0.0f, 1.0f, nullptr, 1.0f);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct ntf_ctx {
    int device = 0;
    int n_cu = 256;
    float* d_w = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_partial = nullptr;
    double* d_ploss = nullptr;
    int64_t partial_blocks = 0;
    int64_t step = 0;
    double* d_chunk = nullptr;
    int64_t chunk_cap = 0;
    float* d_tmp = nullptr;              // kSplits × NTF_N_WEIGHTS first-pass sums
};

namespace {

int32_t hip_status(hipError_t e) { return e == hipSuccess ? NT_OK : NT_ERR_DEVICE; }

int32_t ensure_partials(ntf_ctx* c, int64_t blocks) {
    if (blocks <= c->partial_blocks) return NT_OK;
    if (c->d_partial) (void)hipFree(c->d_partial);
    if (c->d_ploss) (void)hipFree(c->d_ploss);
    c->d_partial = nullptr;
    c->d_ploss = nullptr;
    c->partial_blocks = 0;
    if (hipMalloc(&c->d_partial, (size_t)blocks * kPartW * sizeof(float)) != hipSuccess) return NT_ERR_OOM;
    if (hipMalloc(&c->d_ploss, (size_t)blocks * sizeof(double)) != hipSuccess) return NT_ERR_OOM;
    c->partial_blocks = blocks;
    return NT_OK;
}

bool rows_ok(const float* x, const float* t, const float* cnt, const uint8_t* flags, int64_t n) {
    if (n < 0 || n > (int64_t)1 << 40) return false;
    return n == 0 || (x && t && cnt && flags);
}

int32_t run_grad(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                 const float* aux, int64_t n, float scale, hipStream_t s) {
    const int64_t blocks = (n + kRows - 1) / kRows;
    const int32_t st = ensure_partials(c, blocks);
    if (st != NT_OK) return st;
    hipLaunchKernelGGL(flow_kernel<true>, dim3((unsigned)blocks), dim3(kThreads), 0, s, c->d_w,
                       Rows{x, t, cnt, flags, aux, n}, scale, (float*)nullptr, (float*)nullptr, c->d_partial,
                       c->d_ploss);
    return hip_status(hipGetLastError());
}

}  // namespace

extern "C" {

int32_t ntf_create(int32_t device, ntf_ctx** out) {
    if (!out) return NT_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NT_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return NT_ERR_DEVICE;
    ntf_ctx* c = new (std::nothrow) ntf_ctx();
    if (!c) return NT_ERR_OOM;
    c->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->n_cu = cus;
    if (hipMalloc(&c->d_w, kNW * sizeof(float)) != hipSuccess || hipMalloc(&c->d_m, kNW * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->d_v, kNW * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->d_tmp, (size_t)kSplits * kNW * sizeof(float)) != hipSuccess) {
        ntf_destroy(c);
        return NT_ERR_OOM;
    }
    const int blocks = (kNW + 255) / 256;
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, 0, c->d_w, kNW);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, 0, c->d_m, kNW);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, 0, c->d_v, kNW);
    if (hipDeviceSynchronize() != hipSuccess) {
        ntf_destroy(c);
        return NT_ERR_DEVICE;
    }
    *out = c;
    return NT_OK;
}

void ntf_destroy(ntf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->d_w) (void)hipFree(c->d_w);
    if (c->d_m) (void)hipFree(c->d_m);
    if (c->d_v) (void)hipFree(c->d_v);
    if (c->d_partial) (void)hipFree(c->d_partial);
    if (c->d_ploss) (void)hipFree(c->d_ploss);
    if (c->d_chunk) (void)hipFree(c->d_chunk);
    if (c->d_tmp) (void)hipFree(c->d_tmp);
    delete c;
}

int32_t ntf_set_weights(ntf_ctx* c, const float* weights) {
    if (!c || !weights) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    if (hipMemcpy(c->d_w, weights, kNW * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return NT_ERR_DEVICE;
    const int blocks = (kNW + 255) / 256;
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, 0, c->d_m, kNW);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, 0, c->d_v, kNW);
    c->step = 0;
    return hip_status(hipDeviceSynchronize());
}

int32_t ntf_get_weights(ntf_ctx* c, float* weights) {
    if (!c || !weights) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    if (hipDeviceSynchronize() != hipSuccess) return NT_ERR_DEVICE;
    return hip_status(hipMemcpy(weights, c->d_w, kNW * sizeof(float), hipMemcpyDeviceToHost));
}

int32_t ntf_forward(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                    const float* aux, int64_t n, float* d_params, float* d_row_loss, void* stream) {
    if (!c || !rows_ok(x, t, cnt, flags, n)) return NT_ERR_ARG;
    if (n == 0) return NT_OK;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    const int64_t tiles = (n + kRows - 1) / kRows;
    const int64_t blocks = tiles < c->n_cu ? tiles : c->n_cu;       // one resident workgroup per CU
    hipLaunchKernelGGL(flow_kernel<false>, dim3((unsigned)blocks), dim3(kThreads), 0, (hipStream_t)stream,
                       c->d_w, Rows{x, t, cnt, flags, aux, n}, 1.0f, d_params, d_row_loss, (float*)nullptr,
                       (double*)nullptr);
    return hip_status(hipGetLastError());
}

int32_t ntf_loss_grad(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                      const float* aux, int64_t n, float scale, float* d_grad, double* d_loss, void* stream) {
    if (!c || !d_grad || !rows_ok(x, t, cnt, flags, n) || n == 0) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    int32_t st = run_grad(c, x, t, cnt, flags, aux, n, scale, s);
    if (st != NT_OK) return st;
    const int64_t blocks = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNW + 255) / 256, kSplits), dim3(256), 0, s, c->d_partial,
                       (int)blocks, c->d_tmp);
    hipLaunchKernelGGL(reduce_kernel, dim3((kNW + 255) / 256), dim3(256), 0, s, c->d_tmp, (int)blocks,
                       c->d_ploss, 1.0, d_loss, d_grad, c->d_w, c->d_m, c->d_v, 0.0f, 1.0f, 1.0f);
    return hip_status(hipGetLastError());
}

int32_t ntf_train_step(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                       const float* aux, int64_t n, float lr, double* d_loss, void* stream) {
    if (!c || !rows_ok(x, t, cnt, flags, n) || n == 0 || !(lr >= 0.0f)) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    int32_t st = run_grad(c, x, t, cnt, flags, aux, n, 1.0f / (float)n, s);
    if (st != NT_OK) return st;
    c->step += 1;
    const float c1 = (float)(1.0 - pow(0.9, (double)c->step));
    const float c2 = (float)(1.0 - pow(0.999, (double)c->step));
    const int64_t blocks = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNW + 255) / 256, kSplits), dim3(256), 0, s, c->d_partial,
                       (int)blocks, c->d_tmp);
    hipLaunchKernelGGL(reduce_kernel, dim3((kNW + 255) / 256), dim3(256), 0, s, c->d_tmp, (int)blocks,
                       c->d_ploss, 1.0 / (double)n, d_loss, (float*)nullptr, c->d_w, c->d_m, c->d_v, lr, c1, c2);
    return hip_status(hipGetLastError());
}

int32_t ntf_pair_rows(ntf_ctx* c, const float* d_sources, int32_t n_sources, const float* d_modules,
                      int32_t n_modules, const float* d_pmt_axes, int32_t n_pmts, const float* norm,
                      const float* scale_knots, int32_t n_knots, float log_yield_1gev, float* d_x, float* d_aux,
                      void* stream) {
    if (!c || !d_sources || !d_modules || !d_pmt_axes || !norm || !scale_knots || !d_x || !d_aux) return NT_ERR_ARG;
    if (n_sources < 0 || n_modules <= 0 || n_pmts <= 0 || n_knots < 2 || n_knots > NTF_MAX_SCALE_KNOTS)
        return NT_ERR_ARG;
    if (n_sources == 0) return NT_OK;
    PairArgs a;
    for (int j = 0; j < 5; ++j) {
        if (!(norm[5 + j] > 0.0f)) return NT_ERR_ARG;
        a.mean[j] = norm[j];
        a.inv_std[j] = 1.0f / norm[5 + j];
    }
    for (int j = 0; j < n_knots; ++j) {
        a.knot_d[j] = scale_knots[2 * j];
        a.knot_g[j] = scale_knots[2 * j + 1];
        if (j > 0 && !(a.knot_d[j] > a.knot_d[j - 1])) return NT_ERR_ARG;
    }
    a.n_knots = n_knots;
    a.log_yield_1gev = log_yield_1gev;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    const int64_t n = (int64_t)n_sources * n_modules * n_pmts;
    hipLaunchKernelGGL(pair_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_sources, n_sources, d_modules, n_modules, d_pmt_axes, n_pmts, a, d_x, d_aux);
    return hip_status(hipGetLastError());
}

int32_t ntf_segment_sum(ntf_ctx* c, const float* d_values, const uint8_t* d_skip, const int64_t* d_chunk_offsets,
                        int64_t n_chunks, const int64_t* d_segment_chunks, int64_t n_segments, float sign,
                        int32_t accumulate, double* d_out, void* stream) {
    if (!c || n_segments < 0 || n_chunks < 0) return NT_ERR_ARG;
    if (n_segments == 0) return NT_OK;
    if (!d_values || !d_chunk_offsets || !d_segment_chunks || !d_out || n_chunks > 0x7fffffff) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    if (n_chunks > c->chunk_cap) {
        if (c->d_chunk) (void)hipFree(c->d_chunk);
        c->d_chunk = nullptr;
        c->chunk_cap = 0;
        if (hipMalloc(&c->d_chunk, (size_t)n_chunks * sizeof(double)) != hipSuccess) return NT_ERR_OOM;
        c->chunk_cap = n_chunks;
    }
    if (n_chunks > 0)
        hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)n_chunks), dim3(256), 0, s, d_values, d_skip,
                           d_chunk_offsets, c->d_chunk);
    hipLaunchKernelGGL(segment_sum_kernel, dim3((unsigned)((n_segments + 255) / 256)), dim3(256), 0, s, c->d_chunk,
                       d_segment_chunks, n_segments, sign, (int)accumulate, d_out);
    return hip_status(hipGetLastError());
}

int32_t ntf_time_logpdf(ntf_ctx* c, const float* d_params, int64_t n_pairs, const int32_t* d_pair,
                        const float* d_t, int64_t n, float* d_logpdf, void* stream) {
    if (!c || n < 0 || n_pairs < 0 || (n > 0 && (!d_params || !d_pair || !d_t || !d_logpdf))) return NT_ERR_ARG;
    if (n == 0) return NT_OK;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipLaunchKernelGGL(time_logpdf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_params, n_pairs, d_pair, d_t, n, d_logpdf);
    return hip_status(hipGetLastError());
}

}  // extern "C"
