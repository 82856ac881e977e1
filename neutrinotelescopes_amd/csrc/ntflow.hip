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
// Work decomposition.  A wavefront owns 16 rows.  The layer products run
// transposed (Hᵀ = Wᵀ·Xᵀ): the accumulator fragment of one 16-wide output
// block is, after bias/ReLU/bfloat16, exactly the B operand of the next
// product, so the activations of the forward and the backward chain never
// leave registers; the weights are the A operands, read from LDS as one
// 16-byte fragment each (v_mfma_f32_16x16x32_bf16; the 16-deep input layer
// uses the 16x16x16 form).  The weights arrive as a bfloat16 image in the
// exact LDS layout (rows XOR-swizzled in 16-byte chunks, bank-conflict free,
// and the K index permuted within 32-blocks to match the register order of
// the activation fragments), refreshed by the Adam kernel after each update.
// The per-row flow loss and its gradient run on four lanes per row, one
// spline layer per lane, the layer chain passed along the quad with DPP.
// The training kernel's workgroup (4 wavefronts, 64 rows) then forms the
// weight gradients of its rows with K = 64 MFMA reductions over
// row-contiguous (feature-major) bfloat16 copies of the activations in LDS
// and writes one partial gradient record; the reduction kernels sum the
// records in a fixed order (bitwise reproducible) and apply Adam.  The
// forward kernel keeps only the forward weights in LDS (two workgroups per
// CU) and is persistent per wavefront over 16-row tiles.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <new>

#include "../../include/ntflow.h"

namespace {

constexpr int kIn = NTF_N_IN, kH = NTF_HIDDEN, kP = NTF_P_PAD, kK = NTF_BINS, kL = NTF_LAYERS;
constexpr int kNP = NTF_N_PARAMS;
constexpr int kSpl = 3 * kK + 1;          // parameters of one spline layer
constexpr int kRows = 64;                 // rows per workgroup of the gradient kernel
constexpr int kThreads = 256;             // 4 wavefronts of 16 rows
constexpr int kPLd = 84;                  // FP32 row stride of an output tile in LDS (16-byte rows)
constexpr float kRangeMin = NTF_RANGE_MIN, kRangeMax = NTF_RANGE_MAX;
constexpr float kMinBin = 1e-4f, kMinSlope = 1e-4f;
constexpr float kSlopeOffset = 0.5411666523385311f;   // log(exp(1 − 1e-4) − 1), rq_spline_flow.jl:151
constexpr float kLog2Pi = 1.8378770664093453f;
static_assert(kIn == 16 && kH == 128 && kP == 80, "fragment tiling assumes 16 inputs, 128 hidden, 80 outputs");
static_assert(kL * kSpl + 3 == kNP && kNP <= kP, "outputs");

constexpr int oW1 = 0, ob1 = oW1 + kIn * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kP, kNW = ob3 + kP;
static_assert(kNW == NTF_N_WEIGHTS, "weight layout");

// partial gradient record of one workgroup: dW1 | dW2 | dW3 | 4 × (db1, db2, db3)
constexpr int pW1 = 0, pW2 = pW1 + kIn * kH, pW3 = pW2 + kH * kH, pBias = pW3 + kH * kP;
constexpr int kBiasW = kH + kH + kP;
constexpr int kPartW = pBias + 4 * kBiasW;

// bfloat16 weight image (halfword offsets), the LDS layout of the weights:
//   W1ᵀ [128 h][16 f]      8-byte units XOR-swizzled by bit 3 of the row
//   W2ᵀ [128 out][128 in]  forward A operands    (128-wide rows: 16-byte chunks
//   W3ᵀ [80 c][128 h]      forward A operands     XOR (row & 15), K permuted
//   W2  [128 in][128 out]  backward A operands    within 32-blocks)
//   W3  [128 h][88]        backward A operands (K = c: two permuted 32-blocks, then c = 64..79 in order)
constexpr int hW1T = 0;
constexpr int hW2T = hW1T + kH * kIn;
constexpr int hW3T = hW2T + kH * kH;
constexpr int hFwd = hW3T + kP * kH;      // end of the forward set
constexpr int hW2 = hFwd;
constexpr int kW3Ld = 88;
constexpr int hW3 = hW2 + kH * kH;
constexpr int hImg = hW3 + kH * kW3Ld;
static_assert(hFwd % 8 == 0 && hImg % 8 == 0, "16-byte copies");

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bits(__bf16 b) { return __builtin_bit_cast(uint16_t, b); }
__host__ __device__ constexpr int qperm(int L) { return 8 * ((L & 15) >> 2) + ((L >> 4) << 2) + (L & 3); }
// halfword column of logical index k in a swizzled 128-wide row r
__host__ __device__ constexpr int sw_col(int k, int r) {
    return ((((32 * (k >> 5) + qperm(k & 31)) >> 3) ^ (r & 15)) << 3) + (qperm(k & 31) & 7);
}

// 16×16 MFMA operand layout (wavefront of 64, lane = 16·g + i): A supplies row
// i, B column i, the K slots of lane group g; the accumulator holds C[4g + j][i].
// Activation fragments: element e of lane group g of 32-block kb is feature
// 32·kb + L(8g + e), L(8g + e) = 4g + e (e < 4) or 16 + 4g + e − 4, which is
// where the accumulators of output blocks 2kb and 2kb + 1 put them.
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// the K slots 0..3 / 4..7 of a 32-deep fragment: the two 16-deep blocks 2kb, 2kb+1 in natural order
__device__ __forceinline__ s16x4 half_lo(bf16x8 v) {
    return __builtin_bit_cast(s16x4, __builtin_shufflevector(v, v, 0, 1, 2, 3));
}
__device__ __forceinline__ s16x4 half_hi(bf16x8 v) {
    return __builtin_bit_cast(s16x4, __builtin_shufflevector(v, v, 4, 5, 6, 7));
}
// a 16-deep fragment as a 32-deep one with zero upper K slots.  Every
// accumulation chain keeps one MFMA opcode: this compiler does not pad a
// 16x16x16 MFMA that reads the accumulator of a 16x16x32 one with the wait
// states the hardware needs.
__device__ __forceinline__ bf16x8 widen(s16x4 v) {
    const bf16x4 z = {};
    return __builtin_shufflevector(__builtin_bit_cast(bf16x4, v), z, 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ __forceinline__ s16x4 ld_w1t(const uint16_t* s, int mb, int i, int g) {
    return *reinterpret_cast<const s16x4*>(s + hW1T + (16 * mb + i) * kIn + ((g ^ ((i >> 3) << 1)) << 2));
}
__device__ __forceinline__ bf16x8 ld_sw(const uint16_t* base, int row, int kb, int i, int g) {
    return *reinterpret_cast<const bf16x8*>(base + row * 128 + (((4 * kb + g) ^ i) << 3));
}
__device__ __forceinline__ bf16x8 ld_w3(const uint16_t* s, int row, int kb, int g) {
    return *reinterpret_cast<const bf16x8*>(s + hW3 + row * kW3Ld + 32 * kb + 8 * g);
}
__device__ __forceinline__ s16x4 ld_w3_tail(const uint16_t* s, int row, int g) {
    return *reinterpret_cast<const s16x4*>(s + hW3 + row * kW3Ld + 64 + 4 * g);
}
// feature-major activation copies T[feature][64 rows], 16-byte chunks XOR ((feature >> 1) & 7)
__device__ __forceinline__ int t_idx(int feat, int r) { return feat * 64 + ((((r >> 3) ^ ((feat >> 1) & 7))) << 3) + (r & 7); }
__device__ __forceinline__ bf16x8 ld_t(const uint16_t* T, int feat, int ks, int i, int g) {
    return *reinterpret_cast<const bf16x8*>(T + feat * 64 + (((4 * ks + g) ^ ((i >> 1) & 7)) << 3));
}

template <int Ctrl>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xF, 0xF, false));
}
template <int L>
__device__ __forceinline__ float quad_bcast(float v) { return dpp<L | (L << 2) | (L << 4) | (L << 6)>(v); }
// Σ over the 16 lanes of a DPP row (fixed order; every lane gets the same sum)
__device__ __forceinline__ float row16_sum(float v) {
    v += dpp<0xB1>(v);         // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);         // quad_perm [2,3,0,1]
    v += dpp<0x141>(v);        // row_half_mirror
    v += dpp<0x140>(v);        // row_mirror
    return v;
}
// LDS written by other lanes of this wavefront is visible to the next reads
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Spline of one layer (FP32), register resident: every array index is static
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
__device__ __forceinline__ void make_knots(const float (&u)[kSpl], Knots& k) {
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

// a[j], a[j + 1] by selects (no dynamic register indexing)
__device__ __forceinline__ void pick2(const float (&a)[kK + 1], int j, float& lo, float& hi) {
    lo = a[0];
    hi = a[1];
#pragma unroll
    for (int i = 1; i < kK; ++i) {
        lo = i == j ? a[i] : lo;
        hi = i == j ? a[i + 1] : hi;
    }
}

__device__ __forceinline__ float safe_root(float a, float b, float c) {
    const float disc = b * b - 4.0f * a * c;
    const float s = disc > 0.0f ? sqrtf(disc) : 0.0f;
    return b >= 0.0f ? (2.0f * c) / (-b - s) : (-b + s) / (2.0f * a);
}

// inv_rqs_univariate (rq_spline_flow.jl:334-389) at v: returns x, ld = log|dx/dv|
__device__ __forceinline__ float layer_fwd(const Knots& k, float v, float& ld) {
    if (v <= k.yp[0] || v >= k.yp[kK]) {          // linear tails
        const bool lo = v <= k.yp[0];
        const float de = lo ? k.d[0] : k.d[kK], edge = lo ? k.yp[0] : k.yp[kK], xe = lo ? k.xp[0] : k.xp[kK];
        ld = -logf(de);
        return (v - edge) / de + xe;
    }
    const int j = find_bin(k, v);
    float x0, x1, y0, y1, d0, d1;
    pick2(k.xp, j, x0, x1);
    pick2(k.yp, j, y0, y1);
    pick2(k.d, j, d0, d1);
    const float w = x1 - x0, h = y1 - y0, s = h / w;
    const float wv = fminf(fmaxf((v - y0) / h, 0.0f), 1.0f);
    const float st = d1 + d0 - 2.0f * s;
    const float qc = -s * wv;
    const float qb = d0 - st * wv;
    const float qa = s - qb;
    const float z = fminf(fmaxf(safe_root(qa, qb, qc), 0.0f), 1.0f);
    const float sq_z = z * z, z1mz = z - sq_z, omz = 1.0f - z, sq_1mz = omz * omz;
    const float Q = d1 * sq_z + 2.0f * s * z1mz + d0 * sq_1mz;
    const float D = s + st * z1mz;
    ld = -2.0f * logf(s) - logf(Q) + 2.0f * logf(D);
    return w * z + x0;
}

// Gradient of one inverse layer at v: d/du of (gx·x + gld·ld) and d/dv.  init
// holds everything that does not depend on the incoming gradient gx, so the
// layer chain only passes gv(gx) from lane to lane.
struct LayerBwd {
    bool tail, root;
    int j;
    float de, dv;                                    // tails: edge slope, v − edge
    float w, h, s, z, R, ldz, lds, ldd0, ldd1, Rp, Rs, Rd0, Rd1, sg0, sg1;

    __device__ __forceinline__ void init(const Knots& k, float v) {
        const bool lo = v <= k.yp[0];
        tail = lo || v >= k.yp[kK];
        j = lo ? 0 : kK;
        de = lo ? k.d[0] : k.d[kK];
        dv = v - (lo ? k.yp[0] : k.yp[kK]);
        root = false;
        w = h = s = z = R = ldz = lds = ldd0 = ldd1 = Rp = Rs = Rd0 = Rd1 = sg0 = sg1 = 0.0f;
        if (tail) return;
        j = find_bin(k, v);
        float x0, x1, y0, y1, d0, d1;
        pick2(k.xp, j, x0, x1);
        pick2(k.yp, j, y0, y1);
        pick2(k.d, j, d0, d1);
        pick2(k.sg, j, sg0, sg1);
        w = x1 - x0;
        h = y1 - y0;
        s = h / w;
        const float wv = fminf(fmaxf((v - y0) / h, 0.0f), 1.0f);
        const float st = d1 + d0 - 2.0f * s;
        const float qc = -s * wv;
        const float qb = d0 - st * wv;
        const float qa = s - qb;
        const float zr = safe_root(qa, qb, qc);
        z = fminf(fmaxf(zr, 0.0f), 1.0f);
        root = zr == z;
        const float sq_z = z * z, z1mz = z - sq_z, omz = 1.0f - z, sq_1mz = omz * omz;
        const float Q = d1 * sq_z + 2.0f * s * z1mz + d0 * sq_1mz;
        const float D = s + st * z1mz;
        const float N = s * sq_z + d0 * z1mz;
        R = N / D;
        const float dDz = st * (1.0f - 2.0f * z);
        const float dQz = 2.0f * d1 * z + 2.0f * s * (1.0f - 2.0f * z) - 2.0f * d0 * omz;
        const float dNz = 2.0f * s * z + d0 * (1.0f - 2.0f * z);
        // ld = −2 log s − log Q + 2 log D
        ldz = -dQz / Q + 2.0f * dDz / D;
        lds = -2.0f / s - 2.0f * z1mz / Q + 2.0f * (1.0f - 2.0f * z1mz) / D;
        ldd0 = -sq_1mz / Q + 2.0f * z1mz / D;
        ldd1 = -sq_z / Q + 2.0f * z1mz / D;
        // z solves h·R(z; s, d0, d1) = v − y0 (the forward spline in the bin)
        Rp = (dNz - R * dDz) / D;
        Rs = (sq_z - R * (1.0f - 2.0f * z1mz)) / D;
        Rd0 = (z1mz - R * z1mz) / D;
        Rd1 = (-R * z1mz) / D;
    }

    __device__ __forceinline__ float gv(float gx, float gld) const {
        if (tail) return gx / de;
        if (!root) return 0.0f;
        const float gz = gx * w + gld * ldz;
        return gz / (h * Rp);
    }

    __device__ __forceinline__ void grad(const Knots& k, float gx, float gld, float (&gu)[kSpl]) const {
#pragma unroll
        for (int m = 0; m < kSpl; ++m) gu[m] = 0.0f;
        if (tail) {
            const float gde = -gx * dv / (de * de) - gld / de;
            gu[2 * kK] = j == 0 ? gde * k.sg[0] : 0.0f;
            gu[3 * kK] = j == kK ? gde * k.sg[kK] : 0.0f;
            return;
        }
        const float gz = gx * w + gld * ldz;
        float gw = gx * z, gs = gld * lds, gd0 = gld * ldd0, gd1 = gld * ldd1;
        float gy0 = 0.0f, gh = 0.0f;
        if (root) {
            const float kappa = gz / (h * Rp);
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
        for (int i = 0; i <= kK; ++i) {
            gxp[i] = i == j ? gx - gw : (i == j + 1 ? gw : 0.0f);
            gyp[i] = i == j ? gy0 - gh : (i == j + 1 ? gh : 0.0f);
            gu[2 * kK + i] = i == j ? gd0 * sg0 : (i == j + 1 ? gd1 * sg1 : 0.0f);
        }
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
            gu[m] = k.sw[m] * (gsw[m] - dotw);
            gu[kK + m] = k.sh[m] * (gsh[m] - doth);
        }
    }
};

// Row loss ω·(−[time·logpdf(t) + count·(n log λ − λ − lgamma(n+1))]) of the
// FP32 outputs prow (an LDS row), log λ = p_λ + offset, evaluated by the four
// lanes s = 0..3 of a quad (one row): lane s < 3 owns spline layer s, lane 3
// the tail parameters.  All four lanes return the loss.  With kGrad, the
// quad overwrites prow with scale·d loss / d p and stores its bfloat16 copy
// in the feature-major tile tdp at row r.
template <bool kGrad>
__device__ float quad_row_loss(float* prow, int s, float t, float cnt, int flags, float offset, float weight,
                               float scale, uint16_t* tdp, int r) {
    const bool ft = (flags & NTF_ROW_TIME) != 0, fc = (flags & NTF_ROW_COUNT) != 0;
    const int seg = s < kL ? s : kL - 1;
    const float* tail = prow + kL * kSpl;
    const float mu = tail[0], sraw = tail[1], loglam = tail[2] + offset;
    float loss = 0.0f;
    float g[kSpl];
    float gt0 = 0.0f, gt1 = 0.0f, gt2 = 0.0f;
    if (kGrad) {
#pragma unroll
        for (int m = 0; m < kSpl; ++m) g[m] = 0.0f;
    }
    if (ft) {
        float u[kSpl];
#pragma unroll
        for (int m = 0; m < kSpl; ++m) u[m] = prow[seg * kSpl + m];
        Knots k;
        make_knots(u, k);
        // z_{l+1} = inverse layer l (z_l), layer l on lane l
        float z = t, z_in = t, ld_mine = 0.0f, ld, y;
        y = layer_fwd(k, z, ld);
        if (s == 0) { z_in = z; ld_mine = ld; }
        z = quad_bcast<0>(y);
        y = layer_fwd(k, z, ld);
        if (s == 1) { z_in = z; ld_mine = ld; }
        z = quad_bcast<1>(y);
        y = layer_fwd(k, z, ld);
        if (s == 2) { z_in = z; ld_mine = ld; }
        z = quad_bcast<2>(y);
        const float ldsum = (quad_bcast<0>(ld_mine) + quad_bcast<1>(ld_mine)) + quad_bcast<2>(ld_mine);
        const float sig = 1.0f / (1.0f + expf(-sraw));
        const float S = 100.0f * sig;
        const float q = (z - mu) / S;
        loss -= -logf(S) - 0.5f * (kLog2Pi + q * q) + ldsum;
        if (kGrad) {
            gt0 = -q / S;
            const float gS = 1.0f / S - q * q / S;
            gt1 = gS * S * (1.0f - sig);
            LayerBwd b;
            b.init(k, z_in);
            float gx = q / S, gx_mine = gx;          // d/dz of the last layer's output
            gx = quad_bcast<2>(b.gv(gx, -1.0f));
            if (s == 1) gx_mine = gx;
            gx = quad_bcast<1>(b.gv(gx, -1.0f));
            if (s == 0) gx_mine = gx;
            if (s < kL) b.grad(k, gx_mine, -1.0f, g);
        }
    }
    if (fc) {
        const float lam = expf(loglam);
        loss -= cnt * loglam - lam - lgammaf(cnt + 1.0f);
        if (kGrad) gt2 = lam - cnt;
    }
    if (kGrad) {
        const float ws = weight * scale;
        if (s < kL) {
#pragma unroll
            for (int m = 0; m < kSpl; ++m) {
                const float v = g[m] * ws;
                prow[s * kSpl + m] = v;
                tdp[t_idx(s * kSpl + m, r)] = bits((__bf16)v);
            }
        } else {
            const float v[kP - kL * kSpl] = {gt0 * ws, gt1 * ws, gt2 * ws, 0.0f, 0.0f};
#pragma unroll
            for (int m = 0; m < kP - kL * kSpl; ++m) {
                prow[kL * kSpl + m] = v[m];
                tdp[t_idx(kL * kSpl + m, r)] = bits((__bf16)v[m]);
            }
        }
    }
    return loss * weight;
}

// log-density of time t under the flow outputs p of one row (one lane)
__device__ float time_logpdf_row(const float* p, float t) {
    float z = t, ldsum = 0.0f;
#pragma unroll
    for (int l = 0; l < kL; ++l) {
        float u[kSpl];
#pragma unroll
        for (int m = 0; m < kSpl; ++m) u[m] = p[l * kSpl + m];
        Knots k;
        make_knots(u, k);
        float ld;
        z = layer_fwd(k, z, ld);
        ldsum += ld;
    }
    const float* tail = p + kL * kSpl;
    const float sig = 1.0f / (1.0f + expf(-tail[1]));
    const float S = 100.0f * sig;
    const float q = (z - tail[0]) / S;
    return -logf(S) - 0.5f * (kLog2Pi + q * q) + ldsum;
}

// ---------------------------------------------------------------------------
// MLP on the matrix cores
// ---------------------------------------------------------------------------
struct Rows {
    const float* x;
    const float* t;
    const float* cnt;
    const uint8_t* flags;
    const float* aux;          // [n][2] (log λ offset, row weight) or NULL
    int64_t n;
    const int64_t* idx = nullptr;   // row i is row idx[i] of the arrays (n_src rows); NULL: row i
    int64_t n_src = 0;

    // the arrays' row of row i, or -1 for an index outside them
    __device__ __forceinline__ int64_t src(int64_t i) const {
        if (!idx) return i;
        const int64_t j = idx[i];
        return (j >= 0 && j < n_src) ? j : -1;
    }
};

// global → LDS copy of kHalfs halfwords: every load of a thread in flight at once
template <int kHalfs>
__device__ __forceinline__ void copy_lds(uint16_t* dst, const uint16_t* __restrict__ src, int tid) {
    constexpr int n16 = kHalfs / 8, full = n16 / kThreads, rem = n16 % kThreads;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    uint4 v[full], t = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < full; ++u) v[u] = s[tid + u * kThreads];
    if (rem > 0 && tid < rem) t = s[full * kThreads + tid];
#pragma unroll
    for (int u = 0; u < full; ++u) d[tid + u * kThreads] = v[u];
    if (rem > 0 && tid < rem) d[full * kThreads + tid] = t;
}

__device__ __forceinline__ void load_bias(float* b, const float* __restrict__ w, int tid) {
    static_assert(kH <= kThreads && kP <= kThreads, "one bias per thread");
    if (tid < kH) {
        b[tid] = w[ob1 + tid];
        b[kH + tid] = w[ob2 + tid];
    }
    if (tid < kP) b[2 * kH + tid] = w[ob3 + tid];
}

// Input fragment of row `row` (B operand, K = 16 features): X[row][4g .. 4g+3]
__device__ __forceinline__ s16x4 x_frag(const Rows& rows, int64_t row, int g, uint16_t* tx, int r) {
    f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t sr = row < rows.n ? rows.src(row) : -1;
    if (sr >= 0) v = *reinterpret_cast<const f32x4*>(rows.x + sr * kIn + 4 * g);
    s16x4 f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        f[j] = (short)bits((__bf16)v[j]);
        if (tx) tx[t_idx(4 * g + j, r)] = (uint16_t)f[j];
    }
    return f;
}

// Forward chain of one wavefront's 16 rows: H1ᵀ = relu(W1ᵀXᵀ + b1), H2ᵀ, Pᵀ.
// Lane (i, g) ends with row i's outputs c = 16·mb + 4g + j in pv[mb][j] and
// the ReLU masks of its hidden units 16·mb + 4g + j in bit 4·mb + j of m1, m2
// (set where the bfloat16 activation is nonzero, i.e. > 0); with th1/th2 the
// bfloat16 activations also go to the feature-major tiles at row r.
__device__ __forceinline__ void forward_chain(const uint16_t* sw, const float* bias, s16x4 xf, int i, int g,
                                              uint32_t& m1, uint32_t& m2, f32x4 (&pv)[5], uint16_t* th1,
                                              uint16_t* th2, int r) {
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 h1f[4], h2f[4];
    m1 = m2 = 0u;
#pragma unroll
    for (int mb = 0; mb < kH / 16; ++mb) {
        const f32x4 c = mfma16(ld_w1t(sw, mb, i, g), xf, zero);
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 16 * mb + 4 * g);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const __bf16 v = (__bf16)fmaxf(c[j] + b[j], 0.0f);
            h1f[mb >> 1][(mb & 1) * 4 + j] = v;
            m1 |= (bits(v) != 0 ? 1u : 0u) << (4 * mb + j);
            if (th1) th1[t_idx(16 * mb + 4 * g + j, r)] = bits(v);
        }
    }
    // H2 (rounded to bfloat16 after a 128-deep sum) accumulates in 16-deep
    // steps in natural K order, the same FP32 accumulation sequence as the
    // 16-deep kernels before; the other products use the 32-deep instruction
#pragma unroll
    for (int mb = 0; mb < kH / 16; ++mb) {
        f32x4 c = zero;
#pragma unroll
        for (int kb = 0; kb < kH / 32; ++kb) {
            const bf16x8 a = ld_sw(sw + hW2T, 16 * mb + i, kb, i, g);
            c = mfma16(half_lo(a), half_lo(h1f[kb]), c);
            c = mfma16(half_hi(a), half_hi(h1f[kb]), c);
        }
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + kH + 16 * mb + 4 * g);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const __bf16 v = (__bf16)fmaxf(c[j] + b[j], 0.0f);
            h2f[mb >> 1][(mb & 1) * 4 + j] = v;
            m2 |= (bits(v) != 0 ? 1u : 0u) << (4 * mb + j);
            if (th2) th2[t_idx(16 * mb + 4 * g + j, r)] = bits(v);
        }
    }
#pragma unroll
    for (int mb = 0; mb < kP / 16; ++mb) {
        f32x4 c = zero;
#pragma unroll
        for (int kb = 0; kb < kH / 32; ++kb) c = mfma32(ld_sw(sw + hW3T, 16 * mb + i, kb, i, g), h2f[kb], c);
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 2 * kH + 16 * mb + 4 * g);
#pragma unroll
        for (int j = 0; j < 4; ++j) pv[mb][j] = c[j] + b[j];
    }
}

// LDS of the gradient kernel (bytes).  After the forward pass the forward
// weights' space holds the FP32 output tile and the feature-major dP tile.
constexpr int bB = hFwd * 2;                           // backward weights
constexpr int bTx = hImg * 2;                          // T_x  [16][64]
constexpr int bTh1 = bTx + kIn * 64 * 2;               // T_h1 [128][64], later T_dh1
constexpr int bTh2 = bTh1 + kH * 64 * 2;               // T_h2 [128][64], later T_dh2
constexpr int bBias = bTh2 + kH * 64 * 2;
constexpr int bLoss = bBias + (2 * kH + kP) * 4;
constexpr int bGradEnd = bLoss + kRows * 4;
constexpr int bP = 0;                                  // [64][84] FP32 outputs, then dP
constexpr int bTdp = bP + kRows * kPLd * 4;            // T_dp [80][64]
static_assert(bTdp + kP * 64 * 2 <= bB, "output tiles fit in the forward weights' space");
static_assert(bGradEnd <= 160 * 1024, "LDS");

__global__ __launch_bounds__(kThreads, 1) void flow_grad_kernel(const uint16_t* __restrict__ img,
                                                               const float* __restrict__ w, Rows rows, float scale,
                                                               float* __restrict__ partial,
                                                               double* __restrict__ partial_loss) {
    __shared__ __align__(16) unsigned char sm[bGradEnd];
    uint16_t* sw = reinterpret_cast<uint16_t*>(sm);
    uint16_t* tx = reinterpret_cast<uint16_t*>(sm + bTx);
    uint16_t* th1 = reinterpret_cast<uint16_t*>(sm + bTh1);
    uint16_t* th2 = reinterpret_cast<uint16_t*>(sm + bTh2);
    float* bias = reinterpret_cast<float*>(sm + bBias);
    float* sloss = reinterpret_cast<float*>(sm + bLoss);
    float* P = reinterpret_cast<float*>(sm + bP);
    uint16_t* tdp = reinterpret_cast<uint16_t*>(sm + bTdp);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int r0 = wave * 16;
    const int64_t row0 = (int64_t)blockIdx.x * kRows;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};

    copy_lds<hImg>(sw, img, tid);
    load_bias(bias, w, tid);
    const s16x4 xf = x_frag(rows, row0 + r0 + i, g, tx, r0 + i);
    __syncthreads();

    // ---- forward ------------------------------------------------------------
    uint32_t m1, m2;
    f32x4 pv[5];
    forward_chain(sw, bias, xf, i, g, m1, m2, pv, th1, th2, r0 + i);
    __syncthreads();                                   // forward weights no longer read
#pragma unroll
    for (int mb = 0; mb < kP / 16; ++mb) *reinterpret_cast<f32x4*>(P + (r0 + i) * kPLd + 16 * mb + 4 * g) = pv[mb];
    wave_sync();

    // ---- per-row loss and d loss / d P: four lanes per row ---------------------
    {
        const int s = lane & 3, r = r0 + (lane >> 2);
        const int64_t gr = row0 + r;
        float t = 0.0f, cnt = 0.0f, off = 0.0f, wgt = 1.0f;
        int flags = 0;
        const int64_t sr = gr < rows.n ? rows.src(gr) : -1;
        if (sr >= 0) {
            t = rows.t[sr];
            cnt = rows.cnt[sr];
            flags = rows.flags[sr];
            if (rows.aux) {
                off = rows.aux[2 * sr];
                wgt = rows.aux[2 * sr + 1];
            }
        }
        const float loss = quad_row_loss<true>(P + r * kPLd, s, t, cnt, flags, off, wgt, scale, tdp, r);
        if (s == 0) sloss[r] = loss;
    }
    __syncthreads();
    if (tid == 0 && partial_loss) {
        double acc = 0.0;
        for (int r = 0; r < kRows; ++r) acc += (double)sloss[r];
        partial_loss[blockIdx.x] = acc;
    }

    float* part = partial + (int64_t)blockIdx.x * kPartW;
    float* pb = part + pBias + wave * kBiasW;
    // db3 over this wave's rows from the FP32 dP
    for (int c = lane; c < kP; c += 64) {
        float acc = 0.0f;
        for (int rr = 0; rr < 16; ++rr) acc += P[(r0 + rr) * kPLd + c];
        pb[2 * kH + c] = acc;
    }
    // dW3 = H2ᵀ dP over the 64 rows: this wave's h blocks 2w, 2w+1
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
        const int mb = 2 * wave + mi;
        const bf16x8 a0 = ld_t(th2, 16 * mb + i, 0, i, g), a1 = ld_t(th2, 16 * mb + i, 1, i, g);
#pragma unroll
        for (int nb = 0; nb < kP / 16; ++nb) {
            f32x4 c = mfma32(a0, ld_t(tdp, 16 * nb + i, 0, i, g), zero);
            c = mfma32(a1, ld_t(tdp, 16 * nb + i, 1, i, g), c);
#pragma unroll
            for (int j = 0; j < 4; ++j) part[pW3 + (16 * mb + 4 * g + j) * kP + 16 * nb + i] = c[j];
        }
    }
    // dH2ᵀ = W3 dPᵀ for this wave's rows, masked by H2 > 0; db2
    bf16x8 dh2f[4];
    {
        const float* prow = P + (r0 + i) * kPLd;
        bf16x8 dpf[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const f32x4 lo = *reinterpret_cast<const f32x4*>(prow + 32 * kb + 4 * g);
            const f32x4 hi = *reinterpret_cast<const f32x4*>(prow + 32 * kb + 16 + 4 * g);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                dpf[kb][e] = (__bf16)lo[e];
                dpf[kb][4 + e] = (__bf16)hi[e];
            }
        }
        const f32x4 tl = *reinterpret_cast<const f32x4*>(prow + 64 + 4 * g);
        s16x4 dpt;
#pragma unroll
        for (int e = 0; e < 4; ++e) dpt[e] = (short)bits((__bf16)tl[e]);
#pragma unroll
        for (int mb = 0; mb < kH / 16; ++mb) {
            f32x4 c = mfma32(ld_w3(sw, 16 * mb + i, 0, g), dpf[0], zero);
            c = mfma32(ld_w3(sw, 16 * mb + i, 1, g), dpf[1], c);
            c = mfma32(widen(ld_w3_tail(sw, 16 * mb + i, g)), widen(dpt), c);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = (mb & 1) * 4 + j;
                const float v = (m2 >> (4 * mb + j)) & 1u ? c[j] : 0.0f;
                const float cs = row16_sum(v);
                if (i == 0) pb[kH + 16 * mb + 4 * g + j] = cs;
                dh2f[mb >> 1][e] = (__bf16)v;
            }
        }
    }
    __syncthreads();                                   // dW3 done with T_h2
#pragma unroll
    for (int mb = 0; mb < kH / 16; ++mb)
#pragma unroll
        for (int j = 0; j < 4; ++j) th2[t_idx(16 * mb + 4 * g + j, r0 + i)] = bits(dh2f[mb >> 1][(mb & 1) * 4 + j]);
    // dH1ᵀ = W2 dH2ᵀ masked by H1 > 0; db1
    bf16x8 dh1f[4];
#pragma unroll
    for (int mb = 0; mb < kH / 16; ++mb) {
        f32x4 c = zero;
#pragma unroll
        for (int kb = 0; kb < kH / 32; ++kb) c = mfma32(ld_sw(sw + hW2, 16 * mb + i, kb, i, g), dh2f[kb], c);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = (mb & 1) * 4 + j;
            const float v = (m1 >> (4 * mb + j)) & 1u ? c[j] : 0.0f;
            const float cs = row16_sum(v);
            if (i == 0) pb[16 * mb + 4 * g + j] = cs;
            dh1f[mb >> 1][e] = (__bf16)v;
        }
    }
    __syncthreads();                                   // T_dh2 complete
    // dW2 = H1ᵀ dH2: this wave's h1 blocks 2w, 2w+1
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
        const int mb = 2 * wave + mi;
        const bf16x8 a0 = ld_t(th1, 16 * mb + i, 0, i, g), a1 = ld_t(th1, 16 * mb + i, 1, i, g);
#pragma unroll
        for (int nb = 0; nb < kH / 16; ++nb) {
            f32x4 c = mfma32(a0, ld_t(th2, 16 * nb + i, 0, i, g), zero);
            c = mfma32(a1, ld_t(th2, 16 * nb + i, 1, i, g), c);
#pragma unroll
            for (int j = 0; j < 4; ++j) part[pW2 + (16 * mb + 4 * g + j) * kH + 16 * nb + i] = c[j];
        }
    }
    __syncthreads();                                   // dW2 done with T_h1
#pragma unroll
    for (int mb = 0; mb < kH / 16; ++mb)
#pragma unroll
        for (int j = 0; j < 4; ++j) th1[t_idx(16 * mb + 4 * g + j, r0 + i)] = bits(dh1f[mb >> 1][(mb & 1) * 4 + j]);
    __syncthreads();
    // dW1 = Xᵀ dH1: this wave's h blocks 2w, 2w+1
    {
        const bf16x8 a0 = ld_t(tx, i, 0, i, g), a1 = ld_t(tx, i, 1, i, g);
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            const int nb = 2 * wave + ni;
            f32x4 c = mfma32(a0, ld_t(th1, 16 * nb + i, 0, i, g), zero);
            c = mfma32(a1, ld_t(th1, 16 * nb + i, 1, i, g), c);
#pragma unroll
            for (int j = 0; j < 4; ++j) part[pW1 + (4 * g + j) * kH + 16 * nb + i] = c[j];
        }
    }
}

// LDS of the forward kernel: forward weights, one 16-row output tile per
// wavefront, biases (two workgroups per CU)
constexpr int bfP = hFwd * 2;
constexpr int bfBias = bfP + kRows * kPLd * 4;
constexpr int bFwdEnd = bfBias + (2 * kH + kP) * 4;
static_assert(2 * bFwdEnd <= 160 * 1024, "two forward workgroups per CU");

__global__ __launch_bounds__(kThreads, 2) void flow_fwd_kernel(const uint16_t* __restrict__ img,
                                                              const float* __restrict__ w, Rows rows,
                                                              float* __restrict__ params_out,
                                                              float* __restrict__ loss_out) {
    __shared__ __align__(16) unsigned char sm[bFwdEnd];
    uint16_t* sw = reinterpret_cast<uint16_t*>(sm);
    float* bias = reinterpret_cast<float*>(sm + bfBias);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    float* P = reinterpret_cast<float*>(sm + bfP) + wave * 16 * kPLd;
    copy_lds<hFwd>(sw, img, tid);
    load_bias(bias, w, tid);
    __syncthreads();
    // every wavefront walks its own 16-row tiles; no workgroup barrier below
    const int64_t n_tiles = (rows.n + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < n_tiles; tile += (int64_t)gridDim.x * 4) {
        const int64_t row0 = tile * 16;
        const s16x4 xf = x_frag(rows, row0 + i, g, nullptr, 0);
        uint32_t m1, m2;
        f32x4 pv[5];
        forward_chain(sw, bias, xf, i, g, m1, m2, pv, nullptr, nullptr, 0);
        const bool out_row = row0 + i < rows.n;
#pragma unroll
        for (int mb = 0; mb < kP / 16; ++mb) {
            *reinterpret_cast<f32x4*>(P + i * kPLd + 16 * mb + 4 * g) = pv[mb];
            if (params_out && out_row) *reinterpret_cast<f32x4*>(params_out + (row0 + i) * kP + 16 * mb + 4 * g) = pv[mb];
        }
        wave_sync();
        const int s = lane & 3, rl = lane >> 2;
        const int64_t gr = row0 + rl;
        float t = 0.0f, cnt = 0.0f, off = 0.0f, wgt = 1.0f;
        int flags = 0;
        if (gr < rows.n) {
            t = rows.t[gr];
            cnt = rows.cnt[gr];
            flags = rows.flags[gr];
            if (rows.aux) {
                off = rows.aux[2 * gr];
                wgt = rows.aux[2 * gr + 1];
            }
        }
        const float loss = quad_row_loss<false>(P + rl * kPLd, s, t, cnt, flags, off, wgt, 1.0f, nullptr, 0);
        if (loss_out && s == 0 && gr < rows.n) loss_out[gr] = loss;
        wave_sync();                                   // the next tile overwrites P
    }
}

// image positions of weight q (first, and second for W2/W3); biases have none
__device__ __forceinline__ int img_pos(int q, int& pos2) {
    pos2 = -1;
    if (q < ob1) {
        const int f = q / kH, h = q % kH;
        return hW1T + h * kIn + (((f >> 2) ^ (((h >> 3) & 1) << 1)) << 2) + (f & 3);
    }
    if (q >= oW2 && q < ob2) {
        const int in = (q - oW2) / kH, o = (q - oW2) % kH;
        pos2 = hW2 + in * kH + sw_col(o, in);
        return hW2T + o * kH + sw_col(in, o);
    }
    if (q >= oW3 && q < ob3) {
        const int h = (q - oW3) / kP, c = (q - oW3) % kP;
        pos2 = hW3 + h * kW3Ld + (c < 64 ? 32 * (c >> 5) + qperm(c & 31) : c);
        return hW3T + c * kH + sw_col(h, c);
    }
    return -1;
}

__device__ __forceinline__ void store_img(uint16_t* img, int q, float v) {
    int p2;
    const int p1 = img_pos(q, p2);
    if (p1 < 0) return;
    const uint16_t b = bits((__bf16)v);
    img[p1] = b;
    if (p2 >= 0) img[p2] = b;
}

__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ w, uint16_t* __restrict__ img) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < kNW) store_img(img, q, w[q]);
}

constexpr int kSplits = 16;           // record ranges summed in parallel by the first reduction pass
constexpr int kUnroll = 8;            // records in flight per thread

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

// First pass: tmp[split][q] = Σ over records [split·nb/S, (split+1)·nb/S) in
// order; kUnroll record loads in flight, added in record order.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float* __restrict__ partial, int nblocks,
                                                             float* __restrict__ tmp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int split = blockIdx.y;
    if (q >= kNW) return;
    int nsub;
    const int src = record_source(q, nsub);
    const int b0 = (int)((int64_t)nblocks * split / kSplits), b1 = (int)((int64_t)nblocks * (split + 1) / kSplits);
    float acc = 0.0f;
    int b = b0;
    for (; b + kUnroll <= b1; b += kUnroll) {
        float v[kUnroll][4];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const float* rec = partial + (int64_t)(b + u) * kPartW + src;
#pragma unroll
            for (int s = 0; s < 4; ++s) v[u][s] = s < nsub ? rec[s * kBiasW] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s < nsub) acc += v[u][s];
    }
    for (; b < b1; ++b) {
        const float* rec = partial + (int64_t)b * kPartW + src;
        for (int s = 0; s < nsub; ++s) acc += rec[s * kBiasW];
    }
    tmp[(int64_t)split * kNW + q] = acc;
}

// Second pass: gradient = Σ_split tmp (fixed order), then either the gradient
// or the Adam update (which also refreshes the bfloat16 image); thread 0 also
// sums the per-block losses.
__global__ __launch_bounds__(256) void reduce_kernel(const float* __restrict__ tmp, int nblocks,
                                                    const double* __restrict__ partial_loss, double loss_scale,
                                                    double* __restrict__ loss_out, float* __restrict__ grad_out,
                                                    float* __restrict__ w, float* __restrict__ m,
                                                    float* __restrict__ v, uint16_t* __restrict__ img, float lr,
                                                    float c1, float c2) {
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
    const float wq = w[q] - lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
    w[q] = wq;
    store_img(img, q, wq);
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
    out[q] = time_logpdf_row(p, t[q]);
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
    uint16_t* d_img = nullptr;           // bfloat16 weight image in the kernels' LDS layout
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

int32_t run_grad(ntf_ctx* c, const Rows& rows, float scale, hipStream_t s) {
    const int64_t blocks = (rows.n + kRows - 1) / kRows;
    const int32_t st = ensure_partials(c, blocks);
    if (st != NT_OK) return st;
    hipLaunchKernelGGL(flow_grad_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, c->d_img, c->d_w, rows, scale,
                       c->d_partial, c->d_ploss);
    return hip_status(hipGetLastError());
}

// the gradient of the mean loss of n rows (run_grad's partial records) into
// one Adam step
int32_t adam_update(ntf_ctx* c, int64_t n, float lr, double* d_loss, hipStream_t s) {
    c->step += 1;
    const float c1 = (float)(1.0 - pow(0.9, (double)c->step));
    const float c2 = (float)(1.0 - pow(0.999, (double)c->step));
    const int64_t blocks = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNW + 255) / 256, kSplits), dim3(256), 0, s, c->d_partial,
                       (int)blocks, c->d_tmp);
    hipLaunchKernelGGL(reduce_kernel, dim3((kNW + 255) / 256), dim3(256), 0, s, c->d_tmp, (int)blocks,
                       c->d_ploss, 1.0 / (double)n, d_loss, (float*)nullptr, c->d_w, c->d_m, c->d_v, c->d_img, lr,
                       c1, c2);
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
        hipMalloc(&c->d_tmp, (size_t)kSplits * kNW * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->d_img, (size_t)hImg * sizeof(uint16_t)) != hipSuccess ||
        hipMemset(c->d_img, 0, (size_t)hImg * sizeof(uint16_t)) != hipSuccess) {
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
    if (c->d_img) (void)hipFree(c->d_img);
    delete c;
}

int32_t ntf_set_weights(ntf_ctx* c, const float* weights) {
    if (!c || !weights) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    if (hipMemcpy(c->d_w, weights, kNW * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return NT_ERR_DEVICE;
    const int blocks = (kNW + 255) / 256;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks), dim3(256), 0, 0, c->d_w, c->d_img);
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
    const int64_t groups = (n + 4 * 16 - 1) / (4 * 16);               // four 16-row tiles per workgroup pass
    const int64_t blocks = groups < 2 * c->n_cu ? groups : 2 * c->n_cu; // two resident workgroups per CU
    hipLaunchKernelGGL(flow_fwd_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, (hipStream_t)stream, c->d_img,
                       c->d_w, Rows{x, t, cnt, flags, aux, n}, d_params, d_row_loss);
    return hip_status(hipGetLastError());
}

int32_t ntf_loss_grad(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                      const float* aux, int64_t n, float scale, float* d_grad, double* d_loss, void* stream) {
    if (!c || !d_grad || !rows_ok(x, t, cnt, flags, n) || n == 0) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    int32_t st = run_grad(c, Rows{x, t, cnt, flags, aux, n}, scale, s);
    if (st != NT_OK) return st;
    const int64_t blocks = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNW + 255) / 256, kSplits), dim3(256), 0, s, c->d_partial,
                       (int)blocks, c->d_tmp);
    hipLaunchKernelGGL(reduce_kernel, dim3((kNW + 255) / 256), dim3(256), 0, s, c->d_tmp, (int)blocks,
                       c->d_ploss, 1.0, d_loss, d_grad, c->d_w, c->d_m, c->d_v, c->d_img, 0.0f, 1.0f, 1.0f);
    return hip_status(hipGetLastError());
}

int32_t ntf_train_step(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                       const float* aux, int64_t n, float lr, double* d_loss, void* stream) {
    if (!c || !rows_ok(x, t, cnt, flags, n) || n == 0 || !(lr >= 0.0f)) return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    const int32_t st = run_grad(c, Rows{x, t, cnt, flags, aux, n}, 1.0f / (float)n, s);
    if (st != NT_OK) return st;
    return adam_update(c, n, lr, d_loss, s);
}

int32_t ntf_train_step_gather(ntf_ctx* c, const float* x, const float* t, const float* cnt, const uint8_t* flags,
                              const float* aux, int64_t n_rows, const int64_t* d_idx, int64_t batch, float lr,
                              double* d_loss, void* stream) {
    if (!c || !rows_ok(x, t, cnt, flags, n_rows) || n_rows == 0 || !d_idx || batch <= 0 ||
        batch > (int64_t)1 << 40 || !(lr >= 0.0f))
        return NT_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return NT_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    Rows rows{x, t, cnt, flags, aux, batch};
    rows.idx = d_idx;
    rows.n_src = n_rows;
    const int32_t st = run_grad(c, rows, 1.0f / (float)batch, s);
    if (st != NT_OK) return st;
    return adam_update(c, batch, lr, d_loss, s);
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
