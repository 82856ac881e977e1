// ntprop.hip — persistent photon-propagation kernel for gfx950 and the C ABI
// declared in include/ntprop.h.
//
// Kernel: one photon per lane, a persistent grid that strides over the
// concatenated photon index space of all sources.  Per photon: initial state
// (photon_prop_cuda.jl:435-446), then at most max_steps scattering steps
// (photon_prop_cuda.jl:449-495), unrolled by four so that each step's Philox
// words (three per step, blocks 1, 2, ... of the photon's stream, evaluated
// with the per-photon/per-wave split of PhiloxRun) are fixed registers.  Each
// step finds the first listed target sphere the step reaches: vertical
// strings of targets are culled by a conservative bound first (axis distance,
// then z range), the exact sphere test runs on the survivors.  Target and
// string records are wave-uniform data read through scalar loads; the
// spectrum knots and the FP64 coefficients of the scattering-length fit sit
// in LDS; hits are compacted per wavefront with a ballot + mbcnt prefix into
// an LDS staging area, then written to HBM as contiguous column segments with
// one 64-bit atomic per flush.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "ntprop_device.h"
#include "ntresample.h"

using namespace ntp;

namespace {

constexpr int kThreads = 256;          // 4 wavefronts per workgroup
constexpr int kWaves = kThreads / 64;
constexpr int kStageFields = 16;       // 32-bit words per staged hit
constexpr int kStageSlots = 64;        // staged hits per wavefront

// Targets are read with wave-uniform indices through the constant address
// space, which the compiler turns into scalar (SGPR) loads: one s_load per
// target per step for the whole wavefront, no VGPRs and no LDS traffic.
typedef float f4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4* const_f4_ptr;

// One source of a launch.  Its photons occupy global indices
// [offset, offset + photons); offsets are multiples of 64, so a wavefront's
// 64 consecutive indices always belong to one source, which makes the
// source, its Philox key and its output slot wave-uniform (scalar registers).
struct DevSource {                     // 96 bytes
    int32_t kind;
    uint32_t rng_index;                // Philox counter word 3
    int64_t offset;                    // first global index (multiple of 64)
    int64_t photons;
    uint64_t photon_base;              // RNG photon index of the first photon
    uint32_t key_lo, key_hi;           // XORed into the launch seed
    int32_t slot;                      // output slot
    float pos[3];
    float dir[3];
    float t0, t1, angle;
    float gamma_d, gamma_c, gamma_scale, lrad;
};
static_assert(sizeof(DevSource) == 96, "DevSource layout");
constexpr int kSrcWords = sizeof(DevSource) / 4;

struct HitCols {
    float* pos;
    float* dir;
    float* idir;
    float* wl;
    float* time;
    float* dist;
    uint16_t* module_id;
    uint32_t* source_index;
    uint64_t* photon_id;
    uint8_t* pmt_id;
    float* total_weight;
    float* tres;
    int64_t capacity;
};

// Output of one request: hit columns and counters
// ([0] hits found, [1] photons simulated, [2] sphere tests, [3] string-bound tests).
struct alignas(128) SlotDesc {
    HitCols cols;
    int64_t* counters;
};

// Hit epilogue data of one target (list order), prepared on the host.
struct EpiTarget {
    double pos[3];
    double area_acc;                   // area acceptance (detection.jl:139-146)
    double tgeo;                       // geometric arrival time for tres (ns)
    double cos_open;                   // cos of the PMT opening half-angle
    float radius;                      // m (time residual from another reference position)
    int32_t kind;
    int32_t n_pmts;
    double axes[NT_MAX_PMTS][3];       // PMT axes after the orientation
};

struct Epilogue {
    const EpiTarget* targets;
    const float* abs_x;                // absorption table (FP32, as the reference interpolates it)
    const float* abs_y;
    const double* acc_x;               // PMT acceptance table (FP64)
    const double* acc_y;
    const double* ang_x;               // angular acceptance: cos of incidence → factor (FP64), n_ang = 0: none
    const double* ang_y;
    int n_abs, n_acc, n_ang;
    float abs_len_scale;
    int enabled;
    double c_med;                      // speed of light in the medium at 800 nm (m/ns), for tgeo
};

// Target clusters.  Targets sharing (x, y) form a vertical string; a photon
// step is tested against a string's bounding cylinder first and against its
// spheres only when the step's segment comes within reach.  Per cluster, two
// float4 (one 32-byte scalar load):
//   [2c]     axis x, axis y, lowest centre z, highest centre z
//   [2c + 1] reach r0, first member, member count (int bits), 1/Δz
// with r0 = r_max + 2^-10·(r_max + 1) + 1e-3 m rounded up to FP32.  Members
// of a string are sorted by z; 1/Δz > 0 when their centres sit within 5 % of
// Δz of a uniform ladder z_lo + k·Δz (then a step only visits the members
// near its z range), 0 otherwise (every member is visited).  The
// kernel tests every string with the largest r0 (`cluster_reach`), a larger
// and so still conservative bound, computed once per step.
// Members (targets reordered by cluster, by z inside one): float4
// (x, y, z, r^2) and their list index; the hit kept is the one with the
// lowest list index, so the first listed target still wins.
struct KParams {
    const DevSource* __restrict__ sources;
    const int64_t* __restrict__ offsets;     // n_sources + 1 (multiples of 64)
    const SlotDesc* __restrict__ slots;      // null: single output (hits, counters)
    const float4* __restrict__ targets;      // x, y, z, r^2 in list order
    const uint32_t* __restrict__ module_ids; // per list index
    const float4* __restrict__ clusters;     // two records per cluster
    const float4* __restrict__ members;      // x, y, z, r^2 in cluster order
    const uint32_t* __restrict__ member_idx; // list index of each member
    const float* __restrict__ knots;
    int64_t* counters;                       // [0] hits found, [1] photons, [2] sphere tests, [3] cluster tests
    HitCols hits;
    int64_t total;
    Medium medium;
    Spectrum spectrum;
    WlPoly sca_poly;                         // scattering length over the spectrum's range
    int n_sources;
    int n_targets;
    int n_clusters;                          // 0: test every target in list order
    float cluster_reach;                     // largest cluster reach r0
    uint32_t seed_lo, seed_hi;
};

// Exact reference sphere test (photon_prop_cuda.jl:248-269) on one target.
__device__ __forceinline__ bool sphere_hit(float3 pos, float3 dir, float step, float4 T, float& d) {
    float a;
    const float b = isect_b(pos, dir, T, a);
    if (b >= 0.0f) {
        const float dd = -a - sqrtf(b);
        if ((dd > 0.0f) & (dd < step)) {
            d = dd;
            return true;
        }
    }
    return false;
}

// Kernel arguments used only off the step loop (photon initialisation, hit
// flush), and the target-list bounds of first_hit, are re-read from the
// kernarg segment where they are used, through a pointer the compiler cannot
// match with the loads at kernel entry: their SGPRs are then live only there,
// not across the whole persistent loop (held across it they spill to VGPR
// lanes and cost a readlane per use).  Only kernels whose first argument is
// KParams (propagate_kernel, serve_kernel) may call code that uses kargs().
typedef const __attribute__((address_space(4))) KParams* const_kparams_ptr;
__device__ __forceinline__ const_kparams_ptr kargs() {
#if defined(__HIP_DEVICE_COMPILE__)
    const_kparams_ptr k = (const_kparams_ptr)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return k;
#else
    return nullptr;   // host pass: never called
#endif
}

// First listed target hit by the step [pos, pos + step·dir]: returns its list
// index or -1.  Brute force over the list (n_clusters == 0) or through the
// cluster bounds; `tests` counts the sphere tests behind the bounds (the
// caller counts bound tests and brute-force tests per photon).  The bounds are inflated by m with
// m = 2^-10·(step + r_max + 1) + 1e-3 m: the FP32 discriminant of the sphere
// test can accept rays passing up to ~sqrt(5·2^-23)·|p − c| beyond the radius,
// and |p − c| ≤ step + r for any accepted hit, so a culled target can never be
// one the exact test would accept.  The bound tests are conservative
// geometry, not the reference's arithmetic, and use fused forms; their
// roundings are ~1e-7 relative, far inside the 2^-10 and 1e-3 m slack.
__device__ __forceinline__ int first_hit(const KParams& p, float3 pos, float3 dir, float step, float& dt,
                                         uint32_t& tests) {
    int best = 0x7fffffff;
    const int n_clusters = kargs()->n_clusters, n_targets = kargs()->n_targets;
    if (n_clusters == 0) {
        const const_f4_ptr tgt = (const_f4_ptr)(const void*)p.targets;
        for (int t = 0; t < n_targets; ++t) {
            const f4 tv = tgt[t];
            float d;
            if (sphere_hit(pos, dir, step, make_float4(tv.x, tv.y, tv.z, tv.w), d) && best == 0x7fffffff) {
                best = t;
                dt = d;
            }
        }
        return best == 0x7fffffff ? -1 : best;
    }
    typedef const __attribute__((address_space(4))) uint32_t* const_u32_ptr;
    const const_f4_ptr cl = (const_f4_ptr)(const void*)p.clusters;
    const const_f4_ptr mem = (const_f4_ptr)(const void*)p.members;
    const const_u32_ptr midx = (const_u32_ptr)(const void*)p.member_idx;
    const float dx = dir.x * step, dy = dir.y * step;
    const float dd = __builtin_fmaf(dx, dx, dy * dy);
    const float inv_dd = dd > 0.0f ? __builtin_amdgcn_rcpf(dd) : 0.0f;
    const float re = __builtin_fmaf(step, 0x1p-10f, p.cluster_reach);
    const float re2 = re * re;
    const float re2dd = re2 * dd;
    for (int c = 0; c < n_clusters; ++c) {
        const f4 c0 = cl[2 * c];
        const f4 c1 = cl[2 * c + 1];
        const int first = __float_as_int(c1.y);
        const int count = __float_as_int(c1.z);
        if (first < 0 || count < 0 || first + count > n_targets) continue;  // corrupt table guard
        // distance of the step's xy projection from the string axis first:
        // for strings long compared with a step it is the selective test
        const float wx = c0.x - pos.x, wy = c0.y - pos.y;
        // the step's infinite line first (|w × d|² ≤ re²·|d|²: four
        // instructions, and a segment within re implies a line within re)
        const float cr = __builtin_fmaf(wx, dy, -(wy * dx));
        if (cr * cr > re2dd) continue;
        const float t = fminf(fmaxf(__builtin_fmaf(wx, dx, wy * dy) * inv_dd, 0.0f), 1.0f);
        const float ex = __builtin_fmaf(-t, dx, wx), ey = __builtin_fmaf(-t, dy, wy);
        if (__builtin_fmaf(ex, ex, ey * ey) > re2) continue;
        const float z1 = __builtin_fmaf(dir.z, step, pos.z);
        const float zmin = fminf(pos.z, z1), zmax = fmaxf(pos.z, z1);
        if (zmax < c0.z - re || zmin > c0.w + re) continue;
        int k0 = first, k1 = first + count;
        if (c1.w > 0.0f) {
            // members sorted by z at a uniform spacing 1/c1.w: the lanes still
            // here need only the members whose centres can lie in their
            // [zmin − re, zmax + re]; the index window is widened by a quarter
            // spacing each side and the exact per-member test below still runs
            const float top = (float)(count - 1);
            const int lo = (int)fminf(fmaxf(floorf(__builtin_fmaf(zmin - re - c0.z, c1.w, -0.25f)), 0.0f), top);
            const int hi = (int)fminf(fmaxf(floorf(__builtin_fmaf(zmax + re - c0.z, c1.w, 0.25f)), 0.0f), top);
            // union over the lanes in this branch (wave-uniform)
            uint64_t act = __ballot(1);
            int ulo = count, uhi = -1;
            while (act) {
                const int l = __builtin_ctzll(act);
                act &= act - 1;
                ulo = min(ulo, __builtin_amdgcn_readlane(lo, l));
                uhi = max(uhi, __builtin_amdgcn_readlane(hi, l));
            }
            k0 = first + ulo;
            k1 = first + uhi + 1;
        }
        for (int k = k0; k < k1; ++k) {
            const f4 mv = mem[k];
            if (zmax < mv.z - re || zmin > mv.z + re) continue;
            ++tests;
            float d;
            if (sphere_hit(pos, dir, step, make_float4(mv.x, mv.y, mv.z, mv.w), d)) {
                const int li = (int)midx[k];
                if (li < best) {
                    best = li;
                    dt = d;
                }
            }
        }
    }
    return best == 0x7fffffff ? -1 : best;
}

__device__ __forceinline__ Medium load_medium(const_kparams_ptr k) {
    Medium m;
    m.a01 = k->medium.a01; m.a2 = k->medium.a2; m.a3 = k->medium.a3; m.a4 = k->medium.a4;
    m.vcs = k->medium.vcs; m.vcl = k->medium.vcl; m.g = k->medium.g;
    m.sl_fraction = k->medium.sl_fraction; m.sca_len_scale = k->medium.sca_len_scale;
    m.dispersion_mode = k->medium.dispersion_mode; m.max_steps = k->medium.max_steps;
    m.dir_mode = k->medium.dir_mode;
    m.cherenkov_angle = k->medium.cherenkov_angle;
    m.cherenkov_cos = k->medium.cherenkov_cos;
    m.cherenkov_sin = k->medium.cherenkov_sin;
    m.hg.omg = k->medium.hg.omg; m.hg.twog = k->medium.hg.twog; m.hg.omg2 = k->medium.hg.omg2;
    m.hg.k0 = k->medium.hg.k0; m.hg.inv2g = k->medium.hg.inv2g;
    return m;
}

__device__ __forceinline__ Spectrum load_spectrum(const_kparams_ptr k) {
    Spectrum s;
    s.kind = k->spectrum.kind; s.addressing = k->spectrum.addressing; s.n = k->spectrum.n;
    s.wl = k->spectrum.wl; s.lo = k->spectrum.lo; s.hi = k->spectrum.hi;
    return s;
}

typedef const __attribute__((address_space(4))) int64_t* const_i64_ptr;
typedef const __attribute__((address_space(4))) DevSource* const_src_ptr;

// Largest s with offs[s] <= g, for a wave-uniform g (scalar loads).
__device__ __forceinline__ int find_source(const_i64_ptr offs, int n, int64_t g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// fast_linear_interp(x, xs, ys), utils.jl:24-53, in the table's precision.
template <typename T>
__device__ __forceinline__ T interp_table(T x, const T* xs, const T* ys, int n) {
    const T lo = xs[0], hi = xs[n - 1];
    x = x < lo ? lo : (x > hi ? hi : x);
    if (x == lo) return ys[0];
    if (x == hi) return ys[n - 1];
    int a = 0, b = n - 1;                 // first index with xs[i] >= x lies in (a, b]
    while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (xs[mid] < x) a = mid; else b = mid;
    }
    const T along = (x - xs[a]) / (xs[b] - xs[a]);
    return ys[a] + (ys[b] - ys[a]) * along;
}

// PMT assignment, weight and time residual of one hit (see nt_epilogue);
// `ref` (served requests): the time residual's reference position instead of
// the context's, with the same FP32 distance as the host's tgeo.
__device__ __forceinline__ void hit_epilogue(const Epilogue& ep, int ti, float3 pos, float3 dir, float wl,
                                          float dist, float time, uint8_t& pmt, float& weight, float& tres,
                                          const float* ref = nullptr) {
    const EpiTarget* T = ep.targets + ti;
    // check_pmt_hit, detection.jl:88-104 (FP64 as in the reference)
    if (T->kind == NT_TARGET_SPHERE) {
        pmt = 1;
    } else {
        double r0 = (double)pos.x - T->pos[0], r1 = (double)pos.y - T->pos[1], r2 = (double)pos.z - T->pos[2];
        const double nr = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
        r0 /= nr; r1 /= nr; r2 /= nr;
        pmt = 0;
        for (int j = 0; j < T->n_pmts; ++j) {
            double c = (r0 * T->axes[j][0] + r1 * T->axes[j][1]) + r2 * T->axes[j][2];
            c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
            if (c > T->cos_open) { pmt = (uint8_t)(j + 1); break; }   // acos(c) < opening
        }
    }
    // calc_total_weight!, photon_prop_cuda.jl:850-882
    float abs_len = interp_table(wl, ep.abs_x, ep.abs_y, ep.n_abs);
    if (ep.abs_len_scale != 1.0f) abs_len *= ep.abs_len_scale;
    const double abs_w = (double)expf(-dist / abs_len);
    const double wl_acc = interp_table((double)wl, ep.acc_x, ep.acc_y, ep.n_acc);
    double ang_acc = 1.0;
    if (ep.n_ang > 0 && T->kind == NT_TARGET_MULTI_PMT && pmt > 0) {
        // cos of the angle of incidence on the PMT: −direction · PMT axis
        const double* ax = T->axes[pmt - 1];
        const double ci = -(((double)dir.x * ax[0] + (double)dir.y * ax[1]) + (double)dir.z * ax[2]);
        ang_acc = interp_table(ci, ep.ang_x, ep.ang_y, ep.n_ang);
    }
    weight = (float)(T->area_acc * wl_acc * abs_w * ang_acc);
    // calc_time_residual!, photon_prop_cuda.jl:906-917
    double tgeo = T->tgeo;
    if (ref) {
        const float dx = ref[0] - (float)T->pos[0], dy = ref[1] - (float)T->pos[1], dz = ref[2] - (float)T->pos[2];
        const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
        tgeo = (double)(d - T->radius) / ep.c_med;
    }
    tres = (float)((double)time - tgeo);
}

// Output columns and counters of a slot.
__device__ __forceinline__ HitCols slot_cols(const_kparams_ptr kp, int slot, int64_t*& counters) {
    HitCols h;
    if (kp->slots == nullptr) {
        h.pos = kp->hits.pos; h.dir = kp->hits.dir; h.idir = kp->hits.idir; h.wl = kp->hits.wl;
        h.time = kp->hits.time; h.dist = kp->hits.dist; h.module_id = kp->hits.module_id;
        h.source_index = kp->hits.source_index; h.photon_id = kp->hits.photon_id;
        h.pmt_id = kp->hits.pmt_id; h.total_weight = kp->hits.total_weight; h.tres = kp->hits.tres;
        h.capacity = kp->hits.capacity;
        counters = kp->counters;
    } else {
        const SlotDesc* sd = kp->slots + slot;
        h.pos = sd->cols.pos; h.dir = sd->cols.dir; h.idir = sd->cols.idir; h.wl = sd->cols.wl;
        h.time = sd->cols.time; h.dist = sd->cols.dist; h.module_id = sd->cols.module_id;
        h.source_index = sd->cols.source_index; h.photon_id = sd->cols.photon_id;
        h.pmt_id = sd->cols.pmt_id; h.total_weight = sd->cols.total_weight; h.tres = sd->cols.tres;
        h.capacity = sd->cols.capacity;
        counters = sd->counters;
    }
    return h;
}

// Add a wavefront's photon and test counts to a slot's counters.
__device__ __forceinline__ void commit_counts(int slot, int64_t simulated, uint32_t tests32, uint32_t ctests32,
                                              int lane) {
    uint64_t tests = tests32, ctests = ctests32;      // per-lane counts since the last commit
    for (int off = 32; off > 0; off >>= 1) {
        tests += __shfl_down(tests, off);
        ctests += __shfl_down(ctests, off);
    }
    if (lane == 0) {
        int64_t* counters;
        (void)slot_cols(kargs(), slot, counters);
        if (simulated) atomicAdd((unsigned long long*)&counters[1], (unsigned long long)simulated);
        if (tests) atomicAdd((unsigned long long*)&counters[2], (unsigned long long)tests);
        if (ctests) atomicAdd((unsigned long long*)&counters[3], (unsigned long long)ctests);
    }
}

// Write the staged hits of this wavefront to the output columns of `slot`.
__device__ __forceinline__ void flush_stage(const uint32_t* __restrict__ st, int n, int, int slot) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // the lane index afresh (mbcnt of all lanes): nothing derived from it has to
    // stay live from kernel start to the rare flushes
    const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    int64_t* counters;
    const HitCols h = slot_cols(kargs(), slot, counters);
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long*)&counters[0], (unsigned long long)n);
    base = __shfl(base, 0);
    for (int i = lane; i < n; i += 64) {
        const int64_t idx = (int64_t)base + i;
        if (idx >= h.capacity) continue;
        const float* f = reinterpret_cast<const float*>(st);
        h.pos[3 * idx + 0] = f[0 * kStageSlots + i];
        h.pos[3 * idx + 1] = f[1 * kStageSlots + i];
        h.pos[3 * idx + 2] = f[2 * kStageSlots + i];
        h.dir[3 * idx + 0] = f[3 * kStageSlots + i];
        h.dir[3 * idx + 1] = f[4 * kStageSlots + i];
        h.dir[3 * idx + 2] = f[5 * kStageSlots + i];
        h.idir[3 * idx + 0] = f[6 * kStageSlots + i];
        h.idir[3 * idx + 1] = f[7 * kStageSlots + i];
        h.idir[3 * idx + 2] = f[8 * kStageSlots + i];
        h.wl[idx] = f[9 * kStageSlots + i];
        h.time[idx] = f[10 * kStageSlots + i];
        h.dist[idx] = f[11 * kStageSlots + i];
        h.module_id[idx] = (uint16_t)st[12 * kStageSlots + i];
        if (h.source_index) h.source_index[idx] = st[13 * kStageSlots + i];
        if (h.photon_id)
            h.photon_id[idx] = (uint64_t)st[14 * kStageSlots + i] |
                               ((uint64_t)st[15 * kStageSlots + i] << 32);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Wavefronts per SIMD the kernels are built for: the propagation kernel at
// seven (72 VGPRs, no scratch; 2 % faster than six on one string, 4–5 % on a
// 7-string cluster, eight spill), the serving kernel at six (seven spill).
#ifndef NTP_MIN_WAVES_PER_SIMD
#define NTP_MIN_WAVES_PER_SIMD 7
#endif
#ifndef NTP_SERVE_WAVES_PER_SIMD
#define NTP_SERVE_WAVES_PER_SIMD 6
#endif

// State of one photon at the end of its propagation (the hit columns when `hit`).
struct PhotonOut {
    float3 pos, dir, idir;
    float wl, time, dist;
    uint32_t mid;
    uint32_t steps;                    // steps that tested the targets (the hit step included)
    uint32_t tests;                    // sphere tests behind the string bounds
    bool hit;
};

constexpr int kDirMode = 1;       // DIRM bit: NT_DIR_FP64 (0) or NT_DIR_FP32_REF (1)
constexpr int kDirSL = 2;         // DIRM bit: the medium has a Simplified-Liu share

// One photon of source S from initialisation to its hit or its last step
// (photon_prop_cuda.jl:133-188, 435-506).
template <int KIND, int DIRM, class SrcPtr>
__device__ __forceinline__ void simulate_photon(const KParams& p, const_kparams_ptr kp, SrcPtr S,
                                                const float* s_knots, const double* s_poly, uint64_t pid,
                                                uint32_t sidx, PhotonOut& o) {
    uint32_t tests = 0;
    const Medium& m = p.medium;
    const Medium mi = load_medium(kp);      // initialisation-only fields
    const Rng rng{p.seed_lo ^ S->key_lo, p.seed_hi ^ S->key_hi, (uint32_t)pid, (uint32_t)(pid >> 32), sidx};
    float3 pos, dir, idir;
    float wl, time, dist = 0.0f;
    uint32_t mid = 0;
    bool hit = false;

    // initialize_photon_state, photon_prop_cuda.jl:133-188
    const uint4 r0 = rng.block(0u);
    wl = sample_spectrum(load_spectrum(kp), s_knots, u01(r0.x));
    pos = make_float3(S->pos[0], S->pos[1], S->pos[2]);
    const float3 sdir = make_float3(S->dir[0], S->dir[1], S->dir[2]);
    time = S->t0;
    switch (KIND < 0 ? S->kind : KIND) {
        case NT_SRC_POINTLIKE_ISOTROPIC:
            dir = iso_dir(u01(r0.y), u01(r0.z));
            break;
        case NT_SRC_POINTLIKE_TIME_RANGE:
            dir = iso_dir(u01(r0.y), u01(r0.z));
            time = S->t0 + u01(r0.w) * (S->t1 - S->t0);
            break;
        case NT_SRC_AXICONE:
            dir = rot_from_ez_mode(sdir, sph_to_cart(S->angle, 0.0f + u01(r0.z) * kTwoPi), DIRM & kDirMode);
            break;
        case NT_SRC_PENCIL:
            dir = sdir;
            break;
        case NT_SRC_EXTENDED_CHERENKOV: {
            const float lp = rand_gamma(S->gamma_d, S->gamma_c, S->gamma_scale, rng, nullptr) * S->lrad;
            pos = make_float3(pos.x + lp * sdir.x, pos.y + lp * sdir.y, pos.z + lp * sdir.z);
            time = S->t0 + lp / kCVacMNs;
            dir = (DIRM & kDirMode) == NT_DIR_FP32_REF ? cherenkov_dir32(wl, u01(r0.y), u01(r0.z), u01(r0.w), sdir, mi)
                                                       : cherenkov_dir(wl, u01(r0.y), u01(r0.z), u01(r0.w), sdir, mi);
            break;
        }
        default:  // NT_SRC_POINTLIKE_CHERENKOV
            dir = (DIRM & kDirMode) == NT_DIR_FP32_REF ? cherenkov_dir32(wl, u01(r0.y), u01(r0.z), u01(r0.w), sdir, mi)
                                                       : cherenkov_dir(wl, u01(r0.y), u01(r0.z), u01(r0.w), sdir, mi);
            break;
    }
    idir = dir;
    const float sca_len = scattering_length_fit(wl, &kp->sca_poly, s_poly, mi);
    const float c_grp = group_velocity(wl, mi);

    // Scattering steps read the Philox stream of blocks 1, 2, ... three
    // words per step (four steps per three blocks): step, HG/SL uniform
    // and azimuth from the words' top 24 bits, the mixture uniform from
    // their low bytes.  The loop is unrolled by four so that every
    // step's words are fixed registers; the block index is wave-uniform.
    // A step returns true when the photon hit a target (state at the hit).
    // −ln u · L_sca = log2 u · (−ln 2 · L_sca): the hardware log2 of u ∈ (0, 1]
    // times a per-photon constant (≤ 2 ulp, like −ln u · L_sca itself)
    const float step_scale = -0.693147182f * sca_len;
    auto scatter_step = [&](uint32_t wx, uint32_t wy, uint32_t wz) -> bool {
        const float step = __builtin_amdgcn_logf(u01_oc(wx)) * step_scale;
        // first listed target hit within this step (photon_prop_cuda.jl:457-468)
        float dt = 0.0f;
        const int hit_t = first_hit(p, pos, dir, step, dt, tests);
        if (hit_t < 0) {
            pos = make_float3(pos.x + dir.x * step, pos.y + dir.y * step, pos.z + dir.z * step);
            dist += step;
            // the mixture uniform matters only with a Simplified-Liu share:
            // with none, u = 1 never selects it, and the instances for media
            // without one (kDirSL clear) leave the uniform and the test out
            constexpr bool kSL = (DIRM & kDirSL) != 0;
            float u_mix = 1.0f;
            if (kSL && m.sl_fraction > 0.0f)
                u_mix = u01(((wx & 0xFFu) << 8) | ((wy & 0xFFu) << 16) | ((wz & 0xFFu) << 24));
            if ((DIRM & kDirMode) == NT_DIR_FP32_REF) {
                dir = update_direction32<kSL>(dir, u01(wy), u01(wz), u_mix, m);
            } else {
                const const_kparams_ptr kh = kargs();
                const HgConst hg{kh->medium.hg.omg, kh->medium.hg.twog, kh->medium.hg.omg2, kh->medium.hg.k0,
                                 kh->medium.hg.inv2g};
                dir = update_direction<kSL>(dir, u01(wy), u01(wz), u_mix, m, hg);
            }
            return false;
        }
        pos = make_float3(pos.x + dir.x * dt, pos.y + dir.y * dt, pos.z + dir.z * dt);
        dist += dt;
        // Σ step/c_grp + dt/c_grp of photon_prop_cuda.jl:473,481 taken at
        // once: the same sum up to FP32 rounding of the partial sums
        time += dist / c_grp;
        mid = p.module_ids[hit_t];
        hit = true;
        return true;
    };
    const int max_steps = m.max_steps;
    PhiloxRun run;
    run.init(rng.k0, rng.k1, rng.plo, rng.phi, rng.src);
    int k = 0;
    for (uint32_t blk = 1u; k < max_steps; blk += 3u) {
        const uint4 b0 = run.block(blk);
        if (scatter_step(b0.x, b0.y, b0.z) || ++k >= max_steps) break;
        const uint4 b1 = run.block(blk + 1u);
        if (scatter_step(b0.w, b1.x, b1.y) || ++k >= max_steps) break;
        const uint4 b2 = run.block(blk + 2u);
        if (scatter_step(b1.z, b1.w, b2.x) || ++k >= max_steps) break;
        if (scatter_step(b2.y, b2.z, b2.w) || ++k >= max_steps) break;
    }
    // every step tested all strings (or all targets): k steps without
    // a hit, one more if the photon hit
    o.steps = (uint32_t)k + (hit ? 1u : 0u);
    o.tests = tests;
    o.pos = pos;
    o.dir = dir;
    o.idir = idir;
    o.wl = wl;
    o.time = time;
    o.dist = dist;
    o.mid = mid;
    o.hit = hit;
}

// Write this lane's hit at its compacted position (mask: the wavefront's
// hits, staged: hits already in the stage).
__device__ __forceinline__ void stage_write(uint32_t* st, int staged, uint64_t mask, const PhotonOut& o,
                                            uint32_t sidx, uint64_t pid) {
    const int at = staged + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    float* f = reinterpret_cast<float*>(st);
    f[0 * kStageSlots + at] = o.pos.x;
    f[1 * kStageSlots + at] = o.pos.y;
    f[2 * kStageSlots + at] = o.pos.z;
    f[3 * kStageSlots + at] = o.dir.x;
    f[4 * kStageSlots + at] = o.dir.y;
    f[5 * kStageSlots + at] = o.dir.z;
    f[6 * kStageSlots + at] = o.idir.x;
    f[7 * kStageSlots + at] = o.idir.y;
    f[8 * kStageSlots + at] = o.idir.z;
    f[9 * kStageSlots + at] = o.wl;
    f[10 * kStageSlots + at] = o.time;
    f[11 * kStageSlots + at] = o.dist;
    st[12 * kStageSlots + at] = o.mid;
    st[13 * kStageSlots + at] = sidx;
    st[14 * kStageSlots + at] = (uint32_t)pid;
    st[15 * kStageSlots + at] = (uint32_t)(pid >> 32);
}

// KIND >= 0: every source of the launch has that kind (smaller code, fewer
// registers); KIND < 0: mixed kinds, dispatched per photon.  DIRM: the
// medium's direction arithmetic (NT_DIR_FP64 or NT_DIR_FP32_REF) in bit 0,
// kDirSL when the medium has a Simplified-Liu share (without it the step loop
// carries no mixture test: +5 % on the cascade).
template <int KIND, int DIRM>
__global__ __launch_bounds__(kThreads, NTP_MIN_WAVES_PER_SIMD) void propagate_kernel(KParams p) {
    extern __shared__ float lds[];
    float* s_knots = lds;                                          // n_knots (padded to 4)
    uint32_t* s_stage = reinterpret_cast<uint32_t*>(lds + ((p.spectrum.n + 3) & ~3));

    __shared__ double s_poly[kScaPolyCoefs];                       // scattering-length fit coefficients
    for (int i = threadIdx.x; i < p.spectrum.n; i += kThreads) s_knots[i] = p.knots[i];
    if (threadIdx.x < kScaPolyCoefs) s_poly[threadIdx.x] = p.sca_poly.c[threadIdx.x];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    // wave-uniform by construction; readfirstlane lets the compiler keep the
    // photon base, source range and counts derived from it in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* st = s_stage + wave * (kStageFields * kStageSlots);
    int staged = 0;                     // hits in the stage (wave-uniform)
    int stage_slot = 0;                 // their output slot
    int count_slot = -1;                // slot the counts below belong to
    int64_t simulated = 0;              // wave-uniform
    uint32_t tests = 0, ctests = 0;     // per lane, since the last commit
    bool commit_due = false;            // wave-uniform: a lane's count reached 2^31

    const int64_t stride = (int64_t)gridDim.x * kThreads;
    int src = 0;
    int64_t src_lo = 0, src_hi = 0;     // global index range of `src`

    for (int64_t base = (int64_t)blockIdx.x * kThreads + wave * 64; base < p.total; base += stride) {
        // the wavefront's source (offsets are multiples of 64)
        const const_kparams_ptr kp = kargs();
        if (base < src_lo || base >= src_hi) {
            const const_i64_ptr offs = (const_i64_ptr)(const void*)kp->offsets;
            src = find_source(offs, kp->n_sources, base);
            src_lo = offs[src];
            src_hi = offs[src + 1];
        }
        const const_src_ptr S = (const_src_ptr)(const void*)kp->sources + src;
        const int slot = S->slot;
        if (slot != count_slot || commit_due) {
            if (count_slot >= 0) commit_counts(count_slot, simulated, tests, ctests, lane);
            simulated = tests = ctests = 0;
            count_slot = slot;
        }
        const int64_t pid_local = base - src_lo + lane;
        const bool active = pid_local < S->photons;
        const uint32_t sidx = S->rng_index;
        const uint64_t pid = S->photon_base + (uint64_t)pid_local;
        PhotonOut o;
        o.wl = o.time = o.dist = 0.0f;
        o.mid = o.steps = o.tests = 0;
        o.hit = false;
        if (active) {
            simulate_photon<KIND, DIRM>(p, kp, S, s_knots, s_poly, pid, sidx, o);
            tests += o.tests;
            if (p.n_clusters)
                ctests += o.steps * (uint32_t)p.n_clusters;
            else
                tests += o.steps * (uint32_t)p.n_targets;
        }
        simulated += __popcll(__ballot(active));
        // The per-lane test counts are 32-bit: one photon adds less than
        // (NT_MAX_STEPS + 1)·NT_MAX_TARGETS = 2^31, so they are committed at
        // the next photon once any lane has reached 2^31 and never wrap.
        commit_due = __ballot(((tests | ctests) >> 31) != 0u) != 0;
        // wavefront compaction of this round's hits into the LDS stage
        const uint64_t mask = __ballot(o.hit);
        if (mask) {
            const int cnt = __popcll(mask);
            if (staged > 0 && (staged + cnt > kStageSlots || slot != stage_slot)) {
                flush_stage(st, staged, lane, stage_slot);
                staged = 0;
            }
            stage_slot = slot;
            if (o.hit) stage_write(st, staged, mask, o, sidx, pid);
            staged += cnt;
        }
    }
    if (staged) flush_stage(st, staged, lane, stage_slot);
    if (count_slot >= 0) commit_counts(count_slot, simulated, tests, ctests, lane);
}

// ---------------------------------------------------------------------------
// Serving from a device work list (nt_server_*, NT_SERVER_WORKLIST)
// ---------------------------------------------------------------------------
// The host writes work items (photon ranges of one source of one request:
// DevSource records with `photons` ≤ 2^30 and `slot` = the request's slot)
// and slot records into rings in fine-grained host memory and advances
// `n_pub` in the control block there; no copy engine or kernel launch is
// needed to hand work over.  One wavefront at a time (a device lock) copies
// newly published items into a device copy of the ring and refreshes the
// device copies of n_pub (`n_ready`), `stop` and `epoch`: when a wavefront
// finds nothing to claim (at most every kIdleFetchTicks) and, while all are
// busy, every kBusyFetchTicks.  Claims read device memory only, so the host
// memory traffic does not grow with the number of wavefronts.  Wavefronts of
// a persistent kernel claim tickets of the oldest item that still has
// unclaimed photons: the cursor names that item (only below n_ready, so no
// claim is ever left unprocessed) and a claim is one fetch_add on the item's
// `claimed` count, which cannot fail; an add past the item's end is ignored
// and the cursor moves on with one CAS per item.  Small items come first
// (see claim_ticket).  Tickets are the item's
// photons over the number of wavefronts (multiples of 64, between 64 and
// `ticket`), so one request spreads over every wavefront.  They propagate
// them exactly as
// propagate_kernel propagates a source, write their hits straight into the
// request's pinned host columns, and add the ticket to the slot's finished
// photons; the wavefront that completes a request copies its counters to
// host memory and then writes its sequence number into done[slot].  Every
// ring entry fills whole cache lines and is not reused while a kernel
// instance runs, so no cache can hold a stale copy of it.  A wavefront
// without work sleeps and polls.  The instance leaves as a whole: when the
// host sets `stop`, or when one wavefront has seen neither work nor a new
// `epoch` for `idle_ticks` (100 MHz ticks) and raises `exit_gen` to the
// instance's launch generation; every wavefront then leaves at its next
// claim (claimed tickets are always finished), so work published meanwhile
// waits for a fresh full instance instead of a few remaining wavefronts.
constexpr int kWorkSlots = 1 << 14;        // slot ring (requests in flight per kernel instance)
constexpr int kWorkItems = 1 << 16;        // item ring

struct WorkCtrl {
    uint32_t n_pub;                        // items published (absolute count)
    uint32_t stop;                         // 1: wavefronts exit at their next claim
    uint32_t epoch;                        // advanced by the host while requests are pending
    uint32_t pad;
    uint32_t done[kWorkSlots];             // per slot: sequence number of the completed request
};

struct alignas(64) SlotExtra {
    int64_t total;                         // photons of the request
    int64_t* host_counters;                // where the request's final counters go (host memory)
    float ref[3];                          // time-residual reference position (epilogue)
    int32_t epilogue;                      // 1: fill pmt_id, total_weight, tres
    uint32_t seq;                          // written to done[slot] on completion
};

struct alignas(128) WorkItem {
    DevSource src;
    unsigned long long claimed;            // device ring: photons claimed (0 when published)
};
static_assert(sizeof(WorkItem) == 128 && offsetof(WorkItem, claimed) == sizeof(DevSource), "WorkItem layout");

// Device copy of the host's control words (see fetch_items).
struct alignas(128) DevWork {
    uint32_t n_ready;                      // items copied to the device ring (absolute count)
    uint32_t stop;
    uint32_t epoch;
    uint32_t lock;                         // 1 while a wavefront fetches
    uint64_t last_fetch;                   // s_memrealtime of the last fetch
    uint32_t exit_gen;                     // instances of this generation and older leave
};
constexpr uint32_t kFetchMax = 1024;       // items copied per fetch
constexpr int64_t kSmallItem = int64_t(1) << 21;   // items claimed first (photons)
constexpr uint64_t kIdleFetchTicks = 1000;     // 10 µs
constexpr uint64_t kBusyFetchTicks = 10000;    // 100 µs

struct SrvParams {
    const WorkItem* items;                 // item ring (kWorkItems), host memory
    WorkItem* ditems;                      // its device copy
    DevWork* dw;
    const SlotExtra* extra;                // per slot (kWorkSlots), host memory
    int64_t* slot_done;                    // photons finished per slot
    uint32_t* cursor;                      // per class (small, large): oldest item that may have unclaimed photons
    unsigned long long* busy_ticks;        // Σ ticket durations (100 MHz ticks)
    WorkCtrl* ctrl;
    Epilogue ep;
    const int32_t* target_of_module;       // module id → target list index (epilogue)
    uint32_t ticket;                       // largest ticket (multiple of 64)
    uint32_t n_waves;                      // wavefronts of the kernel
    uint32_t idle_ticks;
    uint32_t gen;                          // launch generation of this instance (from 1)
};

// Copy the items published since the last fetch (at most kFetchMax) from
// the host ring to the device ring, then n_ready, stop and epoch (whole
// wavefront; returns without work unless it takes the lock and the last fetch
// is at least `min_ticks` old).  Returns the number of items copied.
__device__ uint32_t fetch_items(const SrvParams& sp, int lane, uint64_t min_ticks) {
    uint32_t got = 0;
    if (lane == 0) {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        const uint64_t last = __hip_atomic_load(&sp.dw->last_fetch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t free_lock = 0;
        if (now - last >= min_ticks &&
            __hip_atomic_load(&sp.dw->lock, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u &&
            __hip_atomic_compare_exchange_strong(&sp.dw->lock, &free_lock, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            got = 1;
    }
    if (!__builtin_amdgcn_readfirstlane(got)) return 0;
    uint32_t npub = 0, stop = 0, epoch = 0, have = 0;
    if (lane == 0) {
        stop = __hip_atomic_load(&sp.ctrl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        epoch = __hip_atomic_load(&sp.ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        npub = __hip_atomic_load(&sp.ctrl->n_pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        have = __hip_atomic_load(&sp.dw->n_ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    npub = __builtin_amdgcn_readfirstlane(npub);
    stop = __builtin_amdgcn_readfirstlane(stop);
    epoch = __builtin_amdgcn_readfirstlane(epoch);
    have = __builtin_amdgcn_readfirstlane(have);
    const int32_t pending = (int32_t)(npub - have);
    const uint32_t n = pending > 0 ? min((uint32_t)pending, kFetchMax) : 0u;
    // the 7 16-byte words of the source and the (zero) claimed count in each
    // 128-byte entry
    constexpr int kWords = (sizeof(DevSource) + 16) / 16, kEntry = sizeof(WorkItem) / 16;
    const uint4* src = reinterpret_cast<const uint4*>(sp.items);
    uint4* dst = reinterpret_cast<uint4*>(sp.ditems);
    const uint32_t total = n * kWords;
    for (uint32_t w = lane; w < total; w += 64) {
        const uint32_t at = ((have + w / kWords) & (kWorkItems - 1)) * kEntry + w % kWords;
        dst[at] = src[at];
    }
    // the copied entries are visible device-wide before n_ready
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) {
        __hip_atomic_store(&sp.dw->stop, stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&sp.dw->epoch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&sp.dw->n_ready, have + n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&sp.dw->last_fetch, __builtin_amdgcn_s_memrealtime(), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&sp.dw->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    return n;
}

// Claim a ticket (one lane): 0 claimed, 1 nothing to claim now, 2 stop.
// Items of at most kSmallItem photons are claimed first: each class has its
// own cursor over the ring and steps over the other class's items, so a small
// request does not wait behind the rest of a large one.
__device__ int claim_ticket(const SrvParams& sp, uint32_t& item, uint32_t& off, uint32_t& take) {
    int cls = 0;
    for (int tries = 0; tries < 64; ++tries) {
        if (__hip_atomic_load(&sp.dw->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            (int32_t)(__hip_atomic_load(&sp.dw->exit_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - sp.gen) >= 0)
            return 2;
        const uint32_t nready = __hip_atomic_load(&sp.dw->n_ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t k = __hip_atomic_load(sp.cursor + cls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int32_t)(nready - k) <= 0) {
            if (cls == 1) return 1;
            cls = 1;                       // no small item left: the large ones
            continue;
        }
        WorkItem& it = sp.ditems[k & (kWorkItems - 1)];
        const int64_t ph = it.src.photons;
        if ((ph > kSmallItem) == (cls == 0)) {
            // the other class's item: step over it
            __hip_atomic_compare_exchange_strong(sp.cursor + cls, &k, k + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        const int64_t want = min((int64_t)sp.ticket, max((int64_t)64, (ph / sp.n_waves + 63) & ~(int64_t)63));
        const int64_t o = (int64_t)__hip_atomic_fetch_add(&it.claimed, (unsigned long long)want, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
        if (o < ph) {
            item = k;
            off = (uint32_t)o;
            take = (uint32_t)min(want, ph - o);
            return 0;
        }
        // item exhausted: advance the cursor (whoever wins)
        __hip_atomic_compare_exchange_strong(sp.cursor + cls, &k, k + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
    return 1;
}

// Write the staged hits of this wavefront to the columns of `slot`, with the
// epilogue columns (pmt_id, total_weight, tres) when the request asked for them.
__device__ __forceinline__ void flush_stage_srv(const uint32_t* __restrict__ st, int n, int slot,
                                                const SrvParams& sp) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    int64_t* counters;
    const HitCols h = slot_cols(kargs(), slot, counters);
    const SlotExtra& x = sp.extra[slot];
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long*)&counters[0], (unsigned long long)n);
    base = __shfl(base, 0);
    const float* f = reinterpret_cast<const float*>(st);
    for (int i = lane; i < n; i += 64) {
        const int64_t idx = (int64_t)base + i;
        if (idx >= h.capacity) continue;
        const float3 pos = make_float3(f[0 * kStageSlots + i], f[1 * kStageSlots + i], f[2 * kStageSlots + i]);
        const float3 dir = make_float3(f[3 * kStageSlots + i], f[4 * kStageSlots + i], f[5 * kStageSlots + i]);
        const float wl = f[9 * kStageSlots + i], time = f[10 * kStageSlots + i], dist = f[11 * kStageSlots + i];
        const uint32_t mid = st[12 * kStageSlots + i];
        h.pos[3 * idx + 0] = pos.x;
        h.pos[3 * idx + 1] = pos.y;
        h.pos[3 * idx + 2] = pos.z;
        h.dir[3 * idx + 0] = dir.x;
        h.dir[3 * idx + 1] = dir.y;
        h.dir[3 * idx + 2] = dir.z;
        h.idir[3 * idx + 0] = f[6 * kStageSlots + i];
        h.idir[3 * idx + 1] = f[7 * kStageSlots + i];
        h.idir[3 * idx + 2] = f[8 * kStageSlots + i];
        h.wl[idx] = wl;
        h.time[idx] = time;
        h.dist[idx] = dist;
        h.module_id[idx] = (uint16_t)mid;
        h.source_index[idx] = st[13 * kStageSlots + i];
        h.photon_id[idx] = (uint64_t)st[14 * kStageSlots + i] | ((uint64_t)st[15 * kStageSlots + i] << 32);
        if (x.epilogue) {
            const int ti = sp.target_of_module[mid & 0xFFFFu];
            uint8_t pmt = 0;
            float w = 0.0f, tres = 0.0f;
            if (ti >= 0) hit_epilogue(sp.ep, ti, pos, dir, wl, dist, time, pmt, w, tres, x.ref);
            h.pmt_id[idx] = pmt;
            h.total_weight[idx] = w;
            h.tres[idx] = tres;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int DIRM>
__global__ __launch_bounds__(kThreads, NTP_SERVE_WAVES_PER_SIMD) void serve_kernel(KParams p, SrvParams sp) {
    extern __shared__ float lds[];
    float* s_knots = lds;
    uint32_t* s_stage = reinterpret_cast<uint32_t*>(lds + ((p.spectrum.n + 3) & ~3));
    __shared__ double s_poly[kScaPolyCoefs];
    for (int i = threadIdx.x; i < p.spectrum.n; i += kThreads) s_knots[i] = p.knots[i];
    if (threadIdx.x < kScaPolyCoefs) s_poly[threadIdx.x] = p.sca_poly.c[threadIdx.x];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* st = s_stage + wave * (kStageFields * kStageSlots);
    __shared__ __align__(16) uint32_t s_src[(kThreads / 64) * kSrcWords];
    uint32_t last_epoch = 0;
    uint64_t idle_since = 0;
    bool idle = false;
    for (;;) {
        uint32_t item = 0, off = 0, take = 0;
        int state = 0;
        if (lane == 0) state = claim_ticket(sp, item, off, take);
        state = __builtin_amdgcn_readfirstlane(state);
        if (state == 2) break;
        if (state == 1) {
            if (fetch_items(sp, lane, kIdleFetchTicks)) continue;
            uint32_t e = 0;
            if (lane == 0) e = __hip_atomic_load(&sp.dw->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            e = __builtin_amdgcn_readfirstlane(e);
            const uint64_t now = __builtin_amdgcn_s_memrealtime();
            if (!idle || e != last_epoch) {
                idle = true;
                last_epoch = e;
                idle_since = now;
            } else if (now - idle_since > sp.idle_ticks) {
                if (lane == 0)
                    __hip_atomic_fetch_max(&sp.dw->exit_gen, sp.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(127);
            continue;
        }
        idle = false;
        const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
        item = __builtin_amdgcn_readfirstlane(item);
        off = __builtin_amdgcn_readfirstlane(off);
        take = __builtin_amdgcn_readfirstlane(take);
        const const_kparams_ptr kp = kargs();
        // The item's source record is copied once per ticket (one word per lane)
        // into this wavefront's LDS slot, which the photon loop reads.  Earlier
        // reads of the slot are ordered before the rewrite.
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(&sp.ditems[item & (kWorkItems - 1)].src);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < kSrcWords) s_src[wave * kSrcWords + lane] = sw[lane];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const DevSource* S = reinterpret_cast<const DevSource*>(s_src + wave * kSrcWords);
        const int slot = S->slot;
        const uint32_t sidx = S->rng_index;
        int staged = 0;
        int64_t simulated = 0;
        uint32_t tests = 0, ctests = 0;
        for (uint32_t g = 0; g < take; g += 64) {
            const int64_t pid_local = (int64_t)off + g + lane;
            const bool active = pid_local < (int64_t)off + take;
            const uint64_t pid = S->photon_base + (uint64_t)pid_local;
            PhotonOut o;
            o.wl = o.time = o.dist = 0.0f;
            o.mid = o.steps = o.tests = 0;
            o.hit = false;
            if (active) {
                simulate_photon<-1, DIRM>(p, kp, S, s_knots, s_poly, pid, sidx, o);
                tests += o.tests;
                if (p.n_clusters)
                    ctests += o.steps * (uint32_t)p.n_clusters;
                else
                    tests += o.steps * (uint32_t)p.n_targets;
            }
            simulated += __popcll(__ballot(active));
            if (__ballot(((tests | ctests) >> 31) != 0u)) {      // 32-bit per-lane counts never wrap
                commit_counts(slot, simulated, tests, ctests, lane);
                simulated = tests = ctests = 0;
            }
            const uint64_t mask = __ballot(o.hit);
            if (mask) {
                const int cnt = __popcll(mask);
                if (staged + cnt > kStageSlots) {
                    flush_stage_srv(st, staged, slot, sp);
                    staged = 0;
                }
                if (o.hit) stage_write(st, staged, mask, o, sidx, pid);
                staged += cnt;
            }
        }
        if (staged) flush_stage_srv(st, staged, slot, sp);
        commit_counts(slot, simulated, tests, ctests, lane);
        // every lane's hit stores are visible device- and host-wide before the count
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        if (lane == 0) {
            const SlotExtra& x = sp.extra[slot];
            const int64_t done =
                __hip_atomic_fetch_add(sp.slot_done + slot, (int64_t)take, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_SYSTEM) +
                (int64_t)take;
            if (done == x.total) {
                // the request's counters (device atomics) to its host result, then the mark
                int64_t* counters;
                (void)slot_cols(kp, slot, counters);
                for (int q = 0; q < 4; ++q)
                    x.host_counters[q] = __hip_atomic_load(counters + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&sp.ctrl->done[slot], x.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            atomicAdd(sp.busy_ticks, (unsigned long long)(__builtin_amdgcn_s_memrealtime() - t_start));
        }
        (void)fetch_items(sp, lane, kBusyFetchTicks);
    }
}

typedef void (*PropagateFn)(KParams);
constexpr int kKernelVariants = 7;     // one per source kind + mixed

template <int DIRM>
PropagateFn propagate_variant_m(int v) {
    switch (v) {
        case NT_SRC_POINTLIKE_ISOTROPIC: return propagate_kernel<NT_SRC_POINTLIKE_ISOTROPIC, DIRM>;
        case NT_SRC_POINTLIKE_TIME_RANGE: return propagate_kernel<NT_SRC_POINTLIKE_TIME_RANGE, DIRM>;
        case NT_SRC_AXICONE: return propagate_kernel<NT_SRC_AXICONE, DIRM>;
        case NT_SRC_PENCIL: return propagate_kernel<NT_SRC_PENCIL, DIRM>;
        case NT_SRC_EXTENDED_CHERENKOV: return propagate_kernel<NT_SRC_EXTENDED_CHERENKOV, DIRM>;
        case NT_SRC_POINTLIKE_CHERENKOV: return propagate_kernel<NT_SRC_POINTLIKE_CHERENKOV, DIRM>;
        default: return propagate_kernel<-1, DIRM>;
    }
}

// the instance for a source kind, the direction arithmetic and whether the
// medium has a Simplified-Liu share (the instances without one leave its
// mixture uniform and test out of the step loop)
PropagateFn propagate_variant(int v, int dir_mode, bool sl) {
    if (dir_mode == NT_DIR_FP32_REF)
        return sl ? propagate_variant_m<NT_DIR_FP32_REF | kDirSL>(v) : propagate_variant_m<NT_DIR_FP32_REF>(v);
    return sl ? propagate_variant_m<NT_DIR_FP64 | kDirSL>(v) : propagate_variant_m<NT_DIR_FP64>(v);
}

typedef void (*ServeFn)(KParams, SrvParams);
ServeFn serve_variant(int dir_mode, bool sl) {
    if (dir_mode == NT_DIR_FP32_REF)
        return sl ? serve_kernel<NT_DIR_FP32_REF | kDirSL> : serve_kernel<NT_DIR_FP32_REF>;
    return sl ? serve_kernel<NT_DIR_FP64 | kDirSL> : serve_kernel<NT_DIR_FP64>;
}

__global__ void clear_counters(int64_t* c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

// Hit epilogue over the hits the propagation kernel stored (count read from
// its counter, so no host round trip): one thread per hit.
__global__ __launch_bounds__(256) void epilogue_kernel(Epilogue ep, HitCols h, const int64_t* counters,
                                                       const int32_t* __restrict__ target_of_module) {
    const int64_t n = min(counters[0], h.capacity);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int ti = target_of_module[h.module_id[i]];
        uint8_t pmt = 0;
        float w = 0.0f, tres = 0.0f;
        if (ti >= 0)
            hit_epilogue(ep, ti, make_float3(h.pos[3 * i], h.pos[3 * i + 1], h.pos[3 * i + 2]),
                         make_float3(h.dir[3 * i], h.dir[3 * i + 1], h.dir[3 * i + 2]), h.wl[i], h.dist[i],
                         h.time[i], pmt, w, tres);
        if (h.pmt_id) h.pmt_id[i] = pmt;
        if (h.total_weight) h.total_weight[i] = w;
        if (h.tres) h.tres[i] = tres;
    }
}

// ---------------------------------------------------------------------------
// Device-function evaluation kernel (test hook)
// ---------------------------------------------------------------------------
__global__ void eval_kernel(int fn, const float* __restrict__ in, int iw, float* __restrict__ out,
                            int ow, int64_t n, Medium m, Spectrum s, const float* __restrict__ knots,
                            WlPoly sca_poly) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* x = in + i * iw;
    float* y = out + i * ow;
    switch (fn) {
        case NT_FN_INTERSECT: {
            const float3 pos = make_float3(x[0], x[1], x[2]);
            const float3 dir = make_float3(x[3], x[4], x[5]);
            const float4 T = make_float4(x[7], x[8], x[9], x[10] * x[10]);
            float a;
            const float b = isect_b(pos, dir, T, a);
            float isec = 0.0f, d = NAN;
            if (b >= 0.0f) {
                const float dd = -a - sqrtf(b);
                if ((dd > 0.0f) & (dd < x[6])) {
                    isec = 1.0f;
                    d = dd;
                }
            }
            y[0] = isec;
            y[1] = d;
            break;
        }
        case NT_FN_HG_COS: y[0] = m.dir_mode == NT_DIR_FP32_REF ? hg_cos32(x[0], x[1]) : hg_cos(x[0], x[1]); break;
        case NT_FN_SL_COS: y[0] = sl_cos(x[0], x[1]); break;
        case NT_FN_ROT_FROM_EZ: {
            const float3 v = rot_from_ez_mode(make_float3(x[0], x[1], x[2]), make_float3(x[3], x[4], x[5]), m.dir_mode);
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_SPH_TO_CART: {
            const float3 v = sph_to_cart(x[0], x[1]);
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_ISO_DIR: {
            const float3 v = iso_dir(x[0], x[1]);
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_REFRACTIVE_INDEX: y[0] = refractive_index(x[0], m); break;
        case NT_FN_GROUP_VELOCITY: y[0] = group_velocity(x[0], m); break;
        case NT_FN_SCA_LEN: y[0] = scattering_length(x[0], m); break;
        case NT_FN_SCA_LEN_FIT: y[0] = scattering_length_fit(x[0], &sca_poly, m); break;
        case NT_FN_CHERENKOV_DIR: {
            const float3 v = cherenkov_dir_mode(x[0], x[1], x[2], x[3], make_float3(x[4], x[5], x[6]), m);
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_TRACK_DIR: {
            const float3 v = m.dir_mode == NT_DIR_FP32_REF ? track_dir32(x[0], x[1]) : to_f3(track_dir64(x[0], x[1]));
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_UPDATE_DIR: {
            const float3 v = m.dir_mode == NT_DIR_FP32_REF
                                 ? update_direction32(make_float3(x[0], x[1], x[2]), x[3], x[4], x[5], m)
                                 : update_direction(make_float3(x[0], x[1], x[2]), x[3], x[4], x[5], m, m.hg);
            y[0] = v.x; y[1] = v.y; y[2] = v.z;
            break;
        }
        case NT_FN_SPECTRUM: y[0] = sample_spectrum(s, knots, x[0]); break;
        case NT_FN_PHILOX: {
            const uint32_t* u = reinterpret_cast<const uint32_t*>(x);
            const uint4 r = philox(make_uint4(u[0], u[1], u[2], u[3]), u[4], u[5]);
            uint32_t* o = reinterpret_cast<uint32_t*>(y);
            o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
            break;
        }
        case NT_FN_GAMMA: {
            // in: a, b, lrad, seed_lo, seed_hi, photon_lo, photon_hi, source (last five as bits)
            const uint32_t* u = reinterpret_cast<const uint32_t*>(x);
            const float d = (float)((double)x[0] - 1.0 / 3.0);
            const float c = 1.0f / sqrtf(9.0f * d);
            const Rng rng{u[3], u[4], u[5], u[6], u[7]};
            int trials = 0;
            y[0] = rand_gamma(d, c, 1.0f / x[1], rng, &trials) * x[2];
            y[1] = (float)trials;
            break;
        }
        default: break;
    }
}

}  // namespace

// ===========================================================================
// Host side: context and C ABI
// ===========================================================================
struct nt_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    int blocks = 0;                    // grid of the mixed-kind kernel
    int blocks_per_variant[2][7] = {};      // [direction arithmetic][variant]
    int variant = 6;                   // source kind shared by all sources, 6 = mixed
    std::mutex mu;

    bool has_medium = false, has_spectrum = false, has_targets = false, has_sources = false;
    Medium medium{};
    Spectrum spectrum{};
    float spec_lo = 0.0f, spec_hi = 0.0f;  // wavelength range the spectrum produces
    WlPoly sca_poly{};
    float* d_knots = nullptr;
    int knots_cap = 0;

    int n_targets = 0;
    float4* d_targets = nullptr;
    uint32_t* d_module_ids = nullptr;
    int targets_cap = 0;
    // cluster index over the targets (vertical strings)
    bool culling = true;
    int n_clusters = 0;
    float cluster_reach = 0.0f;         // largest reach r0 over the clusters
    float4* d_clusters = nullptr;
    float4* d_members = nullptr;
    uint32_t* d_member_idx = nullptr;
    int64_t clusters_cap = 0, members_cap = 0, member_idx_cap = 0;

    int n_sources = 0;
    DevSource* d_sources = nullptr;
    int64_t* d_offsets = nullptr;
    int sources_cap = 0;
    int64_t total = 0;

    int64_t* d_counters = nullptr;
    char* scratch = nullptr;            // this device's address of the file-scope scratch array
    hipEvent_t last_launch = nullptr;   // recorded after the kernels of every launch (any stream)
    char* d_ws = nullptr;               // resampling workspace
    int64_t ws_cap = 0;
    hipEvent_t last_resample = nullptr; // recorded after the kernels of every nt_resample
    // hit epilogue
    std::vector<nt_target> host_targets;
    bool epi_enabled = false;
    std::vector<float> abs_x, abs_y;
    std::vector<double> acc_x, acc_y, ang_x, ang_y;
    float abs_len_scale = 1.0f;
    double orientation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float epi_source[3] = {0, 0, 0};
    EpiTarget* d_epi_targets = nullptr;
    int32_t* d_target_of_module = nullptr;   // module id -> target list index (65536 entries)
    float* d_abs = nullptr;             // abs_x then abs_y
    double* d_acc = nullptr;            // acc_x then acc_y
    double* d_ang = nullptr;            // ang_x then ang_y
    int64_t epi_targets_cap = 0, abs_cap = 0, acc_cap = 0, ang_cap = 0;
    Epilogue epi{};
    // hit columns owned by the synchronous API
    HitCols d_hits{};
    int64_t hits_cap = 0;
};

namespace {

thread_local char g_err[256];

int32_t fail_hip(hipError_t e, const char* what) {
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? NT_ERR_OOM : NT_ERR_DEVICE;
}

#define HIP_TRY(x)                                   \
    do {                                             \
        hipError_t _e = (x);                         \
        if (_e != hipSuccess) return fail_hip(_e, #x); \
    } while (0)

template <typename T>
int32_t grow(T** p, int64_t need, int64_t* cap_items) {
    if (need <= *cap_items && *p) return NT_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    int64_t n = std::max<int64_t>(need, 1);
    HIP_TRY(hipMalloc((void**)p, sizeof(T) * n));
    *cap_items = n;
    return NT_OK;
}

// ---------------------------------------------------------------------------
// Scratch memory of the allocation-free entry point (nt_propagate_noalloc):
// one file-scope device array holding the counters and the hit columns of
// up to NT_NOALLOC_MAX_HITS hits.  Each GPU holds its own copy until the
// process ends; host code uses only the device address hipGetSymbolAddress
// returns for it (kept per device), never the array's name.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocHits = NT_NOALLOC_MAX_HITS;
struct NoAllocLayout {                 // byte offsets (multiples of 256 for kNoAllocHits = 2^17)
    static constexpr size_t counters = 0;
    static constexpr size_t pos = 256;
    static constexpr size_t dir = pos + 12 * kNoAllocHits;
    static constexpr size_t idir = dir + 12 * kNoAllocHits;
    static constexpr size_t wl = idir + 12 * kNoAllocHits;
    static constexpr size_t time = wl + 4 * kNoAllocHits;
    static constexpr size_t dist = time + 4 * kNoAllocHits;
    static constexpr size_t mid = dist + 4 * kNoAllocHits;
    static constexpr size_t sidx = mid + 2 * kNoAllocHits;
    static constexpr size_t pid = sidx + 4 * kNoAllocHits;
    static constexpr size_t pmt = pid + 8 * kNoAllocHits;
    static constexpr size_t weight = pmt + kNoAllocHits;
    static constexpr size_t tres = weight + 4 * kNoAllocHits;
    static constexpr size_t bytes = tres + 4 * kNoAllocHits;
};
static_assert(NoAllocLayout::pid % 8 == 0 && NoAllocLayout::weight % 4 == 0 && NoAllocLayout::bytes % 256 == 0,
              "scratch column alignment");

__device__ __align__(256) unsigned char scratch[NoAllocLayout::bytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
int32_t scratch_address(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) {
        snprintf(g_err, sizeof g_err, "device %d: at most %d devices", device, kMaxDevices);
        return NT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_addr[device] = static_cast<char*>(p);
    }
    *out = g_scratch_addr[device];
    return NT_OK;
}

void free_hits(HitCols& h) {
    void* ptrs[] = {h.pos, h.dir, h.idir, h.wl, h.time, h.dist, h.module_id, h.source_index, h.photon_id,
                    h.pmt_id, h.total_weight, h.tres};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    h = HitCols{};
}

size_t lds_bytes(int n_knots) {
    return sizeof(float) * ((n_knots + 3) & ~3) + sizeof(uint32_t) * kWaves * kStageFields * kStageSlots;
}

// One launch of the propagation kernel over `n_sources` device sources.
struct LaunchPlan {
    const DevSource* sources;
    const int64_t* offsets;
    int n_sources;
    int64_t total;                     // padded photon index range
    int variant;                       // kernel variant (source kind or mixed)
    const SlotDesc* slots;             // null: single output `cols`, `counters`
    HitCols cols;
    int64_t* counters;
};

int32_t launch_plan(nt_ctx* c, uint64_t seed, const LaunchPlan& lp, hipStream_t s) {
    if (lp.total == 0) return NT_OK;
    KParams kp{};
    kp.sources = lp.sources;
    kp.offsets = lp.offsets;
    kp.slots = lp.slots;
    kp.targets = c->d_targets;
    kp.module_ids = c->d_module_ids;
    kp.clusters = c->d_clusters;
    kp.members = c->d_members;
    kp.member_idx = c->d_member_idx;
    kp.n_clusters = c->culling ? c->n_clusters : 0;
    kp.cluster_reach = c->cluster_reach;
    kp.knots = c->d_knots;
    kp.counters = lp.counters;
    kp.hits = lp.cols;
    kp.total = lp.total;
    kp.medium = c->medium;
    kp.spectrum = c->spectrum;
    kp.sca_poly = c->sca_poly;
    kp.n_sources = lp.n_sources;
    kp.n_targets = c->n_targets;
    kp.seed_lo = (uint32_t)seed;
    kp.seed_hi = (uint32_t)(seed >> 32);
    const int64_t need_blocks = (lp.total + kThreads - 1) / kThreads;
    const int dm = c->medium.dir_mode == NT_DIR_FP32_REF ? 1 : 0;
    const int blocks = (int)std::min<int64_t>(c->blocks_per_variant[dm][lp.variant], need_blocks);
    hipLaunchKernelGGL(propagate_variant(lp.variant, c->medium.dir_mode, c->medium.sl_fraction > 0.0f), dim3(blocks),
                       dim3(kThreads),
                       lds_bytes(c->spectrum.n), s, kp);
    HIP_TRY(hipGetLastError());
    return NT_OK;
}

int32_t launch(nt_ctx* c, uint64_t seed, const HitCols& cols, int64_t* d_counters, hipStream_t s) {
    if (!c->has_medium || !c->has_spectrum || !c->has_targets || !c->has_sources) {
        snprintf(g_err, sizeof g_err, "medium, spectrum, targets and sources must be set first");
        return NT_ERR_STATE;
    }
    hipLaunchKernelGGL(clear_counters, dim3(1), dim3(64), 0, s, d_counters);
    if (c->total == 0) {
        HIP_TRY(hipEventRecord(c->last_launch, s));
        return NT_OK;
    }
    LaunchPlan lp{c->d_sources, c->d_offsets, c->n_sources, c->total, c->variant, nullptr, cols, d_counters};
    const int32_t st = launch_plan(c, seed, lp, s);
    if (st) return st;
    if (c->epi.enabled && (cols.pmt_id || cols.total_weight || cols.tres)) {
        const int eb = (int)std::max<int64_t>(1, std::min<int64_t>((cols.capacity + 255) / 256, 1024));
        hipLaunchKernelGGL(epilogue_kernel, dim3(eb), dim3(256), 0, s, c->epi, cols, d_counters,
                           c->d_target_of_module);
        HIP_TRY(hipGetLastError());
    }
    // the set_* calls wait for this before they rewrite what the kernels read
    HIP_TRY(hipEventRecord(c->last_launch, s));
    return NT_OK;
}

// Wait until the kernels of the last launch (on whatever stream the caller
// gave nt_launch) have finished, so that a set_* call can rewrite the
// device tables they read.
int32_t wait_last_launch(nt_ctx* c) {
    HIP_TRY(hipEventSynchronize(c->last_launch));
    return NT_OK;
}

// Per-target epilogue data (see nt_epilogue), following the precision of the
// reference's host code: PMT axes sph_to_cart in FP32 then rotated in FP64
// (detection.jl:60-70), opening angle asin(sqrt(area/π)/r) in FP32
// (detection.jl:94-95), area acceptance and time residual in FP64
// (detection.jl:139-146, photon_prop_cuda.jl:906-917).
int32_t rebuild_epilogue(nt_ctx* c) {
    if (!c->epi_enabled) {
        c->epi = Epilogue{};
        return NT_OK;
    }
    if (!c->has_medium || !c->has_targets) {
        snprintf(g_err, sizeof g_err, "set the medium and the targets before the epilogue");
        return NT_ERR_STATE;
    }
    const Medium& m = c->medium;
    const float x800 = 1.0f / 800.0f, x800sq = x800 * x800;
    const float n800 = fmaf(x800, m.a2, m.a01) + fmaf(x800sq, m.a3, x800sq * x800 * m.a4);
    const double c_med = 0.299792458 / (double)n800;
    const int n = (int)c->host_targets.size();
    std::vector<EpiTarget> et(std::max(n, 1));
    for (int i = 0; i < n; ++i) {
        const nt_target& t = c->host_targets[i];
        EpiTarget& e = et[i];
        memset(&e, 0, sizeof e);
        for (int k = 0; k < 3; ++k) e.pos[k] = (double)t.position[k];
        e.kind = t.kind;
        e.n_pmts = t.kind == NT_TARGET_MULTI_PMT ? std::min(t.n_pmts, NT_MAX_PMTS) : 1;
        if (t.kind == NT_TARGET_SPHERE) {
            const float num = (float)t.n_pmts * t.pmt_area;
            e.area_acc = (double)num / (4.0 * M_PI * (double)(t.radius * t.radius));
        } else {
            e.area_acc = 1.0;
            const float pmt_radius = sqrtf(t.pmt_area / (float)M_PI);
            e.cos_open = cos((double)asinf(pmt_radius / t.radius));
            for (int j = 0; j < e.n_pmts; ++j) {
                const float th = t.pmt_coordinates[2 * j], ph = t.pmt_coordinates[2 * j + 1];
                const float v[3] = {cosf(ph) * sinf(th), sinf(ph) * sinf(th), cosf(th)};
                for (int r = 0; r < 3; ++r)
                    e.axes[j][r] = (c->orientation[3 * r + 0] * v[0] + c->orientation[3 * r + 1] * v[1]) +
                                   c->orientation[3 * r + 2] * v[2];
            }
        }
        const float dx = c->epi_source[0] - t.position[0], dy = c->epi_source[1] - t.position[1],
                    dz = c->epi_source[2] - t.position[2];
        const float dist = sqrtf((dx * dx + dy * dy) + dz * dz);
        e.tgeo = (double)(dist - t.radius) / c_med;
        e.radius = t.radius;
    }
    int32_t st = grow(&c->d_epi_targets, (int64_t)et.size(), &c->epi_targets_cap);
    if (st) return st;
    // Dict(module_id => target) of photon_prop_cuda.jl:852: the last listed target wins
    std::vector<int32_t> tom(65536, -1);
    for (int i = 0; i < n; ++i) tom[c->host_targets[i].module_id] = i;
    if (!c->d_target_of_module) HIP_TRY(hipMalloc((void**)&c->d_target_of_module, sizeof(int32_t) * 65536));
    HIP_TRY(hipMemcpy(c->d_target_of_module, tom.data(), sizeof(int32_t) * 65536, hipMemcpyHostToDevice));
    st = grow(&c->d_abs, (int64_t)(2 * c->abs_x.size()), &c->abs_cap);
    if (st) return st;
    st = grow(&c->d_acc, (int64_t)(2 * c->acc_x.size()), &c->acc_cap);
    if (st) return st;
    st = grow(&c->d_ang, (int64_t)std::max<size_t>(2 * c->ang_x.size(), 1), &c->ang_cap);
    if (st) return st;
    const size_t na = c->abs_x.size(), nc = c->acc_x.size(), ng = c->ang_x.size();
    if (ng) {
        HIP_TRY(hipMemcpy(c->d_ang, c->ang_x.data(), 8 * ng, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->d_ang + ng, c->ang_y.data(), 8 * ng, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(c->d_epi_targets, et.data(), sizeof(EpiTarget) * et.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_abs, c->abs_x.data(), 4 * na, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_abs + na, c->abs_y.data(), 4 * na, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_acc, c->acc_x.data(), 8 * nc, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_acc + nc, c->acc_y.data(), 8 * nc, hipMemcpyHostToDevice));
    Epilogue& d = c->epi;
    d.targets = c->d_epi_targets;
    d.abs_x = c->d_abs;
    d.abs_y = c->d_abs + na;
    d.acc_x = c->d_acc;
    d.acc_y = c->d_acc + nc;
    d.n_abs = (int)na;
    d.n_acc = (int)nc;
    d.ang_x = c->d_ang;
    d.ang_y = c->d_ang + ng;
    d.n_ang = (int)ng;
    d.abs_len_scale = c->abs_len_scale;
    d.enabled = 1;
    d.c_med = c_med;
    return NT_OK;
}

int64_t round_up64(int64_t n) { return (n + 63) & ~int64_t(63); }

// Every float field of a source is finite (infinite or NaN longitudinal
// parameters would make every Gamma trial fail; the others would propagate
// NaN photons).
__host__ __device__ inline bool source_fields_finite(const nt_source& a) {
    const float v[13] = {a.position[0], a.position[1], a.position[2], a.direction[0], a.direction[1],
                         a.direction[2], a.time, a.time_end, a.angle, a.long_a, a.long_b, a.long_lrad, 0.0f};
    bool ok = true;
    for (int k = 0; k < 12; ++k) ok = ok && __builtin_isfinite(v[k]);
    return ok;
}

// Device record of an nt_source (offset, key, base and slot left at 0).
// nt_set_sources_device: per record the preparation of make_dev_source (same
// FP32 operations, IEEE square root and division), photon counts rounded up
// to multiples of 64 into offs[i + 1]; err[0] = 1 + index of an invalid record.
__global__ __launch_bounds__(256) void prepare_sources_kernel(const nt_source* __restrict__ in, int n,
                                                             DevSource* __restrict__ out, int64_t* offs,
                                                             int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) offs[0] = 0;
    if (i >= n) return;
    const nt_source a = in[i];
    DevSource d{};
    bool ok = a.photons >= 0 && a.photons <= (int64_t(1) << 56) && a.kind >= NT_SRC_POINTLIKE_ISOTROPIC &&
              a.kind <= NT_SRC_POINTLIKE_CHERENKOV && source_fields_finite(a);
    d.kind = a.kind;
    d.rng_index = (uint32_t)i;
    d.photons = a.photons;
    for (int k = 0; k < 3; ++k) { d.pos[k] = a.position[k]; d.dir[k] = a.direction[k]; }
    d.t0 = a.time;
    d.t1 = a.time_end;
    d.angle = a.angle;
    if (ok && a.kind == NT_SRC_EXTENDED_CHERENKOV) {
        ok = a.long_a > 1.0f && a.long_b > 0.0f && a.long_lrad > 0.0f;
        d.gamma_d = (float)((double)a.long_a - 1.0 / 3.0);
        d.gamma_c = 1.0f / sqrtf(9.0f * d.gamma_d);
        d.gamma_scale = 1.0f / a.long_b;
        d.lrad = a.long_lrad;
        ok = ok && isfinite(d.gamma_scale);
    }
    if (!ok) {
        atomicMin(err, i + 1);
        d.photons = 0;
    }
    out[i] = d;
    offs[i + 1] = ok ? ((a.photons + 63) & ~int64_t(63)) : 0;
}

// In-place inclusive scan of offs[1..n] by one block (chunks per thread, a
// block scan of the chunk sums, then the chunks again), then each record's
// offset; total[0] = offs[n].
__global__ __launch_bounds__(1024) void scan_offsets_kernel(int64_t* offs, int n, DevSource* __restrict__ recs,
                                                           int64_t* total) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int lo = 1 + t * chunk, hi = min(n + 1, lo + chunk);
    int64_t sum = 0;
    for (int j = lo; j < hi; ++j) sum += offs[j];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t ? part[t - 1] : 0;
    for (int j = lo; j < hi; ++j) {
        run += offs[j];
        offs[j] = run;
    }
    __syncthreads();
    for (int j = t; j < n; j += 1024) recs[j].offset = offs[j];
    if (t == 0) total[0] = part[1023];
}

// Propagate into the device columns `h` with counters `d_cnt` on the
// context's stream, then copy the stored hits and counts to the caller's host
// columns (the synchronous API's second half).
int32_t launch_and_copy_back(nt_ctx* c, uint64_t seed, const HitCols& h, int64_t* d_cnt, const nt_hits* out,
                             int64_t* n_hits, int64_t* n_sim, int64_t* n_dropped) {
    const int64_t cap = out->capacity;
    int32_t st = launch(c, seed, h, d_cnt, c->stream);
    if (st) return st;
    int64_t cnt[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int64_t stored = std::min(cnt[0], cap);
    if (stored > 0) {
        HIP_TRY(hipMemcpyAsync(out->position, h.pos, 12 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->direction, h.dir, 12 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->initial_direction, h.idir, 12 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->wavelength, h.wl, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->time, h.time, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->dist_travelled, h.dist, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(out->module_id, h.module_id, 2 * stored, hipMemcpyDeviceToHost, c->stream));
        if (out->source_index)
            HIP_TRY(hipMemcpyAsync(out->source_index, h.source_index, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        if (out->photon_id)
            HIP_TRY(hipMemcpyAsync(out->photon_id, h.photon_id, 8 * stored, hipMemcpyDeviceToHost, c->stream));
        if (out->pmt_id && c->epi.enabled)
            HIP_TRY(hipMemcpyAsync(out->pmt_id, h.pmt_id, stored, hipMemcpyDeviceToHost, c->stream));
        if (out->total_weight && c->epi.enabled)
            HIP_TRY(hipMemcpyAsync(out->total_weight, h.total_weight, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        if (out->tres && c->epi.enabled)
            HIP_TRY(hipMemcpyAsync(out->tres, h.tres, 4 * stored, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    *n_hits = stored;
    *n_sim = cnt[1];
    if (n_dropped) *n_dropped = cnt[0] - stored;
    return cnt[0] > cap ? NT_ERR_HIT_OVERFLOW : NT_OK;
}

int32_t make_dev_source(const nt_source& a, int i, DevSource& d) {
    if (a.photons < 0 || a.photons > (int64_t(1) << 56) || a.kind < NT_SRC_POINTLIKE_ISOTROPIC ||
        a.kind > NT_SRC_POINTLIKE_CHERENKOV) {
        snprintf(g_err, sizeof g_err, "source %d: photons must be in [0, 2^56] and kind known", i);
        return NT_ERR_ARG;
    }
    if (!source_fields_finite(a)) {
        snprintf(g_err, sizeof g_err, "source %d: position, direction, times, angle and longitudinal "
                 "parameters must be finite", i);
        return NT_ERR_ARG;
    }
    memset(&d, 0, sizeof d);
    d.kind = a.kind;
    d.rng_index = (uint32_t)i;
    d.photons = a.photons;
    for (int k = 0; k < 3; ++k) { d.pos[k] = a.position[k]; d.dir[k] = a.direction[k]; }
    d.t0 = a.time;
    d.t1 = a.time_end;
    d.angle = a.angle;
    if (a.kind == NT_SRC_EXTENDED_CHERENKOV) {
        if (!(a.long_a > 1.0f && a.long_b > 0.0f && a.long_lrad > 0.0f) || !std::isfinite(1.0f / a.long_b)) {
            snprintf(g_err, sizeof g_err, "source %d: longitudinal parameters need a > 1, b > 0 (1/b finite), "
                     "lrad > 0", i);
            return NT_ERR_ARG;
        }
        // rand_gamma(T(a), T(1/b)) set-up, utils.jl:269-270 and photon_prop_cuda.jl:168
        d.gamma_d = (float)((double)a.long_a - 1.0 / 3.0);
        d.gamma_c = 1.0f / sqrtf(9.0f * d.gamma_d);
        d.gamma_scale = 1.0f / a.long_b;
        d.lrad = a.long_lrad;
    }
    return NT_OK;
}

}  // namespace

extern "C" {

int32_t nt_abi_version(void) { return NTPROP_ABI_VERSION; }

const char* nt_status_string(int32_t st) {
    switch (st) {
        case NT_OK: return "ok";
        case NT_ERR_ARG: return g_err[0] ? g_err : "invalid argument";
        case NT_ERR_OOM: return g_err[0] ? g_err : "out of device memory";
        case NT_ERR_HIT_OVERFLOW: return "hit capacity exceeded";
        case NT_ERR_DEVICE: return g_err[0] ? g_err : "device error";
        case NT_ERR_BUDGET: return g_err[0] ? g_err : "request exceeds the memory budget";
        case NT_ERR_STATE: return g_err[0] ? g_err : "call out of order";
        case NT_ERR_NOT_READY: return "not ready";
        default: return "unknown status";
    }
}

int32_t nt_create(int32_t device, nt_ctx** out) {
    if (!out) return NT_ERR_ARG;
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        snprintf(g_err, sizeof g_err, "device %d out of range (%d devices)", device, n);
        return NT_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    nt_ctx* c = new nt_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return NT_ERR_DEVICE; }
    c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return NT_ERR_DEVICE; }
    if (hipEventCreateWithFlags(&c->last_launch, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(c->last_launch, c->stream) != hipSuccess ||
        hipEventCreateWithFlags(&c->last_resample, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(c->last_resample, c->stream) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NT_ERR_DEVICE;
    }
    for (int dm = 0; dm < 2; ++dm)
        for (int v = 0; v < kKernelVariants; ++v) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, propagate_variant(v, dm ? NT_DIR_FP32_REF : NT_DIR_FP64, true),
                                                             kThreads, lds_bytes(NT_MAX_KNOTS)) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            c->blocks_per_variant[dm][v] = c->num_cus * per_cu;
        }
    c->blocks = c->blocks_per_variant[0][kKernelVariants - 1];
    if (scratch_address(device, &c->scratch) != NT_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NT_ERR_DEVICE;
    }
    if (hipMalloc((void**)&c->d_counters, 4 * sizeof(int64_t)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NT_ERR_OOM;
    }
    *out = c;
    return NT_OK;
}

void nt_destroy(nt_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->d_knots, c->d_targets, c->d_module_ids, c->d_sources, c->d_offsets, c->d_counters,
                    c->d_clusters, c->d_members,
                    c->d_member_idx, c->d_epi_targets, c->d_abs, c->d_acc, c->d_ang, c->d_target_of_module};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    free_hits(c->d_hits);
    if (c->last_launch) (void)hipEventDestroy(c->last_launch);
    if (c->last_resample) (void)hipEventDestroy(c->last_resample);
    if (c->d_ws) (void)hipFree(c->d_ws);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

// FP64 value of the scattering-length formula (medium.jl:331-347) with the
// reference's FP32 constants and concentrations.
static double sca_len_f64(double wl, float vcs, float vcl) {
    const double x = 550.0 / wl;
    const double c = ((double)0.0017f * std::pow(x, (double)4.3f) +
                      (double)1.34f * (double)vcs * std::pow(x, (double)1.7f)) +
                     (double)0.312f * (double)vcl * std::pow(x, (double)0.3f);
    return 1.0 / c;
}

// Polynomial fit of the scattering length over the spectrum's wavelength
// range: Chebyshev interpolation at d + 1 nodes, converted to monomials in t,
// for the smallest degree d <= 20 whose Horner evaluation (fma, as on the
// device) stays within 1e-12 relative of the formula on a 2001-point grid and
// the range ends.  No degree meeting the bound leaves the fit invalid.
static WlPoly fit_scattering_length(const Medium& m, float lo, float hi) {
    WlPoly P{};
    if (!(lo > 0.0f) || !(hi >= lo)) return P;
    // widen by a relative 1e-6 (rounded outward) to take in the FP32 rounding
    // of spectrum samples at the ends
    const double dlo = (double)lo * (1.0 - 1e-6), dhi = (double)hi * (1.0 + 1e-6);
    float flo = (float)dlo, fhi = (float)dhi;
    if ((double)flo > dlo) flo = std::nextafter(flo, 0.0f);
    if ((double)fhi < dhi) fhi = std::nextafter(fhi, INFINITY);
    P.lo = flo;
    P.hi = fhi;
    const double a = flo, b = fhi;
    if (a == b) {
        P.c[0] = sca_len_f64(a, m.vcs, m.vcl);
        P.valid = 1;
        return P;
    }
    P.ta = 2.0 / (b - a);
    P.tb = -(a + b) / (b - a);
    for (int d = 6; d < kScaPolyCoefs; ++d) {
        const int nn = d + 1;
        double f[kScaPolyCoefs], ch[kScaPolyCoefs] = {};
        for (int k = 0; k < nn; ++k) {
            const double t = std::cos(M_PI * (k + 0.5) / nn);
            f[k] = sca_len_f64(0.5 * (a + b) + 0.5 * (b - a) * t, m.vcs, m.vcl);
        }
        for (int j = 0; j < nn; ++j) {
            double sum = 0.0;
            for (int k = 0; k < nn; ++k) sum += f[k] * std::cos(M_PI * j * (k + 0.5) / nn);
            ch[j] = (j == 0 ? 1.0 : 2.0) * sum / nn;
        }
        // Σ ch_j T_j(t) as monomials: T_{j+1} = 2t·T_j − T_{j−1}
        double mono[kScaPolyCoefs] = {}, tm1[kScaPolyCoefs] = {}, tj[kScaPolyCoefs] = {};
        tm1[0] = 1.0;                        // T_0
        mono[0] = ch[0];
        if (nn > 1) { tj[1] = 1.0; mono[1] += ch[1]; }
        for (int j = 1; j + 1 < nn; ++j) {
            double tn[kScaPolyCoefs] = {};
            for (int q = 0; q < kScaPolyCoefs; ++q) {
                tn[q] = -tm1[q] + (q > 0 ? 2.0 * tj[q - 1] : 0.0);
            }
            for (int q = 0; q < kScaPolyCoefs; ++q) {
                mono[q] += ch[j + 1] * tn[q];
                tm1[q] = tj[q];
                tj[q] = tn[q];
            }
        }
        double worst = 0.0;
        for (int k = 0; k <= 2002; ++k) {
            const double wl = k <= 2000 ? a + (b - a) * k / 2000.0 : (k == 2001 ? (double)lo : (double)hi);
            const double t = std::fma(wl, P.ta, P.tb);
            double y = mono[kScaPolyCoefs - 1];
            for (int j = kScaPolyCoefs - 2; j >= 0; --j) y = std::fma(y, t, mono[j]);
            worst = std::max(worst, std::fabs(y / sca_len_f64(wl, m.vcs, m.vcl) - 1.0));
        }
        if (worst <= 1e-12) {
            for (int q = 0; q < kScaPolyCoefs; ++q) P.c[q] = mono[q];
            P.valid = 1;
            return P;
        }
    }
    return WlPoly{};
}

static void refresh_sca_poly(nt_ctx* c) {
    c->sca_poly = (c->has_medium && c->has_spectrum) ? fit_scattering_length(c->medium, c->spec_lo, c->spec_hi)
                                                     : WlPoly{};
}

int32_t nt_set_medium(nt_ctx* c, const nt_medium* m) {
    if (!c || !m) return NT_ERR_ARG;
    if (!(m->mean_scattering_angle > 0.0f && m->mean_scattering_angle < 1.0f) ||
        m->max_steps < 0 || m->max_steps > NT_MAX_STEPS || !(m->sl_fraction >= 0.0f && m->sl_fraction <= 1.0f) ||
        (m->direction_arithmetic != NT_DIR_FP64 && m->direction_arithmetic != NT_DIR_FP32_REF) ||
        !(m->sca_len_scale > 0.0f) ||
        (m->dispersion_mode != NT_DISPERSION_REF && m->dispersion_mode != NT_DISPERSION_QUAN_FRY)) {
        snprintf(g_err, sizeof g_err, "invalid medium parameters");
        return NT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    Medium& d = c->medium;
    d.a01 = m->quan_fry[0]; d.a2 = m->quan_fry[1]; d.a3 = m->quan_fry[2]; d.a4 = m->quan_fry[3];
    d.vcs = m->vol_conc_small_part;
    d.vcl = m->vol_conc_large_part;
    d.g = m->mean_scattering_angle;
    d.sl_fraction = m->sl_fraction;
    d.sca_len_scale = m->sca_len_scale;
    d.dispersion_mode = m->dispersion_mode;
    d.max_steps = m->max_steps;
    d.dir_mode = m->direction_arithmetic;
    d.hg = hg_const(d.g);                // FP64 constants of the scattering function
    c->has_medium = true;
    refresh_sca_poly(c);
    if (c->epi_enabled) {
        HIP_TRY(hipSetDevice(c->device));
        return rebuild_epilogue(c);
    }
    return NT_OK;
}

int32_t nt_set_spectrum(nt_ctx* c, const nt_spectrum* s) {
    if (!c || !s) return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    HIP_TRY(hipSetDevice(c->device));
    Spectrum d{};
    d.kind = s->kind;
    d.addressing = s->addressing;
    d.wl = s->wavelength;
    d.lo = s->wl_range[0];
    d.hi = s->wl_range[1];
    d.n = 0;
    if (s->kind == NT_SPEC_TABLE) {
        if (s->n_knots < 2 || s->n_knots > NT_MAX_KNOTS || !s->knots ||
            s->addressing < NT_TEX_TEXEL_CENTER_CLAMP || s->addressing > NT_TEX_ENDPOINT) {
            snprintf(g_err, sizeof g_err, "spectrum table needs 2..%d knots and a valid addressing", NT_MAX_KNOTS);
            return NT_ERR_ARG;
        }
        c->has_spectrum = false;             // knots change from here on (see nt_set_targets)
        int64_t cap = c->knots_cap;
        int32_t st = grow(&c->d_knots, s->n_knots, &cap);
        if (st) return st;
        c->knots_cap = (int)cap;
        HIP_TRY(hipMemcpy(c->d_knots, s->knots, sizeof(float) * s->n_knots, hipMemcpyHostToDevice));
        d.n = s->n_knots;
    } else if (s->kind == NT_SPEC_MONOCHROMATIC) {
        if (!(s->wavelength > 0.0f)) { snprintf(g_err, sizeof g_err, "wavelength must be > 0"); return NT_ERR_ARG; }
    } else if (s->kind == NT_SPEC_INV_LAMBDA2) {
        if (!(s->wl_range[0] > 0.0f && s->wl_range[1] > s->wl_range[0])) {
            snprintf(g_err, sizeof g_err, "invalid wavelength range");
            return NT_ERR_ARG;
        }
    } else {
        snprintf(g_err, sizeof g_err, "unknown spectrum kind %d", s->kind);
        return NT_ERR_ARG;
    }
    if (!c->d_knots) {                      // keep a valid pointer for the kernel
        int64_t cap = 0;
        int32_t st = grow(&c->d_knots, 1, &cap);
        if (st) return st;
        c->knots_cap = (int)cap;
    }
    c->spectrum = d;
    c->has_spectrum = true;
    if (s->kind == NT_SPEC_TABLE) {
        c->spec_lo = *std::min_element(s->knots, s->knots + s->n_knots);
        c->spec_hi = *std::max_element(s->knots, s->knots + s->n_knots);
    } else if (s->kind == NT_SPEC_MONOCHROMATIC) {
        c->spec_lo = c->spec_hi = s->wavelength;
    } else {
        c->spec_lo = s->wl_range[0];
        c->spec_hi = s->wl_range[1];
    }
    refresh_sca_poly(c);
    return NT_OK;
}

int32_t nt_set_targets(nt_ctx* c, const nt_target* t, int32_t n) {
    if (!c || n < 0 || n > NT_MAX_TARGETS || (n > 0 && !t)) {
        snprintf(g_err, sizeof g_err, "targets: 0..%d entries", NT_MAX_TARGETS);
        return NT_ERR_ARG;
    }
    std::vector<float4> tg(n);
    std::vector<uint32_t> ids(n);
    for (int i = 0; i < n; ++i) {
        if (!(t[i].radius > 0.0f) || !std::isfinite(t[i].radius)) {
            snprintf(g_err, sizeof g_err, "target %d: radius must be finite and > 0", i);
            return NT_ERR_ARG;
        }
        if (!std::isfinite(t[i].position[0]) || !std::isfinite(t[i].position[1]) || !std::isfinite(t[i].position[2])) {
            snprintf(g_err, sizeof g_err, "target %d: position must be finite", i);
            return NT_ERR_ARG;
        }
        tg[i] = make_float4(t[i].position[0], t[i].position[1], t[i].position[2], t[i].radius * t[i].radius);
        ids[i] = t[i].module_id;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    HIP_TRY(hipSetDevice(c->device));
    // buffers change from here on: a failure below leaves the context without
    // targets rather than with freed or half-written tables
    c->has_targets = false;
    int64_t cap = c->targets_cap, cap2 = c->targets_cap;
    int32_t st = grow(&c->d_targets, n, &cap);
    if (st) return st;
    st = grow(&c->d_module_ids, n, &cap2);
    if (st) return st;
    c->targets_cap = (int)std::min(cap, cap2);
    if (n) {
        HIP_TRY(hipMemcpy(c->d_targets, tg.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->d_module_ids, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    }
    // Clusters: targets with identical (x, y) form a vertical string (a
    // single target is a string of one).  Members are sorted by z (then list
    // index) inside a cluster, clusters are ordered by their first member.
    std::vector<int> cluster_of(n, -1);
    std::vector<std::vector<int>> groups;
    for (int i = 0; i < n; ++i) {
        int g = -1;
        for (size_t k = 0; k < groups.size() && g < 0; ++k) {
            const int j = groups[k][0];
            if (tg[j].x == tg[i].x && tg[j].y == tg[i].y) g = (int)k;
        }
        if (g < 0) { groups.push_back({}); g = (int)groups.size() - 1; }
        groups[g].push_back(i);
        cluster_of[i] = g;
    }
    std::vector<float4> cl, mem;
    std::vector<uint32_t> midx;
    float reach_max = 0.0f;
    for (auto& grp : groups) {
        std::stable_sort(grp.begin(), grp.end(), [&](int a, int b) { return tg[a].z < tg[b].z; });
        float zlo = INFINITY, zhi = -INFINITY, rmax = 0.0f;
        for (int i : grp) {
            zlo = std::min(zlo, tg[i].z);
            zhi = std::max(zhi, tg[i].z);
            rmax = std::max(rmax, t[i].radius);
        }
        // uniform ladder test (FP64): |z_k − (z_lo + k·Δz)| ≤ Δz/20 for all k,
        // and Δz ≥ 1e-3·(1 + |z|) so that the kernel's FP32 index arithmetic
        // (errors ≲ 1e-3 of a spacing) stays well inside its quarter-spacing margin
        float inv_dz = 0.0f;
        if (grp.size() >= 2) {
            const double dz = ((double)zhi - (double)zlo) / (double)(grp.size() - 1);
            bool uniform = dz >= 1e-3 * (1.0 + std::max(std::fabs((double)zlo), std::fabs((double)zhi)));
            for (size_t k = 0; k < grp.size() && uniform; ++k)
                uniform = std::fabs((double)tg[grp[k]].z - ((double)zlo + (double)k * dz)) <= 0.05 * dz;
            if (uniform) inv_dz = (float)(1.0 / dz);
        }
        const int first = (int)mem.size(), count = (int)grp.size();
        const double reach = (double)rmax + 0x1p-10 * ((double)rmax + 1.0) + 1e-3;
        float r0 = (float)reach;
        if ((double)r0 < reach) r0 = std::nextafter(r0, INFINITY);
        reach_max = std::max(reach_max, r0);
        float first_bits, count_bits;
        std::memcpy(&first_bits, &first, 4);
        std::memcpy(&count_bits, &count, 4);
        cl.push_back(make_float4(tg[grp[0]].x, tg[grp[0]].y, zlo, zhi));
        cl.push_back(make_float4(r0, first_bits, count_bits, inv_dz));
        for (int i : grp) {
            mem.push_back(tg[i]);
            midx.push_back((uint32_t)i);
        }
    }
    const int nc = (int)groups.size();
    st = grow(&c->d_clusters, std::max(2 * nc, 1), &c->clusters_cap);
    if (st) return st;
    st = grow(&c->d_members, std::max(n, 1), &c->members_cap);
    if (st) return st;
    st = grow(&c->d_member_idx, std::max(n, 1), &c->member_idx_cap);
    if (st) return st;
    if (nc) {
        HIP_TRY(hipMemcpy(c->d_clusters, cl.data(), sizeof(float4) * cl.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->d_members, mem.data(), sizeof(float4) * mem.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->d_member_idx, midx.data(), sizeof(uint32_t) * midx.size(), hipMemcpyHostToDevice));
    }
    // every target set goes through the bounds, single spheres included: a
    // bound costs about what the sphere test's first half does, and it keeps
    // the square root off the steps whose infinite line meets the sphere but
    // whose segment does not (one sphere at 10 m: +2-3 % photons/s)
    c->n_clusters = nc;
    c->cluster_reach = reach_max;
    c->n_targets = n;
    c->host_targets.assign(t, t + n);
    c->has_targets = true;
    if (c->epi_enabled) return rebuild_epilogue(c);
    return NT_OK;
}

int32_t nt_set_sources_device(nt_ctx* c, const nt_source* d_src, int32_t n, int32_t kind_hint, void* stream) {
    if (!c || n < 0 || (n > 0 && !d_src) || kind_hint < -1 || kind_hint > NT_SRC_POINTLIKE_CHERENKOV)
        return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    HIP_TRY(hipSetDevice(c->device));
    c->has_sources = false;                  // records change from here on (see nt_set_targets)
    int64_t cap = c->sources_cap, cap2 = c->sources_cap ? c->sources_cap + 1 : 0;
    int32_t st = grow(&c->d_sources, std::max(n, 1), &cap);
    if (st) return st;
    st = grow(&c->d_offsets, (int64_t)n + 1, &cap2);
    if (st) return st;
    c->sources_cap = (int)std::min(cap, cap2 - 1);
    hipStream_t s = (hipStream_t)stream;
    // error flag and total: the context's counter buffer is free between launches
    int64_t* scratch = c->d_counters;
    int* err = reinterpret_cast<int*>(scratch);
    const int32_t none = 0x7fffffff;
    HIP_TRY(hipMemcpyAsync(err, &none, sizeof none, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(prepare_sources_kernel, dim3((std::max(n, 1) + 255) / 256), dim3(256), 0, s, d_src, n,
                       c->d_sources, c->d_offsets, err);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, c->d_offsets, n, c->d_sources,
                       scratch + 1);
    HIP_TRY(hipGetLastError());
    int64_t back[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(back, scratch, sizeof back, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int32_t bad;
    std::memcpy(&bad, &back[0], sizeof bad);
    if (bad != none) {
        snprintf(g_err, sizeof g_err, "source %d: photons must be in [0, 2^56], kind known, fields finite, "
                 "and extended Cherenkov longitudinal parameters a > 1, b > 0 (1/b finite), lrad > 0", bad - 1);
        c->n_sources = 0;
        c->total = 0;
        c->has_sources = false;
        return NT_ERR_ARG;
    }
    c->n_sources = n;
    c->total = back[1];
    c->variant = kind_hint >= 0 ? kind_hint : kKernelVariants - 1;
    c->has_sources = true;
    return NT_OK;
}

int32_t nt_set_sources(nt_ctx* c, const nt_source* s, int32_t n) {
    if (!c || n < 0 || (n > 0 && !s)) return NT_ERR_ARG;
    std::vector<DevSource> ds(n);
    std::vector<int64_t> offs(n + 1);
    offs[0] = 0;
    for (int i = 0; i < n; ++i) {
        const int32_t st = make_dev_source(s[i], i, ds[i]);
        if (st) return st;
        ds[i].offset = offs[i];
        offs[i + 1] = offs[i] + round_up64(s[i].photons);
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    HIP_TRY(hipSetDevice(c->device));
    c->has_sources = false;                  // records change from here on (see nt_set_targets)
    int64_t cap = c->sources_cap, cap2 = c->sources_cap ? c->sources_cap + 1 : 0;
    int32_t st = grow(&c->d_sources, std::max(n, 1), &cap);
    if (st) return st;
    st = grow(&c->d_offsets, (int64_t)n + 1, &cap2);
    if (st) return st;
    c->sources_cap = (int)std::min(cap, cap2 - 1);
    if (n) HIP_TRY(hipMemcpy(c->d_sources, ds.data(), sizeof(DevSource) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_offsets, offs.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    c->n_sources = n;
    c->total = offs[n];
    c->variant = kKernelVariants - 1;
    if (n > 0 && std::all_of(s, s + n, [&](const nt_source& a) { return a.kind == s[0].kind; }))
        c->variant = s[0].kind;
    c->has_sources = true;
    return NT_OK;
}

int32_t nt_set_epilogue(nt_ctx* c, const nt_epilogue* e) {
    if (!c || !e) return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (int32_t w = wait_last_launch(c)) return w;
    HIP_TRY(hipSetDevice(c->device));
    if (!e->enabled) {
        c->epi_enabled = false;
        return rebuild_epilogue(c);
    }
    auto ascending = [](const auto* x, int n) {
        for (int i = 1; i < n; ++i)
            if (!(x[i] > x[i - 1])) return false;
        return true;
    };
    if (e->n_abs < 2 || e->n_abs > 1024 || !e->abs_wavelength || !e->abs_length || e->n_acc < 2 ||
        e->n_acc > 1024 || !e->acc_wavelength || !e->acc_value || !(e->abs_len_scale > 0.0f) ||
        !ascending(e->abs_wavelength, e->n_abs) || !ascending(e->acc_wavelength, e->n_acc) ||
        (e->n_ang != 0 && (e->n_ang < 2 || e->n_ang > 1024 || !e->ang_cos || !e->ang_value ||
                           !ascending(e->ang_cos, e->n_ang)))) {
        snprintf(g_err, sizeof g_err, "epilogue tables need 2..1024 ascending rows (angular: 0 or 2..1024) "
                 "and abs_len_scale > 0");
        return NT_ERR_ARG;
    }
    c->abs_x.assign(e->abs_wavelength, e->abs_wavelength + e->n_abs);
    c->abs_y.assign(e->abs_length, e->abs_length + e->n_abs);
    c->acc_x.assign(e->acc_wavelength, e->acc_wavelength + e->n_acc);
    c->acc_y.assign(e->acc_value, e->acc_value + e->n_acc);
    c->ang_x.assign(e->ang_cos, e->ang_cos + (e->n_ang > 0 ? e->n_ang : 0));
    c->ang_y.assign(e->ang_value, e->ang_value + (e->n_ang > 0 ? e->n_ang : 0));
    c->abs_len_scale = e->abs_len_scale;
    memcpy(c->orientation, e->orientation, sizeof c->orientation);
    memcpy(c->epi_source, e->source_position, sizeof c->epi_source);
    c->epi_enabled = true;
    return rebuild_epilogue(c);
}

int32_t nt_set_option(nt_ctx* c, int32_t option, int64_t value) {
    if (!c) return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
        case NT_OPT_TARGET_CULLING: c->culling = value != 0; return NT_OK;
        case NT_OPT_CHERENKOV_ANGLE_URAD:
            if (value < 0 || value >= 3141593) {
                snprintf(g_err, sizeof g_err, "Cherenkov angle: 0 (reference) or 1..3141592 microradians");
                return NT_ERR_ARG;
            }
            HIP_TRY(hipSetDevice(c->device));
            if (int32_t w = wait_last_launch(c)) return w;
            c->medium.cherenkov_angle = (double)value * 1e-6;
            // cos and sin once here (the C library's, as the oracle's), not per photon
            c->medium.cherenkov_cos = value ? std::cos(c->medium.cherenkov_angle) : 0.0;
            c->medium.cherenkov_sin = value ? std::sin(c->medium.cherenkov_angle) : 0.0;
            return NT_OK;
        default:
            snprintf(g_err, sizeof g_err, "unknown option %d", option);
            return NT_ERR_ARG;
    }
}

int32_t nt_launch_shape(nt_ctx* c, int32_t* blocks, int32_t* threads) {
    if (!c) return NT_ERR_ARG;
    if (blocks) *blocks = c->blocks_per_variant[c->medium.dir_mode == NT_DIR_FP32_REF ? 1 : 0][c->variant];
    if (threads) *threads = kThreads;
    return NT_OK;
}

int32_t nt_launch(nt_ctx* c, uint64_t seed, const nt_hits* out, int64_t* d_counters, void* stream) {
    if (!c || !out || !d_counters || out->capacity < 0) return NT_ERR_ARG;
    if (out->capacity > 0 && (!out->position || !out->direction || !out->initial_direction ||
                              !out->wavelength || !out->time || !out->dist_travelled || !out->module_id))
        return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    HitCols h{out->position, out->direction, out->initial_direction, out->wavelength, out->time,
              out->dist_travelled, out->module_id, out->source_index, out->photon_id,
              out->pmt_id, out->total_weight, out->tres, out->capacity};
    return launch(c, seed, h, d_counters, (hipStream_t)stream);
}

int32_t nt_propagate(nt_ctx* c, uint64_t seed, const nt_hits* out, int64_t* n_hits, int64_t* n_sim,
                     int64_t* n_dropped) {
    if (!c || !out || out->capacity < 0 || !n_hits || !n_sim) return NT_ERR_ARG;
    if (out->capacity > 0 && (!out->position || !out->direction || !out->initial_direction ||
                              !out->wavelength || !out->time || !out->dist_travelled || !out->module_id))
        return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    const int64_t cap = out->capacity;
    if (cap > c->hits_cap) {
        free_hits(c->d_hits);
        c->hits_cap = 0;
        HitCols& h = c->d_hits;
        const int64_t n = std::max<int64_t>(cap, 1);
        HIP_TRY(hipMalloc((void**)&h.pos, 12 * n));
        HIP_TRY(hipMalloc((void**)&h.dir, 12 * n));
        HIP_TRY(hipMalloc((void**)&h.idir, 12 * n));
        HIP_TRY(hipMalloc((void**)&h.wl, 4 * n));
        HIP_TRY(hipMalloc((void**)&h.time, 4 * n));
        HIP_TRY(hipMalloc((void**)&h.dist, 4 * n));
        HIP_TRY(hipMalloc((void**)&h.module_id, 2 * n));
        HIP_TRY(hipMalloc((void**)&h.source_index, 4 * n));
        HIP_TRY(hipMalloc((void**)&h.photon_id, 8 * n));
        HIP_TRY(hipMalloc((void**)&h.pmt_id, n));
        HIP_TRY(hipMalloc((void**)&h.total_weight, 4 * n));
        HIP_TRY(hipMalloc((void**)&h.tres, 4 * n));
        c->hits_cap = n;
    }
    HitCols h = c->d_hits;
    h.capacity = cap;
    return launch_and_copy_back(c, seed, h, c->d_counters, out, n_hits, n_sim, n_dropped);
}

int32_t nt_propagate_noalloc(nt_ctx* c, uint64_t seed, const nt_hits* out, int64_t* n_hits, int64_t* n_sim,
                             int64_t* n_dropped) {
    if (!c || !out || out->capacity < 0 || !n_hits || !n_sim) return NT_ERR_ARG;
    if (out->capacity > 0 && (!out->position || !out->direction || !out->initial_direction ||
                              !out->wavelength || !out->time || !out->dist_travelled || !out->module_id))
        return NT_ERR_ARG;
    if (out->capacity > kNoAllocHits) {
        snprintf(g_err, sizeof g_err, "nt_propagate_noalloc: capacity %lld above NT_NOALLOC_MAX_HITS (%lld)",
                 (long long)out->capacity, (long long)kNoAllocHits);
        return NT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    using L = NoAllocLayout;
    char* b = c->scratch;
    HitCols h{};
    h.pos = (float*)(b + L::pos); h.dir = (float*)(b + L::dir); h.idir = (float*)(b + L::idir);
    h.wl = (float*)(b + L::wl); h.time = (float*)(b + L::time); h.dist = (float*)(b + L::dist);
    h.module_id = (uint16_t*)(b + L::mid); h.source_index = (uint32_t*)(b + L::sidx);
    h.photon_id = (uint64_t*)(b + L::pid); h.pmt_id = (uint8_t*)(b + L::pmt);
    h.total_weight = (float*)(b + L::weight); h.tres = (float*)(b + L::tres);
    h.capacity = out->capacity;
    return launch_and_copy_back(c, seed, h, (int64_t*)(b + L::counters), out, n_hits, n_sim, n_dropped);
}

int32_t nt_resample(nt_ctx* c, const nt_hits* h, int64_t n, uint64_t seed, double downsample, int32_t per_pmt,
                    uint8_t* d_detected, int64_t* d_n_detected, void* stream) {
    if (!c || !h || n < 0 || n > INT32_MAX || !(downsample >= 0.0) || !std::isfinite(downsample)) {
        snprintf(g_err, sizeof g_err, "nt_resample: 0 <= n <= 2^31 - 1 hits and a finite downsample >= 0");
        return NT_ERR_ARG;
    }
    if (n > 0 && (!h->module_id || !h->total_weight || !h->source_index || !h->photon_id || !d_detected ||
                  (per_pmt && !h->pmt_id))) {
        snprintf(g_err, sizeof g_err, "nt_resample: module_id, total_weight, source_index, photon_id%s and "
                 "the detected column are required", per_pmt ? ", pmt_id" : "");
        return NT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    // the workspace is the context's: the previous call's kernels finish first
    HIP_TRY(hipEventSynchronize(c->last_resample));
    const int64_t need = (int64_t)resample_workspace_bytes(n);
    int32_t st = grow(&c->d_ws, need, &c->ws_cap);
    if (st) return st;
    ResampleArgs a{h->module_id, h->pmt_id, h->total_weight, h->source_index, h->photon_id, n, seed, downsample,
                   per_pmt ? 1 : 0, d_detected, d_n_detected};
    const hipStream_t s = (hipStream_t)stream;
    HIP_TRY(resample_run(a, c->d_ws, (size_t)c->ws_cap, s));
    HIP_TRY(hipEventRecord(c->last_resample, s));
    return NT_OK;
}

int32_t nt_eval(nt_ctx* c, int32_t fn, const float* in, int32_t iw, float* out, int32_t ow, int64_t n) {
    if (!c || fn < 0 || fn >= NT_FN_COUNT || n < 0 || iw < 1 || ow < 1 || (n > 0 && (!in || !out)))
        return NT_ERR_ARG;
    if (n == 0) return NT_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    float *d_in = nullptr, *d_out = nullptr;
    HIP_TRY(hipMalloc((void**)&d_in, sizeof(float) * iw * n));
    hipError_t e = hipMalloc((void**)&d_out, sizeof(float) * ow * n);
    if (e != hipSuccess) { (void)hipFree(d_in); return fail_hip(e, "hipMalloc"); }
    e = hipMemcpy(d_in, in, sizeof(float) * iw * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int tpb = 256;
        const int64_t nb = (n + tpb - 1) / tpb;
        const float* kn = c->d_knots;
        Spectrum sp = c->spectrum;
        if (!c->has_spectrum) { sp = Spectrum{}; sp.kind = NT_SPEC_MONOCHROMATIC; sp.wl = 400.0f; }
        if (!kn) kn = d_in;  // never read unless a table spectrum is set
        hipLaunchKernelGGL(eval_kernel, dim3((unsigned)nb), dim3(tpb), 0, c->stream, fn, d_in, iw, d_out, ow, n,
                           c->medium, sp, kn, c->sca_poly);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(float) * ow * n, hipMemcpyDeviceToHost);
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail_hip(e, "nt_eval");
    return NT_OK;
}

}  // extern "C"

// ===========================================================================
// Serving: continuous batching of propagation requests (nt_server_*)
// ===========================================================================
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over a byte range [0, size).
class Arena {
   public:
    void init(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t n, size_t& off) {
        n = align_up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align_up(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<size_t, size_t> free_;
};

// Column layout of one request's hits inside an arena block (device and pinned host alike).
struct ReqLayout {
    size_t counters, pos, dir, idir, wl, time, dist, mid, sidx, pid, pmt = 0, weight = 0, tres = 0, bytes;
    bool epi;
    explicit ReqLayout(int64_t cap, bool epilogue = false) : epi(epilogue) {
        size_t o = 0;
        auto take = [&](size_t n) { size_t at = o; o += align_up(n); return at; };
        counters = take(4 * sizeof(int64_t));
        pos = take(12 * cap);
        dir = take(12 * cap);
        idir = take(12 * cap);
        wl = take(4 * cap);
        time = take(4 * cap);
        dist = take(4 * cap);
        mid = take(2 * cap);
        sidx = take(4 * cap);
        pid = take(8 * cap);
        if (epi) {
            pmt = take(1 * cap);
            weight = take(4 * cap);
            tres = take(4 * cap);
        }
        bytes = o;
    }
    HitCols cols(char* base, int64_t cap) const {
        HitCols h{};
        h.pos = (float*)(base + pos); h.dir = (float*)(base + dir); h.idir = (float*)(base + idir);
        h.wl = (float*)(base + wl); h.time = (float*)(base + time); h.dist = (float*)(base + dist);
        h.module_id = (uint16_t*)(base + mid); h.source_index = (uint32_t*)(base + sidx);
        h.photon_id = (uint64_t*)(base + pid); h.capacity = cap;
        if (epi) {
            h.pmt_id = (uint8_t*)(base + pmt); h.total_weight = (float*)(base + weight); h.tres = (float*)(base + tres);
        }
        return h;
    }
};

enum ReqState { RQ_QUEUED, RQ_ACTIVE, RQ_SCHEDULED, RQ_COPY_COUNTS, RQ_COPY_HITS, RQ_DONE, RQ_FAILED };

struct Request {
    uint64_t id = 0;
    std::vector<DevSource> src;        // offset, slot, key, base set per round
    std::vector<int64_t> next;         // next photon to schedule, per source
    int64_t photons = 0, scheduled = 0;
    uint64_t seed = 0;
    int64_t capacity = 0;
    size_t bytes = 0, dev_off = 0, host_off = 0;
    int state = RQ_QUEUED;
    uint64_t last_round = 0;
    int64_t n_found = 0, n_sim = 0;
    hipEvent_t ev = nullptr;           // completion of the pending copy
    int32_t status = NT_OK;
    bool epilogue = false;             // work-list mode: fill pmt_id, total_weight, tres
    float ref[3] = {0.0f, 0.0f, 0.0f}; // their time-residual reference position
    int slot = -1;                     // work-list mode: slot ring position
    uint32_t seq = 0;                  // its completion mark
    std::chrono::steady_clock::time_point t_submit, t_done;
};

struct Staging {                       // per-round launch tables (pinned host + device)
    DevSource* h_src = nullptr;
    int64_t* h_off = nullptr;
    SlotDesc* h_slot = nullptr;
    DevSource* d_src = nullptr;
    int64_t* d_off = nullptr;
    SlotDesc* d_slot = nullptr;
    int cap = 0;                       // sources (and slots) it can hold
    hipEvent_t start = nullptr, done = nullptr;
    bool busy = false;
    uint64_t round = 0;
    int64_t photons = 0;
};

}  // namespace

struct nt_server {
    nt_ctx* ctx = nullptr;
    nt_server_config cfg{};
    hipStream_t stream = nullptr, copy_stream = nullptr;
    char* d_arena = nullptr;
    char* h_arena = nullptr;
    Arena dev_alloc, host_alloc;
    std::vector<Staging> staging;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;                        // waiting for memory, FIFO
    std::vector<uint64_t> active;                      // admitted, in admission order
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1, next_round = 1;
    bool stop = false;
    int32_t fatal = NT_OK;
    nt_server_stats stats{};
    std::thread worker;
    // work-list mode (see serve_kernel)
    int32_t mode = NT_SERVER_ROUNDS;
    WorkCtrl* ctrl = nullptr;                          // fine-grained pinned host memory
    WorkItem* h_items = nullptr;                       // rings in fine-grained host memory
    SlotDesc* h_slots = nullptr;
    SlotExtra* h_extra = nullptr;
    int64_t* d_slot_done = nullptr;                    // device: finished photons per slot
    int64_t* d_slot_cnt = nullptr;                     // device: the 4 counters per slot
    unsigned long long* d_claim = nullptr;             // [0] item cursors (small, large), [1] busy ticks
    DevWork* d_work = nullptr;                         // device copies of the control words
    WorkItem* d_items = nullptr;                       // device copy of the item ring
    hipStream_t upload = nullptr;
    hipStream_t kstream[2] = {nullptr, nullptr};
    hipEvent_t kdone[2] = {nullptr, nullptr};
    bool klaunched[2] = {false, false};
    uint32_t items_pub = 0;                            // items published (absolute, = ctrl->n_pub)
    uint64_t slots_pub = 0;                            // slots handed out (absolute)
    uint32_t win_items0 = 0;                           // first item / slot of the running window
    uint64_t win_slots0 = 0;
    bool window_full = false;
    int srv_blocks = 0;
    uint32_t launch_gen = 0;                           // generation of the last launched instance
    std::chrono::steady_clock::time_point last_bump{};
};

namespace {

int32_t ensure_staging(Staging& s, int need) {
    if (need <= s.cap) return NT_OK;
    int cap = std::max(need, 2 * s.cap);
    if (s.h_src) { (void)hipHostFree(s.h_src); (void)hipHostFree(s.h_off); (void)hipHostFree(s.h_slot); }
    if (s.d_src) { (void)hipFree(s.d_src); (void)hipFree(s.d_off); (void)hipFree(s.d_slot); }
    s = Staging{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, s.start, s.done, false, 0, 0};
    HIP_TRY(hipHostMalloc((void**)&s.h_src, sizeof(DevSource) * cap));
    HIP_TRY(hipHostMalloc((void**)&s.h_off, sizeof(int64_t) * (cap + 1)));
    HIP_TRY(hipHostMalloc((void**)&s.h_slot, sizeof(SlotDesc) * cap));
    HIP_TRY(hipMalloc((void**)&s.d_src, sizeof(DevSource) * cap));
    HIP_TRY(hipMalloc((void**)&s.d_off, sizeof(int64_t) * (cap + 1)));
    HIP_TRY(hipMalloc((void**)&s.d_slot, sizeof(SlotDesc) * cap));
    s.cap = cap;
    return NT_OK;
}

// Admit queued requests in order while their buffers fit.
void admit(nt_server* sv) {
    while (!sv->queue.empty()) {
        Request& r = *sv->reqs[sv->queue.front()];
        size_t doff, hoff;
        if (!sv->dev_alloc.alloc(r.bytes, doff)) return;
        if (!sv->host_alloc.alloc(r.bytes, hoff)) { sv->dev_alloc.release(doff, r.bytes); return; }
        r.dev_off = doff;
        r.host_off = hoff;
        r.state = RQ_ACTIVE;
        if (hipMemsetAsync(sv->d_arena + doff + ReqLayout(r.capacity, r.epilogue).counters, 0, 4 * sizeof(int64_t),
                           sv->mode == NT_SERVER_WORKLIST ? sv->upload : sv->stream) != hipSuccess)
            sv->fatal = NT_ERR_DEVICE;
        sv->active.push_back(r.id);
        sv->queue.pop_front();
    }
}

// Pack photon chunks of the active requests into one round and launch it.
int32_t launch_round(nt_server* sv, Staging& stg) {
    std::vector<Request*> want;
    for (uint64_t id : sv->active) {
        Request& r = *sv->reqs[id];
        if (r.state == RQ_ACTIVE && r.scheduled < r.photons) want.push_back(&r);
    }
    if (want.empty()) return NT_ERR_NOT_READY;
    int64_t budget = sv->cfg.round_photons;
    const int64_t share = std::max<int64_t>(budget / (int64_t)want.size(), 1 << 16);
    int n_src_max = 0;
    for (Request* r : want) n_src_max += (int)r->src.size();
    int32_t st = ensure_staging(stg, n_src_max);
    if (st) return st;
    const uint64_t round = sv->next_round++;
    int ns = 0, nslot = 0, kind = -2;
    int64_t off = 0, photons = 0;
    for (Request* r : want) {
        if (budget <= 0) break;
        int64_t take = std::min({share, budget, r->photons - r->scheduled});
        const ReqLayout L(r->capacity);
        SlotDesc& sd = stg.h_slot[nslot];
        sd.cols = L.cols(sv->d_arena + r->dev_off, r->capacity);
        sd.counters = (int64_t*)(sv->d_arena + r->dev_off + L.counters);
        for (size_t j = 0; j < r->src.size() && take > 0; ++j) {
            const int64_t left = r->src[j].photons - r->next[j];
            if (left <= 0) continue;
            const int64_t c = std::min(take, left);
            DevSource d = r->src[j];
            d.photons = c;
            d.photon_base = (uint64_t)r->next[j];
            d.key_lo = (uint32_t)r->seed;
            d.key_hi = (uint32_t)(r->seed >> 32);
            d.rng_index = (uint32_t)j;
            d.slot = nslot;
            d.offset = off;
            stg.h_off[ns] = off;
            stg.h_src[ns++] = d;
            off += round_up64(c);
            r->next[j] += c;
            r->scheduled += c;
            take -= c;
            budget -= c;
            photons += c;
            kind = (kind == -2 || kind == d.kind) ? d.kind : -1;
        }
        ++nslot;
        if (r->scheduled == r->photons) {
            r->state = RQ_SCHEDULED;
            r->last_round = round;
        }
    }
    stg.h_off[ns] = off;
    HIP_TRY(hipMemcpyAsync(stg.d_src, stg.h_src, sizeof(DevSource) * ns, hipMemcpyHostToDevice, sv->stream));
    HIP_TRY(hipMemcpyAsync(stg.d_off, stg.h_off, sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice, sv->stream));
    HIP_TRY(hipMemcpyAsync(stg.d_slot, stg.h_slot, sizeof(SlotDesc) * nslot, hipMemcpyHostToDevice, sv->stream));
    HIP_TRY(hipEventRecord(stg.start, sv->stream));
    {
        std::lock_guard<std::mutex> lk(sv->ctx->mu);
        LaunchPlan lp{stg.d_src, stg.d_off, ns, off, kind < 0 ? kKernelVariants - 1 : kind, stg.d_slot,
                      HitCols{}, nullptr};
        st = launch_plan(sv->ctx, 0, lp, sv->stream);
        if (st) return st;
    }
    HIP_TRY(hipEventRecord(stg.done, sv->stream));
    stg.busy = true;
    stg.round = round;
    stg.photons = photons;
    // counters of the requests this round completes go to pinned memory as soon as it ends
    for (Request* r : want) {
        if (r->state != RQ_SCHEDULED || r->last_round != round) continue;
        const ReqLayout L(r->capacity);
        HIP_TRY(hipStreamWaitEvent(sv->copy_stream, stg.done, 0));
        HIP_TRY(hipMemcpyAsync(sv->h_arena + r->host_off + L.counters, sv->d_arena + r->dev_off + L.counters,
                               4 * sizeof(int64_t), hipMemcpyDeviceToHost, sv->copy_stream));
        HIP_TRY(hipEventRecord(r->ev, sv->copy_stream));
        r->state = RQ_COPY_COUNTS;
    }
    sv->stats.rounds += 1;
    sv->stats.photons += photons;
    return NT_OK;
}

// Advance copies of finished requests; returns true if anything changed.
bool progress(nt_server* sv) {
    bool changed = false;
    for (Staging& s : sv->staging) {
        if (!s.busy || hipEventQuery(s.done) != hipSuccess) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s.start, s.done) == hipSuccess) sv->stats.busy_ms += ms;
        s.busy = false;
        changed = true;
    }
    for (size_t i = 0; i < sv->active.size();) {
        Request& r = *sv->reqs[sv->active[i]];
        const hipError_t q = (r.state == RQ_COPY_COUNTS || r.state == RQ_COPY_HITS) ? hipEventQuery(r.ev)
                                                                                  : hipErrorNotReady;
        if (q != hipSuccess && q != hipErrorNotReady) { sv->fatal = NT_ERR_DEVICE; return true; }
        if (q == hipSuccess && r.state == RQ_COPY_COUNTS) {
            const ReqLayout L(r.capacity, r.epilogue);
            const int64_t* cnt = (const int64_t*)(sv->h_arena + r.host_off + L.counters);
            r.n_found = cnt[0];
            r.n_sim = cnt[1];
            const int64_t n = std::min(r.n_found, r.capacity);
            char* d = sv->d_arena + r.dev_off;
            char* h = sv->h_arena + r.host_off;
            const size_t off[12] = {L.pos, L.dir, L.idir, L.wl, L.time, L.dist, L.mid, L.sidx, L.pid,
                                    L.pmt, L.weight, L.tres};
            const size_t w[12] = {12, 12, 12, 4, 4, 4, 2, 4, 8, 1, 4, 4};
            const int ncol = r.epilogue ? 12 : 9;
            bool ok = true;
            for (int k = 0; k < ncol && n > 0; ++k)
                ok &= hipMemcpyAsync(h + off[k], d + off[k], w[k] * n, hipMemcpyDeviceToHost, sv->copy_stream) ==
                      hipSuccess;
            ok &= hipEventRecord(r.ev, sv->copy_stream) == hipSuccess;
            if (!ok) { sv->fatal = NT_ERR_DEVICE; return true; }
            r.state = RQ_COPY_HITS;
            changed = true;
        } else if (q == hipSuccess && r.state == RQ_COPY_HITS) {
            r.t_done = std::chrono::steady_clock::now();
            r.state = RQ_DONE;
            r.status = r.n_found > r.capacity ? NT_ERR_HIT_OVERFLOW : NT_OK;
            sv->dev_alloc.release(r.dev_off, r.bytes);
            sv->active.erase(sv->active.begin() + i);
            sv->stats.completed += 1;
            changed = true;
            continue;
        }
        ++i;
    }
    if (changed) sv->cv_done.notify_all();
    return changed;
}

constexpr int64_t kItemMax = int64_t(1) << 30;   // photons per work item
constexpr uint32_t kTicket = 4096;               // largest ticket (photons per claim)
constexpr uint32_t kIdleTicks = 200000;          // 2 ms of the 100 MHz real-time counter

int64_t items_of(const Request& r) {
    int64_t n = 0;
    for (const DevSource& d : r.src)
        if (d.photons > 0) n += (d.photons + kItemMax - 1) / kItemMax;
    return n;
}

// Publish the admitted requests that fit the running window: their slot and
// item records go straight into the host rings, then n_pub.
void worklist_publish(nt_server* sv) {
    uint32_t k = sv->items_pub;
    bool any = false;
    for (uint64_t id : sv->active) {
        Request& r = *sv->reqs[id];
        if (r.state != RQ_ACTIVE) continue;
        const int64_t need = items_of(r);
        if ((int64_t)(k - sv->win_items0) + need > kWorkItems ||
            sv->slots_pub + 1 - sv->win_slots0 > (uint64_t)kWorkSlots) {
            sv->window_full = true;
            break;
        }
        const int slot = (int)(sv->slots_pub % kWorkSlots);
        r.slot = slot;
        r.seq = (uint32_t)(sv->slots_pub + 1);
        sv->slots_pub += 1;
        const ReqLayout L(r.capacity, r.epilogue);
        SlotDesc& sd = sv->h_slots[slot];
        sd.cols = L.cols(sv->h_arena + r.host_off, r.capacity);
        sd.counters = sv->d_slot_cnt + 4 * slot;
        SlotExtra& x = sv->h_extra[slot];
        x.total = r.photons;
        x.host_counters = (int64_t*)(sv->h_arena + r.host_off + L.counters);
        for (int q = 0; q < 3; ++q) x.ref[q] = r.ref[q];
        x.epilogue = r.epilogue ? 1 : 0;
        x.seq = r.seq;
        for (size_t j = 0; j < r.src.size(); ++j) {
            for (int64_t at = 0; at < r.src[j].photons; at += kItemMax) {
                DevSource& d = sv->h_items[k % kWorkItems].src;
                d = r.src[j];
                d.photons = std::min(kItemMax, r.src[j].photons - at);
                d.photon_base = (uint64_t)at;
                d.key_lo = (uint32_t)r.seed;
                d.key_hi = (uint32_t)(r.seed >> 32);
                d.rng_index = (uint32_t)j;
                d.slot = slot;
                d.offset = 0;
                sv->h_items[k % kWorkItems].claimed = 0;
                ++k;
            }
        }
        r.state = RQ_SCHEDULED;
        any = true;
    }
    if (!any) return;
    sv->items_pub = k;
    __atomic_store_n(&sv->ctrl->n_pub, k, __ATOMIC_RELEASE);
}

int32_t worklist_launch(nt_server* sv, int i) {
    nt_ctx* c = sv->ctx;
    KParams kp{};
    kp.slots = sv->h_slots;
    kp.targets = c->d_targets;
    kp.module_ids = c->d_module_ids;
    kp.clusters = c->d_clusters;
    kp.members = c->d_members;
    kp.member_idx = c->d_member_idx;
    kp.n_clusters = c->culling ? c->n_clusters : 0;
    kp.cluster_reach = c->cluster_reach;
    kp.knots = c->d_knots;
    kp.medium = c->medium;
    kp.spectrum = c->spectrum;
    kp.sca_poly = c->sca_poly;
    kp.n_targets = c->n_targets;
    SrvParams sp{};
    sp.items = sv->h_items;
    sp.ditems = sv->d_items;
    sp.dw = sv->d_work;
    sp.extra = sv->h_extra;
    sp.slot_done = sv->d_slot_done;
    sp.cursor = reinterpret_cast<uint32_t*>(sv->d_claim);
    sp.busy_ticks = sv->d_claim + 1;
    sp.ctrl = sv->ctrl;
    sp.ep = c->epi;
    sp.target_of_module = c->d_target_of_module;
    sp.ticket = kTicket;
    sp.n_waves = (uint32_t)sv->srv_blocks * (kThreads / 64);
    sp.gen = ++sv->launch_gen;
    sp.idle_ticks = kIdleTicks;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        hipLaunchKernelGGL(serve_variant(c->medium.dir_mode, c->medium.sl_fraction > 0.0f), dim3(sv->srv_blocks),
                           dim3(kThreads),
                           lds_bytes(c->spectrum.n), sv->kstream[i], kp, sp);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(sv->kdone[i], sv->kstream[i]));
    sv->klaunched[i] = true;
    return NT_OK;
}

// Keep a kernel instance running while published requests are pending:
// advance the epoch, launch when none runs (or a second one when the host
// was away long enough for wavefronts to start leaving).  A full window is
// reset once it has drained and no instance runs: the device counts are
// zeroed while nothing reads them.
int32_t worklist_kernels(nt_server* sv, bool pending) {
    for (int i = 0; i < 2; ++i) {
        if (!sv->klaunched[i]) continue;
        const hipError_t q = hipEventQuery(sv->kdone[i]);
        if (q == hipSuccess) sv->klaunched[i] = false;
        else if (q != hipErrorNotReady) return NT_ERR_DEVICE;
    }
    const auto now = std::chrono::steady_clock::now();
    if (pending) {
        const bool stale = now - sv->last_bump > std::chrono::microseconds(kIdleTicks / 200);   // half the idle limit
        __atomic_fetch_add(&sv->ctrl->epoch, 1u, __ATOMIC_RELEASE);
        sv->last_bump = now;
        const int live = (int)sv->klaunched[0] + (int)sv->klaunched[1];
        if (live == 0 || (stale && live < 2)) {
            const int32_t st = worklist_launch(sv, sv->klaunched[0] ? 1 : 0);
            if (st) return st;
            sv->stats.rounds += 1;
        }
    } else if (sv->window_full && !sv->klaunched[0] && !sv->klaunched[1]) {
        HIP_TRY(hipMemsetAsync(sv->d_slot_done, 0, sizeof(int64_t) * kWorkSlots, sv->upload));
        HIP_TRY(hipMemsetAsync(sv->d_slot_cnt, 0, 4 * sizeof(int64_t) * kWorkSlots, sv->upload));
        HIP_TRY(hipStreamSynchronize(sv->upload));
        sv->win_items0 = sv->items_pub;
        sv->win_slots0 = sv->slots_pub;
        sv->window_full = false;
    }
    return NT_OK;
}

// Requests whose completion mark arrived are done: the kernel wrote their
// hits and counters into their pinned host block.
bool worklist_poll(nt_server* sv) {
    bool changed = false;
    for (size_t i = 0; i < sv->active.size();) {
        Request& r = *sv->reqs[sv->active[i]];
        if (r.state != RQ_SCHEDULED || __atomic_load_n(&sv->ctrl->done[r.slot], __ATOMIC_ACQUIRE) != r.seq) {
            ++i;
            continue;
        }
        const ReqLayout L(r.capacity, r.epilogue);
        const volatile int64_t* cnt = (const volatile int64_t*)(sv->h_arena + r.host_off + L.counters);
        r.n_found = cnt[0];
        r.n_sim = cnt[1];
        r.t_done = std::chrono::steady_clock::now();
        r.state = RQ_DONE;
        r.status = r.n_found > r.capacity ? NT_ERR_HIT_OVERFLOW : NT_OK;
        sv->active.erase(sv->active.begin() + i);
        sv->stats.completed += 1;
        sv->stats.photons += r.photons;
        changed = true;
    }
    if (changed) sv->cv_done.notify_all();
    return changed;
}

// Admission for the work list: pinned host memory only (the kernel writes
// the results there directly).
void worklist_admit(nt_server* sv) {
    while (!sv->queue.empty()) {
        Request& r = *sv->reqs[sv->queue.front()];
        size_t hoff;
        if (!sv->host_alloc.alloc(r.bytes, hoff)) return;
        r.host_off = hoff;
        r.state = RQ_ACTIVE;
        sv->active.push_back(r.id);
        sv->queue.pop_front();
    }
}

void worklist_loop(nt_server* sv, std::unique_lock<std::mutex>& lk) {
    while (!sv->stop && sv->fatal == NT_OK) {
        worklist_admit(sv);
        worklist_publish(sv);
        bool pending = false;
        for (uint64_t id : sv->active) pending |= sv->reqs[id]->state == RQ_SCHEDULED;
        const int32_t st = worklist_kernels(sv, pending);
        if (st) {
            sv->fatal = st;
            break;
        }
        if (!worklist_poll(sv)) {
            const bool idle = sv->queue.empty() && sv->active.empty();
            if (idle) sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            else sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

void server_loop(nt_server* sv) {
    (void)hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->mode == NT_SERVER_WORKLIST) worklist_loop(sv, lk);
    while (!sv->stop && sv->fatal == NT_OK) {
        admit(sv);
        bool launched = false;
        for (Staging& s : sv->staging) {
            if (s.busy) continue;
            const int32_t st = launch_round(sv, s);
            if (st == NT_OK) launched = true;
            else if (st != NT_ERR_NOT_READY) sv->fatal = st;
            break;
        }
        const bool changed = progress(sv);
        if (!launched && !changed) {
            const bool idle = sv->queue.empty() && sv->active.empty();
            if (idle) sv->cv_work.wait(lk);
            else sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
    if (sv->fatal != NT_OK) {
        for (auto& kv : sv->reqs)
            if (kv.second->state != RQ_DONE) { kv.second->state = RQ_FAILED; kv.second->status = sv->fatal; }
        sv->cv_done.notify_all();
    }
}

void free_server(nt_server* sv) {
    for (Staging& s : sv->staging) {
        if (s.h_src) { (void)hipHostFree(s.h_src); (void)hipHostFree(s.h_off); (void)hipHostFree(s.h_slot); }
        if (s.d_src) { (void)hipFree(s.d_src); (void)hipFree(s.d_off); (void)hipFree(s.d_slot); }
        if (s.start) (void)hipEventDestroy(s.start);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    for (auto& kv : sv->reqs)
        if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
    if (sv->d_arena) (void)hipFree(sv->d_arena);
    if (sv->h_arena) (void)hipHostFree(sv->h_arena);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    if (sv->copy_stream) (void)hipStreamDestroy(sv->copy_stream);
    if (sv->ctrl) (void)hipHostFree(sv->ctrl);
    if (sv->h_items) (void)hipHostFree(sv->h_items);
    if (sv->h_slots) (void)hipHostFree(sv->h_slots);
    if (sv->h_extra) (void)hipHostFree(sv->h_extra);
    if (sv->d_slot_done) (void)hipFree(sv->d_slot_done);
    if (sv->d_slot_cnt) (void)hipFree(sv->d_slot_cnt);
    if (sv->d_claim) (void)hipFree(sv->d_claim);
    if (sv->d_work) (void)hipFree(sv->d_work);
    if (sv->d_items) (void)hipFree(sv->d_items);
    if (sv->upload) (void)hipStreamDestroy(sv->upload);
    for (int i = 0; i < 2; ++i) {
        if (sv->kstream[i]) (void)hipStreamDestroy(sv->kstream[i]);
        if (sv->kdone[i]) (void)hipEventDestroy(sv->kdone[i]);
    }
    delete sv;
}

}  // namespace

extern "C" {

int32_t nt_server_create(nt_ctx* c, const nt_server_config* cfg, nt_server** out) {
    if (!c || !cfg || !out || cfg->device_budget_bytes < kAlign || cfg->host_budget_bytes < kAlign ||
        cfg->round_photons < 64 || cfg->max_inflight_rounds < 1 || cfg->max_inflight_rounds > 16 ||
        (cfg->mode != NT_SERVER_ROUNDS && cfg->mode != NT_SERVER_WORKLIST)) {
        snprintf(g_err, sizeof g_err,
                 "server config: budgets >= 256 B, round_photons >= 64, 1..16 rounds in flight, a known mode");
        return NT_ERR_ARG;
    }
    if (!c->has_medium || !c->has_spectrum || !c->has_targets) {
        snprintf(g_err, sizeof g_err, "set medium, spectrum and targets before creating a server");
        return NT_ERR_STATE;
    }
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    nt_server* sv = new nt_server();
    sv->ctx = c;
    sv->cfg = *cfg;
    auto fail = [&](hipError_t e, const char* what) { free_server(sv); return fail_hip(e, what); };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    if ((e = hipStreamCreateWithFlags(&sv->copy_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    const bool wl = cfg->mode == NT_SERVER_WORKLIST;
    // work list: the kernel writes results straight into fine-grained pinned memory
    if (!wl && (e = hipMalloc((void**)&sv->d_arena, cfg->device_budget_bytes)) != hipSuccess)
        return fail(e, "device budget");
    if ((e = hipHostMalloc((void**)&sv->h_arena, cfg->host_budget_bytes, wl ? hipHostMallocCoherent : 0)) !=
        hipSuccess)
        return fail(e, "host budget");
    sv->dev_alloc.init(cfg->device_budget_bytes);
    sv->host_alloc.init(cfg->host_budget_bytes);
    sv->staging.resize(cfg->max_inflight_rounds);
    for (Staging& s : sv->staging) {
        if ((e = hipEventCreate(&s.start)) != hipSuccess) return fail(e, "event");
        if ((e = hipEventCreate(&s.done)) != hipSuccess) return fail(e, "event");
    }
    sv->mode = cfg->mode;
    if (sv->mode == NT_SERVER_WORKLIST) {
        if ((e = hipHostMalloc((void**)&sv->ctrl, sizeof(WorkCtrl), hipHostMallocCoherent)) != hipSuccess)
            return fail(e, "control block");
        memset(sv->ctrl, 0, sizeof(WorkCtrl));
        if ((e = hipHostMalloc((void**)&sv->h_items, sizeof(WorkItem) * kWorkItems, hipHostMallocCoherent)) !=
                hipSuccess ||
            (e = hipHostMalloc((void**)&sv->h_slots, sizeof(SlotDesc) * kWorkSlots, hipHostMallocCoherent)) !=
                hipSuccess ||
            (e = hipHostMalloc((void**)&sv->h_extra, sizeof(SlotExtra) * kWorkSlots, hipHostMallocCoherent)) !=
                hipSuccess)
            return fail(e, "work-list rings");
        if ((e = hipMalloc((void**)&sv->d_slot_done, sizeof(int64_t) * kWorkSlots)) != hipSuccess ||
            (e = hipMalloc((void**)&sv->d_slot_cnt, 4 * sizeof(int64_t) * kWorkSlots)) != hipSuccess ||
            (e = hipMalloc((void**)&sv->d_claim, 2 * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMalloc((void**)&sv->d_work, sizeof(DevWork))) != hipSuccess ||
            (e = hipMalloc((void**)&sv->d_items, sizeof(WorkItem) * kWorkItems)) != hipSuccess)
            return fail(e, "work-list counters");
        if ((e = hipMemset(sv->d_claim, 0, 2 * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMemset(sv->d_work, 0, sizeof(DevWork))) != hipSuccess ||
            (e = hipMemset(sv->d_slot_done, 0, sizeof(int64_t) * kWorkSlots)) != hipSuccess ||
            (e = hipMemset(sv->d_slot_cnt, 0, 4 * sizeof(int64_t) * kWorkSlots)) != hipSuccess)
            return fail(e, "work-list counters");
        if ((e = hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
        for (int i = 0; i < 2; ++i) {
            if ((e = hipStreamCreateWithFlags(&sv->kstream[i], hipStreamNonBlocking)) != hipSuccess)
                return fail(e, "stream");
            if ((e = hipEventCreateWithFlags(&sv->kdone[i], hipEventDisableTiming)) != hipSuccess)
                return fail(e, "event");
        }
        int per_cu = 0;
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_variant(c->medium.dir_mode, true), kThreads,
                                                             lds_bytes(NT_MAX_KNOTS))) != hipSuccess)
            return fail(e, "occupancy");
        sv->srv_blocks = std::max(1, per_cu) * c->num_cus;
        if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "work-list rings");
    }
    sv->worker = std::thread(server_loop, sv);
    *out = sv;
    return NT_OK;
}

static int32_t server_submit(nt_server* sv, const nt_source* s, int32_t n, uint64_t seed, int64_t cap,
                             const float* ref, uint64_t* id) {
    if (!sv || n < 1 || !s || cap < 0 || !id) return NT_ERR_ARG;
    auto r = std::make_unique<Request>();
    if (ref) {
        if (sv->mode != NT_SERVER_WORKLIST || !sv->ctx->epi_enabled) {
            snprintf(g_err, sizeof g_err, "epilogue columns need a work-list server and nt_set_epilogue");
            return NT_ERR_STATE;
        }
        if (!std::isfinite(ref[0]) || !std::isfinite(ref[1]) || !std::isfinite(ref[2])) return NT_ERR_ARG;
        r->epilogue = true;
        for (int q = 0; q < 3; ++q) r->ref[q] = ref[q];
    }
    r->src.resize(n);
    r->next.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        const int32_t st = make_dev_source(s[i], i, r->src[i]);
        if (st) return st;
        r->photons += s[i].photons;
    }
    r->seed = seed;
    r->capacity = cap;
    r->bytes = ReqLayout(cap, r->epilogue).bytes;
    if ((sv->mode != NT_SERVER_WORKLIST && r->bytes > sv->cfg.device_budget_bytes) ||
        r->bytes > sv->cfg.host_budget_bytes) {
        snprintf(g_err, sizeof g_err, "request needs %zu bytes of hit buffers, above the budget", r->bytes);
        return NT_ERR_BUDGET;
    }
    if (hipSetDevice(sv->ctx->device) != hipSuccess || hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess)
        return NT_ERR_DEVICE;
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->fatal != NT_OK) { (void)hipEventDestroy(r->ev); return sv->fatal; }
    r->id = sv->next_id++;
    *id = r->id;
    if (r->photons == 0) {                       // nothing to propagate: complete at once
        r->state = RQ_DONE;
        r->t_done = r->t_submit;
        sv->stats.completed += 1;
    } else {
        sv->queue.push_back(r->id);
    }
    sv->reqs[r->id] = std::move(r);
    sv->stats.submitted += 1;
    sv->cv_work.notify_one();
    return NT_OK;
}

int32_t nt_server_submit(nt_server* sv, const nt_source* s, int32_t n, uint64_t seed, int64_t cap, uint64_t* id) {
    return server_submit(sv, s, n, seed, cap, nullptr, id);
}

int32_t nt_server_submit_epilogue(nt_server* sv, const nt_source* s, int32_t n, uint64_t seed, int64_t cap,
                                  const float* ref_position, uint64_t* id) {
    if (!ref_position) return NT_ERR_ARG;
    return server_submit(sv, s, n, seed, cap, ref_position, id);
}

int32_t nt_server_poll(nt_server* sv, uint64_t id) {
    if (!sv) return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return NT_ERR_ARG;
    if (it->second->state == RQ_FAILED) return it->second->status;
    return it->second->state == RQ_DONE ? NT_OK : NT_ERR_NOT_READY;
}

int32_t nt_server_wait(nt_server* sv, uint64_t id, const nt_hits* out, int64_t* n_hits, int64_t* n_sim,
                       int64_t* n_dropped, double* latency_ms) {
    if (!sv || !out) return NT_ERR_ARG;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) return NT_ERR_ARG;
        Request* rp = it->second.get();
        if (out->capacity < rp->capacity) {
            snprintf(g_err, sizeof g_err, "output capacity below the request's hit capacity");
            return NT_ERR_ARG;
        }
        sv->cv_done.wait(lk, [&] { return rp->state == RQ_DONE || rp->state == RQ_FAILED; });
        r = std::move(it->second);
        sv->reqs.erase(it);
    }
    int32_t status = r->state == RQ_FAILED ? r->status : NT_OK;
    const int64_t n = std::min(r->n_found, r->capacity);
    if (status == NT_OK && r->photons > 0 && n > 0) {
        const ReqLayout L(r->capacity);
        const char* h = sv->h_arena + r->host_off;
        memcpy(out->position, h + L.pos, 12 * n);
        memcpy(out->direction, h + L.dir, 12 * n);
        memcpy(out->initial_direction, h + L.idir, 12 * n);
        memcpy(out->wavelength, h + L.wl, 4 * n);
        memcpy(out->time, h + L.time, 4 * n);
        memcpy(out->dist_travelled, h + L.dist, 4 * n);
        memcpy(out->module_id, h + L.mid, 2 * n);
        if (out->source_index) memcpy(out->source_index, h + L.sidx, 4 * n);
        if (out->photon_id) memcpy(out->photon_id, h + L.pid, 8 * n);
        if (r->epilogue) {
            const ReqLayout Le(r->capacity, true);
            if (out->pmt_id) memcpy(out->pmt_id, h + Le.pmt, n);
            if (out->total_weight) memcpy(out->total_weight, h + Le.weight, 4 * n);
            if (out->tres) memcpy(out->tres, h + Le.tres, 4 * n);
        }
    }
    if (n_hits) *n_hits = status == NT_OK ? n : 0;
    if (n_sim) *n_sim = r->n_sim;
    if (n_dropped) *n_dropped = std::max<int64_t>(0, r->n_found - r->capacity);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (status == NT_OK && r->n_found > r->capacity) status = NT_ERR_HIT_OVERFLOW;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (r->photons > 0 && r->state == RQ_DONE) sv->host_alloc.release(r->host_off, r->bytes);
        sv->cv_work.notify_one();                  // freed host memory may admit the next request
    }
    if (r->ev) (void)hipEventDestroy(r->ev);
    return status;
}

int32_t nt_server_stats_get(nt_server* sv, nt_server_stats* out) {
    if (!sv || !out) return NT_ERR_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->stats;
    out->queued = (int64_t)sv->queue.size();
    out->active = (int64_t)sv->active.size();
    if (sv->mode == NT_SERVER_WORKLIST) {
        // Σ ticket durations over all wavefronts ÷ wavefronts of one instance: the
        // time the whole device spent propagating
        unsigned long long ticks = 0;
        if (hipSetDevice(sv->ctx->device) != hipSuccess ||
            hipMemcpyAsync(&ticks, sv->d_claim + 1, sizeof ticks, hipMemcpyDeviceToHost, sv->upload) != hipSuccess ||
            hipStreamSynchronize(sv->upload) != hipSuccess)
            return NT_ERR_DEVICE;
        out->busy_ms = (double)ticks * 1e-5 / (double)((int64_t)sv->srv_blocks * (kThreads / 64));
    }
    return sv->fatal;
}

void nt_server_destroy(nt_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        if (sv->ctrl) __atomic_store_n(&sv->ctrl->stop, 1u, __ATOMIC_RELEASE);   // wavefronts leave at their next claim
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->ctx->device);
    (void)hipStreamSynchronize(sv->stream);
    (void)hipStreamSynchronize(sv->copy_stream);
    for (int i = 0; i < 2; ++i)
        if (sv->kstream[i]) (void)hipStreamSynchronize(sv->kstream[i]);
    if (sv->upload) (void)hipStreamSynchronize(sv->upload);
    free_server(sv);
}

}  // extern "C"
