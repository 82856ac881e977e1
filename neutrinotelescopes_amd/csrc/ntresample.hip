// ntresample.hip — resample_simulation (src/pmt_frontend/pipeline.jl:86-116)
// on the device: weighted photons → one realisation of detected hits.
//
// Per group (module and PMT, or module), the reference draws
//   n = min(pois_rand(Σw · downsample), #photons with w > 0)
// and samples n of the group's photons without replacement with probabilities
// ∝ w (StatsBase sample(...; replace=false) with ProbabilityWeights).  Here:
//   * every hit with w > 0 gets the key ln(u)/w, u in (0, 1] a 53-bit uniform
//     of its photon (Philox4x32-10, key = seed, counter = (0x71000000, photon
//     lo, photon hi, source index)); the n largest keys of a group are a
//     sample of n without replacement ∝ w (Efraimidis & Spirakis, "Weighted
//     random sampling with a reservoir", IPL 97 (2006) 181-185), the same law
//     as n successive weighted draws;
//   * the hits are ordered by key (one radix sort of the keys as 64-bit
//     ordered integers) and then, stably, by group (a second radix sort of
//     the group number), so every group is contiguous with its keys in
//     descending order — no atomics, and the result does not depend on the
//     order the hits were stored in;
//   * per group one thread sums the weights in that order and draws n with the
//     library's Poisson sampler (key = seed, counter (trial, group, 0,
//     0x70000000));
//   * hit p of a group is detected iff its rank in the group is below n.
// The keys and sums are FP64; HBM-bound integer/byte work otherwise.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "ntprop_device.h"
#include "ntresample.h"

namespace ntp {
namespace {

constexpr uint32_t kKeyWord = 0x71000000u;     // counter word 0 of the per-photon key uniform
constexpr uint32_t kCountWord = 0x70000000u;   // counter word 3 of the per-group Poisson draw
constexpr int kGroupBits = 21;                 // module id (16 bits) × 17 PMT slots
constexpr int64_t kGroups = int64_t(65536) * 17;

__device__ __forceinline__ uint32_t group_of(uint16_t module, uint8_t pmt, int per_pmt) {
    return per_pmt ? (uint32_t)module * 17u + (uint32_t)min((int)pmt, 16) : (uint32_t)module;
}

// Descending order of a double as ascending order of a 64-bit integer.
__device__ __forceinline__ uint64_t desc_bits(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint64_t asc = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    return ~asc;
}

__global__ __launch_bounds__(256) void keys_kernel(ResampleArgs a, uint32_t k0, uint32_t k1, uint64_t* key,
                                                   int32_t* idx) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) {
        const float w = a.weight[i];
        double kd = -INFINITY;
        if (w > 0.0f) {
            const uint64_t pid = a.photon_id[i];
            const uint4 r = philox(make_uint4(kKeyWord, (uint32_t)pid, (uint32_t)(pid >> 32), a.source_index[i]),
                                   k0, k1);
            kd = log(1.0 - u53(r.x, r.y)) / (double)w;            // u in (0, 1]
        }
        key[i] = desc_bits(kd);
        idx[i] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void group_keys_kernel(ResampleArgs a, const int32_t* __restrict__ idx,
                                                         uint64_t* gkey) {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * 256) {
        const int32_t i = idx[p];
        gkey[p] = group_of(a.module_id[i], a.per_pmt ? a.pmt_id[i] : (uint8_t)0, a.per_pmt);
    }
}

__global__ __launch_bounds__(256) void clear_groups_kernel(int32_t* start, int32_t* end, int64_t* count) {
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < kGroups; g += (int64_t)gridDim.x * 256) {
        start[g] = 0;
        end[g] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
}

__global__ __launch_bounds__(256) void bounds_kernel(const uint64_t* __restrict__ g, int64_t n, int32_t* start,
                                                     int32_t* end) {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
        const uint64_t gp = g[p];
        if (p == 0 || g[p - 1] != gp) start[gp] = (int32_t)p;
        if (p == n - 1 || g[p + 1] != gp) end[gp] = (int32_t)(p + 1);
    }
}

__global__ __launch_bounds__(256) void draw_kernel(ResampleArgs a, const int32_t* __restrict__ idx,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ end, uint32_t k0, uint32_t k1,
                                                   int32_t* nsel) {
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < kGroups; g += (int64_t)gridDim.x * 256) {
        const int32_t s = start[g], e = end[g];
        if (e <= s) continue;
        double wsum = 0.0;
        int32_t positive = 0;
        for (int32_t p = s; p < e; ++p) {
            const float w = a.weight[idx[p]];
            wsum += (double)w;
            positive += w > 0.0f;
        }
        const int64_t k = poisson_philox(wsum * a.downsample, k0, k1, (uint32_t)g, 0u, kCountWord);
        nsel[g] = (int32_t)min<int64_t>(k, positive);
    }
}

__global__ __launch_bounds__(256) void mark_kernel(ResampleArgs a, const uint64_t* __restrict__ g,
                                                   const int32_t* __restrict__ idx,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ nsel, int64_t* count) {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * 256) {
        const uint64_t gp = g[p];
        const bool sel = (int32_t)(p - start[gp]) < nsel[gp];
        a.detected[idx[p]] = sel ? 1 : 0;
        const uint64_t m = __ballot(sel);
        if (sel && __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) == 0)
            atomicAdd((unsigned long long*)count, (unsigned long long)__popcll(m));
    }
}

struct Parts {
    size_t key_a, key_b, idx_a, idx_b, start, end, nsel, count, cub, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t cub_bytes(int64_t n) {
    size_t t = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, 64);
    return t;
}

Parts parts(int64_t n) {
    Parts p{};
    size_t o = 0;
    auto take = [&](size_t b) { const size_t at = o; o += align256(b); return at; };
    const size_t m = (size_t)std::max<int64_t>(n, 1);
    p.key_a = take(8 * m);
    p.key_b = take(8 * m);
    p.idx_a = take(4 * m);
    p.idx_b = take(4 * m);
    p.start = take(4 * (size_t)kGroups);
    p.end = take(4 * (size_t)kGroups);
    p.nsel = take(4 * (size_t)kGroups);
    p.count = take(8);
    p.cub = take(cub_bytes(n));
    p.total = o;
    return p;
}

}  // namespace

size_t resample_workspace_bytes(int64_t n) { return parts(n).total; }

hipError_t resample_run(const ResampleArgs& a, void* ws, size_t ws_bytes, hipStream_t s) {
    const Parts p = parts(a.n);
    if (ws_bytes < p.total || a.n > INT32_MAX) return hipErrorInvalidValue;
    char* b = static_cast<char*>(ws);
    uint64_t* key_a = (uint64_t*)(b + p.key_a);
    uint64_t* key_b = (uint64_t*)(b + p.key_b);
    int32_t* idx_a = (int32_t*)(b + p.idx_a);
    int32_t* idx_b = (int32_t*)(b + p.idx_b);
    int32_t* start = (int32_t*)(b + p.start);
    int32_t* end = (int32_t*)(b + p.end);
    int32_t* nsel = (int32_t*)(b + p.nsel);
    int64_t* count = a.n_detected ? a.n_detected : (int64_t*)(b + p.count);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const int nb = (int)std::min<int64_t>(std::max<int64_t>((a.n + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(clear_groups_kernel, dim3(1024), dim3(256), 0, s, start, end, count);
    if (a.n == 0) return hipGetLastError();
    hipLaunchKernelGGL(keys_kernel, dim3(nb), dim3(256), 0, s, a, k0, k1, key_a, idx_a);
    size_t cub = cub_bytes(a.n);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(b + p.cub, cub, key_a, key_b, idx_a, idx_b, (int)a.n, 0, 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(group_keys_kernel, dim3(nb), dim3(256), 0, s, a, idx_b, key_a);
    // stable: the keys' descending order survives inside each group
    e = hipcub::DeviceRadixSort::SortPairs(b + p.cub, cub, key_a, key_b, idx_b, idx_a, (int)a.n, 0, kGroupBits, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bounds_kernel, dim3(nb), dim3(256), 0, s, key_b, a.n, start, end);
    hipLaunchKernelGGL(draw_kernel, dim3(1024), dim3(256), 0, s, a, idx_a, start, end, k0, k1, nsel);
    hipLaunchKernelGGL(mark_kernel, dim3(nb), dim3(256), 0, s, a, key_b, idx_a, start, nsel, count);
    return hipGetLastError();
}

}  // namespace ntp
