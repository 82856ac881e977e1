// ntsources.hip — device-side source builder: particles → extended-Cherenkov
// sources with a Poisson photon count (include/ntprop.h, nt_build_cascade_sources).
//
// Reference (paths relative to the reference repository root):
//   ExtendedCherenkovEmitter(particle, medium, wl_range)  src/photon_propagation/lightyield.jl:471-481
//   LongitudinalParameterisation                        lightyield.jl:139-155, 204-213
//   LongitudinalParameters (EMinus, EPlus, Gamma)       lightyield.jl:135-137
//   CherenkovTrackLengthParameters, total_lightyield    lightyield.jl:204-213, 350-366
//
//   VolumeInjector, Cylinder, Cuboid, UniformAngularDistribution
//                                                       src/event_generation/injectors.jl:30-111
//
// Poisson variates: inversion by sequential search for a mean below 10;
// W. Hörmann, "The transformed rejection method for generating Poisson random
// variables", Insurance: Mathematics and Economics 12 (1993) 39-45 (PTRS)
// above, in double precision, driven by Philox4x32-10.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "../../include/ntprop.h"
#include "ntprop_device.h"

namespace {

constexpr uint32_t kPoissonWord = 0x50000000u;   // counter word 3 of the Poisson draws
constexpr uint32_t kInjectWord = 0x60000000u;    // counter word 3 of the injector draws

struct TypeParams {
    double long_alpha, long_beta, long_b;     // LongitudinalParameters
    double track_alpha, track_beta;           // CherenkovTrackLengthParameters (m, E in GeV)
};

__device__ __forceinline__ TypeParams type_params(int type) {
    switch (type) {
        case NT_PARTICLE_EPLUS: return {2.00035, 1.45501, 0.63008, 5.3211320598, 0.99999254};
        case NT_PARTICLE_GAMMA: return {2.83923, 1.34031, 0.64526, 5.3208540905, 0.99999877};
        default: return {2.01849, 1.45469, 0.63207, 5.3207078881, 1.00000211};
    }
}

__device__ __forceinline__ int64_t poisson(double mu, uint32_t k0, uint32_t k1, uint32_t ilo, uint32_t ihi) {
    return ntp::poisson_philox(mu, k0, k1, ilo, ihi, kPoissonWord);
}

__global__ __launch_bounds__(256) void build_sources_kernel(const nt_particle* __restrict__ in, int n,
                                                           nt_cascade_model model, uint32_t k0, uint32_t k1,
                                                           nt_source* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const nt_particle p = in[i];
    const TypeParams tp = type_params(p.type);
    const double e = (double)p.energy;
    nt_source s;
    s.kind = NT_SRC_EXTENDED_CHERENKOV;
    s.reserved = 0;
    const double mean = model.photons_per_m * tp.track_alpha * pow(e, tp.track_beta) * model.oversample;
    s.photons = poisson(mean, k0, k1, (uint32_t)i, 0u);
    for (int q = 0; q < 3; ++q) {
        s.position[q] = p.position[q];
        s.direction[q] = p.direction[q];
    }
    s.time = p.time;
    s.time_end = 0.0f;
    s.angle = 0.0f;
    s.long_a = (float)(tp.long_alpha + tp.long_beta * log10(e));
    s.long_b = (float)tp.long_b;
    s.long_lrad = (float)model.lrad;
    out[i] = s;
}

// 24-bit uniform on [0, 1), exact in FP64
__device__ __forceinline__ double u24(uint32_t w) { return (double)(w >> 8) * 0x1p-24; }

__global__ __launch_bounds__(256) void inject_kernel(nt_injector inj, int64_t n, uint32_t k0, uint32_t k1,
                                                    nt_particle* __restrict__ out) {
    const double two_pi = 6.283185307179586;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t ilo = (uint32_t)i, ihi = (uint32_t)(i >> 32);
        const uint4 a = ntp::philox(make_uint4(0u, ilo, ihi, kInjectWord), k0, k1);
        const uint4 b = ntp::philox(make_uint4(1u, ilo, ihi, kInjectWord), k0, k1);
        double x = inj.center[0], y = inj.center[1], z = inj.center[2];
        if (inj.volume == NT_VOLUME_CYLINDER) {
            z += (u24(a.z) - 0.5) * (double)inj.height;
            const double r = sqrt(u24(a.x) * (double)inj.radius);
            const double phi = two_pi * u24(a.y);
            x += r * cos(phi);
            y += r * sin(phi);
        } else if (inj.volume == NT_VOLUME_CUBOID) {
            x += (u24(a.x) - 0.5) * (double)inj.size[0];
            y += (u24(a.y) - 0.5) * (double)inj.size[1];
            z += (u24(a.z) - 0.5) * (double)inj.size[2];
        }
        const double e = (double)inj.e_min * pow((double)inj.e_max / (double)inj.e_min, u24(a.w));
        const double phi = two_pi * u24(b.x);
        const double ct = 2.0 * u24(b.y) - 1.0;
        const double st = sqrt(1.0 - ct * ct);
        const double ut = u24(b.z);
        const int type = ut < (double)inj.type_cdf[0] ? NT_PARTICLE_EMINUS
                         : ut < (double)inj.type_cdf[1] ? NT_PARTICLE_EPLUS : NT_PARTICLE_GAMMA;
        nt_particle p;
        p.position[0] = (float)x;
        p.position[1] = (float)y;
        p.position[2] = (float)z;
        p.direction[0] = (float)(cos(phi) * st);
        p.direction[1] = (float)(sin(phi) * st);
        p.direction[2] = (float)ct;
        p.time = 0.0f;
        p.energy = (float)e;
        p.type = type;
        p.reserved = 0;
        out[i] = p;
    }
}

}  // namespace

extern "C" int32_t nt_inject_particles(nt_ctx* ctx, const nt_injector* inj, int64_t n, uint64_t seed,
                                       nt_particle* d_particles, void* stream) {
    if (!ctx || !inj || n < 0 || (n > 0 && !d_particles)) return NT_ERR_ARG;
    if (inj->volume < NT_VOLUME_CYLINDER || inj->volume > NT_VOLUME_FIXED) return NT_ERR_ARG;
    if (!(inj->height >= 0.0f) || !(inj->radius >= 0.0f) || !(inj->size[0] >= 0.0f) || !(inj->size[1] >= 0.0f) ||
        !(inj->size[2] >= 0.0f))
        return NT_ERR_ARG;
    if (!(inj->e_min > 0.0f) || !(inj->e_max >= inj->e_min) || !std::isfinite(inj->e_max)) return NT_ERR_ARG;
    if (!(inj->type_cdf[0] >= 0.0f) || !(inj->type_cdf[1] >= inj->type_cdf[0]) ||
        !(inj->type_cdf[2] >= inj->type_cdf[1]) || !(inj->type_cdf[2] > 0.0f) || inj->type_cdf[2] > 1.0f)
        return NT_ERR_ARG;
    if (n == 0) return NT_OK;
    const int64_t want = (n + 255) / 256;
    const int blocks = (int)std::min<int64_t>(want, 65536);
    hipLaunchKernelGGL(inject_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *inj, n, (uint32_t)seed,
                       (uint32_t)(seed >> 32), d_particles);
    return hipGetLastError() == hipSuccess ? NT_OK : NT_ERR_DEVICE;
}

extern "C" int32_t nt_build_cascade_sources(nt_ctx* ctx, const nt_particle* d_particles, int32_t n,
                                            const nt_cascade_model* model, uint64_t seed, nt_source* d_sources,
                                            void* stream) {
    if (!ctx || !model || n < 0 || (n > 0 && (!d_particles || !d_sources))) return NT_ERR_ARG;
    if (!(model->photons_per_m >= 0.0) || !(model->oversample > 0.0) || !(model->lrad > 0.0)) return NT_ERR_ARG;
    if (n == 0) return NT_OK;
    hipLaunchKernelGGL(build_sources_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_particles, n,
                       *model, (uint32_t)seed, (uint32_t)(seed >> 32), d_sources);
    return hipGetLastError() == hipSuccess ? NT_OK : NT_ERR_DEVICE;
}
