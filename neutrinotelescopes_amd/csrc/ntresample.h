// ntresample.h — internal interface of the device resampler (ntresample.hip),
// used by nt_resample in ntprop.hip.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ntp {

struct ResampleArgs {
    const uint16_t* module_id;
    const uint8_t* pmt_id;           // may be null when per_pmt == 0
    const float* weight;
    const uint32_t* source_index;
    const uint64_t* photon_id;
    int64_t n;                       // hits
    uint64_t seed;
    double downsample;
    int per_pmt;
    uint8_t* detected;               // [n] out
    int64_t* n_detected;             // device, out (may be null)
};

// Device workspace the resampler needs for n hits (bytes, 256-aligned parts).
size_t resample_workspace_bytes(int64_t n);

// Enqueue the resampling on `stream`; `ws` holds resample_workspace_bytes(n).
hipError_t resample_run(const ResampleArgs& a, void* ws, size_t ws_bytes, hipStream_t stream);

}  // namespace ntp
