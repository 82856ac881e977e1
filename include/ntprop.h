/*
 * ntprop.h — C ABI of the MI355X photon-propagation library (libntprop.so).
 *
 * The library is a drop-in for the photon-propagation API of the Julia package
 * NeutrinoTelescopes (module PhotonPropagationCuda).  Every entry point below
 * names the reference interface it replaces (paths relative to the reference
 * repository root).  Plain C types only: pointers, sizes, POD structs.
 *
 * Conventions
 *  - Every function returns an int32 status (NT_OK = 0); no exception, device
 *    trap or abort crosses the ABI.  The reference traps the kernel on hit
 *    overflow (src/photon_propagation/photon_prop_cuda.jl:484); here overflow is
 *    NT_ERR_HIT_OVERFLOW with the first `capacity` hits kept and the number of
 *    lost hits reported.
 *  - A context (nt_ctx) owns device-resident state: medium, spectrum table,
 *    targets, sources, counters.  set_* calls are single-writer; each waits
 *    for the kernels of the context's last launch (nt_launch on any stream,
 *    nt_propagate) to finish before it rewrites the tables they read.  A
 *    server (nt_server_*) needs them left unchanged while it runs.
 *  - Random numbers: Philox4x32-10 (Salmon et al., SC'11).  key = (seed_lo32,
 *    seed_hi32); counter = (block, photon_lo32, photon_hi32, source_index),
 *    where photon is the photon's index inside its source.  Block 0 holds the
 *    initial-state draws; scattering steps read the words of blocks 1, 2, ...
 *    in order, three per step (step length, HG/SL uniform and azimuth from the
 *    top 24 bits, the mixture uniform from the three low bytes); block
 *    0x80000000+t holds the draws of Gamma rejection trial t.  A photon's history is
 *    therefore a pure function of (seed, source_index, photon index), whatever
 *    the launch shape or the other sources in the batch.  The reference seeds
 *    CUDA.jl's device RNG with seed + thread index (photon_prop_cuda.jl:423).
 */
#ifndef NTPROP_H
#define NTPROP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NTPROP_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
    NT_OK = 0,
    NT_ERR_ARG = 1,          /* malformed argument (reference: error() in constructors) */
    NT_ERR_OOM = 2,          /* device allocation failed */
    NT_ERR_HIT_OVERFLOW = 3, /* more hits than capacity; first `capacity` kept */
    NT_ERR_DEVICE = 4,       /* HIP runtime error */
    NT_ERR_BUDGET = 5,       /* request does not fit the HBM budget (serving) */
    NT_ERR_STATE = 6,        /* call out of order (e.g. propagate before set_medium) */
    NT_ERR_NOT_READY = 7     /* asynchronous request still running */
};

/* ---- sources: src/photon_propagation/lightyield.jl:421-496 ------------- */
enum {
    NT_SRC_POINTLIKE_ISOTROPIC = 0,  /* PointlikeIsotropicEmitter  lightyield.jl:436-440 */
    NT_SRC_POINTLIKE_TIME_RANGE = 1, /* PointlikeTimeRangeEmitter  lightyield.jl:449-453 */
    NT_SRC_AXICONE = 2,              /* AxiconeEmitter             lightyield.jl:421-427 */
    NT_SRC_PENCIL = 3,               /* PencilEmitter              lightyield.jl:429-434 */
    NT_SRC_EXTENDED_CHERENKOV = 4,   /* ExtendedCherenkovEmitter   lightyield.jl:463-469 */
    NT_SRC_POINTLIKE_CHERENKOV = 5   /* PointlikeCherenkovEmitter  lightyield.jl:491-496 */
    /* CherenkovTrackEmitter (lightyield.jl:519-525) is not supported: its
       device initialiser assigns a scalar to a direction
       (photon_prop_cuda.jl:197) and cannot run in the reference either. */
};

typedef struct nt_source {           /* 64 bytes */
    int32_t kind;                    /* NT_SRC_* */
    int32_t reserved;
    int64_t photons;                 /* >= 0 */
    float position[3];               /* m */
    float direction[3];              /* unit vector (all kinds except isotropic/time-range) */
    float time;                      /* ns; time_range[0] for TIME_RANGE */
    float time_end;                  /* ns; time_range[1] for TIME_RANGE */
    float angle;                     /* rad; AXICONE opening angle */
    float long_a;                    /* LongitudinalParameterisation a  (lightyield.jl:139-143) */
    float long_b;                    /* LongitudinalParameterisation b */
    float long_lrad;                 /* radiation length in m */
} nt_source;

/* ---- medium: WaterProperties, src/photon_propagation/medium.jl:70-125 --- */
enum {
    NT_DISPERSION_REF = 0,       /* dispersion as written in medium.jl:293-298 + group
                                    velocity with λ0 = n·λ (medium.jl:313-318) */
    NT_DISPERSION_QUAN_FRY = 1   /* derivative of the Quan-Fry n(λ) (extension) */
};

typedef struct nt_medium {           /* 48 bytes */
    float quan_fry[4];               /* a01, a2, a3, a4 (medium.jl:174-205) */
    float vol_conc_small_part;       /* ppm */
    float vol_conc_large_part;       /* ppm */
    float mean_scattering_angle;     /* g of Henyey-Greenstein (medium.jl:164-165) */
    float sl_fraction;               /* Simplified-Liu mixture weight; 0 = reference */
    float sca_len_scale;             /* multiplies the scattering length; 1 = reference */
    int32_t dispersion_mode;         /* NT_DISPERSION_* */
    int32_t max_steps;               /* scattering steps per photon; 15 = reference
                                        (photon_prop_cuda.jl:449); 0..NT_MAX_STEPS */
    int32_t direction_arithmetic;    /* NT_DIR_* */
} nt_medium;

/* Arithmetic of the direction sampling and rotations: the scattering
   direction (update_direction, photon_prop_cuda.jl:204-222, with
   rot_from_ez_fast, utils.jl:228-243), the Cherenkov emission direction
   (photon_prop_cuda.jl:107-125, utils.jl:249-260) and the axicone direction.
   NT_DIR_FP64 evaluates the reference formulas in FP64 from the FP32 inputs
   and rounds once: within 2 FP32 ulp of an FP64 evaluation of the reference
   functions (the hardware sine/cosine of the scattering azimuth, <= 1.3e-7,
   is the main remaining error).  NT_DIR_FP32_REF evaluates them in FP32 with
   the reference's operation order, as the reference kernel does; there
   √(1 − cos²θ) near cos θ = ±1 and the track-direction sampler lose up to
   thousands of ulps against FP64, and it runs about a quarter faster. */
enum {
    NT_DIR_FP64 = 0,
    NT_DIR_FP32_REF = 1
};

/* ---- spectrum: src/photon_propagation/spectrum.jl:21-108 --------------- */
enum {
    NT_SPEC_MONOCHROMATIC = 0,   /* Monochromatic (spectrum.jl:21-23) */
    NT_SPEC_TABLE = 1,           /* CherenkovSpectrum: inverse-CDF knots sampled like
                                    the reference's 1-D texture (spectrum.jl:97-105) */
    NT_SPEC_INV_LAMBDA2 = 2      /* closed-form λ ∝ 1/λ² on wl_range (extension) */
};
enum {
    NT_TEX_TEXEL_CENTER_CLAMP = 0,    /* normalized coords, linear filter, clamp:
                                         x = u·N − 0.5, lerp(knots[i], knots[i+1]) */
    NT_TEX_TEXEL_CENTER_CLAMP_Q8 = 1, /* same with the filter weight rounded to 1/256 */
    NT_TEX_ENDPOINT = 2               /* lerp with knots at u = i/(N−1) */
};

typedef struct nt_spectrum {
    int32_t kind;                    /* NT_SPEC_* */
    int32_t addressing;              /* NT_TEX_* (TABLE only) */
    float wavelength;                /* nm (MONOCHROMATIC) */
    float wl_range[2];               /* nm (INV_LAMBDA2, informational for TABLE) */
    int32_t n_knots;                 /* TABLE: 2..NT_MAX_KNOTS */
    const float* knots;              /* TABLE: host pointer, wavelengths in nm */
} nt_spectrum;

#define NT_MAX_KNOTS 1024

/* ---- targets: src/photon_propagation/detection.jl:29-46 ---------------- */
enum {
    NT_TARGET_SPHERE = 0,            /* DetectionSphere  detection.jl:29-35 */
    NT_TARGET_MULTI_PMT = 1          /* MultiPMTDetector detection.jl:39-45 */
};

#define NT_MAX_PMTS 16
#define NT_MAX_TARGETS 4096
/* largest nt_medium.max_steps: one photon then adds at most
   (NT_MAX_STEPS + 1)·NT_MAX_TARGETS = 2^31 to the sphere-test counters */
#define NT_MAX_STEPS 524287

typedef struct nt_target {           /* 160 bytes */
    float position[3];               /* m */
    float radius;                    /* m */
    float pmt_area;                  /* m^2 */
    int32_t n_pmts;
    int32_t kind;                    /* NT_TARGET_* */
    uint16_t module_id;
    uint16_t reserved;
    float pmt_coordinates[2 * NT_MAX_PMTS]; /* (θ, φ) per PMT, column-major like
                                               SMatrix{2,N} */
} nt_target;

/* ---- hit records: PhotonHit, photon_prop_cuda.jl:391-399 --------------- */
typedef struct nt_hits {
    float* position;                 /* [capacity][3]  SVector{3,Float32} column */
    float* direction;                /* [capacity][3]  direction at detection */
    float* initial_direction;        /* [capacity][3] */
    float* wavelength;               /* [capacity] nm */
    float* time;                     /* [capacity] ns */
    float* dist_travelled;           /* [capacity] m */
    uint16_t* module_id;             /* [capacity] */
    uint32_t* source_index;          /* [capacity] optional (NULL: not written) */
    uint64_t* photon_id;             /* [capacity] optional (NULL: not written) */
    /* Written only while the hit epilogue is enabled (nt_set_epilogue); each
       optional.  Replace the reference's host post-processing columns. */
    uint8_t* pmt_id;                 /* [capacity] check_pmt_hit: 1-based PMT, 0 = none
                                        (detection.jl:72-104; DetectionSphere: 1) */
    float* total_weight;             /* [capacity] area_acc·wl_acc·exp(−dist/L_abs(λ))
                                        (photon_prop_cuda.jl:850-882) */
    float* tres;                     /* [capacity] time residual (photon_prop_cuda.jl:906-924) */
    int64_t capacity;
} nt_hits;

/* ---- context ----------------------------------------------------------- */
typedef struct nt_ctx nt_ctx;

int32_t nt_abi_version(void);
const char* nt_status_string(int32_t status);

/* Create a context on HIP device `device`. */
int32_t nt_create(int32_t device, nt_ctx** out_ctx);
void nt_destroy(nt_ctx* ctx);

/* Replaces the `medium` argument of run_photon_prop_no_local_cache
   (photon_prop_cuda.jl:726-731). */
int32_t nt_set_medium(nt_ctx* ctx, const nt_medium* medium);

/* Replaces the `spectrum` argument (a CherenkovSpectrum texture handle,
   spectrum.jl:97-105, or Monochromatic). */
int32_t nt_set_spectrum(nt_ctx* ctx, const nt_spectrum* spectrum);

/* Replaces `targets` (photon_prop_cuda.jl:740: CuVector(targets)).  Order is
   kept: the first listed intersecting target wins (photon_prop_cuda.jl:460-468).
   NT_ERR_ARG for a radius that is not finite and > 0 or a non-finite position
   (the context's targets are then unchanged). */
int32_t nt_set_targets(nt_ctx* ctx, const nt_target* targets, int32_t n_targets);

/* Replaces `sources` (one kernel launch per source in photon_prop_cuda.jl:749-753;
   here one launch for all). */
int32_t nt_set_sources(nt_ctx* ctx, const nt_source* sources, int32_t n_sources);

/* nt_set_sources for records already in device memory (e.g. written by
   nt_build_cascade_sources): the same validation and preparation on the
   device, in stream order, then one 16-byte read-back (total photons and an
   error flag).  kind_hint: the common NT_SRC_* of all records (selects the
   kernel variant), or -1 when kinds are mixed or unknown. */
int32_t nt_set_sources_device(nt_ctx* ctx, const nt_source* d_sources, int32_t n_sources, int32_t kind_hint,
                              void* stream);

/* Synchronous drop-in for run_photon_prop_no_local_cache
   (photon_prop_cuda.jl:726-759): propagate every photon of the sources set by
   nt_set_sources and copy the hits into caller-owned HOST columns.
   *n_hits = hits stored (<= capacity), *n_simulated = photons simulated
   (out_n_ph_simulated), *n_dropped = hits lost to capacity. */
int32_t nt_propagate(nt_ctx* ctx, uint64_t seed, const nt_hits* out_host,
                     int64_t* n_hits, int64_t* n_simulated, int64_t* n_dropped);

/* nt_propagate without any device allocation.  No counterpart in the
   reference: for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
   device hit columns and counters live in a file-scope device array of
   NT_NOALLOC_MAX_HITS hits (9.3 MB, held by every GPU the library runs on
   until the process ends) instead of buffers the context allocates; the
   caller passes host columns as for nt_propagate.  The context's medium,
   spectrum, targets and sources are set beforehand as usual (those calls do
   allocate).  Returns NT_ERR_ARG for a capacity above NT_NOALLOC_MAX_HITS.
   Every call on one GPU shares that GPU's array and it keeps the last call's
   contents: calls through one context are serialised by the context, but
   calls from two contexts on the same GPU that overlap in time race. */
#define NT_NOALLOC_MAX_HITS (1 << 17)
int32_t nt_propagate_noalloc(nt_ctx* ctx, uint64_t seed, const nt_hits* out_host,
                             int64_t* n_hits, int64_t* n_simulated, int64_t* n_dropped);

/* Asynchronous device form: hits go to caller-owned DEVICE columns, counters to
   `d_counters` (device int64[4]: total hits found, photons simulated, sphere
   tests, string-bound tests), which the call clears first.  `stream` is a hipStream_t
   (NULL = default stream).  Hits beyond capacity are counted but not stored. */
int32_t nt_launch(nt_ctx* ctx, uint64_t seed, const nt_hits* out_dev,
                  int64_t* d_counters, void* stream);

/* resample_simulation (src/pmt_frontend/pipeline.jl:86-116) on the device:
   the hits' weighted photons → one realisation of detected hits.  Per group
   of hits — (module_id, pmt_id) with per_pmt != 0, module_id otherwise — draw
   n = min(Poisson(Σ total_weight · downsample), #hits with weight > 0) and
   select n of the group's hits without replacement with probabilities
   ∝ total_weight; d_detected[i] = 1 for the selected hits, 0 otherwise.
   dev_hits are DEVICE columns (module_id, total_weight, source_index,
   photon_id, and pmt_id with per_pmt; e.g. the output of nt_launch with the
   hit epilogue), n_hits their count.  Random numbers: Philox4x32-10 with
   key = seed: each hit's 53-bit uniform u from counter (0x71000000, photon
   lo, photon hi, source index) gives it the key ln(u)/w (the n largest keys
   of a group are the sample; Efraimidis & Spirakis 2006), each group's
   Poisson draw uses counter (trial, group, 0, 0x70000000) with group =
   module_id·17 + min(pmt_id, 16) (per_pmt) or module_id.  The result depends
   on the hits, not on their order.  The reference draws with
   PoissonRandom.pois_rand and StatsBase.sample (unreproducible streams).
   Asynchronous on `stream`; *d_n_detected (device int64, optional) receives
   the number of detected hits.  Calls on one context run one at a time (they
   share the context's workspace). */
int32_t nt_resample(nt_ctx* ctx, const nt_hits* dev_hits, int64_t n_hits, uint64_t seed, double downsample,
                    int32_t per_pmt, uint8_t* d_detected, int64_t* d_n_detected, void* stream);

/* Options (no reference counterpart).
   NT_OPT_TARGET_CULLING (default 1): test a step against the bounding
   cylinder of each vertical string of targets (targets sharing x, y) before
   its spheres.  The bounds are inflated beyond the FP32 rounding of the sphere
   test, so results are identical to testing every target in list order (0). */
/* NT_OPT_CHERENKOV_ANGLE_URAD (default 0): 0 emits Cherenkov photons at
   θc = acos(1/n(λ)) as the reference (cherenkov_angle, medium.jl:309-311);
   a value v in 1..3141592 emits them at the fixed angle v·1e-6 rad (the
   "emission at 41.2°" configuration: 719076), around the same sampled track
   direction.  Kept across nt_set_medium. */
enum { NT_OPT_TARGET_CULLING = 0, NT_OPT_CHERENKOV_ANGLE_URAD = 1 };
int32_t nt_set_option(nt_ctx* ctx, int32_t option, int64_t value);

/* ---- hit epilogue ------------------------------------------------------
   Replaces make_hits_from_photons / calc_total_weight! / calc_time_residual!
   (photon_prop_cuda.jl:850-924), which the reference runs on the host over
   the returned hits.  Computed on the device by an epilogue kernel that runs
   on the launch stream right after the propagation kernel, over the hits it
   stored, into the optional pmt_id / total_weight / tres columns:
     pmt_id       first PMT whose axis (the module's (θ, φ) PMT coordinates
                  rotated by `orientation`) lies within asin(sqrt(area/π)/r)
                  of the hit direction seen from the module centre, in FP64
                  (detection.jl:88-104); DetectionSphere: 1
     total_weight area_acc × wl_acc × exp(−dist/L_abs(λ)) [× angular
                  acceptance, when a table is given]: area_acc =
                  n_pmts·pmt_area/(4πr²) for DetectionSphere, 1 for
                  MultiPMTDetector; L_abs linear in the absorption table
                  (× abs_len_scale); wl_acc linear in the acceptance table
     tres         time − (|source_position − module| − r)/(c/n(800 nm)),
                  as in the reference with its first source's position
   The tables and orientation are copied; call again after nt_set_targets
   or nt_set_medium changes. */
typedef struct nt_epilogue {
    int32_t enabled;                 /* 0 disables the epilogue */
    int32_t n_abs;                   /* absorption table rows (2..1024) */
    const float* abs_wavelength;     /* nm, ascending (medium.jl:392-404) */
    const float* abs_length;         /* m */
    float abs_len_scale;             /* 1 = reference */
    int32_t n_acc;                   /* PMT acceptance rows (2..1024) */
    const double* acc_wavelength;    /* nm, ascending (detection.jl:148-163) */
    const double* acc_value;
    double orientation[9];           /* row-major rotation of the PMT axes */
    float source_position[3];        /* time-residual origin */
    /* Angular acceptance (no reference counterpart; the reference has only
       the opening-angle cap): for a hit on a PMT of a MultiPMTDetector the
       weight is multiplied by this table, linear in the cosine of the angle of
       incidence −direction·(PMT axis), clamped at the ends.  n_ang = 0: none. */
    int32_t n_ang;                   /* 0 or 2..1024 rows */
    const double* ang_cos;           /* ascending */
    const double* ang_value;
} nt_epilogue;
int32_t nt_set_epilogue(nt_ctx* ctx, const nt_epilogue* epilogue);

/* ---- serving: continuous batching of propagation requests --------------
   No reference counterpart (the reference propagates one setup per call).
   A server runs a scheduler thread over the context's medium, spectrum and
   targets (set them before nt_server_create and leave them unchanged while
   it runs).  Requests (sources + seed + hit capacity) are admitted in
   submission order as soon as their hit buffers fit the budgets.
   NT_SERVER_WORKLIST: admitted requests are published to a work list that
   one persistent kernel instance drains; its wavefronts claim photon
   tickets (the request's photons over the number of wavefronts, 64..4096,
   sources of at most 2^21 photons before larger ones), so a request that
   arrives while others run is picked up by the next free wavefront, and
   write hits straight into the request's pinned host buffer (only the host
   budget is used).  The instance leaves after 2 ms without work and is
   launched again for the next request.  NT_SERVER_ROUNDS: every round
   packs photon chunks of all admitted requests into one launch of the
   propagation kernel (at most round_photons photons, shared between
   requests), so requests that arrive while others run join the next round;
   hits are copied to pinned memory on a second stream.  In both modes a
   request's photons use Philox key = its seed, counter source word = its
   source index and photon index = index inside the source, exactly as
   nt_propagate with that seed: a result is the same alone and in any mix.
   Latency is measured from submission until the result is in host memory. */
typedef struct nt_server nt_server;

enum nt_server_mode {
    NT_SERVER_ROUNDS = 0,            /* host-scheduled rounds, one launch each */
    NT_SERVER_WORKLIST = 1           /* persistent kernel pulling from a device work list */
};

typedef struct nt_server_config {
    uint64_t device_budget_bytes;    /* HBM for request hit buffers */
    uint64_t host_budget_bytes;      /* pinned host memory for results */
    int64_t round_photons;           /* NT_SERVER_ROUNDS: photons per batched launch */
    int32_t max_inflight_rounds;     /* NT_SERVER_ROUNDS: launches queued ahead of the device (1..16) */
    int32_t mode;                    /* nt_server_mode */
} nt_server_config;

typedef struct nt_server_stats {
    int64_t submitted, completed, rounds, photons;   /* rounds: launches (work list: kernel instances) */
    int64_t queued, active;          /* requests waiting for memory / admitted */
    double busy_ms;                  /* device time spent propagating: Σ round durations, or (work list)
                                        Σ ticket durations of all wavefronts ÷ wavefronts per instance */
} nt_server_stats;

int32_t nt_server_create(nt_ctx* ctx, const nt_server_config* cfg, nt_server** out);
/* Thread-safe.  NT_ERR_BUDGET if the request can never fit the budgets. */
int32_t nt_server_submit(nt_server* srv, const nt_source* sources, int32_t n_sources, uint64_t seed,
                         int64_t hit_capacity, uint64_t* request_id);
/* As nt_server_submit, and the served hits also get the epilogue columns
   (pmt_id, total_weight, tres; see nt_epilogue) with the time residual
   measured from ref_position[3] instead of the context's epilogue source
   position.  NT_SERVER_WORKLIST only, after nt_set_epilogue; the epilogue
   settings must stay unchanged while the server runs.  No reference
   counterpart (calc_total_weight!/calc_time_residual!, photon_prop_cuda.jl:
   850-917, applied per request). */
int32_t nt_server_submit_epilogue(nt_server* srv, const nt_source* sources, int32_t n_sources, uint64_t seed,
                                  int64_t hit_capacity, const float* ref_position, uint64_t* request_id);
/* NT_OK when the request is complete, NT_ERR_NOT_READY otherwise. */
int32_t nt_server_poll(nt_server* srv, uint64_t request_id);
/* Blocks until the request is complete, copies its hits to host columns
   (capacity >= the request's hit capacity), releases it.  Returns
   NT_ERR_HIT_OVERFLOW when hits were lost (the first `hit_capacity` kept). */
int32_t nt_server_wait(nt_server* srv, uint64_t request_id, const nt_hits* out_host, int64_t* n_hits,
                       int64_t* n_simulated, int64_t* n_dropped, double* latency_ms);
int32_t nt_server_stats_get(nt_server* srv, nt_server_stats* out);
void nt_server_destroy(nt_server* srv);

/* Grid used by nt_launch (blocks, threads per block). */
int32_t nt_launch_shape(nt_ctx* ctx, int32_t* blocks, int32_t* threads);

/* ---- source builder ------------------------------------------------------
   Particles → extended-Cherenkov sources on the device, one thread per
   particle: the reference's ExtendedCherenkovEmitter(particle, medium,
   wl_range; oversample) (lightyield.jl:471-481), i.e.
     long_a   = Float32(alpha + beta·log10 E),  long_b = Float32(b)
                (LongitudinalParameterisation, lightyield.jl:139-155, 204-213;
                 alpha, beta, b of the particle type, lightyield.jl:135-137)
     long_lrad = Float32(model.lrad)
     photons  ~ Poisson(model.photons_per_m · track_alpha · E^track_beta ·
                model.oversample)   (total_lightyield, lightyield.jl:350-366)
   The Poisson draw uses Philox4x32-10 with key = seed, counter =
   (trial, particle index lo, hi, 0x50000000): inversion for a mean below 10,
   Hörmann's transformed rejection (PTRS, 1993) above, in double precision.
   (The reference draws with PoissonRandom.pois_rand, whose streams are not
   reproducible here.)  d_particles and d_sources are device arrays. */
enum { NT_PARTICLE_EMINUS = 0, NT_PARTICLE_EPLUS = 1, NT_PARTICLE_GAMMA = 2 };

typedef struct nt_particle {         /* 40 bytes */
    float position[3];               /* m */
    float direction[3];              /* unit vector */
    float time;                      /* ns */
    float energy;                    /* GeV */
    int32_t type;                    /* NT_PARTICLE_* */
    int32_t reserved;
} nt_particle;

typedef struct nt_cascade_model {
    double photons_per_m;            /* frank_tamm_norm(wl_range, medium) */
    double lrad;                     /* radiation_length / density · 10 (m) */
    double oversample;               /* > 0 */
} nt_cascade_model;

int32_t nt_build_cascade_sources(nt_ctx* ctx, const nt_particle* d_particles, int32_t n,
                                 const nt_cascade_model* model, uint64_t seed, nt_source* d_sources,
                                 void* stream);

/* ---- particle injector ----------------------------------------------------
   VolumeInjector (injectors.jl:93-111) on the device, one thread per
   particle, so that a request generator can feed nt_build_cascade_sources
   without a host round trip.  Particle i draws two Philox4x32-10 blocks with
   key = seed, counter = (block, i lo, i hi, 0x60000000); uniforms are the
   words' top 24 bits, u = (w >> 8)·2^-24, and all arithmetic is FP64, rounded
   to FP32 at the end:
     block 0 word 0-2  position: cylinder (injectors.jl:59-70)
                         z = c_z + (u0 − 1/2)·height, r = sqrt(u1·radius)
                         (as the reference draws it), φ = 2π·u2;
                       cuboid (injectors.jl:72-78) c + (u − 1/2)·size;
                       fixed (injectors.jl:57) the centre
     block 0 word 3    energy, log-uniform: e_min·(e_max/e_min)^u3
     block 1 word 0-1  direction (UniformAngularDistribution,
                       injectors.jl:82-89): φ = 2π·u, cos θ = 2u − 1
     block 1 word 2    type: the first k with u < type_cdf[k]
   time = 0 (injectors.jl:109).  d_particles is a device array of n records. */
enum { NT_VOLUME_CYLINDER = 0, NT_VOLUME_CUBOID = 1, NT_VOLUME_FIXED = 2 };

typedef struct nt_injector {         /* 60 bytes */
    int32_t volume;                  /* NT_VOLUME_* */
    int32_t reserved;
    float center[3];                 /* m */
    float height, radius;            /* cylinder */
    float size[3];                   /* cuboid edge lengths */
    float e_min, e_max;              /* GeV; log-uniform, e_min == e_max: fixed */
    float type_cdf[3];               /* cumulative probabilities of e-, e+, gamma; type_cdf[2] = 1 */
} nt_injector;

int32_t nt_inject_particles(nt_ctx* ctx, const nt_injector* inj, int64_t n, uint64_t seed,
                            nt_particle* d_particles, void* stream);

/* ---- device-function evaluation (test hook) ----------------------------
   Runs one device function of the propagation kernel on n inputs (host
   arrays; row-major, `in_width` floats per row) and writes `out_width` floats
   per row.  Used to check the kernel's math against FP64 evaluations. */
enum {
    NT_FN_INTERSECT = 0,        /* pos3 dir3 step tpos3 r -> isec d  (photon_prop_cuda.jl:233-271) */
    NT_FN_HG_COS = 1,           /* g eta -> cos                     (photon_prop_cuda.jl:70-77) */
    NT_FN_SL_COS = 2,           /* g xi -> cos                      (Simplified-Liu, extension) */
    NT_FN_ROT_FROM_EZ = 3,      /* a3 op3 -> v3                     (utils.jl:228-243) */
    NT_FN_SPH_TO_CART = 4,      /* theta phi -> v3                  (utils.jl:105-115) */
    NT_FN_ISO_DIR = 5,          /* u_cos u_phi -> v3                (photon_prop_cuda.jl:100-104) */
    NT_FN_REFRACTIVE_INDEX = 6, /* wl -> n                          (medium.jl:234-244) */
    NT_FN_GROUP_VELOCITY = 7,   /* wl -> c_grp                      (medium.jl:313-318) */
    NT_FN_SCA_LEN = 8,          /* wl -> L_sca                      (medium.jl:331-347) */
    NT_FN_CHERENKOV_DIR = 9,    /* wl u_phi u_tcos u_tphi dir3 -> v3 (photon_prop_cuda.jl:107-125) */
    NT_FN_UPDATE_DIR = 10,      /* dir3 eta u_phi u_mix -> v3       (photon_prop_cuda.jl:204-222) */
    NT_FN_SPECTRUM = 11,        /* u -> wl                          (photon_prop_cuda.jl:128-130) */
    NT_FN_PHILOX = 12,          /* ctr4 key2 (uint32 bits) -> out4  */
    NT_FN_GAMMA = 13,           /* a b lrad seed_lo seed_hi photon (bits) -> depth trials */
    NT_FN_SCA_LEN_FIT = 14,     /* wl -> L_sca as the kernel takes it: FP64 polynomial fit of the
                                   formula over the spectrum's range, the formula outside it */
    NT_FN_TRACK_DIR = 15,       /* u_cos u_phi -> v3                (utils.jl:249-260) */
    NT_FN_COUNT = 16
};
int32_t nt_eval(nt_ctx* ctx, int32_t fn, const float* in, int32_t in_width,
                float* out, int32_t out_width, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* NTPROP_H */
