"""ctypes binding of libntprop.so (include/ntprop.h).

The library is the product: every propagation call goes through it.  Loading
fails loudly when the shared library has not been built; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("NTPROP_LIBRARY", PKG_DIR / "libntprop.so"))

NT_OK = 0
NT_ERR_ARG = 1
NT_ERR_OOM = 2
NT_ERR_HIT_OVERFLOW = 3
NT_ERR_DEVICE = 4
NT_ERR_BUDGET = 5
NT_ERR_STATE = 6
NT_ERR_NOT_READY = 7

NT_SRC_POINTLIKE_ISOTROPIC = 0
NT_SRC_POINTLIKE_TIME_RANGE = 1
NT_SRC_AXICONE = 2
NT_SRC_PENCIL = 3
NT_SRC_EXTENDED_CHERENKOV = 4
NT_SRC_POINTLIKE_CHERENKOV = 5

NT_DISPERSION_REF = 0
NT_DISPERSION_QUAN_FRY = 1
NT_DIR_FP64 = 0
NT_DIR_FP32_REF = 1

NT_SPEC_MONOCHROMATIC = 0
NT_SPEC_TABLE = 1
NT_SPEC_INV_LAMBDA2 = 2
NT_TEX_TEXEL_CENTER_CLAMP = 0
NT_TEX_TEXEL_CENTER_CLAMP_Q8 = 1
NT_TEX_ENDPOINT = 2

NT_OPT_TARGET_CULLING = 0
NT_OPT_CHERENKOV_ANGLE_URAD = 1

NT_TARGET_SPHERE = 0
NT_TARGET_MULTI_PMT = 1
NT_MAX_PMTS = 16
NT_MAX_TARGETS = 4096
NT_MAX_STEPS = 524287
NT_MAX_KNOTS = 1024
NT_NOALLOC_MAX_HITS = 1 << 17

(NT_FN_INTERSECT, NT_FN_HG_COS, NT_FN_SL_COS, NT_FN_ROT_FROM_EZ, NT_FN_SPH_TO_CART,
 NT_FN_ISO_DIR, NT_FN_REFRACTIVE_INDEX, NT_FN_GROUP_VELOCITY, NT_FN_SCA_LEN,
 NT_FN_CHERENKOV_DIR, NT_FN_UPDATE_DIR, NT_FN_SPECTRUM, NT_FN_PHILOX, NT_FN_GAMMA,
 NT_FN_SCA_LEN_FIT, NT_FN_TRACK_DIR) = range(16)


class nt_source(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("photons", C.c_int64),
        ("position", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("time", C.c_float),
        ("time_end", C.c_float),
        ("angle", C.c_float),
        ("long_a", C.c_float),
        ("long_b", C.c_float),
        ("long_lrad", C.c_float),
    ]


class nt_medium(C.Structure):
    _fields_ = [
        ("quan_fry", C.c_float * 4),
        ("vol_conc_small_part", C.c_float),
        ("vol_conc_large_part", C.c_float),
        ("mean_scattering_angle", C.c_float),
        ("sl_fraction", C.c_float),
        ("sca_len_scale", C.c_float),
        ("dispersion_mode", C.c_int32),
        ("max_steps", C.c_int32),
        ("direction_arithmetic", C.c_int32),
    ]


class nt_spectrum(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("addressing", C.c_int32),
        ("wavelength", C.c_float),
        ("wl_range", C.c_float * 2),
        ("n_knots", C.c_int32),
        ("knots", C.POINTER(C.c_float)),
    ]


class nt_target(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("radius", C.c_float),
        ("pmt_area", C.c_float),
        ("n_pmts", C.c_int32),
        ("kind", C.c_int32),
        ("module_id", C.c_uint16),
        ("reserved", C.c_uint16),
        ("pmt_coordinates", C.c_float * (2 * NT_MAX_PMTS)),
    ]


class nt_hits(C.Structure):
    _fields_ = [
        ("position", C.c_void_p),
        ("direction", C.c_void_p),
        ("initial_direction", C.c_void_p),
        ("wavelength", C.c_void_p),
        ("time", C.c_void_p),
        ("dist_travelled", C.c_void_p),
        ("module_id", C.c_void_p),
        ("source_index", C.c_void_p),
        ("photon_id", C.c_void_p),
        ("pmt_id", C.c_void_p),
        ("total_weight", C.c_void_p),
        ("tres", C.c_void_p),
        ("capacity", C.c_int64),
    ]


class nt_epilogue(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("n_abs", C.c_int32),
        ("abs_wavelength", C.POINTER(C.c_float)),
        ("abs_length", C.POINTER(C.c_float)),
        ("abs_len_scale", C.c_float),
        ("n_acc", C.c_int32),
        ("acc_wavelength", C.POINTER(C.c_double)),
        ("acc_value", C.POINTER(C.c_double)),
        ("orientation", C.c_double * 9),
        ("source_position", C.c_float * 3),
        ("n_ang", C.c_int32),
        ("ang_cos", C.POINTER(C.c_double)),
        ("ang_value", C.POINTER(C.c_double)),
    ]


assert C.sizeof(nt_source) == 64
assert C.sizeof(nt_medium) == 48
assert C.sizeof(nt_target) == 160

class nt_particle(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("time", C.c_float),
                ("energy", C.c_float), ("type", C.c_int32), ("reserved", C.c_int32)]


class nt_cascade_model(C.Structure):
    _fields_ = [("photons_per_m", C.c_double), ("lrad", C.c_double), ("oversample", C.c_double)]


NT_PARTICLE_EMINUS, NT_PARTICLE_EPLUS, NT_PARTICLE_GAMMA = 0, 1, 2
NT_VOLUME_CYLINDER, NT_VOLUME_CUBOID, NT_VOLUME_FIXED = 0, 1, 2


class nt_injector(C.Structure):
    _fields_ = [("volume", C.c_int32), ("reserved", C.c_int32), ("center", C.c_float * 3), ("height", C.c_float),
                ("radius", C.c_float), ("size", C.c_float * 3), ("e_min", C.c_float), ("e_max", C.c_float),
                ("type_cdf", C.c_float * 3)]


class nt_server_config(C.Structure):
    _fields_ = [
        ("device_budget_bytes", C.c_uint64),
        ("host_budget_bytes", C.c_uint64),
        ("round_photons", C.c_int64),
        ("max_inflight_rounds", C.c_int32),
        ("mode", C.c_int32),
    ]


NT_SERVER_ROUNDS = 0
NT_SERVER_WORKLIST = 1


class nt_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("rounds", C.c_int64),
        ("photons", C.c_int64), ("queued", C.c_int64), ("active", C.c_int64),
        ("busy_ms", C.c_double),
    ]


class ntpmt_config(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "tt_shape", "tt_scale", "tt_subtract", "spe_expon_weight", "spe_expon_scale", "spe_norm_mu",
        "spe_norm_sigma", "spe_trunc_low", "pulse_mu", "pulse_beta", "pulse_lo", "pulse_hi", "pulse_amplitude",
        "noise_amp", "sampling_freq", "adc_freq", "unf_pulse_res", "time_range", "lead_time", "min_charge",
        "iir_b0", "iir_b1", "iir_a1", "nnls_tol")] + [
        ("nnls_max_iter", C.c_int64),
        ("decim_taps", C.POINTER(C.c_double)), ("n_decim_taps", C.c_int32),
        ("filt_pulse", C.POINTER(C.c_double)), ("n_filt_pulse", C.c_int32),
        ("filt_pulse_t0", C.c_double), ("filt_pulse_dt", C.c_double), ("lds_factor_rows", C.c_int32)]


class ntpmt_pulses(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("module_id", "pmt_id", "n_hits", "n_pulses", "status", "t0", "time", "charge", "waveform")]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "nt_abi_version": (C.c_int32, []),
    "nt_status_string": (C.c_char_p, [C.c_int32]),
    "nt_create": (C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    "nt_destroy": (None, [C.c_void_p]),
    "nt_set_medium": (C.c_int32, [C.c_void_p, C.POINTER(nt_medium)]),
    "nt_set_spectrum": (C.c_int32, [C.c_void_p, C.POINTER(nt_spectrum)]),
    "nt_set_targets": (C.c_int32, [C.c_void_p, C.POINTER(nt_target), C.c_int32]),
    "nt_set_sources": (C.c_int32, [C.c_void_p, C.POINTER(nt_source), C.c_int32]),
    "nt_set_sources_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "nt_propagate": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(nt_hits), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nt_propagate_noalloc": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(nt_hits), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nt_launch": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(nt_hits), C.c_void_p, C.c_void_p]),
    "nt_resample": (C.c_int32, [C.c_void_p, C.POINTER(nt_hits), C.c_int64, C.c_uint64, C.c_double, C.c_int32,
                                C.c_void_p, C.c_void_p, C.c_void_p]),
    "nt_set_epilogue": (C.c_int32, [C.c_void_p, C.POINTER(nt_epilogue)]),
    "nt_set_option": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int64]),
    "nt_launch_shape": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "nt_server_create": (C.c_int32, [C.c_void_p, C.POINTER(nt_server_config), C.POINTER(C.c_void_p)]),
    "nt_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(nt_source), C.c_int32, C.c_uint64, C.c_int64,
                                     C.POINTER(C.c_uint64)]),
    "nt_server_submit_epilogue": (C.c_int32, [C.c_void_p, C.POINTER(nt_source), C.c_int32, C.c_uint64, C.c_int64,
                                              C.POINTER(C.c_float), C.POINTER(C.c_uint64)]),
    "nt_server_poll": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "nt_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(nt_hits), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "nt_server_stats_get": (C.c_int32, [C.c_void_p, C.POINTER(nt_server_stats)]),
    "nt_server_destroy": (None, [C.c_void_p]),
    "nt_eval": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                            C.c_int64]),
    "nt_inject_particles": (C.c_int32, [C.c_void_p, C.POINTER(nt_injector), C.c_int64, C.c_uint64, C.c_void_p,
                                        C.c_void_p]),
    "nt_build_cascade_sources": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(nt_cascade_model),
                                             C.c_uint64, C.c_void_p, C.c_void_p]),
    # include/ntpmt.h (PMT front end)
    "ntpmt_create": (C.c_int32, [C.c_int32, C.POINTER(ntpmt_config), C.POINTER(C.c_void_p)]),
    "ntpmt_destroy": (None, [C.c_void_p]),
    "ntpmt_sizes": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ntpmt_make_reco_pulses": (C.c_int32, [C.c_void_p, C.POINTER(nt_hits), C.c_void_p, C.c_int64, C.c_int32,
                                           C.c_uint64, C.c_int64, C.POINTER(ntpmt_pulses), C.c_void_p, C.c_void_p]),
    "ntpmt_synchronize": (C.c_int32, [C.c_void_p]),
    "ntpmt_last_error": (C.c_char_p, []),
    # include/ntflow.h (flow surrogate)
    "ntf_create": (C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    "ntf_destroy": (None, [C.c_void_p]),
    "ntf_set_weights": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "ntf_get_weights": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "ntf_forward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ntf_loss_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int64, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ntf_train_step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.c_float, C.c_void_p, C.c_void_p]),
    "ntf_train_step_gather": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_int64, C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p]),
    "ntf_pair_rows": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "ntf_time_logpdf": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                    C.c_void_p, C.c_void_p]),
    "ntf_segment_sum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_int64, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
}

# include/ntflow.h constants
NTF_N_IN = 16
NTF_HIDDEN = 128
NTF_BINS = 8
NTF_LAYERS = 3
NTF_N_PARAMS = NTF_LAYERS * (3 * NTF_BINS + 1) + 3
NTF_P_PAD = 80
NTF_N_WEIGHTS = (NTF_N_IN * NTF_HIDDEN + NTF_HIDDEN + NTF_HIDDEN * NTF_HIDDEN + NTF_HIDDEN
                 + NTF_HIDDEN * NTF_P_PAD + NTF_P_PAD)
NTF_RANGE_MIN, NTF_RANGE_MAX = -20.0, 100.0
NTF_ROW_TIME, NTF_ROW_COUNT = 1, 2
NTF_MAX_SCALE_KNOTS = 64

_lib = None


def load_library(path: Path | None = None) -> C.CDLL:
    """Load libntprop.so; raise if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is not built: run neutrinotelescopes_amd._build.build_library() "
            "(or __graft_entry__.build()) first")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # SONAME).  Loading torch first makes the dynamic loader bind
    # libntprop.so to that runtime, so torch tensors, streams and events and
    # this library share one device context.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.nt_abi_version() != 1:
        raise RuntimeError("libntprop ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


class NTError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = load_library()
        if where.startswith("ntpmt_") and status != NT_OK:
            msg = lib.ntpmt_last_error().decode(errors="replace")
        else:
            msg = lib.nt_status_string(status).decode(errors="replace")
        super().__init__(f"{where}: status {status} ({msg})")
        self.status = status


def check(status: int, where: str) -> None:
    if status != NT_OK:
        raise NTError(status, where)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def as_c_array(ctype, items):
    arr = (ctype * max(len(items), 1))()
    for i, it in enumerate(items):
        arr[i] = it
    return arr
