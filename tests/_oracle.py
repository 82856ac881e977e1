"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker; the product never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from neutrinotelescopes_amd import _abi
from neutrinotelescopes_amd.propagation import PhotonHits

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_SO.exists() or os.access(ROOT / "oracle", os.W_OK):
        # rebuild when stale (header or source newer); a no-op otherwise
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    _lib = _bind(C.CDLL(str(ORACLE_SO)))
    return _lib


def load_native(out_dir):
    """The oracle built for this host (-march=native) into `out_dir`: the CPU
    baseline of the benchmark.  Returns the bound library."""
    so = Path(out_dir) / "liboracle_native.so"
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native", f"NATIVE_OUT={so}"], check=True,
                   capture_output=True)
    return _bind(C.CDLL(str(so)))


def _bind(lib):
    f, d, i, u32, u64, p = C.c_float, C.c_double, C.c_int, C.c_uint32, C.c_uint64, C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    sig = {
        "ntor_philox4x32_10": (None, [p, p, p]),
        "ntor_dippr105": (d, [d]),
        "ntor_quan_fry_params": (None, [d, d, d, dp]),
        "ntor_make_cascadia_medium": (None, [f, C.POINTER(_abi.nt_medium), dp, dp]),
        "ntor_integrate_gauss_quad_pow": (d, [d, d, i]),
        "ntor_integrate_gauss_quad_cos": (d, [d, d]),
        "ntor_frank_tamm": (d, [d, d]),
        "ntor_frank_tamm_norm": (d, [d, d, fp]),
        "ntor_cherenkov_spectrum_knots": (i, [f, f, i, fp, fp]),
        "ntor_refractive_index": (f, [f, fp]),
        "ntor_dispersion": (f, [f, fp, i]),
        "ntor_group_velocity": (f, [f, C.POINTER(_abi.nt_medium)]),
        "ntor_scattering_length": (f, [f, C.POINTER(_abi.nt_medium)]),
        "ntor_cherenkov_angle": (f, [f, fp]),
        "ntor_hg_cos": (f, [f, f]),
        "ntor_sl_cos": (f, [f, f]),
        "ntor_sph_to_cart": (None, [f, f, fp]),
        "ntor_rot_from_ez": (None, [fp, fp, fp]),
        "ntor_rot_to_ez": (None, [fp, fp, fp]),
        "ntor_iso_dir": (None, [f, f, fp]),
        "ntor_cherenkov_dir": (None, [f, f, f, f, fp, fp, fp]),
        "ntor_cherenkov_track_dir": (None, [f, f, fp]),
        "ntor_update_dir": (None, [fp, f, f, f, C.POINTER(_abi.nt_medium), fp]),
        "ntor_intersect": (i, [fp, fp, f, fp, f, fp]),
        "ntor_spectrum_sample": (f, [C.POINTER(_abi.nt_spectrum), f]),
        "ntor_rand_gamma": (f, [f, f, u64, u32, u64, C.POINTER(C.c_int)]),
        "ntor_hg_cos_f64": (d, [d, d]),
        "ntor_sl_cos_f64": (d, [d, d]),
        "ntor_rot_from_ez_f64": (None, [dp, dp, dp]),
        "ntor_cherenkov_track_dir_f64": (None, [d, d, dp]),
        "ntor_set_cherenkov_angle": (None, [d]),
        "ntor_cherenkov_dir_f64": (None, [f, f, f, f, fp, fp, dp]),
        "ntor_update_dir_f64": (None, [fp, f, f, f, C.POINTER(_abi.nt_medium), dp]),
        "ntor_intersect_f64": (i, [dp, dp, d, dp, d, dp]),
        "ntor_group_velocity_f64": (d, [d, dp, i]),
        "ntor_refractive_index_f64": (d, [d, dp]),
        "ntor_scattering_length_f64": (d, [d, d, d]),
        "ntor_propagate": (i, [C.POINTER(_abi.nt_source), i, C.POINTER(_abi.nt_target), i,
                               C.POINTER(_abi.nt_medium), C.POINTER(_abi.nt_spectrum), u64, i,
                               C.POINTER(_abi.nt_hits), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64)]),
        "ntor_propagate_range": (i, [C.POINTER(_abi.nt_source), u32, u64, C.c_int64, C.POINTER(_abi.nt_target), i,
                                     C.POINTER(_abi.nt_medium), C.POINTER(_abi.nt_spectrum), u64, i,
                                     C.POINTER(_abi.nt_hits), C.POINTER(C.c_int64)]),
        "ntor_fast_linear_interp": (d, [d, dp, dp, i]),
        "ntor_check_pmt_hit": (i, [dp, dp, i, d]),
        "ntor_pom_pmt_coordinates": (None, [dp]),
        # PMT front end (oracle/ntpmt_oracle.c)
        "ntor_pmt_normals": (None, [C.POINTER(C.c_uint32), dp]),
        "ntor_pmt_hit_draws": (None, [C.POINTER(_abi.ntpmt_config), C.c_int64, fp, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint64), u64, dp, dp, C.POINTER(C.c_int32)]),
        "ntor_pmt_pulse": (d, [C.POINTER(_abi.ntpmt_config), d]),
        "ntor_pmt_filt_pulse": (d, [C.POINTER(_abi.ntpmt_config), d]),
        "ntor_pmt_iir": (None, [C.POINTER(_abi.ntpmt_config), C.c_int32, dp, dp]),
        "ntor_pmt_decimate": (None, [C.POINTER(_abi.ntpmt_config), C.c_int32, dp, C.c_int32, C.c_int32, dp]),
        "ntor_fnnls": (C.c_int32, [C.c_int32, C.c_int32, dp, dp, d, C.c_int64, dp, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int32)]),
        "ntor_pmt_sizes": (C.c_int32, [C.POINTER(_abi.ntpmt_config)] + [C.POINTER(C.c_int32)] * 4),
        "ntor_pmt_channel": (C.c_int32, [C.POINTER(_abi.ntpmt_config), C.c_int32, dp, dp, u32, u64, dp, dp, dp, dp,
                                         dp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def fptr(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(C.POINTER(C.c_float))


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def dptr(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.ntor_philox4x32_10(c, k, o)
    return list(o)


def propagate_range(setup, first, count, seed=0, capacity=1 << 20, threads=None) -> PhotonHits:
    """Oracle propagation of photons [first, first + count) of the setup's
    single source; hits ordered by photon."""
    lib = load()
    assert len(setup.sources) == 1
    src = setup.sources[0].to_abi()
    tgts = _abi.as_c_array(_abi.nt_target, [t.to_abi() for t in setup.targets])
    med = setup.medium.to_abi()
    spec = setup.spectrum.to_abi()
    hits = PhotonHits.empty(capacity)
    h = hits.to_abi()
    nh = C.c_int64()
    if threads is None:
        threads = min(len(os.sched_getaffinity(0)), 16)
    st = lib.ntor_propagate_range(C.byref(src), 0, C.c_uint64(first), int(count), tgts, len(setup.targets),
                                  C.byref(med), C.byref(spec), C.c_uint64(seed), int(threads), C.byref(h),
                                  C.byref(nh))
    if st not in (0, _abi.NT_ERR_HIT_OVERFLOW):
        raise RuntimeError(f"oracle propagate_range failed: {st}")
    out = hits.truncated(nh.value)
    out.n_simulated = count
    return out


def propagate(setup, seed=0, capacity=1 << 20, threads=None, lib=None) -> PhotonHits:
    """Oracle propagation of a PhotonPropSetup; hits ordered by (source, photon)."""
    lib = lib or load()
    srcs = _abi.as_c_array(_abi.nt_source, [s.to_abi() for s in setup.sources])
    tgts = _abi.as_c_array(_abi.nt_target, [t.to_abi() for t in setup.targets])
    med = setup.medium.to_abi()
    spec = setup.spectrum.to_abi()
    hits = PhotonHits.empty(capacity)
    h = hits.to_abi()
    nh, ns = C.c_int64(), C.c_int64()
    if threads is None:
        threads = min(len(os.sched_getaffinity(0)), 16)
    st = lib.ntor_propagate(srcs, len(setup.sources), tgts, len(setup.targets), C.byref(med),
                            C.byref(spec), C.c_uint64(seed), int(threads), C.byref(h), C.byref(nh),
                            C.byref(ns))
    if st not in (0, _abi.NT_ERR_HIT_OVERFLOW):
        raise RuntimeError(f"oracle propagate failed: {st}")
    out = hits.truncated(nh.value)
    out.n_simulated = ns.value
    return out


# ---------------------------------------------------------------- PMT front end
def pmt_hit_draws(cfg_abi, time, source_index, photon_id, seed):
    """Per hit: (t', charge, status bits) of the transit-time and SPE draws."""
    lib = load()
    n = len(time)
    t = np.ascontiguousarray(time, dtype=np.float32)
    src = np.ascontiguousarray(source_index, dtype=np.uint32)
    pid = np.ascontiguousarray(photon_id, dtype=np.uint64)
    tp, q = np.zeros(n), np.zeros(n)
    fl = np.zeros(n, np.int32)
    lib.ntor_pmt_hit_draws(C.byref(cfg_abi), n, t.ctypes.data_as(C.POINTER(C.c_float)),
                           src.ctypes.data_as(C.POINTER(C.c_uint32)), pid.ctypes.data_as(C.POINTER(C.c_uint64)),
                           C.c_uint64(seed), tp.ctypes.data_as(C.POINTER(C.c_double)),
                           q.ctypes.data_as(C.POINTER(C.c_double)), fl.ctypes.data_as(C.POINTER(C.c_int32)))
    return tp, q, fl


def pmt_sizes(cfg_abi):
    lib = load()
    v = [C.c_int32() for _ in range(4)]
    lib.ntor_pmt_sizes(C.byref(cfg_abi), *[C.byref(x) for x in v])
    return tuple(x.value for x in v)          # n_wave, n_adc, n_pulse, R


def pmt_channel(cfg_abi, t, q, channel, seed, unfold=True):
    """One channel's chain: dict with t0, wave, filt, digi, x (charges on the
    unfolding grid), status, iter, outer."""
    lib = load()
    n_wave, n_adc, n_pulse, _ = pmt_sizes(cfg_abi)
    t = np.ascontiguousarray(t, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    t0 = C.c_double()
    wave, filt, digi, x = np.zeros(n_wave), np.zeros(n_wave), np.zeros(n_adc), np.zeros(n_pulse)
    it, outer = C.c_int64(), C.c_int32()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = lib.ntor_pmt_channel(C.byref(cfg_abi), len(t), P(t), P(q), int(channel), C.c_uint64(seed), C.byref(t0),
                              P(wave), P(filt), P(digi), P(x) if unfold else None, C.byref(it), C.byref(outer))
    return {"t0": t0.value, "wave": wave, "filt": filt, "digi": digi, "x": x if unfold else None, "status": st,
            "iter": it.value, "outer": outer.value}


def fnnls(A, b, tol=1e-8, max_iter=None):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.zeros(n)
    it, outer = C.c_int64(), C.c_int32()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = lib.ntor_fnnls(m, n, P(A), P(b), tol, 30 * n if max_iter is None else max_iter, P(x), C.byref(it),
                        C.byref(outer))
    return x, st
