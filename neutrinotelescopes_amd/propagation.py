"""Host driver of the propagation path: mirror of the public API of module
PhotonPropagationCuda (src/photon_propagation/photon_prop_cuda.jl).

    setup = PhotonPropSetup(sources, targets, medium, spectrum)
    hits = propagate_photons(setup)                   # photon_prop_cuda.jl:804-817
    hits, n_sim = run_photon_prop_no_local_cache(...)  # photon_prop_cuda.jl:726-800
    calc_total_weight(hits, setup)                     # photon_prop_cuda.jl:850-882
    calc_time_residual(hits, setup)                    # photon_prop_cuda.jl:906-924
    make_hits_from_photons(hits, setup, orientation)   # photon_prop_cuda.jl:884-904

All propagation runs in libntprop.so (HIP, gfx950) through the C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .detection import DetectionSphere, MultiPMTDetector, area_acceptance, check_pmt_hit, p_one_pmt_acc
from .medium import WaterProperties, absorption_length, refractive_index
from .spectrum import Spectrum

C_VAC_M_NS = 0.299792458


@dataclass
class PhotonPropSetup:
    """photon_prop_cuda.jl:34-45."""
    sources: list
    targets: list
    medium: WaterProperties
    spectrum: Spectrum
    seed: int = 0

    def __post_init__(self):
        if not isinstance(self.sources, (list, tuple)):
            self.sources = [self.sources]
        if not isinstance(self.targets, (list, tuple)):
            self.targets = [self.targets]


HIT_COLUMNS = ("position", "direction", "initial_direction", "wavelength", "time",
               "dist_travelled", "module_id", "source_index", "photon_id", "pmt_id", "total_weight",
               "tres")


@dataclass
class PhotonHits:
    """Columns of PhotonHit (photon_prop_cuda.jl:391-399) as numpy arrays, plus
    the source index and in-source photon index of each hit."""
    position: np.ndarray
    direction: np.ndarray
    initial_direction: np.ndarray
    wavelength: np.ndarray
    time: np.ndarray
    dist_travelled: np.ndarray
    module_id: np.ndarray
    source_index: np.ndarray
    photon_id: np.ndarray
    pmt_id: np.ndarray = None          # filled by the device epilogue
    total_weight: np.ndarray = None
    tres: np.ndarray = None
    n_simulated: int = 0
    n_dropped: int = 0
    extra: dict = field(default_factory=dict)

    @classmethod
    def empty(cls, capacity: int, epilogue: bool = True) -> "PhotonHits":
        """Unfilled columns for `capacity` hits: the library writes the first
        n rows and callers keep only those (zero-filling hundreds of MB for a
        few thousand hits cost milliseconds per call).  Without `epilogue`
        the PMT id, weight and residual columns are left out (None)."""
        n = int(capacity)
        epi = (np.empty(n, np.uint8), np.empty(n, np.float32), np.empty(n, np.float32)) if epilogue \
            else (None, None, None)
        return cls(np.empty((n, 3), np.float32), np.empty((n, 3), np.float32),
                   np.empty((n, 3), np.float32), np.empty(n, np.float32), np.empty(n, np.float32),
                   np.empty(n, np.float32), np.empty(n, np.uint16), np.empty(n, np.uint32),
                   np.empty(n, np.uint64), *epi)

    def __len__(self):
        return len(self.wavelength)

    def _cols(self, idx):
        return (None if getattr(self, c) is None else getattr(self, c)[idx] for c in HIT_COLUMNS)

    def truncated(self, n: int) -> "PhotonHits":
        return PhotonHits(*self._cols(slice(0, n)), n_simulated=self.n_simulated, n_dropped=self.n_dropped)

    def select(self, mask) -> "PhotonHits":
        out = PhotonHits(*self._cols(mask), n_simulated=self.n_simulated, n_dropped=self.n_dropped)
        out.extra = {k: v[mask] for k, v in self.extra.items()}
        return out

    def sorted(self) -> "PhotonHits":
        """Deterministic order: by (source index, photon index)."""
        order = np.lexsort((self.photon_id, self.source_index))
        return self.select(order)

    def to_abi(self) -> _abi.nt_hits:
        h = _abi.nt_hits()
        for c in HIT_COLUMNS:
            setattr(h, c, None if getattr(self, c) is None else getattr(self, c).ctypes.data)
        h.capacity = len(self)
        return h

    def to_dataframe(self):
        import pandas as pd
        d = {}
        for c in ("position", "direction", "initial_direction"):
            a = getattr(self, c)
            for i, ax in enumerate("xyz"):
                d[f"{c}_{ax}"] = a[:, i]
        for c in ("wavelength", "time", "dist_travelled", "module_id", "source_index", "photon_id",
                  "pmt_id", "total_weight", "tres"):
            if getattr(self, c) is not None:
                d[c] = getattr(self, c)
        d.update(self.extra)
        return pd.DataFrame(d)


class DeviceHitColumns:
    """Hit columns in device memory (torch tensors) for `PropagationEngine.launch`
    (nt_launch): the PhotonHit columns, source and photon index, and with
    `epilogue` the PMT id, weight and residual; `counters` (device int64[4]:
    hits found, photons simulated, sphere tests, string-bound tests) is
    cleared by each launch.  Hits beyond `capacity` are counted, not stored."""

    def __init__(self, capacity: int, device: int = 0, epilogue: bool = True):
        import torch
        dev = torch.device(f"cuda:{device}")
        n = int(capacity)
        f = lambda *shape: torch.empty(*shape, dtype=torch.float32, device=dev)
        self.capacity = n
        self.cols = dict(position=f(n, 3), direction=f(n, 3), initial_direction=f(n, 3), wavelength=f(n),
                         time=f(n), dist_travelled=f(n),
                         module_id=torch.empty(n, dtype=torch.int16, device=dev),        # uint16 bits
                         source_index=torch.empty(n, dtype=torch.int32, device=dev),     # uint32 bits
                         photon_id=torch.empty(n, dtype=torch.int64, device=dev))        # uint64 bits
        if epilogue:
            self.cols.update(pmt_id=torch.empty(n, dtype=torch.uint8, device=dev), total_weight=f(n), tres=f(n))
        self.counters = torch.zeros(4, dtype=torch.int64, device=dev)
        self.abi = _abi.nt_hits()
        for k, v in self.cols.items():
            setattr(self.abi, k, v.data_ptr())
        self.abi.capacity = n

    def __getattr__(self, name):
        cols = self.__dict__.get("cols", {})
        if name in cols:
            return cols[name]
        raise AttributeError(name)


class PropagationEngine:
    """A libntprop context: device-resident medium, spectrum, targets and sources."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load_library()
        self.device = int(device)
        h = C.c_void_p()
        _abi.check(self.lib.nt_create(int(device), C.byref(h)), "nt_create")
        self.handle = h
        self._keep = []
        self._epilogue = False          # device epilogue enabled (set_epilogue succeeded)

    def close(self):
        if self.handle:
            self.lib.nt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_medium(self, medium: WaterProperties):
        m = medium.to_abi()
        _abi.check(self.lib.nt_set_medium(self.handle, C.byref(m)), "nt_set_medium")

    def set_spectrum(self, spectrum: Spectrum):
        s = spectrum.to_abi()
        _abi.check(self.lib.nt_set_spectrum(self.handle, C.byref(s)), "nt_set_spectrum")

    def set_targets(self, targets):
        arr = _abi.as_c_array(_abi.nt_target, [t.to_abi() for t in targets])
        _abi.check(self.lib.nt_set_targets(self.handle, arr, len(targets)), "nt_set_targets")

    def set_sources(self, sources):
        arr = _abi.as_c_array(_abi.nt_source, [s.to_abi() for s in sources])
        _abi.check(self.lib.nt_set_sources(self.handle, arr, len(sources)), "nt_set_sources")

    def setup(self, setup: PhotonPropSetup):
        self.set_medium(setup.medium)
        self.set_spectrum(setup.spectrum)
        self.set_targets(setup.targets)
        self.set_sources(setup.sources)

    def set_epilogue(self, medium: WaterProperties | None = None, orientation=None,
                     source_position=(0.0, 0.0, 0.0), enabled: bool = True, angular_acceptance=None):
        """Enable the device hit epilogue (PMT id, total weight, time residual)
        with the absorption and PMT-acceptance tables of this package."""
        e = _abi.nt_epilogue()
        e.enabled = int(enabled)
        self._epilogue = False
        if enabled:
            from .detection import pmt_acceptance_table
            from .medium import absorption_table
            ax, ay = absorption_table()
            self._abs = (np.ascontiguousarray(ax, np.float32), np.ascontiguousarray(ay, np.float32))
            cx, cy = pmt_acceptance_table()
            self._acc = (np.ascontiguousarray(cx, np.float64), np.ascontiguousarray(cy, np.float64))
            e.n_abs = len(self._abs[0])
            e.abs_wavelength = self._abs[0].ctypes.data_as(C.POINTER(C.c_float))
            e.abs_length = self._abs[1].ctypes.data_as(C.POINTER(C.c_float))
            e.abs_len_scale = float(medium.abs_len_scale) if medium is not None else 1.0
            e.n_acc = len(self._acc[0])
            e.acc_wavelength = self._acc[0].ctypes.data_as(C.POINTER(C.c_double))
            e.acc_value = self._acc[1].ctypes.data_as(C.POINTER(C.c_double))
            o = np.eye(3) if orientation is None else np.asarray(orientation, np.float64)
            for i, v in enumerate(o.reshape(-1)):
                e.orientation[i] = float(v)
            for i, v in enumerate(np.asarray(source_position, np.float32).reshape(3)):
                e.source_position[i] = float(v)
            if angular_acceptance is not None:
                cx, cy = angular_acceptance
                self._ang = (np.ascontiguousarray(cx, np.float64), np.ascontiguousarray(cy, np.float64))
                e.n_ang = len(self._ang[0])
                e.ang_cos = self._ang[0].ctypes.data_as(C.POINTER(C.c_double))
                e.ang_value = self._ang[1].ctypes.data_as(C.POINTER(C.c_double))
        _abi.check(self.lib.nt_set_epilogue(self.handle, C.byref(e)), "nt_set_epilogue")
        self._epilogue = bool(enabled)

    def set_cherenkov_angle(self, angle_rad: float | None = None):
        """Emit Cherenkov photons at a fixed angle (rad) instead of
        acos(1/n(λ)) (NT_OPT_CHERENKOV_ANGLE_URAD); None restores the
        reference behaviour."""
        v = 0 if angle_rad is None else int(round(float(angle_rad) * 1e6))
        _abi.check(self.lib.nt_set_option(self.handle, _abi.NT_OPT_CHERENKOV_ANGLE_URAD, v), "nt_set_option")

    def launch_shape(self):
        b, t = C.c_int32(), C.c_int32()
        _abi.check(self.lib.nt_launch_shape(self.handle, C.byref(b), C.byref(t)), "nt_launch_shape")
        return b.value, t.value

    def propagate(self, seed: int = 0, capacity: int = 1 << 20, raise_on_overflow: bool = True):
        """Synchronous propagation into host columns (nt_propagate)."""
        hits = PhotonHits.empty(capacity, epilogue=self._epilogue)
        h = hits.to_abi()
        nh, ns, nd = C.c_int64(), C.c_int64(), C.c_int64()
        st = self.lib.nt_propagate(self.handle, C.c_uint64(seed & (2**64 - 1)), C.byref(h), C.byref(nh),
                                   C.byref(ns), C.byref(nd))
        if st != _abi.NT_OK and not (st == _abi.NT_ERR_HIT_OVERFLOW and not raise_on_overflow):
            raise _abi.NTError(st, "nt_propagate")
        out = hits.truncated(nh.value)
        if not self._epilogue:
            # without the device epilogue these columns read as zeros
            out.pmt_id = np.zeros(nh.value, np.uint8)
            out.total_weight = np.zeros(nh.value, np.float32)
            out.tres = np.zeros(nh.value, np.float32)
        out.n_simulated = ns.value
        out.n_dropped = nd.value
        return out

    def launch(self, seed: int, out: DeviceHitColumns, stream=None):
        """Asynchronous propagation into device columns (nt_launch) on `stream`
        (a torch stream; default: the current stream of the engine's
        device).  The hit count is out.counters[0] once the stream reaches
        this point; with the epilogue enabled (set_epilogue) the launch
        fills the PMT id, weight and residual columns too."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(torch.device(f"cuda:{self.device}"))
        if self._epilogue and out.cols.get("pmt_id") is None:
            raise ValueError("the epilogue is enabled: the columns need pmt_id, total_weight and tres")
        _abi.check(self.lib.nt_launch(self.handle, C.c_uint64(seed & (2**64 - 1)), C.byref(out.abi),
                                      C.c_void_p(out.counters.data_ptr()), C.c_void_p(stream.cuda_stream)),
                   "nt_launch")

    def resample(self, hits: PhotonHits, seed: int, downsample: float = 1.0, per_pmt: bool = True) -> np.ndarray:
        """resample_simulation (pipeline.jl:86-116) of host hits on the device
        (nt_resample): the hits are uploaded, and the detected flags (bool per
        hit) come back."""
        import torch
        dev = torch.device(f"cuda:{self.device}")
        n = len(hits)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt, copy=False)).to(dev)
        cols = {"module_id": t(hits.module_id, np.uint16).view(torch.int16),
                "total_weight": t(hits.total_weight, np.float32),
                "source_index": t(hits.source_index, np.uint32).view(torch.int32),
                "photon_id": t(hits.photon_id, np.uint64).view(torch.int64)}
        if per_pmt:
            cols["pmt_id"] = t(hits.pmt_id, np.uint8)
        h = _abi.nt_hits()
        for k, v in cols.items():
            setattr(h, k, v.data_ptr())
        h.capacity = n
        flags = torch.zeros(max(n, 1), dtype=torch.uint8, device=dev)
        count = torch.zeros(1, dtype=torch.int64, device=dev)
        stream = torch.cuda.current_stream(dev)
        _abi.check(self.lib.nt_resample(self.handle, C.byref(h), n, C.c_uint64(seed & (2**64 - 1)), float(downsample),
                                        int(per_pmt), C.c_void_p(flags.data_ptr()), C.c_void_p(count.data_ptr()),
                                        C.c_void_p(stream.cuda_stream)), "nt_resample")
        out = flags[:n].cpu().numpy().astype(bool)
        assert int(count.item()) == int(out.sum())
        return out

    def eval(self, fn: int, inputs: np.ndarray, out_width: int) -> np.ndarray:
        """Run one device function of the kernel on rows of `inputs`."""
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros((x.shape[0], out_width), dtype=np.float32)
        _abi.check(self.lib.nt_eval(self.handle, int(fn), x.ctypes.data, x.shape[1], y.ctypes.data,
                                    out_width, x.shape[0]), "nt_eval")
        return y


_ENGINES: dict = {}


def get_engine(device: int = 0) -> PropagationEngine:
    if device not in _ENGINES:
        _ENGINES[device] = PropagationEngine(device)
    return _ENGINES[device]


def run_photon_prop_no_local_cache(sources, targets, medium, spectrum, *, seed: int = 0,
                                   capacity: int = 1 << 22, device: int = 0, epilogue: bool = False,
                                   orientation=None):
    """photon_prop_cuda.jl:726-800: (hits, n_photons_simulated).  One launch
    covers all sources; the hit capacity is `capacity` (the reference reserves
    0.7 of device memory on every call, photon_prop_cuda.jl:732-736)."""
    eng = get_engine(device)
    setup = PhotonPropSetup(list(sources) if isinstance(sources, (list, tuple)) else [sources],
                            list(targets) if isinstance(targets, (list, tuple)) else [targets],
                            medium, spectrum)
    eng.setup(setup)
    eng.set_epilogue(medium, orientation, setup.sources[0].position if setup.sources else (0, 0, 0),
                     enabled=epilogue)
    hits = eng.propagate(seed=seed, capacity=capacity)
    if not epilogue:
        hits.pmt_id = hits.total_weight = hits.tres = None
    return hits, hits.n_simulated


def propagate_photons(setup: PhotonPropSetup, *, capacity: int = 1 << 22, device: int = 0,
                      single_target_swap: bool = True, epilogue: bool = False,
                      orientation=None) -> PhotonHits:
    """photon_prop_cuda.jl:804-817.

    With one target the reference takes its single-target kernel, which stores
    the initial direction in the `direction` column and the final direction in
    `initial_direction` (photon_prop_cuda.jl:590-592); `single_target_swap`
    (default True) reproduces that column order."""
    hits, _ = run_photon_prop_no_local_cache(setup.sources, setup.targets, setup.medium,
                                             setup.spectrum, seed=setup.seed, capacity=capacity,
                                             device=device, epilogue=epilogue,
                                             orientation=orientation)
    if single_target_swap and len(setup.targets) == 1:
        hits.direction, hits.initial_direction = hits.initial_direction, hits.direction
    return hits


def _target_map(setup: PhotonPropSetup):
    return {int(t.module_id): t for t in setup.targets}


def calc_total_weight(hits: PhotonHits, setup: PhotonPropSetup, angular_acceptance=None,
                      orientation=None) -> PhotonHits:
    """photon_prop_cuda.jl:850-882: area_acc × wl_acc × exp(−dist/L_abs(λ)),
    added as extra columns area_acc, abs_weight, wl_acc, ref_ix, total_weight.
    With `angular_acceptance` = (cos of incidence, factor) the weights of hits
    on a PMT of a MultiPMTDetector (under `orientation`) are multiplied by the
    table at −direction·(PMT axis) (no reference counterpart)."""
    tmap = _target_map(setup)
    area = np.array([area_acceptance(None, tmap[int(m)]) for m in hits.module_id], dtype=np.float64)
    abs_len = absorption_length(hits.wavelength, setup.medium)
    abs_weight = np.exp(-hits.dist_travelled / abs_len).astype(np.float64)
    wl_acc = p_one_pmt_acc(hits.wavelength)
    hits.extra["area_acc"] = area
    hits.extra["abs_weight"] = abs_weight
    hits.extra["wl_acc"] = wl_acc
    hits.extra["ref_ix"] = refractive_index(hits.wavelength, setup.medium)
    ang = np.ones(len(hits))
    if angular_acceptance is not None:
        from .detection import MultiPMTDetector, get_pmt_positions
        cx, cy = (np.asarray(a, np.float64) for a in angular_acceptance)
        o = np.eye(3) if orientation is None else np.asarray(orientation, np.float64)
        for mid in np.unique(hits.module_id):
            t = tmap[int(mid)]
            if not isinstance(t, MultiPMTDetector):
                continue
            sel = np.flatnonzero(hits.module_id == mid)
            pmt = check_pmt_hit(hits.position[sel], t, orientation)
            axes = get_pmt_positions(t, o)
            on = pmt > 0
            ci = -np.sum(hits.direction[sel][on].astype(np.float64) * axes[pmt[on] - 1], axis=1)
            ang[sel[on]] = np.interp(ci, cx, cy)
    hits.extra["angular_acceptance"] = ang
    hits.extra["total_weight"] = area * wl_acc * abs_weight * ang
    return hits


def calc_time_residual(hits: PhotonHits, setup: PhotonPropSetup) -> PhotonHits:
    """photon_prop_cuda.jl:906-924: t − (|src₁ − target| − r)/(c/n(800 nm)),
    with the first source's position for every hit, as in the reference."""
    tmap = _target_map(setup)
    src = np.asarray(setup.sources[0].position, dtype=np.float32)
    n800 = float(refractive_index(np.float32(800.0), setup.medium))
    tres = np.empty(len(hits), dtype=np.float64)
    for mid in np.unique(hits.module_id):
        t = tmap[int(mid)]
        sel = hits.module_id == mid
        d = src - np.asarray(t.position, dtype=np.float32)
        dist = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])          # float32 norm
        tgeo = float(dist - np.float32(t.radius)) / (C_VAC_M_NS / n800)    # float64
        tres[sel] = hits.time[sel].astype(np.float64) - tgeo
    hits.extra["tres"] = tres
    return hits


def make_hits_from_photons(hits: PhotonHits, setup: PhotonPropSetup, orientation=None) -> PhotonHits:
    """photon_prop_cuda.jl:884-904: keep photons that land on a PMT and add pmt_id."""
    tmap = _target_map(setup)
    pmt = np.zeros(len(hits), dtype=np.int64)
    for mid in np.unique(hits.module_id):
        sel = hits.module_id == mid
        pmt[sel] = check_pmt_hit(hits.position[sel], tmap[int(mid)], orientation)
    hits.extra["pmt_id"] = pmt
    return hits.select(pmt > 0)


def resample_simulation(times, module_id, total_weight, pmt_id=None, downsample: float = 1.0,
                        rng: np.random.Generator | None = None):
    """Weighted photons → one realisation of detected hits, pipeline.jl:86-116:
    per (module, PMT) group (per module when `pmt_id` is None) draw
    n = min(Poisson(Σw·downsample), #photons) and sample n of the group's
    times without replacement with probabilities ∝ weight (photons with
    weight 0 never).  Returns (times, module_id, pmt_id or None)."""
    rng = rng or np.random.default_rng()
    times = np.asarray(times)
    module_id = np.asarray(module_id)
    w = np.asarray(total_weight, dtype=np.float64)
    pmt = None if pmt_id is None else np.asarray(pmt_id)
    key = module_id.astype(np.int64) * 65536 + (0 if pmt is None else pmt.astype(np.int64))
    order = np.argsort(key, kind="stable")
    ks = key[order]
    starts = np.r_[0, np.flatnonzero(np.diff(ks)) + 1] if ks.size else np.zeros(0, np.int64)
    ends = np.r_[starts[1:], ks.size] if ks.size else np.zeros(0, np.int64)
    out_t, out_m, out_p = [], [], []
    for a, b in zip(starts, ends):
        idx = order[a:b]
        idx = idx[w[idx] > 0]
        if idx.size == 0:
            continue
        ws = w[idx]
        n = min(int(rng.poisson(ws.sum() * downsample)), idx.size)
        if n == 0:
            continue
        pick = idx[rng.choice(idx.size, size=n, replace=False, p=ws / ws.sum())]
        out_t.append(times[pick])
        out_m.append(module_id[pick])
        if pmt is not None:
            out_p.append(pmt[pick])
    cat = lambda parts, like: np.concatenate(parts) if parts else np.zeros(0, like.dtype)
    return cat(out_t, times), cat(out_m, module_id), (cat(out_p, pmt) if pmt is not None else None)
