"""Serving form: continuous batching of propagation requests (nt_server_* in
libntprop.so).

    eng = PropagationEngine(0); eng.set_medium(...); eng.set_spectrum(...); eng.set_targets(...)
    srv = PhotonServer(eng, device_budget=64 << 30)
    rid = srv.submit([source, ...], seed=7, capacity=100_000)
    hits, latency_ms = srv.wait(rid)

Every request is propagated with Philox key = its seed, exactly as
`PropagationEngine.propagate(seed)` would propagate its sources alone: the
result does not depend on what else is being served.  There is no reference
counterpart (the reference propagates one setup per call).
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _abi
from .propagation import PhotonHits, PropagationEngine


class PhotonServer:
    """mode "worklist": a persistent kernel pulls photon tickets of the admitted
    requests from a device work list (requests join at once); "rounds": the
    scheduler thread packs the requests into one launch per round of at most
    `round_photons` photons (`max_inflight_rounds` queued ahead)."""

    def __init__(self, engine: PropagationEngine, device_budget: int = 8 << 30,
                 host_budget: int = 4 << 30, round_photons: int = 20_000_000,
                 max_inflight_rounds: int = 2, mode: str = "worklist"):
        if mode not in ("worklist", "rounds"):
            raise ValueError("mode must be 'worklist' or 'rounds'")
        self.engine = engine
        self.lib = engine.lib
        self.mode = mode
        cfg = _abi.nt_server_config(int(device_budget), int(host_budget), int(round_photons),
                                    int(max_inflight_rounds),
                                    _abi.NT_SERVER_WORKLIST if mode == "worklist" else _abi.NT_SERVER_ROUNDS)
        h = C.c_void_p()
        _abi.check(self.lib.nt_server_create(engine.handle, C.byref(cfg), C.byref(h)), "nt_server_create")
        self.handle = h
        self._capacity = {}

    def close(self):
        if self.handle:
            self.lib.nt_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, sources, seed: int, capacity: int, epilogue_position=None) -> int:
        """`epilogue_position` (work-list mode, after `engine.set_epilogue`): the
        hits also get pmt_id, total_weight and tres, the time residual measured
        from this position."""
        arr = _abi.as_c_array(_abi.nt_source, [s.to_abi() for s in sources])
        rid = C.c_uint64()
        if epilogue_position is None:
            _abi.check(self.lib.nt_server_submit(self.handle, arr, len(sources), C.c_uint64(seed & (2**64 - 1)),
                                                 int(capacity), C.byref(rid)), "nt_server_submit")
        else:
            ref = (C.c_float * 3)(*[float(v) for v in np.asarray(epilogue_position, np.float32)])
            _abi.check(self.lib.nt_server_submit_epilogue(self.handle, arr, len(sources),
                                                          C.c_uint64(seed & (2**64 - 1)), int(capacity), ref,
                                                          C.byref(rid)), "nt_server_submit_epilogue")
        self._capacity[rid.value] = (int(capacity), epilogue_position is not None)
        return rid.value

    def poll(self, rid: int) -> bool:
        st = self.lib.nt_server_poll(self.handle, C.c_uint64(rid))
        if st == _abi.NT_ERR_NOT_READY:
            return False
        _abi.check(st, "nt_server_poll")
        return True

    def wait(self, rid: int, raise_on_overflow: bool = True):
        cap, epi = self._capacity.pop(rid)
        hits = PhotonHits.empty(max(cap, 1), epilogue=epi)
        h = hits.to_abi()
        nh, ns, nd, lat = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        st = self.lib.nt_server_wait(self.handle, C.c_uint64(rid), C.byref(h), C.byref(nh), C.byref(ns),
                                     C.byref(nd), C.byref(lat))
        if st != _abi.NT_OK and not (st == _abi.NT_ERR_HIT_OVERFLOW and not raise_on_overflow):
            raise _abi.NTError(st, "nt_server_wait")
        out = hits.truncated(nh.value)
        out.n_simulated, out.n_dropped = ns.value, nd.value
        return out, lat.value

    def stats(self) -> dict:
        s = _abi.nt_server_stats()
        _abi.check(self.lib.nt_server_stats_get(self.handle, C.byref(s)), "nt_server_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}


# ---------------------------------------------------------------------------
# Request stream of the serving configuration
# ---------------------------------------------------------------------------
def cluster_targets(n_strings: int = 7, per_string: int = 20, spacing_h: float = 50.0,
                    spacing_v: float = 50.0, radius: float = 0.21, pmt_area: float | None = None):
    """A centre string plus a hexagon of strings at `spacing_h`, `per_string`
    16-PMT modules each at `spacing_v` vertical spacing, centred on z = 0."""
    from .detection import MultiPMTDetector, make_pom_pmt_coordinates
    pmt_area = pmt_area if pmt_area is not None else (75e-3 / 2) ** 2 * math.pi
    coords = make_pom_pmt_coordinates(np.float32)
    xy = [(0.0, 0.0)] + [(spacing_h * math.cos(k * math.pi / 3), spacing_h * math.sin(k * math.pi / 3))
                         for k in range(n_strings - 1)]
    z0 = -spacing_v * (per_string - 1) / 2
    out = []
    for x, y in xy:
        for k in range(per_string):
            out.append(MultiPMTDetector(np.array([x, y, z0 + spacing_v * k], np.float32), radius, pmt_area,
                                        coords, len(out) + 1))
    return out


@dataclass
class CascadeRequest:
    sources: list
    seed: int
    capacity: int
    energy: float


def cascade_requests(n: int, medium, rng: np.random.Generator, radius: float = 150.0,
                     height: float = 300.0, e_range=(1e2, 1e6), photons_per_100gev: float = 1e4,
                     hit_fraction: float = 0.02):
    """Cascades uniform in a cylinder (area-uniform radius), isotropic
    direction, log-uniform energy; photons = photons_per_100gev·E/100 GeV."""
    from .lightyield import ExtendedCherenkovEmitter, LongitudinalParameterisation, PEMinus
    out = []
    for _ in range(n):
        r = radius * math.sqrt(rng.random())
        ph = 2 * math.pi * rng.random()
        pos = np.array([r * math.cos(ph), r * math.sin(ph), height * (rng.random() - 0.5)], np.float32)
        ct = 2 * rng.random() - 1
        st = math.sqrt(1 - ct * ct)
        ph = 2 * math.pi * rng.random()
        d = np.array([st * math.cos(ph), st * math.sin(ph), ct], np.float32)
        e = math.exp(math.log(e_range[0]) + rng.random() * math.log(e_range[1] / e_range[0]))
        photons = int(photons_per_100gev * e / 100.0)
        lp = LongitudinalParameterisation.from_energy(e, medium, PEMinus)
        src = ExtendedCherenkovEmitter(pos, d, 0.0, photons, lp)
        out.append(CascadeRequest([src], int(rng.integers(0, 2**63)), max(1024, int(photons * hit_fraction)), e))
    return out


def run_poisson_stream(server: PhotonServer, requests, rate_hz: float, rng: np.random.Generator,
                       keep_hits: bool = False):
    """Submit `requests` at Poisson arrival times (rate `rate_hz`) from one
    thread while another collects results; returns per-request latencies
    (ms), photons, wall time and hit counts (and the hits themselves with
    `keep_hits`)."""
    latencies = [None] * len(requests)
    hits = [0] * len(requests)
    results = [None] * len(requests)
    ids = [None] * len(requests)
    submitted = threading.Semaphore(0)
    errors = []

    def collector():
        try:
            for i in range(len(requests)):
                submitted.acquire()
                h, lat = server.wait(ids[i], raise_on_overflow=False)
                latencies[i] = lat
                hits[i] = len(h)
                if keep_hits:
                    results[i] = h
        except Exception as exc:   # pragma: no cover - reported to the caller
            errors.append(exc)

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i, r in enumerate(requests):
        t_next += rng.exponential(1.0 / rate_hz)
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids[i] = server.submit(r.sources, r.seed, r.capacity)
        submitted.release()
    th.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    lat = np.array(latencies)
    photons = sum(s.photons for r in requests for s in r.sources)
    out = {"latency_ms": lat, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "wall_s": wall, "photons": photons, "photons_per_s": photons / wall, "hits": int(sum(hits)),
           "requests": len(requests)}
    if keep_hits:
        out["results"] = results
    return out
