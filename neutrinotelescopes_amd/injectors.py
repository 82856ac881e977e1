"""Particle injection and the device-side source builder.

Reference (paths relative to the reference repository root):
  Cylinder, Cuboid, FixedPosition, VolumeInjector, UniformAngularDistribution
                                               src/event_generation/injectors.jl:30-111
                                               (host sampling, and nt_inject_particles on the device)
  ExtendedCherenkovEmitter(particle, medium, wl_range; oversample)
                                               src/photon_propagation/lightyield.jl:471-481
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from .lightyield import (ExtendedCherenkovEmitter, LongitudinalParameterisation, Particle, PEMinus, PEPlus,
                         PGamma)
from .medium import WaterProperties, material_density, radiation_length
from .spectrum import _n64, frank_tamm_norm

_TYPE_CODE = {PEMinus: _abi.NT_PARTICLE_EMINUS, PEPlus: _abi.NT_PARTICLE_EPLUS, PGamma: _abi.NT_PARTICLE_GAMMA}


@dataclass
class Cylinder:
    """injectors.jl:30-34, 59-70: centre, height, radius.  The radial
    coordinate is sqrt(u · radius), as the reference draws it (uniform in
    area only for radius 1); kept so that injected samples follow the same
    distribution."""
    center: np.ndarray
    height: float
    radius: float

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        z = rng.uniform(-0.5 * self.height, 0.5 * self.height, n)
        r = np.sqrt(rng.random(n) * self.radius)
        a = rng.random(n) * 2.0 * math.pi
        return np.asarray(self.center, np.float64) + np.column_stack([r * np.cos(a), r * np.sin(a), z])


@dataclass
class Cuboid:
    """injectors.jl:41-46, 72-78."""
    center: np.ndarray
    l_x: float
    l_y: float
    l_z: float

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        h = 0.5 * np.array([self.l_x, self.l_y, self.l_z])
        return np.asarray(self.center, np.float64) + rng.uniform(-h, h, (n, 3))


@dataclass
class FixedPosition:
    """injectors.jl:48-57."""
    position: np.ndarray

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        return np.tile(np.asarray(self.position, np.float64), (n, 1))


@dataclass
class LogUniform:
    lo: float
    hi: float

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        return np.exp(rng.uniform(math.log(self.lo), math.log(self.hi), n))


class UniformAngularDistribution:
    """injectors.jl:82-89: φ uniform, θ = acos(2u − 1)."""

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        phi = rng.random(n) * 2.0 * math.pi
        theta = np.arccos(2.0 * rng.random(n) - 1.0)
        return np.column_stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)])


@dataclass
class VolumeInjector:
    """injectors.jl:93-111: position in a volume, energy from a distribution,
    particle type from a categorical distribution, direction from an angular
    distribution; time 0."""
    volume: Cylinder
    e_dist: LogUniform
    types: tuple = (PEMinus,)
    type_probs: tuple = (1.0,)
    angular_dist: UniformAngularDistribution = UniformAngularDistribution()

    def sample(self, n: int, rng: np.random.Generator) -> list:
        pos = self.volume.sample(n, rng)
        e = self.e_dist.sample(n, rng)
        t = rng.choice(len(self.types), size=n, p=np.asarray(self.type_probs) / np.sum(self.type_probs))
        d = self.angular_dist.sample(n, rng)
        return [Particle(pos[i], d[i], 0.0, float(e[i]), self.types[t[i]]) for i in range(n)]


def injector_to_abi(inj: VolumeInjector) -> _abi.nt_injector:
    """The injector as the device injector's record (nt_inject_particles):
    its volume, a log-uniform energy range and the cumulative type
    probabilities in the order e-, e+, gamma."""
    if not isinstance(inj.e_dist, LogUniform):
        raise ValueError("the device injector draws log-uniform energies")
    r = _abi.nt_injector()
    v = inj.volume
    if isinstance(v, Cylinder):
        r.volume, r.height, r.radius = _abi.NT_VOLUME_CYLINDER, float(v.height), float(v.radius)
        c = v.center
    elif isinstance(v, Cuboid):
        r.volume = _abi.NT_VOLUME_CUBOID
        r.size[0], r.size[1], r.size[2] = float(v.l_x), float(v.l_y), float(v.l_z)
        c = v.center
    elif isinstance(v, FixedPosition):
        r.volume, c = _abi.NT_VOLUME_FIXED, v.position
    else:
        raise ValueError(f"unsupported volume {type(v).__name__}")
    for q in range(3):
        r.center[q] = float(c[q])
    r.e_min, r.e_max = float(inj.e_dist.lo), float(inj.e_dist.hi)
    p = np.zeros(3)
    w = np.asarray(inj.type_probs, np.float64) / np.sum(inj.type_probs)
    for t, pw in zip(inj.types, w):
        p[_TYPE_CODE[t]] += pw
    cdf = np.cumsum(p)
    cdf[-1] = 1.0
    for q in range(3):
        r.type_cdf[q] = float(cdf[q])
    return r


def inject_particles_device(inj: VolumeInjector, n: int, seed: int, engine, device: int = 0, stream=None):
    """n particles of `inj` drawn on the device (nt_inject_particles): a torch
    uint8 tensor of n nt_particle records in HBM, ready for
    build_sources_device."""
    import torch
    rec = injector_to_abi(inj)
    dev = torch.device(f"cuda:{device}")
    out = torch.empty(max(n, 1) * C.sizeof(_abi.nt_particle), dtype=torch.uint8, device=dev)
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    _abi.check(engine.lib.nt_inject_particles(engine.handle, C.byref(rec), int(n), C.c_uint64(seed & (2 ** 64 - 1)),
                                              C.c_void_p(out.data_ptr()), C.c_void_p(st.cuda_stream)),
               "nt_inject_particles")
    return out[:n * C.sizeof(_abi.nt_particle)]


def cascade_model(medium: WaterProperties, wl_range=(300.0, 800.0), oversample: float = 1.0) -> _abi.nt_cascade_model:
    """Photons per metre of Cherenkov track (frank_tamm_norm) and the radiation
    length in m of `medium` (lightyield.jl:147-155, 356-366)."""
    m = _abi.nt_cascade_model()
    m.photons_per_m = float(frank_tamm_norm(wl_range, _n64(medium)))
    m.lrad = float(radiation_length(medium) / material_density(medium) * 10)
    m.oversample = float(oversample)
    return m


PARTICLE_DTYPE = np.dtype([("position", "<f4", 3), ("direction", "<f4", 3), ("time", "<f4"), ("energy", "<f4"),
                           ("type", "<i4"), ("reserved", "<i4")])
SOURCE_DTYPE = np.dtype([("kind", "<i4"), ("reserved", "<i4"), ("photons", "<i8"), ("position", "<f4", 3),
                         ("direction", "<f4", 3), ("time", "<f4"), ("time_end", "<f4"), ("angle", "<f4"),
                         ("long_a", "<f4"), ("long_b", "<f4"), ("long_lrad", "<f4")])
assert PARTICLE_DTYPE.itemsize == C.sizeof(_abi.nt_particle) and SOURCE_DTYPE.itemsize == C.sizeof(_abi.nt_source)


def pack_particles(position, direction, time, energy, type_code) -> np.ndarray:
    """Columns (n×3, n×3, n, n, n) as an nt_particle record array."""
    position = np.asarray(position, np.float64).reshape(-1, 3)
    out = np.zeros(len(position), PARTICLE_DTYPE)
    out["position"] = position
    out["direction"] = np.asarray(direction, np.float64).reshape(-1, 3)
    out["time"] = time
    out["energy"] = energy
    out["type"] = type_code
    return out


def particles_to_abi(particles) -> np.ndarray:
    """Particle objects as an nt_particle record array."""
    return pack_particles([p.position for p in particles], [p.direction for p in particles],
                          [p.time for p in particles], [p.energy for p in particles],
                          [_TYPE_CODE[p.type] for p in particles])


def build_sources_device(d_particles, n: int, model: _abi.nt_cascade_model, seed: int, engine, stream=None):
    """nt_build_cascade_sources on device memory: `d_particles` a torch uint8
    tensor of n nt_particle records; returns a torch uint8 tensor of n
    nt_source records on the same device (no host round trip)."""
    import torch
    if d_particles.dtype != torch.uint8 or d_particles.numel() != n * C.sizeof(_abi.nt_particle):
        raise ValueError("d_particles must hold n nt_particle records as uint8")
    d_out = torch.empty(max(n, 1) * C.sizeof(_abi.nt_source), dtype=torch.uint8, device=d_particles.device)
    st = stream if stream is not None else torch.cuda.current_stream(d_particles.device)
    _abi.check(engine.lib.nt_build_cascade_sources(engine.handle, C.c_void_p(d_particles.data_ptr()), int(n),
                                                   C.byref(model), C.c_uint64(seed & (2 ** 64 - 1)),
                                                   C.c_void_p(d_out.data_ptr()), C.c_void_p(st.cuda_stream)),
               "nt_build_cascade_sources")
    return d_out[:n * C.sizeof(_abi.nt_source)]


def set_sources_device(engine, d_sources, n: int, kind_hint: int = _abi.NT_SRC_EXTENDED_CHERENKOV, stream=None):
    """Make n nt_source records in device memory (a torch uint8 tensor, e.g.
    from build_sources_device) the engine's sources without a host copy
    (nt_set_sources_device)."""
    import torch
    if d_sources.dtype != torch.uint8 or d_sources.numel() != n * C.sizeof(_abi.nt_source):
        raise ValueError("d_sources must hold n nt_source records as uint8")
    st = stream if stream is not None else torch.cuda.current_stream(d_sources.device)
    _abi.check(engine.lib.nt_set_sources_device(engine.handle, C.c_void_p(d_sources.data_ptr()), int(n),
                                                int(kind_hint), C.c_void_p(st.cuda_stream)),
               "nt_set_sources_device")


def build_cascade_sources(particles, medium: WaterProperties, wl_range=(300.0, 800.0), oversample: float = 1.0,
                          seed: int = 0, device: int = 0, engine=None) -> list:
    """ExtendedCherenkovEmitter for every particle, built on the GPU
    (nt_build_cascade_sources): Poisson photon count around the full light
    yield × oversample and the longitudinal profile of the particle type."""
    import torch

    from .propagation import get_engine
    eng = engine if engine is not None else get_engine(device)
    n = len(particles)
    model = cascade_model(medium, wl_range, oversample)
    if n == 0:
        return []
    dev = torch.device(f"cuda:{device}")
    d_in = torch.from_numpy(particles_to_abi(particles).view(np.uint8)).to(dev)
    rec = build_sources_device(d_in, n, model, seed, eng).cpu().numpy().view(SOURCE_DTYPE)
    out = []
    for r in rec:
        lp = LongitudinalParameterisation(float(r["long_a"]), float(r["long_b"]), float(r["long_lrad"]))
        out.append(ExtendedCherenkovEmitter(r["position"].copy(), r["direction"].copy(), float(r["time"]),
                                            int(r["photons"]), lp))
    return out
