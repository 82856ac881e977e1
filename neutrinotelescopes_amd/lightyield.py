"""Photon sources: mirror of module LightYield (src/photon_propagation/lightyield.jl)
and module Types (src/types.jl).

Each emitter is a plain dataclass with the reference's field names and order
(lightyield.jl:421-496) and converts to the C-ABI `nt_source`.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np

from . import _abi
from .medium import WaterProperties, material_density, radiation_length, refractive_index
from .spectrum import frank_tamm_norm, _n64


# ---- particle types, src/types.jl:8-24 -----------------------------------
class ParticleType:
    pass


class PEPlus(ParticleType):
    pdg = -11


class PEMinus(ParticleType):
    pdg = 11


class PGamma(ParticleType):
    pdg = 22


@dataclass
class Particle:
    """types.jl:18-24."""
    position: np.ndarray
    direction: np.ndarray
    time: float
    energy: float
    type: type


# ---- longitudinal profile, lightyield.jl:129-213 --------------------------
@dataclass(frozen=True)
class LongitudinalParameters:
    alpha: float
    beta: float
    b: float


LongitudinalParametersEMinus = LongitudinalParameters(2.01849, 1.45469, 0.63207)
LongitudinalParametersEPlus = LongitudinalParameters(2.00035, 1.45501, 0.63008)
LongitudinalParametersGamma = LongitudinalParameters(2.83923, 1.34031, 0.64526)

_LONG = {PEMinus: LongitudinalParametersEMinus, PEPlus: LongitudinalParametersEPlus,
         PGamma: LongitudinalParametersGamma}


@dataclass(frozen=True)
class CherenkovTrackLengthParameters:
    alpha: float
    beta: float
    alpha_dev: float
    beta_dev: float


_TRACK = {
    PEMinus: CherenkovTrackLengthParameters(5.3207078881, 1.00000211, 0.0578170887, 0.5),
    PEPlus: CherenkovTrackLengthParameters(5.3211320598, 0.99999254, 0.0573419669, 0.5),
    PGamma: CherenkovTrackLengthParameters(5.3208540905, 0.99999877, 0.0578170887, 5.66586567),
}


def get_longitudinal_params(ptype) -> LongitudinalParameters:
    return _LONG[ptype]


def long_parameter_a_edep(energy: float, long_pars: LongitudinalParameters) -> float:
    """lightyield.jl:204-209."""
    return long_pars.alpha + long_pars.beta * np.log10(energy)


@dataclass
class LongitudinalParameterisation:
    """lightyield.jl:139-155: Γ(a, b) shower-depth profile in radiation lengths."""
    a: float
    b: float
    lrad: float

    @classmethod
    def from_energy(cls, energy: float, medium: WaterProperties, long_pars_or_ptype):
        lp = long_pars_or_ptype if isinstance(long_pars_or_ptype, LongitudinalParameters) \
            else get_longitudinal_params(long_pars_or_ptype)
        b = np.float32(lp.b)
        a = np.float32(long_parameter_a_edep(energy, lp))
        lrad = np.float32(radiation_length(medium) / material_density(medium) * 10)
        return cls(float(a), float(b), float(lrad))


def cherenkov_track_length(energy: float, ptype) -> float:
    """Track length in m, lightyield.jl:350-353."""
    p = _TRACK[ptype]
    return p.alpha * energy ** p.beta


def total_lightyield(particle: Particle, medium: WaterProperties, wl_range) -> float:
    """Expected Cherenkov photons, lightyield.jl:356-366."""
    return frank_tamm_norm(wl_range, _n64(medium)) * cherenkov_track_length(particle.energy,
                                                                            particle.type)


# ---- emitters, lightyield.jl:369-530 ---------------------------------------
def _v3(x):
    return np.asarray(x, dtype=np.float32).reshape(3)


class PhotonSource:
    photons: int

    def to_abi(self) -> _abi.nt_source:  # pragma: no cover - abstract
        raise NotImplementedError


@dataclass
class AxiconeEmitter(PhotonSource):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    angle: float

    def to_abi(self):
        return _src(_abi.NT_SRC_AXICONE, self.photons, self.position, self.direction, self.time,
                    angle=self.angle)


@dataclass
class PencilEmitter(PhotonSource):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int

    def to_abi(self):
        return _src(_abi.NT_SRC_PENCIL, self.photons, self.position, self.direction, self.time)


@dataclass
class PointlikeIsotropicEmitter(PhotonSource):
    position: np.ndarray
    time: float
    photons: int

    def to_abi(self):
        return _src(_abi.NT_SRC_POINTLIKE_ISOTROPIC, self.photons, self.position, (0, 0, 1), self.time)


@dataclass
class PointlikeTimeRangeEmitter(PhotonSource):
    position: np.ndarray
    time_range: tuple
    photons: int

    def to_abi(self):
        return _src(_abi.NT_SRC_POINTLIKE_TIME_RANGE, self.photons, self.position, (0, 0, 1),
                    self.time_range[0], time_end=self.time_range[1])


class CherenkovEmitter(PhotonSource):
    pass


@dataclass
class ExtendedCherenkovEmitter(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    long_param: LongitudinalParameterisation

    @classmethod
    def from_particle(cls, particle: Particle, medium: WaterProperties, wl_range,
                      oversample: float = 1.0, rng: np.random.Generator | None = None):
        """lightyield.jl:471-481: Poisson-distributed photon count."""
        rng = rng or np.random.default_rng()
        lp = LongitudinalParameterisation.from_energy(particle.energy, medium, particle.type)
        photons = int(rng.poisson(total_lightyield(particle, medium, wl_range) * oversample))
        return cls(particle.position, particle.direction, particle.time, photons, lp)

    def to_abi(self):
        return _src(_abi.NT_SRC_EXTENDED_CHERENKOV, self.photons, self.position, self.direction,
                    self.time, long=(self.long_param.a, self.long_param.b, self.long_param.lrad))


@dataclass
class PointlikeCherenkovEmitter(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int

    @classmethod
    def from_particle(cls, particle: Particle, medium: WaterProperties | None = None,
                      wl_range=None, photons: int | None = None,
                      rng: np.random.Generator | None = None):
        """lightyield.jl:498-509."""
        if photons is None:
            rng = rng or np.random.default_rng()
            photons = int(rng.poisson(total_lightyield(particle, medium, wl_range)))
        return cls(particle.position, particle.direction, particle.time, photons)

    def to_abi(self):
        return _src(_abi.NT_SRC_POINTLIKE_CHERENKOV, self.photons, self.position, self.direction,
                    self.time)


def _src(kind, photons, position, direction, time, time_end=0.0, angle=0.0, long=(0.0, 0.0, 0.0)):
    s = _abi.nt_source()
    s.kind = kind
    s.photons = int(photons)
    p, d = _v3(position), _v3(direction)
    for i in range(3):
        s.position[i] = float(p[i])
        s.direction[i] = float(d[i])
    s.time = float(time)
    s.time_end = float(time_end)
    s.angle = float(angle)
    s.long_a, s.long_b, s.long_lrad = (float(v) for v in long)
    return s


def split_source(source: PhotonSource, parts: int) -> list:
    """lightyield.jl:372-401: `parts` copies; the last takes the remainder."""
    if source.photons < parts:
        raise ValueError("Cannot split source. Fewer photons than parts")
    ph_split, remainder = divmod(int(source.photons), int(parts))
    out = []
    for i in range(parts):
        nph = ph_split + remainder if i == parts - 1 else ph_split
        out.append(dataclasses.replace(source, photons=nph))
    return out


def oversample_source(source: PhotonSource, factor: float) -> PhotonSource:
    """lightyield.jl:410-420."""
    return dataclasses.replace(source, photons=int(source.photons * factor))
