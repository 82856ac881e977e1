"""Emission spectra: mirror of module Spectral, src/photon_propagation/spectrum.jl.

`CherenkovSpectrum(wl_range, interp_steps, medium)` builds the same inverse-CDF
knot table as the reference (Frank-Tamm CDF at 1-nm steps, each a 10-point
Gauss-Legendre integral, interpolated at range(0, 1, interp_steps)); the
kernel samples it the way the reference's CUDA texture does (normalized
coordinates, linear filter, clamp addressing).  The knots are computed in
float64 here (the reference accumulates partly in float32; the difference is
below 1e-6 nm).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .medium import WaterProperties, refractive_index

FINE_STRUCTURE = 7.2973525693e-3
GL10_NODES, GL10_WEIGHTS = np.polynomial.legendre.leggauss(10)


def integrate_gauss_quad(f, a: float, b: float):
    """10-point Gauss-Legendre quadrature, utils.jl:82-103."""
    ba_half = (b - a) / 2.0
    mid = (a + b) / 2.0
    return float(np.dot(GL10_WEIGHTS, f(ba_half * GL10_NODES + mid) * ba_half))


def frank_tamm(wavelength, ref_index):
    """Photons per unit length per unit wavelength, spectrum.jl:32-34."""
    wavelength = np.asarray(wavelength, dtype=np.float64)
    ref_index = np.asarray(ref_index, dtype=np.float64)
    return 2 * np.pi * FINE_STRUCTURE / wavelength ** 2 * (1 - 1 / ref_index ** 2)


def _n64(medium: WaterProperties):
    a01, a2, a3, a4 = (float(v) for v in medium.quan_fry_params)

    def n(wl):
        x = 1.0 / wl
        return a01 + x * a2 + x * x * a3 + x * x * x * a4
    return n


def frank_tamm_norm(wl_range, ref_index_func) -> float:
    """Cherenkov photons per metre in `wl_range`, spectrum.jl:42-45."""
    return integrate_gauss_quad(lambda x: frank_tamm(x, ref_index_func(x)),
                                float(wl_range[0]), float(wl_range[1])) * 1e9


def frank_tamm_inverted_cdf(wl_range, medium: WaterProperties, step_size: float = 1.0):
    """(norms, wl_steps) of the Frank-Tamm CDF at `step_size` nm, spectrum.jl:55-71."""
    n = _n64(medium)
    lo, hi = float(wl_range[0]), float(wl_range[1])
    wl_steps = lo + step_size * np.arange(int(np.floor((hi - lo) / step_size)) + 1)
    full = frank_tamm_norm((lo, hi), n)
    norms = np.array([0.0] + [frank_tamm_norm((lo, s), n) / full for s in wl_steps[1:]])
    order = np.argsort(norms, kind="stable")
    return norms[order], wl_steps[order]


def cherenkov_spectrum_knots(wl_range, interp_steps: int, medium: WaterProperties) -> np.ndarray:
    """Inverse-CDF knots at range(0, 1, interp_steps), spectrum.jl:97-100."""
    norms, wl_steps = frank_tamm_inverted_cdf(wl_range, medium)
    u = np.linspace(0.0, 1.0, int(interp_steps))
    return np.interp(u, norms, wl_steps).astype(np.float32)


class Spectrum:
    def to_abi(self) -> _abi.nt_spectrum:  # pragma: no cover - abstract
        raise NotImplementedError


@dataclass
class Monochromatic(Spectrum):
    """spectrum.jl:21-23."""
    wavelength: float

    def to_abi(self):
        s = _abi.nt_spectrum()
        s.kind = _abi.NT_SPEC_MONOCHROMATIC
        s.wavelength = float(self.wavelength)
        return s


@dataclass
class CherenkovSpectrum(Spectrum):
    """spectrum.jl:89-108.  `addressing` selects how the knot table is sampled:
    "texture" (the reference's texture fetch), "texture_q8" (same with the
    filter weight rounded to 1/256, as texture hardware stores it) or
    "endpoint" (knots at u = i/(N-1))."""
    wl_range: tuple
    interp_steps: int = 30
    medium: WaterProperties | None = None
    knots: np.ndarray | None = None
    addressing: str = "texture"
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.knots is None:
            if self.medium is None:
                raise ValueError("CherenkovSpectrum needs a medium or explicit knots")
            self.knots = cherenkov_spectrum_knots(self.wl_range, self.interp_steps, self.medium)
        self.knots = np.ascontiguousarray(self.knots, dtype=np.float32)
        if not 2 <= len(self.knots) <= _abi.NT_MAX_KNOTS:
            raise ValueError("knot count out of range")

    def to_abi(self):
        s = _abi.nt_spectrum()
        s.kind = _abi.NT_SPEC_TABLE
        s.addressing = {"texture": _abi.NT_TEX_TEXEL_CENTER_CLAMP,
                        "texture_q8": _abi.NT_TEX_TEXEL_CENTER_CLAMP_Q8,
                        "endpoint": _abi.NT_TEX_ENDPOINT}[self.addressing]
        s.wl_range[0], s.wl_range[1] = float(self.wl_range[0]), float(self.wl_range[1])
        s.n_knots = len(self.knots)
        s.knots = self.knots.ctypes.data_as(_abi.C.POINTER(_abi.C.c_float))
        return s


@dataclass
class InvLambda2Spectrum(Spectrum):
    """Extension: wavelengths drawn ∝ 1/λ² on `wl_range` by the closed-form
    inverse CDF λ(u) = 1/((1−u)/λ_lo + u/λ_hi)."""
    wl_range: tuple

    def to_abi(self):
        s = _abi.nt_spectrum()
        s.kind = _abi.NT_SPEC_INV_LAMBDA2
        s.wl_range[0], s.wl_range[1] = float(self.wl_range[0]), float(self.wl_range[1])
        return s
