"""Water medium (host side): mirror of module Medium, src/photon_propagation/medium.jl.

`WaterProperties` keeps the reference's fields (salinity in permille, pressure
in atm, temperature in °C, particle concentrations in ppm, radiation length in
g/cm², density in kg/m³, mean scattering cosine, Quan-Fry coefficients), all as
float32 like `WaterProperties{Float32}`.  Host evaluations below run in
numpy float32 with the reference's operation order; the propagation kernel
carries its own device implementations.

Extensions with no reference counterpart (defaults reproduce the reference):
  sl_fraction      Simplified-Liu share of the mixed phase function
  sca_len_scale    multiplier on the scattering length
  dispersion_mode  "ref" keeps medium.jl:293-318 as written, "quan_fry"
                   differentiates the Quan-Fry polynomial
  max_steps        scattering steps per photon (15, photon_prop_cuda.jl:449)
  abs_len_scale    multiplier on the absorption-length table (weights only)
  direction_arithmetic
                   "fp64": direction sampling and rotations evaluated in FP64
                   and rounded once (within 2 FP32 ulp of the reference
                   formulas in FP64); "fp32": the reference's FP32 arithmetic
"""
from __future__ import annotations

import csv
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _abi

f32 = np.float32
C_VAC_M_NS = 0.299792458
DATA_DIR = Path(__file__).resolve().parent / "data"


def DIPPR105(temperature: float, A=0.14395, B=0.0112, C=649.727, D=0.05107) -> float:
    """Water density (kg/m³) at `temperature` in K.  medium.jl:26-48."""
    return A / (B ** (1 + (1 - temperature / C) ** D))


def _calc_quan_fry_params(salinity: float, temperature: float, pressure: float):
    """Quan-Fry coefficients (a01, a2, a3, a4).  medium.jl:174-205."""
    n0, n1, n2, n3, n4 = 1.31405, 1.45e-5, 1.779e-4, 1.05e-6, 1.6e-8
    n5, n6, n7, n8, n9, n10 = 2.02e-6, 15.868, 0.01155, 0.00423, 4382.0, 1.1455e6
    t, s, p = float(temperature), float(salinity), float(pressure)
    a01 = n0 + (n2 - n3 * t + n4 * t ** 2) * s - n5 * t ** 2 + n1 * p
    a2 = n6 + n7 * s - n8 * t
    return a01, a2, -n9, n10


@dataclass
class WaterProperties:
    """medium.jl:70-125 (constructed from plain SI-like numbers, not Unitful)."""
    salinity: float          # permille
    pressure: float          # atm
    temperature: float       # °C
    vol_conc_small_part: float   # ppm
    vol_conc_large_part: float   # ppm
    radiation_length: float  # g/cm^2
    mean_scattering_angle: float
    density: float = field(init=False)
    quan_fry_params: tuple = field(init=False)
    # extensions (defaults = reference behaviour)
    sl_fraction: float = 0.0
    sca_len_scale: float = 1.0
    dispersion_mode: str = "ref"
    max_steps: int = 15
    abs_len_scale: float = 1.0
    direction_arithmetic: str = "fp64"

    def __post_init__(self):
        for name in ("salinity", "pressure", "temperature", "vol_conc_small_part",
                     "vol_conc_large_part", "radiation_length", "mean_scattering_angle"):
            setattr(self, name, f32(getattr(self, name)))
        qf = _calc_quan_fry_params(self.salinity, self.temperature, self.pressure)
        self.quan_fry_params = tuple(f32(v) for v in qf)
        self.density = f32(DIPPR105(float(self.temperature) + 273.15))
        if self.dispersion_mode not in ("ref", "quan_fry"):
            raise ValueError("dispersion_mode must be 'ref' or 'quan_fry'")
        if self.direction_arithmetic not in ("fp64", "fp32"):
            raise ValueError("direction_arithmetic must be 'fp64' or 'fp32'")
        if not 0 <= int(self.max_steps) <= _abi.NT_MAX_STEPS:
            raise ValueError(f"max_steps must be in 0..{_abi.NT_MAX_STEPS}")

    def to_abi(self) -> _abi.nt_medium:
        m = _abi.nt_medium()
        for i, v in enumerate(self.quan_fry_params):
            m.quan_fry[i] = float(v)
        m.vol_conc_small_part = float(self.vol_conc_small_part)
        m.vol_conc_large_part = float(self.vol_conc_large_part)
        m.mean_scattering_angle = float(self.mean_scattering_angle)
        m.sl_fraction = float(self.sl_fraction)
        m.sca_len_scale = float(self.sca_len_scale)
        m.dispersion_mode = (_abi.NT_DISPERSION_REF if self.dispersion_mode == "ref"
                             else _abi.NT_DISPERSION_QUAN_FRY)
        m.max_steps = int(self.max_steps)
        m.direction_arithmetic = _abi.NT_DIR_FP64 if self.direction_arithmetic == "fp64" else _abi.NT_DIR_FP32_REF
        return m


MediumProperties = WaterProperties


def make_cascadia_medium_properties(mean_scattering_angle: float, **extensions) -> WaterProperties:
    """Cascadia Basin water, medium.jl:131-138 (269.44088 bar converted to atm)."""
    pressure_atm = float(f32(269.44088)) * 1e5 / 101325.0
    return WaterProperties(34.82, pressure_atm, 1.8, 0.0075, 0.0075, 36.08,
                           mean_scattering_angle, **extensions)


def salinity(m): return m.salinity
def pressure(m): return m.pressure
def temperature(m): return m.temperature
def material_density(m): return m.density
def vol_conc_small_part(m): return m.vol_conc_small_part
def vol_conc_large_part(m): return m.vol_conc_large_part
def radiation_length(m): return m.radiation_length
def mean_scattering_angle(m): return m.mean_scattering_angle


def refractive_index(wavelength, medium: WaterProperties):
    """refractive_index_fry, medium.jl:234-244 (float32)."""
    a01, a2, a3, a4 = medium.quan_fry_params
    wl = np.asarray(wavelength, dtype=f32)
    x = f32(1) / wl
    x2 = x * x
    # fma(x, a2, a01) + fma(x2, a3, x2*x*a4), fma evaluated exactly then rounded
    t1 = (x.astype(np.float64) * np.float64(a2) + np.float64(a01)).astype(f32)
    t2 = (x2.astype(np.float64) * np.float64(a3) + np.asarray(x2 * x * a4, dtype=np.float64)).astype(f32)
    return t1 + t2


def dispersion(wavelength, medium: WaterProperties):
    """dispersion_fry, medium.jl:293-298 (reference destructuring in 'ref' mode)."""
    a01, a2, a3, a4 = medium.quan_fry_params
    c = (a01, a2, a3) if medium.dispersion_mode == "ref" else (a2, a3, a4)
    wl = np.asarray(wavelength, dtype=f32)
    x = f32(1) / wl
    s = (c[0] + f32(2) * x * c[1]) + f32(3) * (x * x) * c[2]
    return (s * f32(-1)) / (wl * wl)


def group_velocity(wavelength, medium: WaterProperties):
    """medium.jl:313-318."""
    wl = np.asarray(wavelength, dtype=f32)
    n = refractive_index(wl, medium)
    lam0 = n * wl if medium.dispersion_mode == "ref" else wl
    return f32(C_VAC_M_NS) / (n - lam0 * dispersion(wl, medium))


def cherenkov_angle(wavelength, medium: WaterProperties):
    """medium.jl:309-311."""
    return np.arccos(f32(1) / refractive_index(wavelength, medium))


def scattering_length(wavelength, medium: WaterProperties):
    """_sca_len_part_conc, medium.jl:331-347 (× sca_len_scale)."""
    wl = np.asarray(wavelength, dtype=f32)
    x = f32(550) / wl
    c = (f32(0.0017) * x ** f32(4.3) + f32(1.34) * medium.vol_conc_small_part * x ** f32(1.7)) \
        + f32(0.312) * medium.vol_conc_large_part * x ** f32(0.3)
    l = f32(1) / c
    return l if medium.sca_len_scale == 1.0 else l * f32(medium.sca_len_scale)


def _read_table(name: str):
    xs, ys = [], []
    with open(DATA_DIR / name) as fh:
        rows = [r for r in fh if not r.startswith("#")]
    for row in csv.DictReader(rows):
        vals = list(row.values())
        xs.append(float(vals[0]))
        ys.append(float(vals[1]))
    return np.array(xs), np.array(ys)


_ABS_TABLE = None


def absorption_table():
    """STRAW absorption-length fit (medium.jl:392-404), wavelength nm -> m."""
    global _ABS_TABLE
    if _ABS_TABLE is None:
        _ABS_TABLE = _read_table("attenuation_length_straw_fit.csv")
    return _ABS_TABLE


def fast_linear_interp(x_eval, xs, ys):
    """utils.jl:24-53, vectorised: clamp to the table, exact end values, linear
    between the bracketing knots (searchsortedfirst)."""
    x = np.clip(np.asarray(x_eval), xs[0], xs[-1])
    iu = np.searchsorted(xs, x, side="left")
    iu = np.clip(iu, 1, len(xs) - 1)
    il = iu - 1
    along = (x - xs[il]) / (xs[iu] - xs[il])
    out = ys[il] + (ys[iu] - ys[il]) * along
    out = np.where(x == xs[0], ys[0], out)
    out = np.where(x == xs[-1], ys[-1], out)
    return out


def absorption_length(wavelength, medium: WaterProperties):
    """medium.jl:412-414: table interpolated in the wavelength's precision."""
    wl = np.asarray(wavelength)
    xs, ys = absorption_table()
    dt = np.dtype(wl.dtype if wl.dtype in (np.float32, np.float64) else np.float64)
    out = fast_linear_interp(wl.astype(dt), xs.astype(dt), ys.astype(dt))
    return out if medium.abs_len_scale == 1.0 else out * dt.type(medium.abs_len_scale)
