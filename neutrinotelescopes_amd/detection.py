"""Detection targets: mirror of module Detection (src/photon_propagation/detection.jl)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .medium import _read_table, fast_linear_interp


@dataclass
class DetectionSphere:
    """detection.jl:29-35."""
    position: np.ndarray
    radius: float
    n_pmts: int
    pmt_area: float
    module_id: int = 1

    def to_abi(self) -> _abi.nt_target:
        t = _abi.nt_target()
        _fill_target(t, self.position, self.radius, self.pmt_area, self.module_id)
        t.n_pmts = int(self.n_pmts)
        t.kind = _abi.NT_TARGET_SPHERE
        return t


@dataclass
class MultiPMTDetector:
    """detection.jl:39-45; `pmt_coordinates` is 2×N (θ, φ) per PMT."""
    position: np.ndarray
    radius: float
    pmt_area: float
    pmt_coordinates: np.ndarray
    module_id: int = 1

    @property
    def n_pmts(self) -> int:
        return int(np.asarray(self.pmt_coordinates).shape[1])

    def to_abi(self) -> _abi.nt_target:
        t = _abi.nt_target()
        _fill_target(t, self.position, self.radius, self.pmt_area, self.module_id)
        pc = np.asarray(self.pmt_coordinates, dtype=np.float32)
        if pc.shape[0] != 2 or pc.shape[1] > _abi.NT_MAX_PMTS:
            raise ValueError("pmt_coordinates must be 2 x N with N <= 16")
        t.n_pmts = pc.shape[1]
        t.kind = _abi.NT_TARGET_MULTI_PMT
        flat = pc.T.reshape(-1)       # column-major (θ1, φ1, θ2, φ2, ...)
        for i, v in enumerate(flat):
            t.pmt_coordinates[i] = float(v)
        return t


PhotonTarget = (DetectionSphere, MultiPMTDetector)


def _fill_target(t, position, radius, pmt_area, module_id):
    p = np.asarray(position, dtype=np.float32).reshape(3)
    for i in range(3):
        t.position[i] = float(p[i])
    t.radius = float(np.float32(radius))
    t.pmt_area = float(np.float32(pmt_area))
    if not 0 <= int(module_id) <= 0xFFFF:
        raise ValueError("module_id must fit UInt16")
    t.module_id = int(module_id)


def get_pmt_count(target) -> int:
    return 1 if isinstance(target, DetectionSphere) else target.n_pmts


def sph_to_cart(theta, phi):
    """utils.jl:105-115 (float64)."""
    theta, phi = np.asarray(theta, dtype=np.float64), np.asarray(phi, dtype=np.float64)
    return np.stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)],
                    axis=-1)


def cart_to_sph(x, y, z):
    """utils.jl:118-133."""
    if z == 1:
        return 0.0, 0.0
    if z == -1:
        return float(np.pi), 0.0
    theta = float(np.arccos(z))
    if x == 0 and y > 0:
        phi = np.pi / 2
    elif x == 0 and y < 0:
        phi = -np.pi / 2
    else:
        phi = float(np.arctan2(y, x))
    return theta, phi


def calc_rot_matrix(a, b):
    """Rotation taking a to b, utils.jl:173-184."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    if np.array_equal(a, b):
        return np.eye(3)
    v = np.cross(a, b)
    s = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + s + s @ s * (1 - a @ b) / (v @ v)


def make_pom_pmt_coordinates(dtype=np.float64) -> np.ndarray:
    """16-PMT POM layout, detection.jl:107-134 (2×16, θ then φ)."""
    coords = np.empty((2, 16))
    coords[0, 0:4] = np.deg2rad(90 - 57.5)
    coords[1, 0:4] = np.pi / 4 + np.arange(4) * np.pi / 2
    coords[0, 4:8] = np.deg2rad(90 - 25)
    coords[1, 4:8] = np.arange(4) * np.pi / 2
    coords[0, 8:12] = np.deg2rad(90 + 25)
    coords[1, 8:12] = np.arange(4) * np.pi / 2
    coords[0, 12:16] = np.deg2rad(90 + 57.5)
    coords[1, 12:16] = np.pi / 4 + np.arange(4) * np.pi / 2
    R = calc_rot_matrix([0.0, 0.0, 1.0], [1.0, 0.0, 0.0])
    for j in range(16):
        cart = R @ sph_to_cart(coords[0, j], coords[1, j])
        coords[:, j] = cart_to_sph(*cart)
    return coords.astype(dtype)


def get_pmt_positions(target: MultiPMTDetector, orientation: np.ndarray) -> np.ndarray:
    """PMT axes (N×3) after `orientation`, detection.jl:60-70: sph_to_cart in
    the coordinates' precision, then the FP64 rotation."""
    pc = np.asarray(target.pmt_coordinates)
    if pc.dtype == np.float32:
        th, ph = pc[0].astype(np.float64), pc[1].astype(np.float64)
        s = lambda x: x.astype(np.float32)            # correctly rounded float32 sin/cos
        v = np.stack([s(np.cos(ph)) * s(np.sin(th)), s(np.sin(ph)) * s(np.sin(th)), s(np.cos(th))], -1)
        v = v.astype(np.float64)
    else:
        v = sph_to_cart(pc[0], pc[1])
    return (np.asarray(orientation, dtype=np.float64) @ v.T).T


def opening_angle(target: MultiPMTDetector):
    """asin(sqrt(pmt_area/π)/radius), detection.jl:94-95, in the target's precision."""
    if np.asarray(target.pmt_coordinates).dtype == np.float32:
        r = np.sqrt(np.float32(target.pmt_area) / np.float32(np.pi))
        return float(np.float32(np.arcsin(np.float64(r / np.float32(target.radius)))))
    return float(np.arcsin(np.sqrt(float(target.pmt_area) / np.pi) / float(target.radius)))


def check_pmt_hit(hit_positions, target, orientation=None) -> np.ndarray:
    """1-based PMT index per hit (0 = between PMTs), detection.jl:72-104."""
    hit_positions = np.asarray(hit_positions, dtype=np.float64).reshape(-1, 3)
    if isinstance(target, DetectionSphere):
        return np.ones(len(hit_positions), dtype=np.int64)
    pmt_pos = get_pmt_positions(target, np.eye(3) if orientation is None else orientation)
    opening = opening_angle(target)
    rel = hit_positions - np.asarray(target.position, dtype=np.float64)
    rel = rel / np.linalg.norm(rel, axis=1, keepdims=True)
    ang = np.arccos(np.clip(rel @ pmt_pos.T, -1.0, 1.0))    # hits × PMTs
    inside = ang < opening
    first = np.argmax(inside, axis=1)
    return np.where(inside.any(axis=1), first + 1, 0)


def area_acceptance(position, target) -> float:
    """detection.jl:139-146."""
    if isinstance(target, DetectionSphere):
        num = np.float32(np.float32(target.n_pmts) * np.float32(target.pmt_area))   # Int·Float32
        r = np.float32(target.radius)
        return float(num) / (4 * np.pi * float(r * r))                              # Float64
    return 1.0


_PMT_ACC = None


def pmt_acceptance_table():
    """assets/PMTAcc.csv (wavelength nm, single-PMT acceptance), detection.jl:161-163."""
    global _PMT_ACC
    if _PMT_ACC is None:
        _PMT_ACC = _read_table("pmt_acceptance.csv")
    return _PMT_ACC


def p_one_pmt_acc(wavelength):
    """Single-PMT wavelength acceptance, detection.jl:148-163 (linear interpolation
    of assets/PMTAcc.csv)."""
    xs, ys = pmt_acceptance_table()
    wl = np.asarray(wavelength, dtype=np.float64)
    if np.any((wl < xs[0]) | (wl > xs[-1])):
        raise ValueError("wavelength outside the PMT acceptance table")
    return np.interp(wl, xs, ys)


def make_detector_cube(nx, ny, nz, spacing_vert, spacing_hor):
    """detection.jl:167-178."""
    out = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                out.append(np.array([-0.5 * spacing_hor * nx + i * spacing_hor,
                                     -0.5 * spacing_hor * ny + j * spacing_hor,
                                     -0.5 * spacing_vert * nz + k * spacing_vert]))
    # LinearIndices((1:nx,1:ny,1:nz)) order: i fastest
    return out


def make_targets(positions, n_pmts, pmt_area):
    """detection.jl:208-210."""
    return [DetectionSphere(p, 0.21, n_pmts, pmt_area) for p in positions]


def lambertian_angular_acceptance(n: int = 21):
    """A stand-in angular acceptance table (the configuration names a
    tabulated angular × wavelength acceptance and the reference provides
    none): the PMT's projected area, factor cos η for incidence angles η below
    90°, 0 beyond, at n cosines in [−1, 1]."""
    c = np.linspace(-1.0, 1.0, n)
    return c, np.clip(c, 0.0, None)
