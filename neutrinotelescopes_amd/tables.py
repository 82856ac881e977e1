"""Photon tables in the reference's layout: scripts/photon_tables.jl.

`run_sim` follows run_sim (photon_tables.jl:52-169) for one table entry: a
cascade (or point-like Cherenkov emitter) at distance d on the z axis of a
16-PMT module at the origin, propagated with the oversampling retry loop,
written as one `photons` dataset [tres, pos_x, pos_y, pos_z, total_weight];
then `n_resample` random module orientations, each giving a `pmt_hits`
dataset [tres, pmt_id] of resampled detected hits with the source's
direction and position angles in the rotated module frame.  `run_sims`
scans the parameters on a Sobol sequence (photon_tables.jl:172-230).
`export_tables` writes both groups in the reference's HDF5 group/dataset/
attribute layout as an .npz archive (no HDF5 library in this environment).

The propagation, the weights and time residuals (device hit epilogue) and the
resampling (nt_resample) run on the device; the PMT assignment under each
random orientation is the host's FP64 check_pmt_hit, as the reference's
make_hits_from_photons (photon_prop_cuda.jl:884-904).
"""
from __future__ import annotations

import dataclasses
import json
import math

import numpy as np

from .detection import MultiPMTDetector, cart_to_sph, check_pmt_hit, make_pom_pmt_coordinates
from .lightyield import (ExtendedCherenkovEmitter, Particle, PEMinus, PointlikeCherenkovEmitter,
                         oversample_source)
from .propagation import PhotonHits, PhotonPropSetup

WL_RANGE = (300.0, 800.0)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """A rotation matrix uniform on SO(3) (rand(RotMatrix3)): the unit
    quaternion of a normalised 4-d Gaussian."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def table_target(radius: float = 0.21) -> MultiPMTDetector:
    """The module of photon_tables.jl:178-186: 16 PMTs of 75 mm at the origin."""
    pmt_area = float(np.float32((75e-3 / 2) ** 2 * math.pi))
    return MultiPMTDetector(np.zeros(3, np.float32), np.float32(radius), pmt_area,
                            make_pom_pmt_coordinates(np.float32), 1)


def run_sim(engine, energy: float, distance: float, dir_costheta: float, dir_phi: float, target, spectrum,
            medium, n_resample: int = 100, extended: bool = True, seed: int = 0, rng=None,
            max_photons: float = 1e13, min_hits: int = 100):
    """One table entry (photon_tables.jl:52-169).  Returns None when the
    oversampled source would exceed `max_photons` (the reference skips it),
    else {"photons": (data, attrs), "pmt_hits": [(data, attrs), ...],
    "oversample": factor}."""
    rng = rng or np.random.default_rng(seed)
    sim_attrs = {"energy": float(energy), "distance": float(distance), "dir_costheta": float(dir_costheta),
                 "dir_phi": float(dir_phi)}
    theta = np.float32(math.acos(dir_costheta))
    direction = np.array([np.cos(np.float32(dir_phi)) * np.sin(theta), np.sin(np.float32(dir_phi)) * np.sin(theta),
                          np.cos(theta)], np.float32)                      # sph_to_cart in Float32
    ppos = np.array([0.0, 0.0, distance], np.float32)
    particle = Particle(ppos, direction, 0.0, float(np.float32(energy)), PEMinus)
    if extended:
        source = ExtendedCherenkovEmitter.from_particle(particle, medium, WL_RANGE, rng=rng)
    else:
        source = PointlikeCherenkovEmitter.from_particle(particle, medium, WL_RANGE, rng=rng)
    oversample = 1.0
    # oversample ×10 until more than `min_hits` photons arrive (photon_tables.jl:93-109)
    while True:
        if source.photons > max_photons:
            return None
        setup = PhotonPropSetup([source], [target], medium, spectrum, seed=seed)
        engine.setup(setup)
        engine.set_epilogue(medium, source_position=source.position)
        hits = engine.propagate(seed=seed, capacity=1 << 20, raise_on_overflow=False)
        if hits.n_dropped:
            hits = engine.propagate(seed=seed, capacity=len(hits) + hits.n_dropped)
        if len(hits) > min_hits:
            break
        source = oversample_source(source, 10)
        oversample *= 10
    engine.set_epilogue(enabled=False)
    # calc_time_residual!, calc_total_weight! (device epilogue), weights ÷ oversample
    weight = hits.total_weight.astype(np.float64) / oversample
    tres = hits.tres.astype(np.float64)
    photons = (np.column_stack([tres, hits.position.astype(np.float64), weight]), dict(sim_attrs))
    pmt_hits = []
    for j in range(n_resample):
        orientation = random_rotation(rng)
        pmt = check_pmt_hit(hits.position, target, orientation)           # make_hits_from_photons
        on = pmt > 0
        if on.sum() < 10:
            continue
        sub = hits.select(on)
        sub.pmt_id = pmt[on].astype(np.uint8)
        sub.total_weight = weight[on].astype(np.float32)
        det = engine.resample(sub, seed=(seed * 1000003 + j + 1) & (2 ** 64 - 1), per_pmt=True)
        if not det.any():
            continue
        # the source in the frame where the rotated module is upright again
        direction_rot = orientation.T @ direction.astype(np.float64)
        position_rot = orientation.T @ ppos.astype(np.float64)
        position_rot_normed = position_rot / np.linalg.norm(position_rot)
        d_theta, d_phi = cart_to_sph(*direction_rot)
        p_theta, p_phi = cart_to_sph(*position_rot_normed)
        if not math.isclose(float(ppos @ direction) / float(np.linalg.norm(ppos)),
                            float(position_rot_normed @ direction_rot), rel_tol=1e-6, abs_tol=1e-6):
            raise RuntimeError("relative angle not preserved")
        # the reference updates one attribute dictionary in place, so the
        # rotated dir_phi replaces the sampled one from here on
        sim_attrs.update(dir_theta=d_theta, dir_phi=d_phi, pos_theta=p_theta, pos_phi=p_phi)
        pmt_hits.append((np.column_stack([tres[on][det], pmt[on][det].astype(np.float64)]), dict(sim_attrs)))
    return {"photons": photons, "pmt_hits": pmt_hits, "oversample": oversample}


def sobol_parameters(n_sims: int, n_skip: int = 0, extended: bool = True) -> np.ndarray:
    """The parameter scan of photon_tables.jl:196-224 on an unscrambled Sobol
    sequence: (log10 E in [2, 5], log10 d in [1, log10 150], cos θ in [−1, 1],
    φ in [0, 2π]) for cascades, (log10 d, cos θ) at E = 1e5 GeV, φ = 0 for
    point-like emitters."""
    from scipy.stats import qmc
    lo, hi = ([2.0, 1.0, -1.0, 0.0], [5.0, math.log10(150.0), 1.0, 2 * math.pi]) if extended else \
        ([1.0, -1.0], [math.log10(150.0), 1.0])
    s = qmc.Sobol(d=len(lo), scramble=False)
    pts = s.random(n_sims + n_skip)[n_skip:]
    return qmc.scale(pts, lo, hi)


def run_sims(engine, medium, spectrum, n_sims: int, n_skip: int = 0, extended: bool = True,
             n_resample: int = 100, seed: int = 0, **kw) -> list:
    """photon_tables.jl:172-230: one run_sim per Sobol point."""
    target = table_target()
    out = []
    for i, p in enumerate(sobol_parameters(n_sims, n_skip, extended)):
        if extended:
            energy, distance, ct, phi = 10 ** p[0], float(np.float32(10 ** p[1])), p[2], p[3]
        else:
            energy, distance, ct, phi = 1e5, float(np.float32(10 ** p[0])), p[1], 0.0
        r = run_sim(engine, energy, distance, ct, phi, target, spectrum, medium, n_resample, extended,
                    seed=seed + i, **kw)
        if r is not None:
            out.append(r)
    return out


def export_tables(path, runs: list) -> None:
    """Groups 'photons' and 'pmt_hits' (save_hdf!, photon_tables.jl:16-49):
    datasets 'dataset_<k>' (float64 matrices) with their attributes, group
    attribute 'nsims' = the number of datasets, as an .npz archive: keys
    '<group>/dataset_<k>', '<group>/dataset_<k>/attrs' (JSON), '<group>/nsims'."""
    arrays = {}
    for group in ("photons", "pmt_hits"):
        k = 0
        for r in runs:
            items = [r["photons"]] if group == "photons" else r["pmt_hits"]
            for data, attrs in items:
                k += 1
                arrays[f"{group}/dataset_{k}"] = np.asarray(data, np.float64)
                arrays[f"{group}/dataset_{k}/attrs"] = np.array(json.dumps(attrs))
        arrays[f"{group}/nsims"] = np.array(k)
    np.savez_compressed(path, **arrays)


def load_tables(path) -> dict:
    z = np.load(path, allow_pickle=False)
    out = {}
    for group in ("photons", "pmt_hits"):
        n = int(z[f"{group}/nsims"]) if f"{group}/nsims" in z else 0
        out[group] = [(z[f"{group}/dataset_{k}"], json.loads(str(z[f"{group}/dataset_{k}/attrs"])))
                      for k in range(1, n + 1)]
    return out
