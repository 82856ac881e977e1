"""Flow surrogate pipeline: photon tables generated by the propagation
engine, training rows, GPU training, event likelihood with a direct-
propagation fallback near modules, and the surrogate's error against direct
propagation.

Reference (paths relative to the reference repository root):
  table generation        scripts/photon_tables.jl:52-230 (run_sim, run_sims)
  flow inputs / labels    src/surrogate_models/extended_cascade_model.jl:298-478
                          (preproc_labels, calc_flow_inputs, create_pmt_table)
  training                extended_cascade_model.jl:198-295
  flow vs direct          scripts/fisher_extended_casc_model.jl:41-83

Conditioning of one (cascade, module, PMT) pair — five numbers that fix the
PMT's response in a homogeneous medium exactly (the module sits in the
cascade's photon field; its PMT caps are symmetric about their axes):
  ln d        distance cascade vertex → module centre (m)
  cos θ       cascade direction · ŝ, ŝ = unit vector module → cascade
              (the reference's frame: source on the module's +z axis,
              scripts/photon_tables.jl:68-76)
  cos β       PMT axis · ŝ
  cos φ       azimuth of the PMT axis about ŝ, measured from the plane of ŝ
              and the cascade direction (mirror symmetric, hence the cosine)
  log10 E     cascade energy (GeV)
The reference feeds [ln d, ln E, direction xyz, position xyz] + one-hot PMT
(24 inputs, extended_cascade_model.jl:328-364); the five numbers above carry
the same information for any module orientation.
"""
from __future__ import annotations

import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .detection import MultiPMTDetector, make_pom_pmt_coordinates, sph_to_cart
from .lightyield import (ExtendedCherenkovEmitter, LongitudinalParameterisation, Particle, PEMinus,
                         total_lightyield)

PMT_AREA = (75e-3 / 2) ** 2 * math.pi
MODULE_RADIUS = 0.21
WL_RANGE = (300.0, 800.0)
N_FEATURES = 5
FALLBACK_DISTANCE = 5.0        # m: pairs closer than this use direct propagation


def pom_axes() -> np.ndarray:
    """PMT axes of the 16-PMT module in its own frame (16 × 3)."""
    c = make_pom_pmt_coordinates(np.float64)
    return sph_to_cart(c[0], c[1])


def cos_opening(radius: float = MODULE_RADIUS, pmt_area: float = PMT_AREA) -> float:
    """cos of the PMT cap half-angle asin(sqrt(area/π)/r) (detection.jl:94-95)."""
    return math.cos(math.asin(math.sqrt(pmt_area / math.pi) / radius))


def random_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation matrices (n × 3 × 3) from normalised Gaussian
    quaternions (the reference draws rand(RotMatrix3), photon_tables.jl:124)."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def isotropic_directions(n: int, rng: np.random.Generator) -> np.ndarray:
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def pmt_features(src_pos, src_dir, log10e, mod_pos, axes) -> np.ndarray:
    """The five conditioning numbers (module docstring) for broadcastable
    arrays of cascade position/direction, log10 energy, module position and
    world-frame PMT axis."""
    src_pos, src_dir, mod_pos, axes = (np.asarray(a, np.float64) for a in (src_pos, src_dir, mod_pos, axes))
    r = mod_pos - src_pos
    d = np.linalg.norm(r, axis=-1)
    s = -r / d[..., None]
    cth = np.sum(src_dir * s, axis=-1)
    perp = src_dir - cth[..., None] * s
    nperp = np.linalg.norm(perp, axis=-1)
    # direction along ŝ: any perpendicular (the response is then symmetric about ŝ)
    alt = np.cross(s, np.where(np.abs(s[..., :1]) < 0.9, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]))
    e1 = np.where((nperp > 1e-9)[..., None], perp / np.maximum(nperp, 1e-300)[..., None], alt)
    e1 /= np.linalg.norm(e1, axis=-1, keepdims=True)
    e2 = np.cross(s, e1)
    cb = np.sum(axes * s, axis=-1)
    a1 = np.sum(axes * e1, axis=-1)
    a2 = np.sum(axes * e2, axis=-1)
    na = np.hypot(a1, a2)
    cphi = np.where(na > 1e-12, a1 / np.maximum(na, 1e-300), 1.0)
    cols = np.broadcast_arrays(np.log(d), cth, cb, cphi, np.asarray(log10e, np.float64))
    return np.stack(cols, axis=-1)


_LOG_YIELD_1GEV = None


def log_yield(log10e) -> np.ndarray:
    """ln of the cascade's expected photon count (∝ E, lightyield.jl:356-366)."""
    global _LOG_YIELD_1GEV
    if _LOG_YIELD_1GEV is None:
        from .medium import make_cascadia_medium_properties
        _LOG_YIELD_1GEV = math.log(cascade_photons(1.0, make_cascadia_medium_properties(0.95)))
    return _LOG_YIELD_1GEV + math.log(10.0) * np.asarray(log10e, np.float64)


@dataclass
class CountScale:
    """Reference expected count per PMT: light yield × a radial profile g(d)
    fitted to the tables (mean count per emitted photon in bins of ln d,
    interpolated in ln d).  The network's log λ output is relative to it (row
    offset) and count rows are weighted by its inverse in training, so the
    Poisson gradient (λ − n)/scale stays of order one over the ~10 decades
    the counts span, while its zero stays at λ = E[n]."""
    log_d: np.ndarray          # bin centres (ln m)
    log_g: np.ndarray          # ln mean count per photon

    @classmethod
    def fit(cls, features: np.ndarray, count: np.ndarray, n_bins: int = 24) -> "CountScale":
        ld = features[:, 0]
        ratio = count / np.exp(log_yield(features[:, 4]))
        edges = np.linspace(ld.min(), ld.max() + 1e-9, n_bins + 1)
        idx = np.clip(np.digitize(ld, edges) - 1, 0, n_bins - 1)
        s = np.bincount(idx, weights=ratio, minlength=n_bins)
        c = np.bincount(idx, minlength=n_bins)
        ok = (c > 0) & (s > 0)
        centres = 0.5 * (edges[1:] + edges[:-1])
        return cls(centres[ok], np.log(s[ok] / c[ok]))

    def __call__(self, features: np.ndarray) -> np.ndarray:
        f = np.asarray(features, np.float64)
        return log_yield(f[..., 4]) + np.interp(f[..., 0], self.log_d, self.log_g)

    def to_json(self) -> dict:
        return {"log_d": self.log_d.tolist(), "log_g": self.log_g.tolist()}


@dataclass
class FeatureNorm:
    """z-scoring of the features (the reference's Normalizer per label,
    extended_cascade_model.jl:313-326)."""
    mean: np.ndarray
    std: np.ndarray

    @classmethod
    def fit(cls, f: np.ndarray) -> "FeatureNorm":
        return cls(f.mean(axis=0), np.maximum(f.std(axis=0), 1e-6))

    def __call__(self, f: np.ndarray) -> np.ndarray:
        return ((f - self.mean) / self.std).astype(np.float32)

    def to_json(self) -> dict:
        return {"mean": self.mean.tolist(), "std": self.std.tolist()}


# ---------------------------------------------------------------------------
# photon tables
# ---------------------------------------------------------------------------
def table_detector(n_strings: int = 16, per_string: int = 12, d_range=(1.0, 300.0),
                   rho_range=(1.0, 150.0)) -> np.ndarray:
    """Module positions around a cascade at the origin: vertical strings at
    log-spaced horizontal distances (golden-angle azimuths), modules at
    log-spaced distances from the origin along each string.  Strings keep the
    engine's per-string culling effective, so one launch fills ~200 table
    entries (the reference propagates one module per launch)."""
    rhos = np.geomspace(*rho_range, n_strings)
    out = []
    for s, rho in enumerate(rhos):
        az = s * math.pi * (3.0 - math.sqrt(5.0))
        ds = np.geomspace(max(d_range[0], rho * 1.05), d_range[1], per_string)
        z = np.sqrt(np.maximum(ds ** 2 - rho ** 2, 0.0))
        z *= np.where(np.arange(per_string) % 2 == 0, 1.0, -1.0)
        # keep modules on one string apart (≥ 2.5 module radii)
        order = np.argsort(z)
        zs = z[order]
        for k in range(1, per_string):
            zs[k] = max(zs[k], zs[k - 1] + 2.5 * MODULE_RADIUS)
        z[order] = zs
        for zz in z:
            out.append((rho * math.cos(az), rho * math.sin(az), zz))
    return np.array(out)


@dataclass
class TableSet:
    """Training rows built from engine photon tables (raw features)."""
    features: np.ndarray       # [n, 5]
    t: np.ndarray              # [n] time residual (time rows)
    count: np.ndarray          # [n] expected hits (count rows)
    flags: np.ndarray          # [n] uint8 (1 time row, 2 count row)
    info: dict = field(default_factory=dict)

    def concat(self, other: "TableSet") -> "TableSet":
        return TableSet(np.concatenate([self.features, other.features]), np.concatenate([self.t, other.t]),
                        np.concatenate([self.count, other.count]), np.concatenate([self.flags, other.flags]),
                        {**self.info, **other.info})

    def aux(self, scale: "CountScale") -> np.ndarray:
        """(log λ offset, row weight) per row: count rows carry the reference
        scale and its inverse as weight; time rows (0, 1)."""
        off = np.where(self.flags & 2, scale(self.features), 0.0)
        return np.column_stack([off, np.where(self.flags & 2, np.exp(-off), 1.0)]).astype(np.float32)


def thin_hits(hits_mod: np.ndarray, n_mod: int, max_hits: int, rng, pos, tres, w):
    """About `max_hits` hits per module: each hit of a module with n > max_hits
    hits is kept with probability max_hits/n and its weight divided by that
    probability (unbiased for every weighted sum)."""
    counts = np.bincount(hits_mod, minlength=n_mod)
    if counts.max(initial=0) <= max_hits:
        return pos, tres, w, hits_mod
    p = np.minimum(1.0, max_hits / np.maximum(counts, 1))
    keep = rng.random(hits_mod.size) < p[hits_mod]
    m = hits_mod[keep]
    return pos[keep], tres[keep], np.asarray(w[keep], np.float64) / p[m], m


def rows_from_hits(hits_pos, hits_tres, hits_w, hits_mod, mod_pos, src_pos, src_dir, log10e, scale, rng,
                   orientations: int = 4, time_rows: int = 32, max_hits: int = 4000):
    """Per-PMT rows of one cascade from its weighted module hits: each module
    is turned into `orientations` random orientations of the 16-PMT layout
    (photon_tables.jl:122-169); a hit belongs to the first PMT whose cap
    contains it (detection.jl:88-104).  Count rows hold scale·Σ weights of the
    PMT (expected hits; also for PMTs without hits); time rows hold up to
    `time_rows` time residuals drawn ∝ weight from the PMT's hits
    (resample_simulation per PMT)."""
    n_mod = len(mod_pos)
    axes0 = pom_axes()
    cosop = cos_opening()
    hits_mod = np.asarray(hits_mod, np.int64)
    w = np.asarray(hits_w, np.float64)
    hits_pos, hits_tres, w, hits_mod = thin_hits(hits_mod, n_mod, max_hits, rng, hits_pos, hits_tres, w)
    rel = hits_pos - mod_pos[hits_mod]
    rel /= np.linalg.norm(rel, axis=1, keepdims=True)
    rots = random_rotations(n_mod * orientations, rng).reshape(n_mod, orientations, 3, 3)
    axes = np.einsum("mokl,pl->mopk", rots, axes0)                  # [mod, orient, pmt, 3]
    n_caps = n_mod * orientations * 16
    cap_feat = pmt_features(src_pos, src_dir, log10e, mod_pos[:, None, None, :], axes).reshape(n_caps, N_FEATURES)
    cap_w = np.zeros(n_caps)
    t_parts, f_parts = [], []
    for o in range(orientations):
        dots = np.einsum("hk,hpk->hp", rel, axes[hits_mod, o])      # [hits, 16]
        inside = dots > cosop
        has = inside.any(axis=1)
        pmt = np.argmax(inside, axis=1)
        h_idx = np.flatnonzero(has)
        cap = (hits_mod[h_idx] * orientations + o) * 16 + pmt[h_idx]
        cap_w += np.bincount(cap, weights=w[h_idx], minlength=n_caps)
        if h_idx.size == 0:
            continue
        # weighted resampling of time rows per cap (inverse CDF on grouped cumsums)
        order = np.argsort(cap, kind="stable")
        c_s, h_s = cap[order], h_idx[order]
        w_s = w[h_s]
        cs = np.cumsum(w_s)
        starts = np.r_[0, np.flatnonzero(np.diff(c_s)) + 1]
        ends = np.r_[starts[1:], c_s.size]
        n_in = ends - starts
        n_draw = np.minimum(n_in, time_rows)
        base = np.where(starts > 0, cs[starts - 1], 0.0)
        tot = cs[ends - 1] - base
        which = np.repeat(np.arange(starts.size), n_draw)
        u = base[which] + rng.random(which.size) * tot[which]
        pos = np.searchsorted(cs, u, side="right")
        pos = np.clip(pos, starts[which], ends[which] - 1)
        t_parts.append(hits_tres[h_s[pos]])
        f_parts.append(cap_feat[c_s[starts][which]])
    count_rows = (cap_feat, np.zeros(n_caps), cap_w * scale, np.full(n_caps, 2, np.uint8))
    if t_parts:
        tf = np.concatenate(f_parts)
        tt = np.concatenate(t_parts)
        time_rows_ = (tf, tt, np.zeros(tt.size), np.full(tt.size, 1, np.uint8))
    else:
        time_rows_ = (np.zeros((0, N_FEATURES)), np.zeros(0), np.zeros(0), np.zeros(0, np.uint8))
    return TableSet(*(np.concatenate([a, b]) for a, b in zip(count_rows, time_rows_)))


def make_targets(positions) -> list:
    coords = make_pom_pmt_coordinates(np.float32)
    return [MultiPMTDetector(np.asarray(p, np.float32), MODULE_RADIUS, PMT_AREA, coords, k + 1)
            for k, p in enumerate(positions)]


def cascade_photons(energy: float, medium) -> float:
    """Expected Cherenkov photons of an electron cascade, 300-800 nm
    (total_lightyield, lightyield.jl:356-366)."""
    p = Particle(np.zeros(3), np.array([0.0, 0.0, 1.0]), 0.0, energy, PEMinus)
    return total_lightyield(p, medium, WL_RANGE)


def generate_tables(engine, medium, spectrum, n_runs: int, photons_per_run: int, seed: int = 0,
                    log10e_range=(2.0, 6.0), orientations: int = 4, time_rows: int = 32,
                    max_hits: int = 4000, capacity: int = 1 << 24, export: list | None = None) -> TableSet:
    """Photon tables from the engine (photon_tables.jl:52-230, Sobol draws
    replaced by uniform ones): per run one electron cascade at the origin
    with log10 E uniform in `log10e_range` and an isotropic direction,
    propagated with `photons_per_run` photons to the table detector; hit
    weights (absorption × PMT acceptance) and time residuals come from the
    device epilogue; counts are scaled to the cascade's full light yield.
    `export` (a list) receives per-module photon tables in the reference's
    layout (see `export_reference_tables`)."""
    rng = np.random.default_rng(seed)
    positions = table_detector()
    engine.set_medium(medium)
    engine.set_spectrum(spectrum)
    engine.set_targets(make_targets(positions))
    parts = []
    t0 = time.perf_counter()
    t_prop = 0.0
    photons_total = 0
    for run in range(n_runs):
        le = rng.uniform(*log10e_range)
        e = 10.0 ** le
        u = isotropic_directions(1, rng)[0]
        lp = LongitudinalParameterisation.from_energy(e, medium, PEMinus)
        src = ExtendedCherenkovEmitter(np.zeros(3, np.float32), u.astype(np.float32), 0.0, int(photons_per_run), lp)
        engine.set_sources([src])
        engine.set_epilogue(medium, None, (0.0, 0.0, 0.0))
        tp = time.perf_counter()
        hits = engine.propagate(seed=seed * 100003 + run + 1, capacity=capacity)
        t_prop += time.perf_counter() - tp
        photons_total += hits.n_simulated
        scale = cascade_photons(e, medium) / hits.n_simulated
        # hits arrive in device append order; photon order makes the thinning,
        # the time-row draws and so the whole table a function of the seed
        order = np.argsort(hits.photon_id)
        mod = hits.module_id[order].astype(np.int64) - 1
        hpos, htres, hw = hits.position[order], hits.tres[order], hits.total_weight[order]
        pos, tres, w, mod_k = thin_hits(mod, len(positions), max_hits, rng, hpos, htres, hw)
        ts = rows_from_hits(pos.astype(np.float64), tres.astype(np.float64), w, mod_k, positions, np.zeros(3), u, le,
                            scale, rng, orientations, time_rows, max_hits)
        parts.append(ts)
        if export is not None:
            export.extend(_reference_layout(hpos, htres, hw, mod, positions, u, e, scale))
    out = TableSet(*(np.concatenate([getattr(p, k) for p in parts]) for k in ("features", "t", "count", "flags")))
    out.info = {"runs": n_runs, "photons_per_run": photons_per_run, "photons": photons_total,
                "seconds": time.perf_counter() - t0, "propagate_seconds": t_prop,
                "modules_per_run": len(positions)}
    return out


def _reference_layout(hpos, htres, hw, mod, positions, u, energy, scale):
    """Per-module 'photons' datasets of photon_tables.jl:111-120 in the
    reference's frame (module at the origin, cascade on its +z axis)."""
    from .detection import cart_to_sph
    out = []
    for m in np.unique(mod):
        sel = mod == m
        r = -positions[m]                                  # module → cascade
        d = float(np.linalg.norm(r))
        s = r / d
        # rotation taking ŝ to e_z (Rodrigues, utils.jl:173-184)
        v = np.cross(s, [0.0, 0.0, 1.0])
        c = float(s[2])
        if np.linalg.norm(v) < 1e-12:
            R = np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
        else:
            vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
            R = np.eye(3) + vx + vx @ vx * (1 - c) / (v @ v)
        pos = (hpos[sel] - positions[m]) @ R.T
        ur = R @ u
        th, ph = cart_to_sph(*ur)
        out.append({"attrs": {"energy": energy, "distance": d, "dir_costheta": float(math.cos(th)),
                              "dir_phi": float(ph)},
                    "data": np.column_stack([htres[sel], pos, hw[sel] * scale])})
    return out


def export_reference_tables(path, datasets: list) -> None:
    """Write photon tables in the reference's HDF5 layout (photon_tables.jl:
    16-49: group 'photons', datasets 'dataset_<k>' of columns [tres, pos_x,
    pos_y, pos_z, total_weight] with attributes, group attribute 'nsims') as
    an .npz archive (no HDF5 library in this environment): keys
    'photons/dataset_<k>' and 'photons/dataset_<k>/attrs' (JSON), plus
    'photons/nsims'."""
    arrays = {}
    for k, ds in enumerate(datasets, start=1):
        arrays[f"photons/dataset_{k}"] = np.asarray(ds["data"], np.float64)
        arrays[f"photons/dataset_{k}/attrs"] = np.array(json.dumps(ds["attrs"]))
    arrays["photons/nsims"] = np.array(len(datasets))
    np.savez_compressed(path, **arrays)


def load_reference_tables(path) -> list:
    z = np.load(path, allow_pickle=False)
    n = int(z["photons/nsims"])
    return [{"data": z[f"photons/dataset_{k}"], "attrs": json.loads(str(z[f"photons/dataset_{k}/attrs"]))}
            for k in range(1, n + 1)]


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------
@dataclass
class Surrogate:
    """Trained flow + its feature normaliser and count scale."""
    model: object              # flow.FlowModel
    norm: FeatureNorm
    scale: CountScale

    def rows(self, features, t=None, count=None, flags=None):
        """Evaluation rows: log λ offset from the features, weight 1."""
        from .flow import FlowRows
        aux = np.column_stack([self.scale(features), np.ones(len(features))])
        return FlowRows.from_numpy(self.norm(features), t, count, flags, aux)


def scale_only_surrogate(sur: Surrogate) -> Surrogate:
    """The method without the learned part: a network whose weights are all
    zero, so that every output is 0 and log λ of a PMT is the count scale's
    offset ln(yield(E)·g(d)) alone (the time density is the untransformed
    base normal)."""
    from .flow import FlowModel
    m = FlowModel(sur.model.device, n_in=N_FEATURES)
    m.set_weights(np.zeros_like(m.get_weights()))
    return Surrogate(m, sur.norm, sur.scale)


def train_surrogate(tables: TableSet, steps: int = 3000, batch: int = 16384, lr: float = 1e-3,
                    seed: int = 0, model=None, log_every: int = 0, holdout: float = 0.1,
                    count_weight: float = 1.0, average_every: int = 64, count_power: float = 1.0):
    """Adam on the mean row loss (train_model!, extended_cascade_model.jl:
    237-295; lr decays ×0.1 over the last third).  With `average_every` > 0
    the returned weights are the mean of snapshots taken every that many steps
    during the decayed phase (Polyak-Ruppert averaging: the jitter of the
    last Adam steps averages out instead of landing in the expected counts).
    Count rows are weighted by count_weight · λ_ref^-count_power (λ_ref the
    CountScale reference): 1 makes every row's Poisson gradient (λ − n)/λ_ref
    of order one, 0.5 scales it like the row's statistical significance
    (λ − n)/√λ_ref, so bright rows that are known precisely count for more than
    faint ones dominated by Poisson noise.
    Returns the surrogate and a history dict (train/holdout loss)."""
    import torch

    from .flow import FlowModel, FlowRows
    rng = np.random.default_rng(seed)
    n = len(tables.t)
    perm = rng.permutation(n)
    n_hold = int(n * holdout)
    hold, train = perm[:n_hold], perm[n_hold:]
    norm = FeatureNorm.fit(tables.features[train])
    cnt = train[tables.flags[train] & 2 != 0]
    scale = CountScale.fit(tables.features[cnt], tables.count[cnt])
    model = model or FlowModel(0, seed=seed, n_in=N_FEATURES)
    dev = model.dev
    aux = tables.aux(scale)
    is_cnt = (tables.flags & 2) != 0
    aux[:, 1] = np.where(is_cnt, count_weight * np.exp(-count_power * aux[:, 0].astype(np.float64)),
                         aux[:, 1]).astype(np.float32)
    rows = FlowRows.from_numpy(norm(tables.features[train]), tables.t[train], tables.count[train],
                               tables.flags[train], aux[train], device=dev)
    hold_rows = FlowRows.from_numpy(norm(tables.features[hold]), tables.t[hold], tables.count[hold],
                                    tables.flags[hold], aux[hold], device=dev)
    gen = torch.Generator(device=dev).manual_seed(seed)
    loss_buf = torch.zeros(1, dtype=torch.float64, device=dev)
    hist = {"step": [], "train": [], "holdout": []}

    def holdout_loss():
        _, l = model.forward(hold_rows, want_params=False)
        return float(l.double().mean())

    hist["initial_holdout"] = holdout_loss()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    avg, n_avg = None, 0
    for step in range(steps):
        frac = step / max(steps, 1)
        decayed = frac >= 2 / 3
        lr_s = lr * 0.1 if decayed else lr
        idx = torch.randint(0, rows.n, (min(batch, rows.n),), generator=gen, device=dev)
        model.train_step_gather(rows, idx, lr_s, loss_buf)
        if average_every and decayed and (steps - 1 - step) % average_every == 0:
            w = model.get_weights().astype(np.float64)
            avg = w if avg is None else avg + w
            n_avg += 1
        if log_every and (step % log_every == 0 or step == steps - 1):
            hist["step"].append(step)
            hist["train"].append(float(loss_buf.cpu()[0]))
            hist["holdout"].append(holdout_loss())
    torch.cuda.synchronize(dev)
    hist["seconds"] = time.perf_counter() - t0
    hist["steps"] = steps
    if n_avg > 1:
        hist["last_holdout"] = holdout_loss()
        model.set_weights((avg / n_avg).astype(np.float32))
        hist["averaged_snapshots"] = n_avg
    hist["final_holdout"] = holdout_loss()
    return Surrogate(model, norm, scale), hist


# ---------------------------------------------------------------------------
# evaluation: expected counts / densities, direct propagation, likelihood
# ---------------------------------------------------------------------------
def array_detector(n_strings: int = 10, per_string: int = 20, spacing_h: float = 60.0,
                   spacing_v: float = 20.0) -> np.ndarray:
    """10-string × 20-module array: a centre string and rings of strings at
    `spacing_h`, modules every `spacing_v` m (centred on z = 0)."""
    xy = [(0.0, 0.0)]
    ring, k = 1, 0
    while len(xy) < n_strings:
        n_ring = 6 * ring
        for j in range(n_ring):
            if len(xy) >= n_strings:
                break
            a = 2 * math.pi * j / n_ring
            xy.append((ring * spacing_h * math.cos(a), ring * spacing_h * math.sin(a)))
        ring += 1
    zs = (np.arange(per_string) - (per_string - 1) / 2) * spacing_v
    return np.array([(x, y, z) for x, y in xy for z in zs])


@dataclass
class Cascades:
    pos: np.ndarray            # [n, 3]
    dir: np.ndarray            # [n, 3]
    energy: np.ndarray         # [n] GeV


def random_cascades(n: int, rng, log10e_range=(3.0, 6.0), radius: float = 100.0, half_height: float = 200.0,
                    min_distance: float = 0.0, modules: np.ndarray | None = None) -> Cascades:
    """Vertices uniform in a cylinder, isotropic directions, log-uniform energies."""
    r = radius * np.sqrt(rng.random(n))
    a = rng.uniform(0, 2 * math.pi, n)
    pos = np.column_stack([r * np.cos(a), r * np.sin(a), rng.uniform(-half_height, half_height, n)])
    return Cascades(pos, isotropic_directions(n, rng), 10.0 ** rng.uniform(*log10e_range, n))


def pair_features(casc: Cascades, modules: np.ndarray, axes: np.ndarray | None = None) -> np.ndarray:
    """Features of every (cascade, module, PMT) pair: [n_casc, n_mod, 16, 5]."""
    axes = pom_axes() if axes is None else axes
    return pmt_features(casc.pos[:, None, None, :], casc.dir[:, None, None, :],
                        np.log10(casc.energy)[:, None, None], modules[None, :, None, :], axes[None, None, :, :])


def surrogate_expectations(sur: Surrogate, casc: Cascades, modules: np.ndarray) -> np.ndarray:
    """Expected hits per (cascade, module, PMT) from the flow's log λ head."""
    f = pair_features(casc, modules)
    shape = f.shape[:-1]
    fr = f.reshape(-1, N_FEATURES)
    rows = sur.rows(fr, flags=np.full(len(fr), 2, np.uint8))
    p, _ = sur.model.forward(rows)
    return np.exp(p[:, 77].double().cpu().numpy() + sur.scale(fr)).reshape(shape)


def surrogate_time_density(sur: Surrogate, features: np.ndarray, times: np.ndarray) -> np.ndarray:
    """pdf of each pair's time residual at `times`: [n_pairs, n_times]."""
    npairs = features.shape[0]
    f = np.repeat(features, times.size, axis=0)
    t = np.tile(times, npairs)
    rows = sur.rows(f, t, None, np.full(t.size, 1, np.uint8))
    _, loss = sur.model.forward(rows, want_params=False)
    return np.exp(-loss.double().cpu().numpy()).reshape(npairs, times.size)


def direct_propagation(engine, medium, spectrum, casc: Cascades, modules: np.ndarray, photons: int,
                       seed: int = 0, capacity: int = 1 << 23):
    """Direct propagation of each cascade (photons simulated ≤ `photons`,
    weights scaled to the full light yield): per-PMT expected hits
    [n_casc, n_mod, 16] and per-hit (cascade, module, pmt, tres, weight)."""
    engine.set_medium(medium)
    engine.set_spectrum(spectrum)
    engine.set_targets(make_targets(modules))
    n_mod = len(modules)
    lam = np.zeros((len(casc.energy), n_mod, 16))
    hit_list = []
    sim = 0
    t0 = time.perf_counter()
    for i in range(len(casc.energy)):
        e = float(casc.energy[i])
        n_real = cascade_photons(e, medium)
        n_sim = int(min(photons, n_real))
        lp = LongitudinalParameterisation.from_energy(e, medium, PEMinus)
        src = ExtendedCherenkovEmitter(casc.pos[i].astype(np.float32), casc.dir[i].astype(np.float32), 0.0, n_sim, lp)
        engine.set_sources([src])
        engine.set_epilogue(medium, None, casc.pos[i])
        hits = engine.propagate(seed=seed * 7919 + i + 1, capacity=capacity)
        sim += hits.n_simulated
        scale = n_real / hits.n_simulated
        m = hits.module_id.astype(np.int64) - 1
        k = hits.pmt_id.astype(np.int64) - 1
        ok = k >= 0
        np.add.at(lam[i], (m[ok], k[ok]), hits.total_weight[ok] * scale)
        hit_list.append(np.column_stack([np.full(ok.sum(), i), m[ok], k[ok], hits.tres[ok],
                                         hits.total_weight[ok] * scale]))
    return lam, np.concatenate(hit_list) if hit_list else np.zeros((0, 5)), {
        "photons": sim, "seconds": time.perf_counter() - t0}


def module_errors(lam_sur: np.ndarray, lam_dir: np.ndarray, n_sim_hits: np.ndarray, min_hits: int = 400,
                  dir_var: np.ndarray | None = None):
    """Per-module expected-count error of the surrogate against direct
    propagation (Σ over the module's PMTs), over cascade-module pairs whose
    direct estimate rests on ≥ min_hits simulated hits (≤ 5 % statistical
    error).  With `dir_var` (the direct estimate's variance per pair, Σw² of
    its weighted hits) also the direct estimate's median relative statistical
    error and the surrogate's own RMS error with that variance taken out: for
    δ = (s − d)/s and a truth d = μ(1 + η) independent of the surrogate
    s = μ(1 + ε), E[δ²] − E[Var d / s²] = E[ε²/(1 + ε)²] ≈ ε_rms² (dividing by
    s rather than by the noisy d keeps the subtraction unbiased)."""
    s = lam_sur.sum(-1)
    d = lam_dir.sum(-1)
    sel = n_sim_hits >= min_hits
    rel = (s[sel] - d[sel]) / d[sel]
    extra = {}
    if dir_var is not None and rel.size:
        v, ss, dd = dir_var[sel], s[sel], d[sel]
        delta = (ss - dd) / ss
        extra = {"truth_rel_stat_err_weighted_median": float(np.median(np.sqrt(v) / dd)),
                 "rms_rel_err": float(np.sqrt(np.mean(rel ** 2))),
                 "rms_rel_err_intrinsic": float(np.sqrt(max(0.0, np.mean(delta ** 2) - np.mean(v / ss ** 2))))}
    return {**extra, "pairs": int(sel.sum()), "median_abs_rel_err": float(np.median(np.abs(rel))) if rel.size else None,
            "mean_abs_rel_err": float(np.mean(np.abs(rel))) if rel.size else None,
            "bias": float(np.sum(s[sel]) / np.sum(d[sel]) - 1) if rel.size else None,
            "rel_err": rel}


def time_kl(sur: Surrogate, casc: Cascades, modules: np.ndarray, lam_sur: np.ndarray, hits: np.ndarray,
            pairs: list, bins: np.ndarray) -> np.ndarray:
    """KL(direct ‖ surrogate) of each listed (cascade, module) pair's time
    residual distribution (PMTs pooled, weighted by expected hits), on
    `bins`."""
    centres = 0.5 * (bins[1:] + bins[:-1])
    width = np.diff(bins)
    out = []
    axes = pom_axes()
    for (i, m) in pairs:
        sel = (hits[:, 0] == i) & (hits[:, 1] == m)
        h, _ = np.histogram(hits[sel, 3], bins=bins, weights=hits[sel, 4])
        f = pmt_features(casc.pos[i], casc.dir[i], math.log10(casc.energy[i]), modules[m], axes)
        dens = surrogate_time_density(sur, f, centres)                  # [16, bins]
        w = lam_sur[i, m] / lam_sur[i, m].sum()
        q = (w[:, None] * dens).sum(0) * width
        p = h / h.sum()
        q = q / q.sum()
        ok = p > 0
        out.append(float(np.sum(p[ok] * np.log(p[ok] / np.maximum(q[ok], 1e-300)))))
    return np.array(out)


# ---------------------------------------------------------------------------
# event likelihood
# ---------------------------------------------------------------------------
@dataclass
class Observation:
    """Observed hits of a set of events: per-PMT counts [n_ev, n_mod, 16] and
    hit list (event, module, pmt, time residual)."""
    counts: np.ndarray
    hits: np.ndarray           # [n_hits, 4]


def observe(lam: np.ndarray, direct_hits: np.ndarray, rng) -> Observation:
    """Poisson counts around `lam`; each observed hit's time drawn ∝ weight
    from the direct-propagation hits of its PMT."""
    counts = rng.poisson(lam)
    key = (direct_hits[:, 0].astype(np.int64) * lam.shape[1] + direct_hits[:, 1].astype(np.int64)) * 16 \
        + direct_hits[:, 2].astype(np.int64)
    order = np.argsort(key, kind="stable")
    ks, w = key[order], direct_hits[order, 4]
    cs = np.cumsum(w)
    uniq, start = np.unique(ks, return_index=True)
    end = np.r_[start[1:], ks.size]
    flat = counts.reshape(-1)
    want = flat[uniq]
    which = np.repeat(np.arange(uniq.size), want)
    base = np.where(start > 0, cs[np.maximum(start - 1, 0)], 0.0)
    u = base[which] + rng.random(which.size) * (cs[end - 1] - base)[which]
    pos = np.clip(np.searchsorted(cs, u, side="right"), start[which], end[which] - 1)
    k = uniq[which]
    ev, rem = np.divmod(k, lam.shape[1] * 16)
    m, p = np.divmod(rem, 16)
    counts_fixed = np.zeros_like(flat)
    np.add.at(counts_fixed, k, 1)            # PMTs without simulated hits observe none
    return Observation(counts_fixed.reshape(lam.shape),
                       np.column_stack([ev, m, p, direct_hits[order[pos], 3]]))


def _bin_index(bins: np.ndarray, x: np.ndarray) -> np.ndarray:
    """np.searchsorted(bins, x, side="right") − 1 clipped to [0, len(bins) − 2];
    for evenly spaced bins by arithmetic (one floor per value, then one step
    of correction at the edges, so the result equals the search exactly)."""
    nb = bins.size - 1
    d = np.diff(bins)
    if nb > 0 and np.all(d == d[0]):
        i = np.clip(np.floor((x - bins[0]) / d[0]), -1, nb).astype(np.int64)
        i = np.where((i + 1 <= nb) & (bins[np.clip(i + 1, 0, nb)] <= x), i + 1, i)
        i = np.where((i >= 0) & (bins[np.clip(i, 0, nb)] > x), i - 1, i)
        return np.clip(i, 0, nb - 1)
    return np.clip(np.searchsorted(bins, x, side="right") - 1, 0, nb - 1)


def fallback_terms(engine, medium, spectrum, casc: Cascades, modules: np.ndarray, counts: np.ndarray,
                   near: np.ndarray, near_hits: np.ndarray, photons: int = 10 ** 7, seed: int = 0,
                   bins=np.arange(-20.0, 2001.0, 2.0)) -> np.ndarray:
    """Log-likelihood terms of the (cascade, module) pairs in `near` from
    direct propagation, all such cascades in one launch (one source each,
    photons ≤ `photons`, weights scaled to the full light yield): Poisson
    counts with λ = Σ weights per PMT and a histogram density of the time
    residual per PMT (floor 1e-6/ns).  `counts` are the observed per-PMT
    counts [n_ev, n_mod, 16], `near_hits` the observed hits in near pairs
    (pair key, event, time).  Returns the terms per event."""
    from .medium import refractive_index
    from .propagation import C_VAC_M_NS
    n_ev = len(casc.energy)
    out = np.zeros(n_ev)
    evs = np.flatnonzero(near.any(axis=1))
    if evs.size == 0:
        return out
    mods = np.flatnonzero(near[evs].any(axis=0))
    srcs, scales = [], []
    for i in evs:
        e = float(casc.energy[i])
        n_real = cascade_photons(e, medium)
        n_sim = int(min(photons, n_real))
        lp = LongitudinalParameterisation.from_energy(e, medium, PEMinus)
        srcs.append(ExtendedCherenkovEmitter(casc.pos[i].astype(np.float32), casc.dir[i].astype(np.float32), 0.0,
                                             n_sim, lp))
        scales.append(n_real / n_sim)
    engine.set_targets(make_targets(modules[mods]))
    engine.set_sources(srcs)
    engine.set_epilogue(medium, None, casc.pos[evs[0]])
    # host hit columns sized for the usual case; an overflow repeats the
    # launch with room for all of its hits
    h = engine.propagate(seed=seed, capacity=1 << 18, raise_on_overflow=False)
    if h.n_dropped:
        h = engine.propagate(seed=seed, capacity=len(h) + h.n_dropped)
    si = h.source_index.astype(np.int64)
    ev = evs[si]
    m = mods[h.module_id.astype(np.int64) - 1]
    k = h.pmt_id.astype(np.int64) - 1
    keep = (k >= 0) & near[ev, m]
    ev, m, k, si = ev[keep], m[keep], k[keep], si[keep]
    w = h.total_weight[keep].astype(np.float64) * np.asarray(scales)[si]
    n800 = float(refractive_index(np.float32(800.0), medium))
    dist = np.linalg.norm(casc.pos[ev] - modules[m], axis=1)
    tres = h.time[keep].astype(np.float64) - (dist - MODULE_RADIUS) / (C_VAC_M_NS / n800)
    n_mod = len(modules)
    pairs = np.argwhere(near)
    all_keys = ((pairs[:, 0] * n_mod + pairs[:, 1])[:, None] * 16 + np.arange(16)).reshape(-1)
    # (event, module) → index of the pair among the near pairs, then × 16 + PMT
    pair_of = np.full(near.shape, -1, np.int64)
    pair_of[pairs[:, 0], pairs[:, 1]] = np.arange(len(pairs))
    lam = np.zeros(all_keys.size)
    pos = pair_of[ev, m] * 16 + k
    np.add.at(lam, pos, w)
    n_obs = counts.reshape(-1)[all_keys]
    with np.errstate(divide="ignore"):
        pois = n_obs * np.log(np.where(lam > 0, lam, 1.0)) - lam - np.vectorize(math.lgamma)(n_obs + 1.0)
    pois = np.where((lam == 0) & (n_obs > 0), -1e30, pois)
    np.add.at(out, all_keys // (n_mod * 16), pois)
    # time densities: per-PMT weighted histograms, evaluated at the observed
    # times only (sparse: one entry per occupied (PMT, bin))
    if len(near_hits):
        nb = bins.size - 1
        hb = _bin_index(bins, tres)
        cell = pos * nb + hb
        cells, inv = np.unique(cell, return_inverse=True)
        cell_w = np.bincount(inv, weights=w)
        ob = _bin_index(bins, near_hits[:, 2])
        hk = near_hits[:, 0].astype(np.int64)
        op = pair_of[hk // (16 * n_mod), (hk // 16) % n_mod] * 16 + hk % 16
        oc = op * nb + ob
        if cells.size:
            j = np.clip(np.searchsorted(cells, oc), 0, cells.size - 1)
            w_cell = np.where(cells[j] == oc, cell_w[j], 0.0)
        else:
            w_cell = np.zeros(oc.size)
        w_pmt = lam[op]
        dens = w_cell / np.maximum(w_pmt, 1e-300) / np.diff(bins)[ob] + 1e-6
        np.add.at(out, near_hits[:, 1].astype(np.int64), np.log(dens))
    return out


class NearFallback:
    """fallback_terms on the device, split in two so that the surrogate part
    of the likelihood runs while the direct propagation does.  `launch`
    sets the near modules as targets and one source per event with near
    pairs (photons ≤ `photons`, weights scaled to the full light yield) and
    enqueues the propagation with the hit epilogue (nt_launch) into device
    columns on `stream`; `finish` waits for the hit count and reduces on the
    same stream: Poisson terms with λ = Σ weights per PMT and histogram time
    densities per PMT (floor 1e-6/ns) at the observed times, summed per
    event.  The terms equal fallback_terms' for the same seed (the same
    photons; sums in another order).  The observed side (counts, near hits,
    bins) is staged in HBM once, here."""

    def __init__(self, engine, medium, casc: Cascades, modules: np.ndarray, ev: "DeviceEvents",
                 photons: int = 10 ** 7, bins=np.arange(-20.0, 2001.0, 2.0), capacity: int = 1 << 20):
        import torch

        from .medium import refractive_index
        from .propagation import C_VAC_M_NS, DeviceHitColumns
        self.engine, self.medium, self.modules = engine, medium, np.asarray(modules, np.float64)
        dev = torch.device(f"cuda:{engine.device}")
        self.dev = dev
        self.n_ev = len(casc.energy)
        near = ev.near
        n_mod = len(modules)
        self.evs = np.flatnonzero(near.any(axis=1))
        self.mods = np.flatnonzero(near[self.evs].any(axis=0)) if self.evs.size else np.zeros(0, np.int64)
        self.srcs, scales = [], []
        for i in self.evs:
            e = float(casc.energy[i])
            n_real = cascade_photons(e, medium)
            n_sim = int(min(photons, n_real))
            lp = LongitudinalParameterisation.from_energy(e, medium, PEMinus)
            self.srcs.append(ExtendedCherenkovEmitter(casc.pos[i].astype(np.float32), casc.dir[i].astype(np.float32),
                                                      0.0, n_sim, lp))
            scales.append(n_real / n_sim)
        self.src_pos = casc.pos[self.evs[0]] if self.evs.size else None
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        pairs = np.argwhere(near)
        pair_of = np.full(near.shape, -1, np.int64)
        pair_of[pairs[:, 0], pairs[:, 1]] = np.arange(len(pairs))
        # per launch source s / target t: the event and module they stand for
        self.ev_of_src = T(self.evs, torch.int64)
        self.mod_of_tgt = T(self.mods, torch.int64)
        self.scale = T(np.asarray(scales, np.float64), torch.float64)
        self.pair_of = T(pair_of, torch.int64)
        self.casc_pos = T(casc.pos, torch.float64)
        self.mod_pos = T(self.modules, torch.float64)
        n800 = float(refractive_index(np.float32(800.0), medium))
        self.inv_c = n800 / C_VAC_M_NS
        all_keys = ((pairs[:, 0] * n_mod + pairs[:, 1])[:, None] * 16 + np.arange(16)).reshape(-1)
        self.n_pmt = all_keys.size
        self.n_obs = ev.counts[T(all_keys, torch.int64)].double()          # observed counts of the near PMTs
        self.pmt_event = T(all_keys // (n_mod * 16), torch.int64)
        self.bins = T(np.asarray(bins, np.float64), torch.float64)
        self.nb = len(bins) - 1
        nh = ev.near_hits
        if len(nh):
            hk = nh[:, 0].astype(np.int64)
            op = pair_of[hk // (16 * n_mod), (hk // 16) % n_mod] * 16 + hk % 16
            ob = _bin_index(np.asarray(bins, np.float64), nh[:, 2])
            self.obs_cell = T(op * self.nb + ob, torch.int64)
            self.obs_pmt = T(op, torch.int64)
            self.obs_event = T(nh[:, 1].astype(np.int64), torch.int64)
            self.obs_width = T(np.diff(bins)[ob], torch.float64)
        else:
            self.obs_cell = None
        self.cols = DeviceHitColumns(capacity, engine.device, epilogue=True)
        self.stream = torch.cuda.Stream(dev)              # default priority
        self.seed = 0

    def _setup_engine(self):
        self.engine.set_targets(make_targets(self.modules[self.mods]))
        self.engine.set_sources(self.srcs)
        self.engine.set_epilogue(self.medium, None, self.src_pos)

    def launch(self, seed: int, stream) -> None:
        """Sets the engine up and enqueues the propagation on `stream`."""
        self.seed = seed
        if not self.evs.size:
            return
        self._setup_engine()
        self.engine.launch(seed, self.cols, stream)

    def finish(self, stream):
        """Per-event terms (device float64 [n_ev]), computed on `stream`."""
        import torch

        from .propagation import DeviceHitColumns
        with torch.cuda.stream(stream):
            out = torch.zeros(self.n_ev, dtype=torch.float64, device=self.dev)
            if not self.evs.size:
                return out
            n = int(self.cols.counters[0].item())
            if n > self.cols.capacity:              # room for every hit, then the same launch again
                self.cols = DeviceHitColumns(n, self.engine.device, epilogue=True)
                self.engine.launch(self.seed, self.cols, stream)
                n = int(self.cols.counters[0].item())
            c = self.cols
            si = c.source_index[:n].long()
            ev = self.ev_of_src[si]
            m = self.mod_of_tgt[(c.module_id[:n].long() & 0xFFFF) - 1]
            k = c.pmt_id[:n].long() - 1
            p = self.pair_of[ev, m]
            keep = (k >= 0) & (p >= 0)
            ev, m, k, si, p = ev[keep], m[keep], k[keep], si[keep], p[keep]
            w = c.total_weight[:n][keep].double() * self.scale[si]
            dist = torch.linalg.vector_norm(self.casc_pos[ev] - self.mod_pos[m], dim=1)
            tres = c.time[:n][keep].double() - (dist - MODULE_RADIUS) * self.inv_c
            pos = p * 16 + k
            lam = torch.zeros(self.n_pmt, dtype=torch.float64, device=self.dev).index_add_(0, pos, w)
            pois = self.n_obs * torch.log(torch.where(lam > 0, lam, torch.ones_like(lam))) - lam \
                - torch.lgamma(self.n_obs + 1.0)
            pois = torch.where((lam == 0) & (self.n_obs > 0), torch.full_like(lam, -1e30), pois)
            out.index_add_(0, self.pmt_event, pois)
            if self.obs_cell is not None:
                hb = (torch.bucketize(tres, self.bins, right=True) - 1).clamp_(0, self.nb - 1)
                cells, inv = torch.unique(pos * self.nb + hb, return_inverse=True)
                cell_w = torch.zeros(cells.numel(), dtype=torch.float64, device=self.dev).index_add_(0, inv, w)
                if cells.numel():
                    j = torch.searchsorted(cells, self.obs_cell).clamp_(0, cells.numel() - 1)
                    w_cell = torch.where(cells[j] == self.obs_cell, cell_w[j], torch.zeros_like(cell_w[j]))
                else:
                    w_cell = torch.zeros(self.obs_cell.numel(), dtype=torch.float64, device=self.dev)
                dens = w_cell / torch.clamp(lam[self.obs_pmt], min=1e-300) / self.obs_width + 1e-6
                out.index_add_(0, self.obs_event, torch.log(dens))
        return out


def event_llh(sur: Surrogate, ev: "DeviceEvents", fallback: NearFallback | None, seed: int = 0):
    """The event log-likelihood (device float64 [n_ev]): the surrogate for
    the pairs beyond FALLBACK_DISTANCE on the current stream while the near
    pairs' direct propagation and its reduction run on the fallback's own
    stream, joined at the end."""
    import torch
    main = torch.cuda.current_stream(ev.sources.device)
    if fallback is None or not fallback.evs.size:
        return surrogate_llh(sur, ev)
    side = fallback.stream
    side.wait_stream(main)
    fallback.launch(seed, side)
    llh = surrogate_llh(sur, ev)
    fb = fallback.finish(side)
    main.wait_stream(side)
    fb.record_stream(main)
    return llh + fb


def direct_likelihood(engine, medium, spectrum, casc: Cascades, modules: np.ndarray, obs: Observation,
                      photons: int, seed: int = 0) -> np.ndarray:
    """The event log-likelihood from direct propagation alone: every
    (cascade, module) pair takes the terms fallback_terms computes (Poisson
    count per PMT from the summed weights, histogram time density per PMT),
    all cascades in one launch of ≤ `photons` photons each."""
    n_mod = len(modules)
    allp = np.ones((len(casc.energy), n_mod), bool)
    hp = ((obs.hits[:, 0].astype(np.int64) * n_mod + obs.hits[:, 1].astype(np.int64)) * 16
          + obs.hits[:, 2].astype(np.int64))
    hits = np.column_stack([hp, obs.hits[:, 0].astype(np.int64), obs.hits[:, 3]])
    return fallback_terms(engine, medium, spectrum, casc, modules, obs.counts, allp, hits, photons=photons,
                          seed=seed)


@dataclass
class DeviceEvents:
    """Events and their observed hits resident in HBM for the likelihood."""
    n_ev: int
    n_mod: int
    sources: object            # [n_ev, 8] float32
    modules: object            # [n_mod, 3] float32
    axes: object               # [16, 3] float32
    counts: object             # [n_pairs] float32
    flags: object              # [n_pairs] uint8
    zeros: object              # [n_pairs] float32
    hit_pair: object           # [n_hits] int32 (grouped by event)
    hit_t: object              # [n_hits] float32
    pair_segments: object      # flow.Segments: pairs of each event
    hit_segments: object       # flow.Segments: hits of each event
    pair_near: object          # [n_pairs] uint8 (pairs within the fallback distance)
    hit_near: object           # [n_hits] uint8
    near: np.ndarray           # [n_ev, n_mod] host
    near_hits: np.ndarray      # [k, 3] host: (pair key, event, time) of hits in near pairs


def to_device(casc: Cascades, modules: np.ndarray, obs: Observation, device: str = "cuda:0") -> DeviceEvents:
    import torch

    from .flow import Segments
    dev = torch.device(device)
    n_ev, n_mod = len(casc.energy), len(modules)
    src = np.zeros((n_ev, 8), np.float32)
    src[:, :3], src[:, 3:6], src[:, 6] = casc.pos, casc.dir, np.log10(casc.energy)
    near = np.linalg.norm(modules[None, :, :] - casc.pos[:, None, :], axis=-1) < FALLBACK_DISTANCE
    n_pairs = n_ev * n_mod * 16
    hp = ((obs.hits[:, 0].astype(np.int64) * n_mod + obs.hits[:, 1].astype(np.int64)) * 16
          + obs.hits[:, 2].astype(np.int64))
    order = np.argsort(hp, kind="stable")                  # group hits by event (pair order)
    hp = hp[order]
    T = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    hit_off = np.searchsorted(hp // (n_mod * 16), np.arange(n_ev + 1))
    hit_near = near.reshape(-1)[hp // 16]
    th = obs.hits[order, 3]
    near_hits = np.column_stack([hp[hit_near], hp[hit_near] // (n_mod * 16), th[hit_near]])
    return DeviceEvents(n_ev, n_mod, T(src), T(modules, np.float32), T(pom_axes(), np.float32),
                        T(obs.counts.reshape(-1), np.float32), torch.full((n_pairs,), 2, dtype=torch.uint8, device=dev),
                        torch.zeros(n_pairs, dtype=torch.float32, device=dev), T(hp, np.int32),
                        T(th, np.float32), Segments.from_offsets(np.arange(n_ev + 1) * n_mod * 16, device=dev),
                        Segments.from_offsets(hit_off, device=dev), T(np.repeat(near.reshape(-1), 16), np.uint8),
                        T(hit_near, np.uint8), near, near_hits)


def surrogate_llh(sur: Surrogate, ev: DeviceEvents, fallback: np.ndarray | None = None):
    """Log-likelihood of every event (device tensor [n_ev], float64): rows of
    all (cascade, module, PMT) pairs built on the device, one forward pass of
    the network over the pairs (Poisson term of each PMT's observed count),
    the time logpdf of every observed hit from its pair's flow parameters;
    pairs within FALLBACK_DISTANCE are masked and their direct-propagation
    terms (`fallback`, per event, from fallback_terms) added."""
    import torch

    from .flow import FlowRows
    m = sur.model
    x, aux = m.pair_rows(ev.sources, ev.modules, ev.axes, sur.norm.mean, sur.norm.std,
                         np.column_stack([sur.scale.log_d, sur.scale.log_g]), float(log_yield(0.0)))
    rows = FlowRows(x, ev.zeros, ev.counts, ev.flags, aux)
    params, loss_c = m.forward(rows)
    lp = m.time_logpdf(params, ev.hit_pair, ev.hit_t)
    llh = torch.empty(ev.n_ev, dtype=torch.float64, device=loss_c.device)
    m.segment_sum(loss_c, ev.pair_segments, llh, -1.0, ev.pair_near)
    m.segment_sum(lp, ev.hit_segments, llh, 1.0, ev.hit_near, accumulate=True)
    if fallback is not None:
        llh += torch.from_numpy(fallback).to(llh.device)
    return llh


def direct_rate(engine, medium, spectrum, modules, energy: float = 3e4, photons: int = 10 ** 9, seed: int = 5):
    """Photons/s of direct propagation of one cascade into the array (host
    hit columns included, as the reference's propagate_photons returns them)."""
    engine.set_medium(medium)
    engine.set_spectrum(spectrum)
    engine.set_targets(make_targets(modules))
    lp = LongitudinalParameterisation.from_energy(energy, medium, PEMinus)
    src = ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([0, 0, 1.0], np.float32), 0.0, photons, lp)
    engine.set_sources([src])
    engine.set_epilogue(medium, None, (0.0, 0.0, 0.0))
    engine.propagate(seed=seed, capacity=1 << 23)        # warm-up
    t0 = time.perf_counter()
    h = engine.propagate(seed=seed + 1, capacity=1 << 23)
    return h.n_simulated / (time.perf_counter() - t0)


def surrogate_study(n_runs: int = 128, photons_per_run: int = 2 * 10 ** 8, train_steps: int = 192000,
                    n_events: int = 1024, n_error_events: int = 128, error_photons: int = 2 * 10 ** 8,
                    data_photons: int = 10 ** 7, seed: int = 0, device: int = 0, log=None, lr: float = 1e-3,
                    count_weight: float = 2.0, count_power: float = 1.0, batch: int = 16384,
                    orientations: int = 4) -> dict:
    """Tables → training → error against direct propagation → event
    likelihood throughput against direct propagation."""
    from .flow import FlowModel
    from .medium import make_cascadia_medium_properties
    from .propagation import PropagationEngine
    from .spectrum import CherenkovSpectrum
    say = log or (lambda *a: None)
    med = make_cascadia_medium_properties(0.95)
    spec = CherenkovSpectrum(WL_RANGE, 30, med)
    eng = PropagationEngine(device)
    tables = generate_tables(eng, med, spec, n_runs, photons_per_run, seed, orientations=orientations)
    say(f"tables: {len(tables.t)} rows from {tables.info['photons']:.3g} photons in {tables.info['seconds']:.1f} s"
        f" ({tables.info['propagate_seconds']:.1f} s propagation)")
    sur, hist = train_surrogate(tables, steps=train_steps, seed=seed, lr=lr, count_weight=count_weight,
                                count_power=count_power, batch=batch)
    say(f"training: {train_steps} steps in {hist['seconds']:.1f} s, holdout loss {hist['initial_holdout']:.3f}"
        f" -> {hist['final_holdout']:.3f}")
    modules = array_detector()
    rng = np.random.default_rng(seed + 1)
    casc = random_cascades(n_events, rng)
    # --- error against high-statistics direct propagation
    ne = min(n_error_events, n_events)
    ev = Cascades(casc.pos[:ne], casc.dir[:ne], casc.energy[:ne])
    # selection on one direct run, truth from an independent one (no selection bias)
    _, hits_sel, _ = direct_propagation(eng, med, spec, ev, modules, error_photons, seed + 12)
    lam_dir, hits_dir, dinfo = direct_propagation(eng, med, spec, ev, modules, 3 * error_photons, seed + 2)
    nh = np.zeros((ne, len(modules)))
    np.add.at(nh, (hits_sel[:, 0].astype(int), hits_sel[:, 1].astype(int)), 1)
    nh_truth = np.zeros((ne, len(modules)))
    np.add.at(nh_truth, (hits_dir[:, 0].astype(int), hits_dir[:, 1].astype(int)), 1)
    # the truth's variance per pair: Σw² of its weighted hits
    truth_var = np.zeros((ne, len(modules)))
    np.add.at(truth_var, (hits_dir[:, 0].astype(int), hits_dir[:, 1].astype(int)), hits_dir[:, 4] ** 2)
    dist = np.linalg.norm(modules[None] - ev.pos[:, None], axis=-1)
    nh_far = np.where(dist >= FALLBACK_DISTANCE, nh, 0)
    lam_sur = surrogate_expectations(sur, ev, modules)
    err = module_errors(lam_sur, lam_dir, nh_far, dir_var=truth_var)
    err50 = module_errors(lam_sur, lam_dir, nh_far, min_hits=50, dir_var=truth_var)
    base = scale_only_surrogate(sur)
    lam_base = surrogate_expectations(base, ev, modules)
    base.model.close()
    err_base = module_errors(lam_base, lam_dir, nh_far)
    err50_base = module_errors(lam_base, lam_dir, nh_far, min_hits=50)
    sel_pairs = nh_far >= 400
    truth_stat = float(np.median(1.0 / np.sqrt(np.maximum(nh_truth[sel_pairs], 1)))) if sel_pairs.any() else None
    sel50 = nh_far >= 50
    truth_stat50 = float(np.median(1.0 / np.sqrt(np.maximum(nh_truth[sel50], 1)))) if sel50.any() else None
    untrained = Surrogate(FlowModel(device, seed=seed + 99, n_in=N_FEATURES), sur.norm, sur.scale)
    err0 = module_errors(surrogate_expectations(untrained, ev, modules), lam_dir, nh_far)
    sel = np.argwhere(nh_far >= 400)
    pick = sel[np.random.default_rng(seed + 3).permutation(len(sel))[:64]]
    bins = np.arange(-20.0, 1000.0, 4.0)
    kl = time_kl(sur, ev, modules, lam_sur, hits_dir, [tuple(p) for p in pick], bins)
    kl0 = time_kl(untrained, ev, modules, lam_sur, hits_dir, [tuple(p) for p in pick], bins)
    untrained.model.close()
    say(f"error: {err['pairs']} module pairs, median |rel| {err['median_abs_rel_err']}, bias {err['bias']};"
        f" untrained {err0['median_abs_rel_err']}; KL median {np.median(kl) if kl.size else None}")
    # --- likelihood throughput (observed data resident in HBM before timing)
    import torch
    lam_data, hits_data, _ = direct_propagation(eng, med, spec, casc, modules, data_photons, seed + 4)
    obs = observe(lam_data, hits_data, np.random.default_rng(seed + 5))
    dev_ev = to_device(casc, modules, obs, str(sur.model.dev))
    # the whole likelihood (event_llh: the surrogate and the near pairs'
    # direct propagation on two streams), and each part alone; medians of 5
    nf = NearFallback(eng, med, casc, modules, dev_ev)
    main = torch.cuda.current_stream(sur.model.dev)

    def timed(fn):
        ts = []
        for rep in range(7):
            torch.cuda.synchronize(sur.model.dev)
            t1 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize(sur.model.dev)
            if rep >= 2:                        # two warm-up calls
                ts.append(time.perf_counter() - t1)
        return float(np.median(ts)), r

    def fallback_only():
        nf.launch(seed + 6, main)
        return nf.finish(main)

    t_all, llh = timed(lambda: event_llh(sur, dev_ev, nf, seed=seed + 6))
    t_flow, _ = timed(lambda: surrogate_llh(sur, dev_ev))
    t_fb, _ = timed(fallback_only)
    llh = llh.cpu().numpy()
    linfo = {"seconds": t_all, "flow_seconds": t_flow, "fallback_seconds": t_fb,
             "rows": int(dev_ev.counts.shape[0]), "hits": int(dev_ev.hit_t.shape[0]),
             "fallback_pairs": int(dev_ev.near.sum())}
    # direct likelihood at matching statistics: error_photons per cascade is
    # what gives the compared pairs >= 400 hits (<= 5 % statistical error, the
    # surrogate's count error); measured on a subset of the same events
    n_direct = min(32, n_events)
    sub = Cascades(casc.pos[:n_direct], casc.dir[:n_direct], casc.energy[:n_direct])
    hsub = obs.hits[obs.hits[:, 0] < n_direct]
    obs_sub = Observation(obs.counts[:n_direct], hsub)
    direct_likelihood(eng, med, spec, Cascades(sub.pos[:2], sub.dir[:2], sub.energy[:2]), modules,
                      Observation(obs.counts[:2], hsub[hsub[:, 0] < 2]), min(error_photons, 10 ** 7),
                      seed=seed + 7)                                  # warm-up
    t0 = time.perf_counter()
    llh_direct = direct_likelihood(eng, med, spec, sub, modules, obs_sub, error_photons, seed=seed + 8)
    direct_llh_s = time.perf_counter() - t0
    direct_events_per_s = n_direct / direct_llh_s
    rate = direct_rate(eng, med, spec, modules)
    out = {
        "config": "tables_log10E_2-6_d_1-300m; RQ flow 8 bins x 3 layers, MLP 5-128-128-78, bf16 MFMA; "
                  "likelihood: 10 strings x 20 modules x 16 PMTs, cascades log-uniform 1e3-1e6 GeV",
        "table_rows": int(len(tables.t)), "table_photons": int(tables.info["photons"]),
        "table_seconds": tables.info["seconds"], "table_propagate_seconds": tables.info["propagate_seconds"],
        "train_steps": train_steps, "train_seconds": hist["seconds"],
        "train_step_ms": hist["seconds"] / max(train_steps, 1) * 1e3,
        "holdout_loss_initial": hist["initial_holdout"], "holdout_loss_final": hist["final_holdout"],
        "holdout_loss_last_step": hist.get("last_holdout"),
        "error_pairs": err["pairs"], "count_err_median": err["median_abs_rel_err"],
        "count_err_mean": err["mean_abs_rel_err"], "count_bias": err["bias"],
        "error_pairs_50hits": err50["pairs"], "count_err_median_50hits": err50["median_abs_rel_err"],
        "count_err_mean_50hits": err50["mean_abs_rel_err"], "count_bias_50hits": err50["bias"],
        "count_err_median_untrained": err0["median_abs_rel_err"],
        "count_err_median_scale_only": err_base["median_abs_rel_err"],
        "count_err_median_50hits_scale_only": err50_base["median_abs_rel_err"],
        "truth_rel_stat_err_median": truth_stat, "truth_rel_stat_err_median_50hits": truth_stat50,
        "truth_rel_stat_err_weighted_median": err.get("truth_rel_stat_err_weighted_median"),
        "truth_rel_stat_err_weighted_median_50hits": err50.get("truth_rel_stat_err_weighted_median"),
        "count_err_rms": err.get("rms_rel_err"), "count_err_rms_intrinsic": err.get("rms_rel_err_intrinsic"),
        "count_err_rms_50hits": err50.get("rms_rel_err"),
        "count_err_rms_intrinsic_50hits": err50.get("rms_rel_err_intrinsic"),
        "time_kl_median": float(np.median(kl)) if kl.size else None,
        "time_kl_median_untrained": float(np.median(kl0)) if kl0.size else None,
        "events": n_events, "llh_seconds": linfo["seconds"], "llh_flow_seconds": linfo["flow_seconds"],
        "llh_fallback_seconds": linfo["fallback_seconds"],
        "llh_pairs": linfo["rows"], "llh_hits": linfo["hits"], "fallback_pairs": linfo["fallback_pairs"],
        "surrogate_events_per_s": n_events / linfo["seconds"],
        "direct_llh_events": n_direct, "direct_llh_photons_per_event": int(error_photons),
        "direct_llh_seconds": direct_llh_s, "direct_events_per_s": direct_events_per_s,
        "direct_llh_finite": bool(np.all(np.isfinite(llh_direct))),
        "speedup": (n_events / linfo["seconds"]) / direct_events_per_s,
        "direct_photons_per_s": rate,
        "llh_finite": bool(np.all(np.isfinite(llh))),
    }
    eng.close()
    sur.model.close()
    return out
