"""The reference flow surrogate's own conditioning layout and trained-model
file (src/surrogate_models/extended_cascade_model.jl).

The reference conditions its RQNormFlow on 24 numbers per (cascade, PMT):
eight z-scored labels [ln d, ln E, direction xyz, normalised position xyz]
and a 16-wide one-hot PMT index (`preproc_labels`, :328-364), built from
particles and modules by `_calc_flow_inputs`/`calc_flow_inputs` (:367-422)
or from the `pmt_hits` table datasets by `create_pmt_table`/`read_hdf`
(:298-311, :440-478).  This module restates those host steps (FP64, as the
reference computes them) and reads the reference's trained model file
(`assets/rq_spline_model.bson`: Dense 24→768 ReLU, Dense 768→34, K = 10,
range [−20, 100] ns, the eight label normalisers) with a data-only BSON
decoder: nothing in the file is executed.

The network and flow of this package (`flow.FlowModel`) use five conditioning
features instead; the functions here let a user of the reference carry its
tables, normalisers and trained weights over.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

LABELS = ("log_distance", "log_energy", "dir_x", "dir_y", "dir_z", "pos_x", "pos_y", "pos_z")
N_PMT_LEVELS = 16                      # categorical(pmt_id, levels=1:16), extended_cascade_model.jl:353
HPARAM_FIELDS = ("K", "batch_size", "mlp_layers", "mlp_layer_size", "lr", "epochs", "dropout",
                 "non_linearity", "seed", "use_l2_norm")   # RQNormFlowHParams, :153-164


@dataclasses.dataclass
class Normalizer:
    """Normalizer (extended_cascade_model.jl:313-326): (x − mean)/σ, fitted with
    the sample standard deviation (Julia's `std`, n − 1)."""
    mean: float
    sigma: float

    @classmethod
    def fit(cls, x) -> "Normalizer":
        x = np.asarray(x, dtype=np.float64)
        return cls(float(np.mean(x)), float(np.std(x, ddof=1)))

    def __call__(self, x) -> np.ndarray:
        return (np.asarray(x, dtype=np.float64) - self.mean) / self.sigma


def _sph_to_cart(theta, phi) -> np.ndarray:
    """sph_to_cart, utils.jl:105-115 (FP64 columns)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    st = np.sin(theta)
    return np.stack([np.cos(phi) * st, np.sin(phi) * st, np.cos(theta)], axis=-1)


def _cart_to_sph32(x, y, z) -> tuple[float, float]:
    """cart_to_sph, utils.jl:118-133, on FP32 components (the reference's
    positions and directions are Float32, so acos/atan run in FP32)."""
    x, y, z = np.float32(x), np.float32(y), np.float32(z)
    if z == 1:
        return 0.0, 0.0
    if z == -1:
        return math.pi, 0.0
    theta = np.arccos(z)
    if x == 0 and y > 0:
        phi = math.pi / 2
    elif x == 0 and y < 0:
        phi = -math.pi / 2
    else:
        phi = np.arctan2(y, x)
    return float(theta), float(phi)


def preproc_labels(table: dict, norm_dict: dict | None = None) -> tuple[np.ndarray, dict]:
    """preproc_labels (extended_cascade_model.jl:328-364).  `table` holds the
    columns distance, energy, dir_theta, dir_phi, pos_theta, pos_phi and
    optionally pmt_id (1..16).  Returns the conditioning rows [N, 8 (+16)] in
    the reference's column order and the normalisers (fitted per label when
    `norm_dict` is None, else applied)."""
    cols = {"log_distance": np.log(np.asarray(table["distance"], dtype=np.float64)),
            "log_energy": np.log(np.asarray(table["energy"], dtype=np.float64))}
    d = _sph_to_cart(table["dir_theta"], table["dir_phi"])
    p = _sph_to_cart(table["pos_theta"], table["pos_phi"])
    for k, name in enumerate(("x", "y", "z")):
        cols[f"dir_{name}"] = d[:, k]
        cols[f"pos_{name}"] = p[:, k]
    fitted = norm_dict is None
    norm_dict = {} if fitted else dict(norm_dict)
    out = []
    for name in LABELS:
        if fitted:
            norm_dict[name] = Normalizer.fit(cols[name])
        out.append(norm_dict[name](cols[name]))
    cond = np.stack(out, axis=1)
    if "pmt_id" in table:
        pmt = np.asarray(table["pmt_id"]).astype(np.int64)
        if np.any((pmt < 1) | (pmt > N_PMT_LEVELS)):
            raise ValueError("pmt_id must lie in 1..16")
        one_hot = (pmt[:, None] == np.arange(1, N_PMT_LEVELS + 1)[None, :]).astype(np.float64)
        cond = np.concatenate([cond, one_hot], axis=1)
    return cond, norm_dict


def flow_input_table(particles, targets) -> dict:
    """_calc_flow_inputs (extended_cascade_model.jl:367-410): one row per
    (particle i, module j, PMT k) in the reference's LinearIndices order
    (i fastest, then j, then k).  The reference takes the position angles
    from the particle's own position vector (normalised), as written."""
    n_p, n_t = len(particles), len(targets)
    n_pmt = targets[0].n_pmts if n_t else 0
    n = n_p * n_t * n_pmt
    t = {k: np.empty(n, np.float64) for k in ("distance", "energy", "dir_theta", "dir_phi", "pos_theta", "pos_phi")}
    t["pmt_id"] = np.empty(n, np.int64)
    for j, tg in enumerate(targets):
        if tg.n_pmts != n_pmt:
            raise ValueError("all modules need the same PMT count")
        tpos = np.asarray(tg.position, dtype=np.float32)
        for i, pa in enumerate(particles):
            ppos = np.asarray(pa.position, dtype=np.float32)
            dv = ppos - tpos
            distance = np.sqrt(np.float32(dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
            npos = ppos / np.sqrt(np.float32(ppos[0] * ppos[0] + ppos[1] * ppos[1]) + ppos[2] * ppos[2])
            dth, dph = _cart_to_sph32(*np.asarray(pa.direction, dtype=np.float32))
            pth, pph = _cart_to_sph32(*npos)
            for k in range(n_pmt):
                q = i + n_p * (j + n_t * k)
                t["distance"][q] = float(distance)
                t["energy"][q] = float(pa.energy)
                t["dir_theta"][q], t["dir_phi"][q] = dth, dph
                t["pos_theta"][q], t["pos_phi"][q] = pth, pph
                t["pmt_id"][q] = k + 1
    return t


def calc_flow_inputs(particles, targets, norm_dict: dict) -> np.ndarray:
    """calc_flow_inputs (extended_cascade_model.jl:412-422): the conditioning
    matrix [24, N] with the given normalisers."""
    cond, _ = preproc_labels(flow_input_table(particles, targets), norm_dict)
    return cond.T.copy()


def create_pmt_table(data, attrs: dict, limit: bool = True) -> dict:
    """create_pmt_table (extended_cascade_model.jl:298-311): the [tres, pmt_id]
    rows of one `pmt_hits` dataset with its attributes as columns and
    hits_per_pmt (rows of the dataset with the same PMT), cut to the first 200
    rows when `limit`."""
    data = np.asarray(data, dtype=np.float64)
    tres, pmt = data[:, 0], data[:, 1]
    ids, inverse, counts = np.unique(pmt, return_inverse=True, return_counts=True)
    table = {"tres": tres, "pmt_id": pmt, "hits_per_pmt": counts[inverse].astype(np.float64)}
    for k, v in attrs.items():
        table[k] = np.full(len(tres), v, dtype=np.float64)
    if limit and len(tres) > 200:
        table = {k: v[:200] for k, v in table.items()}
    return table


def read_tables(tables, nsel_frac: float = 0.8, rng: np.random.Generator | None = None):
    """read_hdf (extended_cascade_model.jl:440-478) over `tables.export_tables`
    archives (a path, a list of paths, or already loaded `tables.load_tables`
    dicts): the `pmt_hits` datasets in name order (HDF5's), shuffled with `rng`
    when given, the first ceil(n·nsel_frac) of each file, the tables in random
    order.  Returns tres, hits_per_pmt, the conditioning [24, N] with normalisers
    fitted on these rows, and the normalisers."""
    from .tables import load_tables
    rng_tables = rng if rng is not None else np.random.default_rng()
    items = tables if isinstance(tables, (list, tuple)) else [tables]
    all_hits = []
    for item in items:
        loaded = load_tables(item) if not isinstance(item, dict) else item
        named = [(f"dataset_{k + 1}", d) for k, d in enumerate(loaded["pmt_hits"])]
        named.sort(key=lambda x: x[0])                        # HDF5 lists group members by name
        if rng is not None:
            named = [named[q] for q in rng.permutation(len(named))]
        end = len(named) if nsel_frac == 1 else int(math.ceil(len(named) * nsel_frac))
        for _, (data, attrs) in named[:end]:
            all_hits.append(create_pmt_table(data, attrs))
    if not all_hits:
        raise ValueError("no pmt_hits datasets")
    all_hits = [all_hits[q] for q in rng_tables.permutation(len(all_hits))]
    keys = set.intersection(*(set(h) for h in all_hits))
    hits = {k: np.concatenate([h[k] for h in all_hits]) for k in keys}
    cond, tf_dict = preproc_labels(hits)
    return hits["tres"], hits["hits_per_pmt"], cond.T.copy(), tf_dict


# ---------------------------------------------------------------------------
# The trained model file (Julia BSON.jl layout, read as data)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ReferenceFlow:
    """RQNormFlow as saved by the reference (`@save model hparams opt tf_dict`):
    Dense weights in Flux's [out, in] layout (FP32)."""
    w1: np.ndarray
    b1: np.ndarray
    w2: np.ndarray
    b2: np.ndarray
    n_bins: int
    range_min: float
    range_max: float
    has_log_expec: bool
    normalizers: dict
    hparams: dict

    @property
    def n_params(self) -> int:
        return int(self.w2.shape[0])

    def to_npz(self, path) -> None:
        np.savez_compressed(path, w1=self.w1, b1=self.b1, w2=self.w2, b2=self.b2,
                            meta=np.array([self.n_bins, self.range_min, self.range_max, float(self.has_log_expec)]),
                            norm_names=np.array(list(self.normalizers)),
                            norm_values=np.array([[v.mean, v.sigma] for v in self.normalizers.values()]),
                            hparam_names=np.array(list(self.hparams)),
                            hparam_values=np.array([str(v) for v in self.hparams.values()]))

    @classmethod
    def from_npz(cls, path) -> "ReferenceFlow":
        z = np.load(path, allow_pickle=False)
        meta = z["meta"]
        norms = {str(k): Normalizer(float(v[0]), float(v[1])) for k, v in zip(z["norm_names"], z["norm_values"])}
        hp = {str(k): str(v) for k, v in zip(z["hparam_names"], z["hparam_values"])}
        return cls(z["w1"], z["b1"], z["w2"], z["b2"], int(meta[0]), float(meta[1]), float(meta[2]),
                   bool(meta[3]), norms, hp)


_JL_FLOATS = {"Float32": "<f4", "Float64": "<f8"}


def _jl_type_name(t) -> str:
    return t["name"][-1] if isinstance(t, dict) and "name" in t else ""


def _jl_value(v):
    """A BSON.jl scalar: plain BSON value, or a bits struct (e.g. Float32) as raw bytes."""
    if isinstance(v, dict) and v.get("tag") == "struct" and isinstance(v.get("data"), (bytes, bytearray)):
        return float(np.frombuffer(bytes(v["data"]), _JL_FLOATS[_jl_type_name(v["type"])])[0])
    if isinstance(v, dict) and v.get("tag") == "symbol":
        return v["name"]
    return v


def _jl_array(a) -> np.ndarray:
    """A BSON.jl array: column-major raw data of `size`."""
    if not (isinstance(a, dict) and a.get("tag") == "array"):
        raise ValueError("not a BSON.jl array")
    dt = _JL_FLOATS.get(_jl_type_name(a["type"]))
    if dt is None:
        raise ValueError(f"unsupported element type {_jl_type_name(a['type'])}")
    size = [int(s) for s in a["size"]]
    flat = np.frombuffer(bytes(a["data"]), dt)
    return flat.reshape(size[::-1]).T.copy()


def load_reference_flow(path) -> ReferenceFlow:
    """Read the reference's `rq_spline_model.bson` (BSON.jl: model, hparams,
    opt, tf_dict) with a data-only decoder.  Supports the saved architecture:
    Chain(Dense(24 => H, relu), Dropout, Dense(H => 3K + 4))."""
    import bson                      # pymongo's BSON decoder: parses data, executes nothing
    with open(path, "rb") as f:
        doc = bson.decode(f.read())
    model = doc["model"]
    if _jl_type_name(model["type"]) != "RQNormFlow":
        raise ValueError("the file holds no RQNormFlow")
    chain, K, rmin, rmax, has_le = model["data"]
    layers = chain["data"][0]["data"]
    dense = [ly for ly in layers if _jl_type_name(ly.get("type")) == "Dense"]
    if len(dense) != 2:
        raise ValueError(f"expected two Dense layers, found {len(dense)}")
    w1, b1 = _jl_array(dense[0]["data"][0]), _jl_array(dense[0]["data"][1])
    w2, b2 = _jl_array(dense[1]["data"][0]), _jl_array(dense[1]["data"][1])
    # activations are singleton function structs: NNlib.#relu, then Base.#identity
    acts = [_jl_type_name(d["data"][2].get("type")) if isinstance(d["data"][2], dict) else "" for d in dense]
    if acts != ["#relu", "#identity"]:
        raise ValueError(f"unexpected activations {acts}")
    if w1.shape[1] != 24 or w2.shape[1] != w1.shape[0] or w2.shape[0] != 3 * int(K) + 4:
        raise ValueError(f"unexpected shapes {w1.shape} {w2.shape} for K = {K}")
    tf = doc["tf_dict"]["data"]
    norms = {str(k): Normalizer(float(_jl_value(v["data"][0])), float(_jl_value(v["data"][1])))
             for k, v in zip(tf[0], tf[1])}
    hp = {name: _jl_value(v) for name, v in zip(HPARAM_FIELDS, doc["hparams"]["data"])}
    return ReferenceFlow(w1.astype(np.float32), b1.astype(np.float32), w2.astype(np.float32),
                         b2.astype(np.float32), int(K), float(rmin), float(rmax), bool(has_le), norms, hp)
