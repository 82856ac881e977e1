"""resample.py — FP64 restatement of the device resampler (checker of
nt_resample, include/ntprop.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the library.

resample_simulation (src/pmt_frontend/pipeline.jl:86-116): per group of hits
((module, PMT) or module) draw n = min(Poisson(Σw·downsample), #hits with
w > 0) and take n of the group's hits without replacement with
probabilities ∝ w.  The restatement follows the device's random-number
layout exactly: each hit's key ln(u)/w with u = 1 − u53 of Philox4x32-10
(key = seed, counter (0x71000000, photon lo, photon hi, source index)); the n
largest keys of a group are taken (Efraimidis & Spirakis, IPL 97 (2006)
181-185); the group's weights are summed in descending-key order; the
Poisson draw is oracle/poisson.py's with counter (trial, group, 0,
0x70000000).
"""
from __future__ import annotations

import numpy as np

from . import poisson as P

KEY_WORD = 0x71000000
COUNT_WORD = 0x70000000
_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox_np(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on arrays of counters (values < 2^32 in uint64)."""
    c = [np.asarray(x, np.uint64) & _MASK for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0 & 0xFFFFFFFF), np.uint64(k1 & 0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_M0) * c[0]
        p1 = np.uint64(_M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & _MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & _MASK]
        k0 = (k0 + np.uint64(_W0)) & _MASK
        k1 = (k1 + np.uint64(_W1)) & _MASK
    return c


def _philox_scalar(ctr, key):
    r = philox_np(*[np.array([x]) for x in ctr], key[0], key[1])
    return [int(v[0]) for v in r]


def group_ids(module_id, pmt_id, per_pmt: bool):
    m = np.asarray(module_id, np.int64)
    return m * 17 + np.minimum(np.asarray(pmt_id, np.int64), 16) if per_pmt else m


def keys(weight, source_index, photon_id, seed: int):
    """ln(u)/w per hit (−inf for w <= 0)."""
    w = np.asarray(weight, np.float32).astype(np.float64)
    pid = np.asarray(photon_id, np.uint64)
    r = philox_np(np.full(pid.shape, KEY_WORD, np.uint64), pid & _MASK, pid >> np.uint64(32),
                  np.asarray(source_index, np.uint64), seed & 0xFFFFFFFF, seed >> 32)
    u53 = ((r[0] >> np.uint64(5)).astype(np.float64) * 67108864.0 + (r[1] >> np.uint64(6)).astype(np.float64)) \
        * 2.0 ** -53
    with np.errstate(divide="ignore"):
        k = np.log(1.0 - u53) / np.where(w > 0, w, 1.0)
    return np.where(w > 0, k, -np.inf)


def select_top(group, key, weight, counts_fn):
    """Order hits by group, then descending key (ties: hit order); per group
    n = counts_fn(group, Σw in that order, #positive); the first n are
    selected.  Returns the boolean selection and {group: n}."""
    order = np.lexsort((-key, group))
    g = group[order]
    w = np.asarray(weight, np.float32).astype(np.float64)[order]
    sel = np.zeros(len(group), bool)
    starts = np.r_[0, np.flatnonzero(np.diff(g)) + 1] if g.size else np.zeros(0, np.int64)
    ends = np.r_[starts[1:], g.size] if g.size else np.zeros(0, np.int64)
    drawn = {}
    for a, b in zip(starts, ends):
        wsum = float(np.cumsum(w[a:b])[-1])          # sequential sum, as one device thread adds them
        pos = int(np.count_nonzero(w[a:b] > 0))
        n = min(counts_fn(int(g[a]), wsum), pos)
        drawn[int(g[a])] = n
        sel[order[a:a + n]] = True
    return sel, drawn


def resample(module_id, pmt_id, weight, source_index, photon_id, seed: int, downsample: float = 1.0,
              per_pmt: bool = True):
    """The detected flags nt_resample computes for these hits."""
    g = group_ids(module_id, pmt_id, per_pmt)
    k = keys(weight, source_index, photon_id, seed)
    return select_top(g, k, weight,
                      lambda grp, wsum: P.poisson(wsum * downsample, _philox_scalar, seed, grp, COUNT_WORD))
