"""poisson.py — FP64 restatement of the source builder's Poisson draws and
source records (checker of nt_build_cascade_sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the library.

Poisson variates: inversion by sequential search for a mean below 10;
W. Hörmann, "The transformed rejection method for generating Poisson random
variables", Insurance: Mathematics and Economics 12 (1993) 39-45 (PTRS)
above.  Uniforms: 53 bits from two Philox4x32-10 words, key = seed, counter
= (trial, particle index lo, hi, 0x50000000).  The source record follows
ExtendedCherenkovEmitter(particle, medium, wl_range; oversample)
(src/photon_propagation/lightyield.jl:471-481) with the type constants of
lightyield.jl:135-137 and the track-length parameters used for
total_lightyield (lightyield.jl:350-366).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammaln

POISSON_WORD = 0x50000000

TYPE_PARAMS = {   # long alpha, beta, b; track alpha (m), beta
    0: (2.01849, 1.45469, 0.63207, 5.3207078881, 1.00000211),   # e-
    1: (2.00035, 1.45501, 0.63008, 5.3211320598, 0.99999254),   # e+
    2: (2.83923, 1.34031, 0.64526, 5.3208540905, 0.99999877),   # gamma
}


def u53(hi: int, lo: int) -> float:
    return ((hi >> 5) * 67108864.0 + (lo >> 6)) * 2.0 ** -53


def ptrs_trial(mu: float, U, V):
    """One PTRS trial (Hörmann 1993, algorithm PTRS) for uniforms U in
    [-0.5, 0.5) and V in [0, 1): the variate, or -1 when rejected.  Works
    elementwise on numpy arrays, so the law of the algorithm can be
    integrated over a (U, V) grid."""
    smu = math.sqrt(mu)
    b = 0.931 + 2.53 * smu
    a = -0.059 + 0.02483 * b
    inv_alpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    U = np.asarray(U, np.float64)
    V = np.asarray(V, np.float64)
    us = 0.5 - np.abs(U)
    kd = np.floor((2.0 * a / us + b) * U + mu + 0.43)
    quick = (us >= 0.07) & (V <= vr)
    skip = (kd < 0) | ((us < 0.013) & (V > us))
    with np.errstate(divide="ignore", invalid="ignore"):
        full = (np.log(V) + math.log(inv_alpha) - np.log(a / (us * us) + b)
                <= -mu + kd * math.log(mu) - gammaln(np.maximum(kd, 0.0) + 1.0))
    return np.where(quick | (~skip & full), kd, -1.0)


def poisson(mu: float, philox, seed: int, index: int, word: int = POISSON_WORD) -> int:
    """philox(ctr[4], key[2]) -> 4 words; counter (trial, index lo, hi, word)."""
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    ilo, ihi = index & 0xFFFFFFFF, (index >> 32) & 0xFFFFFFFF
    if not mu > 0.0:
        return 0
    if mu < 10.0:
        r = philox((0, ilo, ihi, word), key)
        u = u53(r[0], r[1])
        p = math.exp(-mu)
        f = p
        k = 0
        while u > f and k < 1000:
            k += 1
            p *= mu / k
            f += p
        return k
    for trial in range(1000):
        r = philox((trial, ilo, ihi, word), key)
        k = int(ptrs_trial(mu, u53(r[0], r[1]) - 0.5, u53(r[2], r[3])))
        if k >= 0:
            return k
    return int(math.floor(mu + 0.5))


def cascade_source(energy: float, ptype: int, photons_per_m: float, oversample: float, philox, seed: int,
                   index: int):
    """(photons, long_a, long_b) of one particle; long_a/long_b are rounded to
    FP32 by the caller."""
    la, lb, b, ta, tb = TYPE_PARAMS.get(ptype, TYPE_PARAMS[0])
    e = float(energy)
    mean = photons_per_m * ta * e ** tb * oversample
    return poisson(mean, philox, seed, index), la + lb * math.log10(e), b
