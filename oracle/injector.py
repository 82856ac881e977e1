"""injector.py — FP64 restatement of the device particle injector
(checker of nt_inject_particles).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the library.

VolumeInjector (src/event_generation/injectors.jl:93-111) with Cylinder
(:59-70, radius drawn as sqrt(u·radius) as the reference does), Cuboid
(:72-78), FixedPosition (:57), UniformAngularDistribution (:82-89), a
log-uniform energy and a categorical type, from the Philox4x32-10 words
key = seed, counter = (block, i lo, i hi, 0x60000000), u = (w >> 8)·2^-24.
"""
from __future__ import annotations

import math

import numpy as np

INJECT_WORD = 0x60000000


def u24(w: int) -> float:
    return (w >> 8) * 2.0 ** -24


def inject(rec: dict, n: int, seed: int, philox) -> np.ndarray:
    """rec: volume (0 cylinder, 1 cuboid, 2 fixed), center[3], height,
    radius, size[3], e_min, e_max, type_cdf[3] (FP32 values).  Returns an
    [n, 9] FP64 array: position xyz, direction xyz, time, energy, type, each
    before the final FP32 rounding."""
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    f = lambda v: float(np.float32(v))
    c = [f(v) for v in rec["center"]]
    out = np.zeros((n, 9))
    for i in range(n):
        lo, hi = i & 0xFFFFFFFF, i >> 32
        a = philox((0, lo, hi, INJECT_WORD), key)
        b = philox((1, lo, hi, INJECT_WORD), key)
        x, y, z = c
        if rec["volume"] == 0:
            z += (u24(a[2]) - 0.5) * f(rec["height"])
            r = math.sqrt(u24(a[0]) * f(rec["radius"]))
            phi = 2 * math.pi * u24(a[1])
            x += r * math.cos(phi)
            y += r * math.sin(phi)
        elif rec["volume"] == 1:
            x += (u24(a[0]) - 0.5) * f(rec["size"][0])
            y += (u24(a[1]) - 0.5) * f(rec["size"][1])
            z += (u24(a[2]) - 0.5) * f(rec["size"][2])
        e = f(rec["e_min"]) * (f(rec["e_max"]) / f(rec["e_min"])) ** u24(a[3])
        phi = 2 * math.pi * u24(b[0])
        ct = 2.0 * u24(b[1]) - 1.0
        st = math.sqrt(1.0 - ct * ct)
        ut = u24(b[2])
        cdf = [f(v) for v in rec["type_cdf"]]
        t = 0 if ut < cdf[0] else (1 if ut < cdf[1] else 2)
        out[i] = (x, y, z, math.cos(phi) * st, math.sin(phi) * st, ct, 0.0, e, t)
    return out
