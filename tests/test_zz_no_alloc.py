"""nt_propagate_noalloc (the allocation-free entry point) against nt_propagate
through the C ABI: the same hits, counts and epilogue columns on two calls in
a row, the second with a smaller input; capacities above the scratch array's
size are refused.  GPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd import _abi
from neutrinotelescopes_amd.propagation import PhotonHits, PropagationEngine

pytestmark = pytest.mark.gpu
PMT_AREA = (75e-3 / 2) ** 2 * math.pi


def call(eng, fn, seed, capacity):
    hits = PhotonHits.empty(capacity, epilogue=True)
    h = hits.to_abi()
    nh, ns, nd = C.c_int64(), C.c_int64(), C.c_int64()
    st = fn(eng.handle, C.c_uint64(seed), C.byref(h), C.byref(nh), C.byref(ns), C.byref(nd))
    out = hits.truncated(nh.value)
    out.n_simulated, out.n_dropped = ns.value, nd.value
    return st, out


def setup(n, r):
    med = nt.make_cascadia_medium_properties(0.95)
    coords = nt.make_pom_pmt_coordinates(np.float32)
    tgts = [nt.MultiPMTDetector(np.array([0, 0, 6.0], np.float32), r, PMT_AREA, coords, 3),
            nt.DetectionSphere(np.array([4.0, 0, 0], np.float32), r, 16, PMT_AREA, 4)]
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, n)
    return nt.PhotonPropSetup([src], tgts, med, nt.CherenkovSpectrum((300.0, 600.0), 30, med))


def assert_same(a, b):
    a, b = a.sorted(), b.sorted()
    assert len(a) == len(b) and a.n_simulated == b.n_simulated and a.n_dropped == b.n_dropped
    for col in ("position", "direction", "initial_direction", "wavelength", "time", "dist_travelled",
                "module_id", "source_index", "photon_id", "pmt_id", "total_weight", "tres"):
        assert np.array_equal(getattr(a, col), getattr(b, col)), col


def test_noalloc_equals_propagate_two_calls():
    eng = PropagationEngine(0)
    try:
        for n, r, seed in ((2_000_000, 1.0, 3), (200_000, 0.5, 4)):       # second call: smaller input
            eng.setup(setup(n, r))
            eng.set_epilogue()
            st0, ref = call(eng, eng.lib.nt_propagate, seed, 1 << 16)
            st1, got = call(eng, eng.lib.nt_propagate_noalloc, seed, 1 << 16)
            assert st0 == st1 == _abi.NT_OK
            assert len(ref) > 100
            assert_same(ref, got)
        # overflow keeps the first `capacity` hits and reports the rest, as nt_propagate
        st0, ref = call(eng, eng.lib.nt_propagate, 4, 10)
        st1, got = call(eng, eng.lib.nt_propagate_noalloc, 4, 10)
        assert st0 == st1 == _abi.NT_ERR_HIT_OVERFLOW
        assert len(got) == 10 and got.n_dropped == ref.n_dropped > 0
        # capacities above the scratch array are refused before any launch
        st, _ = call(eng, eng.lib.nt_propagate_noalloc, 4, _abi.NT_NOALLOC_MAX_HITS + 1)
        assert st == _abi.NT_ERR_ARG
        assert "NT_NOALLOC_MAX_HITS" in eng.lib.nt_status_string(st).decode()
    finally:
        eng.close()
