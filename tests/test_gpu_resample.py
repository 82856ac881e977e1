"""nt_resample (resample_simulation, pipeline.jl:86-116, on the device)
against the FP64 oracle (oracle/resample.py) on the same Philox draws: the
same hits detected, per (module, PMT) and per module, at downsample 1 and
0.5; independent of the order the hits are given in; empty input."""
import math

import numpy as np
import pytest

import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd import _abi
from oracle import resample as R

pytestmark = pytest.mark.gpu
PMT_AREA = (75e-3 / 2) ** 2 * math.pi


@pytest.fixture(scope="module")
def string_hits(engine):
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
    src = nt.ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), 0.0, 20_000_000, lp)
    coords = nt.make_pom_pmt_coordinates(np.float32)
    tgts = [nt.MultiPMTDetector(np.array([10.0, 0.0, -95.0 + 10.0 * k], np.float32), 0.21, PMT_AREA, coords, k + 1)
            for k in range(20)]
    engine.setup(nt.PhotonPropSetup([src], tgts, med, spec))
    engine.set_epilogue(med, source_position=src.position)
    h = engine.propagate(seed=3, capacity=1 << 20)
    engine.set_epilogue(enabled=False)
    return h


@pytest.mark.parametrize("per_pmt,downsample", [(True, 1.0), (False, 1.0), (True, 0.5)])
def test_device_resampling_equals_oracle(engine, string_hits, per_pmt, downsample):
    h = string_hits
    assert len(h) > 2000 and np.count_nonzero(h.total_weight > 0) > 1000
    got = engine.resample(h, seed=11, downsample=downsample, per_pmt=per_pmt)
    want, drawn = R.resample(h.module_id, h.pmt_id, h.total_weight, h.source_index, h.photon_id, seed=11,
                             downsample=downsample, per_pmt=per_pmt)
    print("hits", len(h), "detected", int(got.sum()), "groups", len(drawn))
    assert got.sum() > 50
    assert np.array_equal(got, want)
    assert not got[h.total_weight <= 0].any()


def test_device_resampling_order_and_empty(engine, string_hits):
    h = string_hits
    a = engine.resample(h, seed=12)
    perm = np.random.default_rng(1).permutation(len(h))
    b = engine.resample(h.select(perm), seed=12)
    assert np.array_equal(a[perm], b)
    assert engine.resample(h.select(np.zeros(len(h), bool)), seed=12).size == 0
    hh = _abi.nt_hits()
    assert engine.lib.nt_resample(engine.handle, hh, 5, 1, -1.0, 1, None, None, None) == _abi.NT_ERR_ARG
