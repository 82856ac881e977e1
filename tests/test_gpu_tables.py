"""Photon-table entries in the reference's layout (scripts/photon_tables.jl:
52-169) from the engine: the `photons` dataset, the `pmt_hits` datasets of
resampled hits under random orientations with the rotated-frame angles,
the ×10 oversampling retry and the photon-count cut."""
import math

import numpy as np
import pytest

import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd import tables as T

pytestmark = pytest.mark.gpu


def test_table_entry_layout(engine):
    med = nt.make_cascadia_medium_properties(0.99)
    spec = nt.CherenkovSpectrum(T.WL_RANGE, 30, med)
    tgt = T.table_target()
    r = T.run_sim(engine, 1e4, 20.0, 0.3, 1.0, tgt, spec, med, n_resample=6, seed=2)
    data, attrs = r["photons"]
    assert r["oversample"] == 1.0 and data.shape[1] == 5 and len(data) > 100
    assert attrs == {"energy": 1e4, "distance": 20.0, "dir_costheta": 0.3, "dir_phi": 1.0}
    # hit positions on the module surface, weights in (0, 1], time residuals from the cascade
    # FP32 intersections of photons arriving from tens of metres: grazing
    # hits carry the cancellation of the sphere test's discriminant
    dev = np.abs(np.linalg.norm(data[:, 1:4], axis=1) - 0.21)
    print("|pos| - r: median", np.median(dev), "99 %", np.percentile(dev, 99), "max", dev.max())
    assert np.median(dev) < 1e-4 and np.percentile(dev, 99) < 2e-3 and dev.max() < 0.05
    assert np.all((data[:, 4] >= 0) & (data[:, 4] <= 1))
    assert np.median(data[:, 0]) > -5
    assert 1 <= len(r["pmt_hits"]) <= 6
    ct = 0.3
    for ph, a in r["pmt_hits"]:
        assert ph.shape[1] == 2 and np.all((ph[:, 1] >= 1) & (ph[:, 1] <= 16)) and len(ph) > 0
        d = np.array([math.sin(a["dir_theta"]) * math.cos(a["dir_phi"]),
                      math.sin(a["dir_theta"]) * math.sin(a["dir_phi"]), math.cos(a["dir_theta"])])
        p = np.array([math.sin(a["pos_theta"]) * math.cos(a["pos_phi"]),
                      math.sin(a["pos_theta"]) * math.sin(a["pos_phi"]), math.cos(a["pos_theta"])])
        assert float(d @ p) == pytest.approx(ct, abs=1e-5)        # the angle source-dir vs module-source kept
        # resampled counts are about Σ weight of the photons landing on PMTs
        assert set(a) >= {"energy", "distance", "dir_costheta", "dir_theta", "dir_phi", "pos_theta", "pos_phi"}


def test_oversampling_retry_and_cut(engine):
    med = nt.make_cascadia_medium_properties(0.99)
    spec = nt.CherenkovSpectrum(T.WL_RANGE, 30, med)
    tgt = T.table_target()
    # 100 GeV at 120 m: the real light yield (~2e7 photons) leaves too few hits
    r = T.run_sim(engine, 100.0, 120.0, -0.5, 0.0, tgt, spec, med, n_resample=2, seed=5)
    assert r["oversample"] >= 10
    data, _ = r["photons"]
    assert len(data) > 100
    w_raw = data[:, 4] * r["oversample"]
    assert np.all(w_raw <= 1.0 + 1e-6) and np.max(w_raw) > 1e-4       # weights divided by the oversampling
    assert T.run_sim(engine, 100.0, 120.0, -0.5, 0.0, tgt, spec, med, n_resample=0, seed=5,
                     max_photons=1e8) is None
