"""Host side of the serving workload (no GPU): the request stream and the
cluster it is served against."""
import math

import numpy as np

import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd.serving import cascade_requests, cluster_targets


def test_cluster_geometry():
    tg = cluster_targets()
    assert len(tg) == 140
    pos = np.array([t.position for t in tg], np.float64)
    xy = np.unique(np.round(pos[:, :2], 4), axis=0)
    assert len(xy) == 7
    r = np.linalg.norm(xy, axis=1)
    assert np.isclose(r, 0).sum() == 1 and np.allclose(np.sort(r)[1:], 50.0, atol=1e-4)
    # neighbouring outer strings are 50 m apart (a regular hexagon)
    outer = xy[r > 1]
    d = np.linalg.norm(outer[:, None] - outer[None], axis=-1)
    assert np.allclose(np.sort(d, axis=1)[:, 1], 50.0, atol=1e-3)
    z = np.unique(np.round(pos[:, 2], 4))
    assert len(z) == 20 and np.allclose(np.diff(z), 50.0) and np.isclose(z.mean(), 0.0)
    assert sorted(t.module_id for t in tg) == list(range(1, 141))


def test_request_stream_law():
    med = nt.make_cascadia_medium_properties(0.95)
    reqs = cascade_requests(4000, med, np.random.default_rng(5))
    pos = np.array([r.sources[0].position for r in reqs], np.float64)
    dirs = np.array([r.sources[0].direction for r in reqs], np.float64)
    e = np.array([r.energy for r in reqs])
    ph = np.array([r.sources[0].photons for r in reqs])
    rr = np.hypot(pos[:, 0], pos[:, 1])
    assert rr.max() <= 150.0 and np.abs(pos[:, 2]).max() <= 150.0
    # area-uniform radius: P(r < 150/√2) = 1/2
    assert abs((rr < 150 / math.sqrt(2)).mean() - 0.5) < 0.03
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-6)
    assert abs(dirs[:, 2].mean()) < 0.05 and abs((dirs[:, 2] > 0).mean() - 0.5) < 0.03
    # log-uniform energy in [1e2, 1e6] GeV, photons = 1e4·E/100 GeV
    assert e.min() >= 1e2 and e.max() <= 1e6
    assert abs(np.mean(np.log10(e)) - 4.0) < 0.06
    np.testing.assert_array_equal(ph, (1e4 * e / 100.0).astype(np.int64))
    assert all(r.capacity >= 1024 for r in reqs)
    assert len({r.seed for r in reqs}) == len(reqs)
