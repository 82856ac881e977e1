"""Host side of the hit record and its post-processing (no GPU): the
PhotonHits columns, and calc_total_weight / calc_time_residual /
make_hits_from_photons (photon_prop_cuda.jl:850-924) on hits placed where
the geometry fixes the answer."""
import math

import numpy as np

import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd.propagation import PhotonHits

PMT_AREA = (75e-3 / 2) ** 2 * math.pi


def hits_at(positions, wavelength, dist, time, module_id):
    n = len(positions)
    h = PhotonHits.empty(n, epilogue=False)
    h.position[:] = np.asarray(positions, np.float32)
    h.direction[:] = 0.0
    h.initial_direction[:] = 0.0
    h.wavelength[:] = wavelength
    h.dist_travelled[:] = dist
    h.time[:] = time
    h.module_id[:] = module_id
    h.source_index[:] = 0
    h.photon_id[:] = np.arange(n)
    return h


def test_columns_truncate_select_sort():
    h = PhotonHits.empty(6, epilogue=False)
    assert h.pmt_id is None and h.total_weight is None and h.tres is None
    assert PhotonHits.empty(6).pmt_id.shape == (6,)
    h.source_index[:] = [1, 0, 1, 0, 2, 0]
    h.photon_id[:] = [5, 9, 2, 1, 0, 4]
    h.time[:] = np.arange(6, dtype=np.float32)
    t = h.truncated(4)
    assert len(t) == 4 and t.tres is None
    s = h.sorted()
    assert list(zip(s.source_index, s.photon_id)) == [(0, 1), (0, 4), (0, 9), (1, 2), (1, 5), (2, 0)]
    assert list(s.time) == [3, 5, 1, 2, 0, 4]


def test_total_weight_closed_forms():
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    sphere = nt.DetectionSphere(np.array([0, 0, 10.0], np.float32), 0.21, 16, PMT_AREA, 1)
    pom = nt.MultiPMTDetector(np.array([0, 20.0, 0], np.float32), 0.21, PMT_AREA,
                              nt.make_pom_pmt_coordinates(np.float32), 2)
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 10)
    setup = nt.PhotonPropSetup([src], [sphere, pom], med, spec)
    L400 = 14.5                                    # STRAW fit at 400 nm (medium.jl:386-387)
    h = hits_at([[0, 0, 9.79]] * 3 + [[0, 19.79, 0]], 400.0, [0.0, L400, 2 * L400, 0.0], 1.0, [1, 1, 1, 2])
    nt.calc_total_weight(h, setup)
    np.testing.assert_allclose(h.extra["abs_weight"], [1.0, math.exp(-1), math.exp(-2), 1.0], rtol=1e-6)
    # DetectionSphere: n_pmts·area/(4πr²); MultiPMTDetector: 1 (detection.jl:139-146)
    sphere_acc = 16 * PMT_AREA / (4 * math.pi * 0.21 ** 2)
    np.testing.assert_allclose(h.extra["area_acc"], [sphere_acc] * 3 + [1.0], rtol=1e-6)
    # one wavelength acceptance for all four hits, and the product of the three factors
    assert np.all(h.extra["wl_acc"] == h.extra["wl_acc"][0]) and 0 < h.extra["wl_acc"][0] < 1
    np.testing.assert_allclose(h.extra["total_weight"],
                               h.extra["area_acc"] * h.extra["wl_acc"] * h.extra["abs_weight"], rtol=1e-12)
    # zero acceptance at the table's first knot (200 nm)
    z = hits_at([[0, 0, 9.79]], 200.0, 0.0, 1.0, [1])
    assert nt.calc_total_weight(z, setup).extra["total_weight"][0] == 0.0


def test_time_residual_of_a_direct_photon():
    """A photon that flies straight from the first source to the module's
    surface at c/n(800 nm) has residual 0; a later one its delay."""
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    tgt = nt.DetectionSphere(np.array([3.0, 4.0, 12.0], np.float32), 0.5, 16, PMT_AREA, 7)
    srcs = [nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 10),
            nt.PointlikeIsotropicEmitter(np.array([50.0, 0, 0], np.float32), 0.0, 10)]
    setup = nt.PhotonPropSetup(srcs, [tgt], med, spec)
    n800 = float(nt.refractive_index(np.float32(800.0), med))
    t_direct = (13.0 - 0.5) / (0.299792458 / n800)        # |(3, 4, 12)| = 13
    h = hits_at([[3.0, 4.0, 11.5]] * 2, 450.0, 12.5, [t_direct, t_direct + 7.25], [7, 7])
    nt.calc_time_residual(h, setup)
    np.testing.assert_allclose(h.extra["tres"], [0.0, 7.25], atol=1e-5)


def test_make_hits_keeps_photons_on_pmts():
    """Hits on a PMT axis land on that PMT; a hit between caps is dropped."""
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    coords = nt.make_pom_pmt_coordinates(np.float32)
    c = np.array([0.0, 0.0, 5.0])
    tgt = nt.MultiPMTDetector(c.astype(np.float32), 0.21, PMT_AREA, coords, 3)
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 10)
    setup = nt.PhotonPropSetup([src], [tgt], med, spec)
    th, ph = coords.astype(np.float64)
    axes = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    on_axis = c + 0.21 * axes[[0, 5, 11]]
    # a direction far from every PMT axis: the farthest of many random directions
    rng = np.random.default_rng(3)
    v = rng.normal(size=(20000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    far = v[np.argmin((v @ axes.T).max(axis=1))]
    h = hits_at(np.vstack([on_axis, c + 0.21 * far]), 450.0, 5.0, 1.0, [3, 3, 3, 3])
    kept = nt.make_hits_from_photons(h, setup)
    assert list(kept.extra["pmt_id"]) == [1, 6, 12]
    assert len(kept) == 3
