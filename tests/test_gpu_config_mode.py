"""The cascade configuration's CONFIG-mode options (no reference
counterpart): Cherenkov emission at a fixed angle (41.2°) and an angular
acceptance table in the hit epilogue.  Both default off (the reference's
acos(1/n(λ)) and opening-angle cap)."""
import math

import numpy as np
import pytest

import _oracle as O
import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd import _abi
from test_gpu_functions import _rot64, _sph, _track64, _u24, errs_in_ulps
from test_gpu_propagation import PMT_AREA, assert_hits_match, run_gpu, string_targets_r

pytestmark = pytest.mark.gpu
THETA_C = 719076e-6          # 41.2° in microradians, the option's unit


def test_fixed_angle_emission_direction(engine):
    """FP64 arithmetic: with the fixed angle the photon sits at exactly θ from
    rot_from_ez(source, track): the FP64 evaluation with cos θ, sin θ in place
    of 1/n, √(n²−1)/n, within 2 FP32 ulps."""
    med = nt.make_cascadia_medium_properties(0.95)
    engine.set_medium(med)
    engine.set_spectrum(nt.Monochromatic(400.0))
    rng = np.random.default_rng(12)
    n = 50000
    wl = rng.uniform(300, 800, n).astype(np.float32)
    up, ut, utp = _u24(rng, n), _u24(rng, n), _u24(rng, n)
    src = rng.normal(size=(n, 3))
    src = (src / np.linalg.norm(src, axis=1, keepdims=True)).astype(np.float32)
    try:
        engine.set_cherenkov_angle(THETA_C)
        y = engine.eval(_abi.NT_FN_CHERENKOV_DIR, np.column_stack([wl, up, ut, utp, src]), 3)
    finally:
        engine.set_cherenkov_angle(None)
    ph = _sph(np.full(n, math.cos(THETA_C)), 2 * np.pi * up.astype(np.float64))
    ref = _rot64(src.astype(np.float64), _rot64(_track64(ut, utp), ph))
    assert errs_in_ulps(y, ref, 1.0).max() <= 2.0
    y0 = engine.eval(_abi.NT_FN_CHERENKOV_DIR, np.column_stack([wl, up, ut, utp, src]), 3)
    assert not np.allclose(y0, y, atol=1e-3)                          # off again: acos(1/n(λ))


def test_fixed_angle_propagation_matches_oracle(engine, oracle):
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
    src = nt.ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), 0.0, 2_000_000, lp)
    s = nt.PhotonPropSetup([src], string_targets_r(1.0), med, spec)
    engine.setup(s)
    try:
        engine.set_cherenkov_angle(THETA_C)
        oracle.ntor_set_cherenkov_angle(THETA_C)
        g = engine.propagate(seed=4, capacity=1 << 20)
        o = O.propagate(s, seed=4)
    finally:
        engine.set_cherenkov_angle(None)
        oracle.ntor_set_cherenkov_angle(0.0)
    assert len(o) > 200
    assert_hits_match(g, o)
    ref = run_gpu(engine, s, seed=4)
    assert not np.array_equal(np.sort(ref.photon_id), np.sort(g.photon_id))


def test_angular_acceptance_in_epilogue(engine):
    """Device weights with the angular table against the host mirror on the
    same hits, and against the same epilogue without the table."""
    from scipy.spatial.transform import Rotation
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
    src = nt.ExtendedCherenkovEmitter(np.array([1.0, 2.0, 3.0], np.float32), np.array([1, 0, 0], np.float32),
                                      5.0, 30_000_000, lp)
    tg = string_targets_r(0.5) + [nt.DetectionSphere(np.array([-15.0, 0.0, 3.0], np.float32), 1.0, 16,
                                                     PMT_AREA, 99)]
    s = nt.PhotonPropSetup([src], tg, med, spec)
    orient = Rotation.random(random_state=4).as_matrix()
    table = nt.lambertian_angular_acceptance()
    engine.setup(s)
    engine.set_epilogue(med, orient, src.position, angular_acceptance=table)
    h = engine.propagate(seed=31, capacity=1 << 21)
    engine.set_epilogue(med, orient, src.position)
    plain = engine.propagate(seed=31, capacity=1 << 21)
    engine.set_epilogue(enabled=False)
    ref = h.select(np.arange(len(h)))
    nt.calc_total_weight(ref, s, angular_acceptance=table, orientation=orient)
    on = (h.pmt_id > 0) & (h.module_id != 99)
    assert on.sum() > 50
    assert np.allclose(h.total_weight, ref.extra["total_weight"], rtol=2e-6, atol=1e-12)
    hs, ps = h.sorted(), plain.sorted()
    ratio = hs.total_weight[ps.total_weight > 0] / ps.total_weight[ps.total_weight > 0]
    assert np.all(ratio <= 1 + 1e-6) and np.any(ratio < 0.9)          # the table only lowers weights here
    assert np.array_equal(hs.total_weight[hs.module_id == 99], ps.total_weight[ps.module_id == 99])
