"""Propagation parity: the HIP kernel (through the C ABI) against the CPU
oracle.

Same seed: both sides draw the same Philox streams, so a photon's history is
the same up to last-bit differences of the transcendental functions (ocml on
the GPU, glibc in the oracle).  Hits are matched by (source, photon index):
at least 99.5 % of each side's hits must match and matched hits agree within
1e-4 relative.  Different seeds: independent samples, compared as the
reference's Monte Carlo would be: per-module counts within 3σ Poisson and
two-sample KS tests on arrival time and wavelength at p > 0.01.
"""
import dataclasses
import math

import numpy as np
import pytest
from scipy import stats

import _oracle as O
import neutrinotelescopes_amd as nt
from neutrinotelescopes_amd import _abi

pytestmark = pytest.mark.gpu

PMT_AREA = (75e-3 / 2) ** 2 * math.pi


def key(h):
    return (h.source_index.astype(np.uint64) << np.uint64(40)) | h.photon_id.astype(np.uint64)


def assert_hits_match(g, o, min_frac=0.995):
    kg, ko = key(g), key(o)
    assert len(np.unique(kg)) == len(kg), "a photon produced two hits"
    common, ig, io = np.intersect1d(kg, ko, return_indices=True)
    if len(kg) + len(ko) == 0:
        return
    assert len(common) >= min_frac * len(ko) - 1, (len(common), len(ko), len(kg))
    assert len(common) >= min_frac * len(kg) - 1, (len(common), len(ko), len(kg))
    assert np.array_equal(g.module_id[ig], o.module_id[io])
    assert np.array_equal(g.wavelength[ig], o.wavelength[io]) or \
        np.allclose(g.wavelength[ig], o.wavelength[io], rtol=1e-6)
    # Rounding differences of the transcendental chains (the GPU math library
    # vs glibc, a few ulp; the Cherenkov emission direction agrees to 5e-5)
    # persist through the scattering rotations, so a hit position may move by
    # that relative amount times the path length: the position tolerance is
    # relative to |position| + distance travelled.
    path = o.dist_travelled[io].astype(np.float64)[:, None]
    for col, tol in (("position", 1e-4), ("direction", 1e-4), ("initial_direction", 1e-5),
                     ("time", 1e-4), ("dist_travelled", 1e-4)):
        a, b = getattr(g, col)[ig], getattr(o, col)[io]
        scale = np.maximum(np.abs(b), 1.0)
        if col == "position":
            scale = scale + path
        bad = np.any(np.abs(a - b) > tol * scale, axis=-1) if a.ndim == 2 else np.abs(a - b) > tol * scale
        # a photon whose path grazes a sphere or whose scattering angle is
        # decided by a last-bit difference may diverge; allow a few
        assert bad.sum() <= max(3, 0.002 * len(common)), (col, int(bad.sum()), len(common))


def run_gpu(engine, setup, seed=0, capacity=1 << 20):
    engine.setup(setup)
    h = engine.propagate(seed=seed, capacity=capacity)
    return h


def iso_sphere_setup(n, g=0.95, r=0.21, **medium_ext):
    """Reference mode: isotropic point source, one DetectionSphere at
    10 m, Cascadia water, Frank-Tamm 300-600 nm table of 30 knots."""
    med = nt.make_cascadia_medium_properties(g, **medium_ext)
    spec = nt.CherenkovSpectrum((300.0, 600.0), 30, med)
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, n)
    tgt = nt.DetectionSphere(np.array([0, 0, 10.0], np.float32), r, 16, PMT_AREA, 1)
    return nt.PhotonPropSetup([src], [tgt], med, spec)


def string_targets(n=20, spacing=50.0, x=20.0):
    coords = nt.make_pom_pmt_coordinates(np.float32)
    return [nt.MultiPMTDetector(np.array([x, 0.0, -475.0 + spacing * k], np.float32), 0.21, PMT_AREA,
                                coords, k + 1) for k in range(n)]


def string_targets_r(radius, n=20, spacing=50.0, x=20.0):
    coords = nt.make_pom_pmt_coordinates(np.float32)
    return [nt.MultiPMTDetector(np.array([x, 0.0, -475.0 + spacing * k], np.float32), radius, PMT_AREA,
                                coords, k + 1) for k in range(n)]


def cascade_string_setup(n, energy=1e4, radius=0.21):
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(energy, med, nt.PEMinus)
    src = nt.ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), 0.0, n, lp)
    return nt.PhotonPropSetup([src], string_targets_r(radius), med, spec)


def test_isotropic_sphere_per_photon(engine):
    s = iso_sphere_setup(1_000_000, r=1.0)
    g = run_gpu(engine, s, seed=5)
    o = O.propagate(s, seed=5)
    assert g.n_simulated == o.n_simulated == 1_000_000
    assert len(o) > 1000
    assert_hits_match(g, o)


def test_isotropic_sphere_statistics_independent_seeds(engine):
    s = iso_sphere_setup(2_000_000, r=1.0)
    g = run_gpu(engine, s, seed=11)
    o = O.propagate(s, seed=12)
    ng, no = len(g), len(o)
    assert abs(ng - no) <= 3 * math.sqrt(ng + no)
    assert stats.ks_2samp(g.time, o.time).pvalue > 0.01
    assert stats.ks_2samp(g.wavelength, o.wavelength).pvalue > 0.01
    assert stats.ks_2samp(g.dist_travelled, o.dist_travelled).pvalue > 0.01


def test_cascade_string_per_photon(engine):
    # modules enlarged to r = 1 m so that 1e6 photons give ~10^3 hits to match
    s = cascade_string_setup(1_000_000, radius=1.0)
    g = run_gpu(engine, s, seed=3)
    o = O.propagate(s, seed=3)
    assert len(o) > 300
    assert_hits_match(g, o)


def test_cascade_string_module_counts_independent_seeds(engine):
    # the configuration itself (r = 0.21 m, 1e7 photons): per-module counts
    # of independent GPU and oracle runs agree within 3σ Poisson
    s = cascade_string_setup(10_000_000)
    g = run_gpu(engine, s, seed=4)
    o = O.propagate(s, seed=5)
    assert len(o) > 100
    for mid in range(1, 21):
        a, b = int(np.sum(g.module_id == mid)), int(np.sum(o.module_id == mid))
        assert abs(a - b) <= 3 * math.sqrt(a + b) + 1, (mid, a, b)
    assert stats.ks_2samp(g.time, o.time).pvalue > 0.01


@pytest.mark.parametrize("kind", ["time_range", "axicone", "pencil", "pencil_vertical", "pointlike_cherenkov",
                                  "extended_cherenkov_vertical"])
def test_source_kinds(engine, kind):
    med = nt.make_cascadia_medium_properties(0.9)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    p = np.array([0.5, -0.3, 0.2], np.float32)
    d = np.array([0.0, 0.6, 0.8], np.float32)
    n = 300_000
    src = {
        "time_range": nt.PointlikeTimeRangeEmitter(p, (5.0, 25.0), n),
        "axicone": nt.AxiconeEmitter(p, d, 1.0, n, 0.7),
        "pencil": nt.PencilEmitter(p, d, 2.0, n),
        # every photon's first scattering rotates about exactly +e_z (Q2 in the step loop)
        "pencil_vertical": nt.PencilEmitter(p, np.array([0, 0, 1.0], np.float32), 2.0, n),
        "pointlike_cherenkov": nt.PointlikeCherenkovEmitter(p, d, 0.0, n),
        "extended_cherenkov_vertical": nt.ExtendedCherenkovEmitter(
            p, np.array([0, 0, 1.0], np.float32), 0.0, n, nt.LongitudinalParameterisation(6.0, 0.63, 0.36)),
    }[kind]
    tgts = [nt.DetectionSphere(np.array([0, 6.0, 8.0], np.float32), 1.0, 16, PMT_AREA, 7),
            nt.DetectionSphere(np.array([0, -5.0, 3.0], np.float32), 1.0, 16, PMT_AREA, 8)]
    s = nt.PhotonPropSetup([src], tgts, med, spec)
    g = run_gpu(engine, s, seed=9)
    o = O.propagate(s, seed=9)
    assert len(o) > 50
    assert_hits_match(g, o)


def test_multi_source_ragged(engine):
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.Monochromatic(450.0)
    srcs = [nt.PointlikeIsotropicEmitter(np.array([0, 0, float(i)], np.float32), float(i), nph)
            for i, nph in enumerate([0, 1, 123457, 0, 77777, 5])]
    tgt = [nt.DetectionSphere(np.array([0, 0, 4.0], np.float32), 0.8, 16, PMT_AREA, 3)]
    s = nt.PhotonPropSetup(srcs, tgt, med, spec)
    g = run_gpu(engine, s, seed=21)
    o = O.propagate(s, seed=21)
    assert g.n_simulated == sum(x.photons for x in srcs)
    assert set(np.unique(g.source_index)) <= {1, 2, 4, 5}
    assert_hits_match(g, o)


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
def test_extensions_sl_invlambda2_quanfry(engine, arith):
    med = nt.make_cascadia_medium_properties(0.95, sl_fraction=0.2, dispersion_mode="quan_fry",
                                             max_steps=25, direction_arithmetic=arith)
    med.sca_len_scale = float(50.0 / nt.scattering_length(np.float32(400.0), med))
    spec = nt.InvLambda2Spectrum((300.0, 600.0))
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 500_000)
    tgt = [nt.DetectionSphere(np.array([0, 0, 10.0], np.float32), 1.0, 16, PMT_AREA, 1)]
    s = nt.PhotonPropSetup([src], tgt, med, spec)
    g = run_gpu(engine, s, seed=2)
    o = O.propagate(s, seed=2)
    assert_hits_match(g, o)


@pytest.mark.parametrize("max_steps", [0, 1, 2, 3, 6])
def test_step_caps_per_photon(engine, max_steps):
    """The step loop is unrolled by four: caps that end it inside a group
    (1, 2, 3, 6), and no scattering step at all (0), give the oracle's hits
    photon by photon (HG medium; a 2 m sphere 3 m away so that photons hit
    within the first steps)."""
    med = nt.make_cascadia_medium_properties(0.95, max_steps=max_steps)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 400_000)
    tgt = [nt.DetectionSphere(np.array([0, 0, 3.0], np.float32), 2.0, 16, PMT_AREA, 1)]
    s = nt.PhotonPropSetup([src], tgt, med, spec)
    g = run_gpu(engine, s, seed=3)
    o = O.propagate(s, seed=3)
    if max_steps == 0:
        assert len(g) == len(o) == 0
        return
    assert len(o) > 1000
    assert_hits_match(g, o)


def test_no_targets_and_zero_photons(engine):
    med = nt.make_cascadia_medium_properties(0.95)
    s = nt.PhotonPropSetup([nt.PointlikeIsotropicEmitter(np.zeros(3), 0.0, 10000)], [], med,
                           nt.Monochromatic(400.0))
    g = run_gpu(engine, s)
    assert len(g) == 0 and g.n_simulated == 10000
    s = nt.PhotonPropSetup([nt.PointlikeIsotropicEmitter(np.zeros(3), 0.0, 0)],
                           [nt.DetectionSphere(np.array([0, 0, 2.0]), 1.0, 1, 1.0)], med, nt.Monochromatic(400.0))
    g = run_gpu(engine, s)
    assert len(g) == 0 and g.n_simulated == 0


def test_overflow_keeps_prefix_and_counts(engine):
    s = iso_sphere_setup(400_000, r=2.0)
    full = run_gpu(engine, s, seed=8)
    engine.setup(s)
    part = engine.propagate(seed=8, capacity=100, raise_on_overflow=False)
    assert len(part) == 100 and part.n_dropped == len(full) - 100
    with pytest.raises(_abi.NTError) as e:
        engine.propagate(seed=8, capacity=100)
    assert e.value.status == _abi.NT_ERR_HIT_OVERFLOW
    kf = dict(zip(key(full).tolist(), range(len(full))))
    idx = [kf[k] for k in key(part).tolist()]          # every stored hit is a real hit
    assert np.array_equal(full.time[idx], part.time)


def test_invalid_arguments(engine):
    lib = engine.lib
    bad = nt.make_cascadia_medium_properties(0.95).to_abi()
    bad.mean_scattering_angle = 1.5
    assert lib.nt_set_medium(engine.handle, bad) == _abi.NT_ERR_ARG
    src = nt.PointlikeIsotropicEmitter(np.zeros(3), 0.0, 10).to_abi()
    src.photons = -1
    arr = _abi.as_c_array(_abi.nt_source, [src])
    assert lib.nt_set_sources(engine.handle, arr, 1) == _abi.NT_ERR_ARG
    src.photons, src.kind = 10, 42
    arr = _abi.as_c_array(_abi.nt_source, [src])
    assert lib.nt_set_sources(engine.handle, arr, 1) == _abi.NT_ERR_ARG
    # targets: non-finite positions or radii are refused
    for pos, r in (((0.0, 0.0, float("nan")), 0.2), ((float("inf"), 0.0, 0.0), 0.2), ((0.0, 0.0, 0.0), float("inf"))):
        t = nt.DetectionSphere(np.array(pos, np.float32), r, 16, PMT_AREA, 1).to_abi()
        assert lib.nt_set_targets(engine.handle, _abi.as_c_array(_abi.nt_target, [t]), 1) == _abi.NT_ERR_ARG


def test_sources_reject_non_finite_fields(engine):
    """Infinite or NaN source fields are refused by both source paths and by
    the server (an infinite longitudinal parameter would make every Γ trial
    fail), so no kernel is launched with them."""
    torch = pytest.importorskip("torch")
    lib = engine.lib
    engine.set_medium(nt.make_cascadia_medium_properties(0.95))
    engine.set_spectrum(nt.Monochromatic(400.0))
    engine.set_targets([nt.DetectionSphere(np.array([0, 0, 5.0], np.float32), 0.5, 16, PMT_AREA, 1)])
    lp = nt.LongitudinalParameterisation(6.0, 0.63, 0.36)
    base = nt.ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), 0.0, 100, lp)
    good = base.to_abi()
    arr = _abi.as_c_array(_abi.nt_source, [good])
    assert lib.nt_set_sources(engine.handle, arr, 1) == _abi.NT_OK
    cases = [("long_a", float("inf")), ("long_b", float("nan")), ("long_lrad", float("inf")),
             ("long_b", 1e-45), ("time", float("nan")), ("time_end", float("inf")), ("angle", float("nan"))]
    cases += [("position", i) for i in range(3)] + [("direction", i) for i in range(3)]
    for field, v in cases:
        src = base.to_abi()
        if field in ("position", "direction"):
            getattr(src, field)[v] = float("inf")
        else:
            setattr(src, field, v)
        arr = _abi.as_c_array(_abi.nt_source, [good, src])
        assert lib.nt_set_sources(engine.handle, arr, 2) == _abi.NT_ERR_ARG, field
        raw = np.frombuffer(bytes(arr), np.uint8)
        d = torch.from_numpy(raw.copy()).to("cuda:0")
        st = lib.nt_set_sources_device(engine.handle, d.data_ptr(), 2, -1, None)
        assert st == _abi.NT_ERR_ARG, field
        assert "source 1" in lib.nt_status_string(st).decode()
    engine.set_sources([base])                            # the context still works
    assert engine.propagate(seed=1).n_simulated == 100


def test_single_target_column_swap(engine):
    """Q4 (photon_prop_cuda.jl:590-592, taken for one target at :806-808): with
    one target propagate_photons returns the initial direction in `direction`
    and the final direction in `initial_direction`; with a second (never hit)
    target the same photons come back with the columns as named."""
    s1 = iso_sphere_setup(300_000, r=1.0)
    far = nt.DetectionSphere(np.array([0, 0, 1e4], np.float32), 0.1, 16, PMT_AREA, 2)
    s2 = nt.PhotonPropSetup(s1.sources, s1.targets + [far], s1.medium, s1.spectrum)
    one = nt.propagate_photons(s1).sorted()
    two = nt.propagate_photons(s2).sorted()
    assert len(one) > 100 and np.array_equal(one.photon_id, two.photon_id)
    assert set(two.module_id) == {1}
    assert np.array_equal(one.direction, two.initial_direction)
    assert np.array_equal(one.initial_direction, two.direction)
    assert not np.array_equal(two.direction, two.initial_direction)       # scattered photons
    plain = nt.propagate_photons(s1, single_target_swap=False).sorted()
    assert np.array_equal(plain.direction, two.direction)


@pytest.mark.parametrize("arith", ["fp32", "fp64"])
def test_scattered_history_brute_force_device(engine, arith):
    """The kernel against a brute-force FP64 history of one photon (its Philox
    words, two scatterings, a small sphere centred on the third leg), as the
    oracle is pinned in test_oracle.py."""
    from test_oracle import _brute_force_history
    seed, photon = 0x1234_5678_9ABC, 77
    med = nt.make_cascadia_medium_properties(0.95, sca_len_scale=0.1, direction_arithmetic=arith)
    wl = np.float32(400.0)
    L = float(nt.scattering_length(wl, med))
    cg = float(nt.group_velocity(wl, med))
    legs = _brute_force_history(seed, photon, L, float(np.float32(0.95)), 3)
    p3, d3, s3 = legs[2]
    r = 0.02
    tgt = nt.DetectionSphere((p3 + d3 * (s3 / 2)).astype(np.float32), r, 1, 1.0, 5)
    s = nt.PhotonPropSetup([nt.PointlikeIsotropicEmitter(np.zeros(3), 0.0, photon + 1)], [tgt], med,
                           nt.Monochromatic(400.0))
    h = run_gpu(engine, s, seed=seed)
    k = np.flatnonzero(h.photon_id == photon)
    assert len(k) == 1
    i = int(k[0])
    dist = legs[0][2] + legs[1][2] + (s3 / 2 - r)
    tol = 2e-4 if arith == "fp32" else 2e-5
    assert float(h.dist_travelled[i]) == pytest.approx(dist, rel=tol)
    assert np.allclose(h.position[i], p3 + d3 * (s3 / 2 - r), atol=tol * (1 + dist))
    assert float(h.time[i]) == pytest.approx(dist / cg, rel=tol)
    assert np.allclose(h.direction[i], d3, atol=tol)


def test_fp32_reference_arithmetic_per_photon(engine):
    """The FP32 reference direction arithmetic against the oracle's FP32
    restatement, photon for photon (cascade on the 20-module string and an
    axicone source)."""
    s = cascade_string_setup(2_000_000, radius=1.0)
    s = nt.PhotonPropSetup(s.sources, s.targets, dataclasses.replace(s.medium, direction_arithmetic="fp32"),
                           s.spectrum)
    g = run_gpu(engine, s, seed=6)
    o = O.propagate(s, seed=6)
    assert len(o) > 200
    assert_hits_match(g, o)
    med = nt.make_cascadia_medium_properties(0.9, direction_arithmetic="fp32")
    src = nt.AxiconeEmitter(np.array([0.5, -0.3, 0.2], np.float32), np.array([0.0, 0.6, 0.8], np.float32),
                            1.0, 300_000, 0.7)
    tg = [nt.DetectionSphere(np.array([0, 6.0, 8.0], np.float32), 1.0, 16, PMT_AREA, 7)]
    s = nt.PhotonPropSetup([src], tg, med, nt.CherenkovSpectrum((300.0, 800.0), 30, med))
    assert_hits_match(run_gpu(engine, s, seed=9), O.propagate(s, seed=9))


def test_isotropic_sphere_1e8_properties(engine):
    """Isotropic source, one sphere at 10 m, 1e8 photons: size-independent properties."""
    n = 100_000_000
    s = iso_sphere_setup(n)
    g = run_gpu(engine, s, seed=1, capacity=1 << 22)
    assert g.n_simulated == n and g.n_dropped == 0
    # every hit lies on the sphere surface, inside the table's wavelengths,
    # not earlier than light from the origin could arrive
    # radial error of an FP32 intersection: ~ulp(D^2)/(2 sqrt(b)), large only
    # for rays grazing the sphere
    rad = np.linalg.norm(g.position.astype(np.float64) - np.array([0, 0, 10.0]), axis=1)
    assert np.median(np.abs(rad - 0.21)) < 2e-5 and np.abs(rad - 0.21).max() < 5e-3
    assert g.wavelength.min() >= 300.0 and g.wavelength.max() <= 600.0
    cmax = float(np.max(nt.group_velocity(np.linspace(300, 600, 301).astype(np.float32), s.medium)))
    assert g.time.min() >= (10.0 - 0.21) / cmax * (1 - 1e-5)
    assert np.all(g.dist_travelled >= 10.0 - 0.21 - 1e-4)
    assert len(np.unique(g.photon_id)) == len(g)
    # hit count agrees with the oracle's rate measured on 2e6 photons (3σ)
    o = O.propagate(iso_sphere_setup(2_000_000), seed=99)
    rate = len(o) / 2e6
    expect = rate * n
    sigma = math.sqrt(expect + (n / 2e6) ** 2 * len(o))
    assert abs(len(g) - expect) < 3 * sigma, (len(g), expect)
    # unscattered photons alone give at least the solid-angle fraction × survival
    # to 9.79 m without scattering
    assert len(g) > n * (1 - math.sqrt(1 - 0.021 ** 2)) / 2 * math.exp(-9.79 / 18.0) * 0.9


def test_device_launch_with_torch_buffers(engine):
    torch = pytest.importorskip("torch")
    s = iso_sphere_setup(500_000, r=1.0)
    host = run_gpu(engine, s, seed=4)
    cap = 1 << 16
    dev = torch.device("cuda:0")
    cols = dict(position=torch.zeros(cap, 3, device=dev), direction=torch.zeros(cap, 3, device=dev),
                initial_direction=torch.zeros(cap, 3, device=dev), wavelength=torch.zeros(cap, device=dev),
                time=torch.zeros(cap, device=dev), dist_travelled=torch.zeros(cap, device=dev),
                module_id=torch.zeros(cap, dtype=torch.int16, device=dev),
                source_index=torch.zeros(cap, dtype=torch.int32, device=dev),
                photon_id=torch.zeros(cap, dtype=torch.int64, device=dev))
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    h = _abi.nt_hits()
    for k, v in cols.items():
        setattr(h, k, v.data_ptr())
    h.capacity = cap
    stream = torch.cuda.current_stream()
    _abi.check(engine.lib.nt_launch(engine.handle, 4, h, counters.data_ptr(), stream.cuda_stream), "nt_launch")
    torch.cuda.synchronize()
    c = counters.cpu().numpy()
    assert c[0] == len(host)
    assert c[1] == 500_000
    assert c[2] > 0 or c[3] > 0
    n = int(c[0])
    pid = cols["photon_id"][:n].cpu().numpy().astype(np.uint64)
    assert np.array_equal(np.sort(pid), np.sort(host.photon_id))


def test_setters_wait_for_async_launch(engine):
    """set_* right after an asynchronous nt_launch waits for its kernels: the
    launch's hits are those of the targets it was launched with."""
    torch = pytest.importorskip("torch")
    s = iso_sphere_setup(30_000_000, r=1.0)
    ref = run_gpu(engine, s, seed=12).sorted()
    engine.setup(s)
    cap = 1 << 18
    dev = torch.device("cuda:0")
    cols = dict(position=torch.zeros(cap, 3, device=dev), direction=torch.zeros(cap, 3, device=dev),
                initial_direction=torch.zeros(cap, 3, device=dev), wavelength=torch.zeros(cap, device=dev),
                time=torch.zeros(cap, device=dev), dist_travelled=torch.zeros(cap, device=dev),
                module_id=torch.zeros(cap, dtype=torch.int16, device=dev),
                source_index=torch.zeros(cap, dtype=torch.int32, device=dev),
                photon_id=torch.zeros(cap, dtype=torch.int64, device=dev))
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    h = _abi.nt_hits()
    for k, v in cols.items():
        setattr(h, k, v.data_ptr())
    h.capacity = cap
    stream = torch.cuda.Stream()
    _abi.check(engine.lib.nt_launch(engine.handle, 12, h, counters.data_ptr(), stream.cuda_stream), "nt_launch")
    # immediately: other targets (the launch must still see the sphere at 10 m)
    engine.set_targets([nt.DetectionSphere(np.array([0, 0, -3.0], np.float32), 2.0, 16, PMT_AREA, 9)])
    stream.synchronize()
    n = int(counters[0].item())
    assert n == len(ref)
    pid = np.sort(cols["photon_id"][:n].cpu().numpy().astype(np.uint64))
    assert np.array_equal(pid, ref.photon_id)
    assert set(cols["module_id"][:n].cpu().numpy().tolist()) == {1}


def cluster_layout(n_strings=7, per_string=20, spacing=50.0, radius=0.21):
    """A centre string plus a hexagon at 50 m, 20 modules each (the serving cluster)."""
    coords = nt.make_pom_pmt_coordinates(np.float32)
    xy = [(0.0, 0.0)] + [(50.0 * math.cos(k * math.pi / 3), 50.0 * math.sin(k * math.pi / 3))
                         for k in range(n_strings - 1)]
    tg = []
    for s, (x, y) in enumerate(xy):
        for k in range(per_string):
            tg.append(nt.MultiPMTDetector(np.array([x, y, -475.0 + spacing * k], np.float32), radius,
                                          PMT_AREA, coords, len(tg) + 1))
    return tg


def ladder_layouts(kind):
    """Strings whose list order is not their z order.  "dense_ladder": two
    uniform ladders (5 m and 7 m spacing, r = 2 m) listed in shuffled order,
    so a step spans several members; "irregular": a string with random z
    spacing and overlapping spheres (r = 3 m, gaps 2-9 m) next to a uniform
    one, so several members can be hit in one step and the first listed must
    win."""
    rng = np.random.default_rng(17)
    coords = nt.make_pom_pmt_coordinates(np.float32)
    if kind == "dense_ladder":
        pts = [(8.0, 0.0, -60.0 + 5.0 * k, 2.0) for k in range(30)]
        pts += [(-6.0, 9.0, -40.0 + 7.0 * k, 2.0) for k in range(20)]
    else:
        z = -50.0 + np.cumsum(rng.uniform(2.0, 9.0, 25))
        pts = [(7.0, 2.0, float(zk), 3.0) for zk in z]
        pts += [(-9.0, -4.0, -60.0 + 6.0 * k, 2.0) for k in range(20)]
    order = rng.permutation(len(pts))
    return [nt.MultiPMTDetector(np.array(pts[i][:3], np.float32), pts[i][3], PMT_AREA, coords, j + 1)
            for j, i in enumerate(order)]


@pytest.mark.parametrize("layout", ["string", "cluster", "dense_ladder", "irregular"])
def test_culling_identical_to_brute_force(engine, layout):
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(1e5, med, nt.PEMinus)
    if layout == "string":
        tg = string_targets_r(2.0)
        pos = np.zeros(3, np.float32)
    elif layout == "cluster":
        tg = cluster_layout(radius=1.5)
        pos = np.array([10.0, 5.0, 20.0], np.float32)
    else:
        tg = ladder_layouts(layout)
        pos = np.array([0.0, 1.0, -5.0], np.float32)
    src = nt.ExtendedCherenkovEmitter(pos, np.array([0.6, 0.0, 0.8], np.float32), 0.0, 1_000_000, lp)
    s = nt.PhotonPropSetup([src], tg, med, spec)
    engine.setup(s)
    _abi.check(engine.lib.nt_set_option(engine.handle, _abi.NT_OPT_TARGET_CULLING, 1), "opt")
    a = engine.propagate(seed=13).sorted()
    _abi.check(engine.lib.nt_set_option(engine.handle, _abi.NT_OPT_TARGET_CULLING, 0), "opt")
    b = engine.propagate(seed=13).sorted()
    _abi.check(engine.lib.nt_set_option(engine.handle, _abi.NT_OPT_TARGET_CULLING, 1), "opt")
    assert len(a) == len(b) > 100
    for col in ("position", "direction", "initial_direction", "wavelength", "time", "dist_travelled",
                "module_id", "photon_id"):
        assert np.array_equal(getattr(a, col), getattr(b, col)), col


def test_epilogue_matches_host_postprocessing(engine):
    """Device PMT assignment, weights and time residuals against the host
    mirror of make_hits_from_photons / calc_total_weight! / calc_time_residual!
    evaluated on the same hits."""
    from scipy.spatial.transform import Rotation
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
    lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
    src = nt.ExtendedCherenkovEmitter(np.array([1.0, 2.0, 3.0], np.float32), np.array([1, 0, 0], np.float32),
                                      5.0, 1_000_000, lp)
    tg = string_targets_r(1.0) + [nt.DetectionSphere(np.array([-15.0, 0.0, 3.0], np.float32), 1.0, 16,
                                                     PMT_AREA, 99)]
    s = nt.PhotonPropSetup([src], tg, med, spec)
    orient = Rotation.random(random_state=3).as_matrix()
    engine.setup(s)
    engine.set_epilogue(med, orient, src.position)
    h = engine.propagate(seed=31)
    engine.set_epilogue(enabled=False)
    assert len(h) > 200 and set(np.unique(h.module_id)) >= {99}
    ref = h.select(np.arange(len(h)))
    nt.calc_total_weight(ref, s)
    nt.calc_time_residual(ref, s)
    tmap = {t.module_id: t for t in tg}
    pmt = np.zeros(len(h), np.int64)
    for mid in np.unique(h.module_id):
        sel = h.module_id == mid
        pmt[sel] = nt.check_pmt_hit(h.position[sel], tmap[int(mid)], orient)
    assert np.sum(h.pmt_id != pmt) <= 1                  # FP64 boundary ties only
    assert 0 < np.mean(pmt[h.module_id != 99] > 0) < 0.5
    assert np.allclose(h.total_weight, ref.extra["total_weight"], rtol=2e-6, atol=1e-12)
    assert np.allclose(h.tres, ref.extra["tres"], rtol=0, atol=2e-5 * np.maximum(1, np.abs(ref.extra["tres"])).max())
    # the reference's PMT filter keeps only hits that land on a PMT
    kept = nt.make_hits_from_photons(ref, s, orient)
    assert len(kept) == int(np.sum(pmt > 0))


def test_photon_indices_above_2_32(engine):
    """One source of 2^32 + 2^20 photons (the reference's photon counts are
    Int64): the hits of the last 2^20 photons, whose 64-bit index has a
    nonzero high word, match the oracle photon for photon."""
    import _oracle
    first, count = 1 << 32, 1 << 20
    med = nt.make_cascadia_medium_properties(0.95)
    spec = nt.CherenkovSpectrum((300.0, 600.0), 30, med)
    src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, first + count)
    tgt = nt.DetectionSphere(np.array([0, 0, 10.0], np.float32), 0.21, 16, PMT_AREA, 1)
    setup = nt.PhotonPropSetup([src], [tgt], med, spec, seed=5)
    print("photons", first + count, flush=True)
    engine.setup(setup)
    g = engine.propagate(seed=5, capacity=1 << 21)
    assert g.n_simulated == first + count and g.n_dropped == 0
    high = g.select(g.photon_id >= first).sorted()
    o = _oracle.propagate_range(setup, first, count, seed=5).sorted()
    print("hits above 2^32:", len(high), "oracle", len(o), "of", len(g), "in total")
    assert len(o) > 50
    assert_hits_match(high, o)
