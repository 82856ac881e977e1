"""Source builder checker and injectors on the CPU: the FP64 Poisson oracle
against the Poisson law (moments, binned χ² against scipy's pmf, both
branches), the light-yield and profile terms against the host mirror, and
the injector's sampling distributions (injectors.jl)."""
import math

import numpy as np
import pytest
from scipy import stats

from neutrinotelescopes_amd import (LongitudinalParameterisation, Particle, PEMinus, PEPlus, PGamma,
                                    make_cascadia_medium_properties, total_lightyield)
from neutrinotelescopes_amd import injectors as I
from oracle import poisson as P

from _oracle import philox


def draws(mu, n, seed=11):
    return np.array([P.poisson(mu, philox, seed, i) for i in range(n)])


@pytest.mark.parametrize("mu,n", [(0.7, 20000), (4.0, 20000), (9.99, 20000), (10.0, 30000), (37.5, 30000),
                                  (400.0, 20000), (2.5e5, 4000)])
def test_poisson_oracle_matches_poisson_law(mu, n):
    k = draws(mu, n)
    assert abs(k.mean() - mu) < 5 * math.sqrt(mu / n)
    # variance of the sample variance ≈ (μ + 2μ²)/n
    assert abs(k.var() - mu) < 5 * math.sqrt((mu + 2 * mu * mu) / n)
    if mu > 1e4:
        return
    # χ² over bins holding ≥ 20 expected counts
    lo, hi = int(stats.poisson.ppf(1e-4, mu)), int(stats.poisson.ppf(1 - 1e-4, mu))
    edges = np.arange(lo, hi + 2)
    pmf = stats.poisson.pmf(edges[:-1], mu)
    pmf[0] += stats.poisson.cdf(lo - 1, mu)
    pmf[-1] += stats.poisson.sf(hi, mu)
    obs = np.bincount(np.clip(k, lo, hi) - lo, minlength=len(pmf))
    # merge bins to ≥ 20 expected
    e_m, o_m, acc_e, acc_o = [], [], 0.0, 0
    for e, o in zip(pmf * n, obs):
        acc_e += e
        acc_o += o
        if acc_e >= 20:
            e_m.append(acc_e)
            o_m.append(acc_o)
            acc_e, acc_o = 0.0, 0
    e_m[-1] += acc_e
    o_m[-1] += acc_o
    chi2 = np.sum((np.array(o_m) - np.array(e_m)) ** 2 / np.array(e_m))
    assert stats.chi2.sf(chi2, len(e_m) - 1) > 1e-4, (mu, chi2, len(e_m))


@pytest.mark.parametrize("mu", [10.0, 37.5, 400.0])
def test_ptrs_law_by_grid_integration(mu):
    """The law of one PTRS trial, integrated on a (U, V) midpoint grid and
    normalised over acceptances, is the Poisson pmf: no sampling noise, so a
    hat that fails to dominate the pmf or a squeeze that leaves it shows up
    (total variation; the grid alone leaves ≈1e-3 at μ = 400)."""
    nu, nv = 16384, 4096
    V = (np.arange(nv) + 0.5) / nv
    hist = np.zeros(int(mu + 40 * math.sqrt(mu) + 50))
    accepted = 0
    for s in range(0, nu, 1024):
        U = ((np.arange(s, min(s + 1024, nu)) + 0.5) / nu - 0.5)[:, None]
        k = P.ptrs_trial(mu, U, V[None, :]).ravel()
        k = k[k >= 0].astype(np.int64)
        accepted += len(k)
        hist += np.bincount(k, minlength=len(hist))[:len(hist)]
    p = hist / hist.sum()
    tv = 0.5 * np.abs(p - stats.poisson.pmf(np.arange(len(p)), mu)).sum()
    assert tv < 3e-3, tv
    assert accepted / (nu * nv) > 0.7          # acceptance per trial: 0.75 at μ = 10, rising with μ


def test_poisson_oracle_edge_cases_and_determinism():
    assert P.poisson(0.0, philox, 1, 0) == 0
    assert P.poisson(-3.0, philox, 1, 0) == 0
    assert P.poisson(float("nan"), philox, 1, 0) == 0
    a = [P.poisson(50.0, philox, 5, i) for i in range(50)]
    assert a == [P.poisson(50.0, philox, 5, i) for i in range(50)]
    assert a != [P.poisson(50.0, philox, 6, i) for i in range(50)]
    # the two branches meet: means just below and above the switch agree in law
    assert abs(draws(9.999, 20000).mean() - draws(10.0, 20000, seed=12).mean()) < 5 * math.sqrt(20 / 20000)


@pytest.mark.parametrize("ptype,code", [(PEMinus, 0), (PEPlus, 1), (PGamma, 2)])
def test_cascade_source_terms_match_host_mirror(ptype, code):
    med = make_cascadia_medium_properties(0.95)
    model = I.cascade_model(med, (300.0, 800.0), oversample=1.0)
    for e in (0.3, 17.0, 1e4, 3.7e6):
        e32 = float(np.float32(e))
        part = Particle(np.zeros(3), np.array([0, 0, 1.0]), 0.0, e32, ptype)
        lp = LongitudinalParameterisation.from_energy(e32, med, ptype)
        _, a, b = P.cascade_source(e32, code, model.photons_per_m, 1.0, philox, 0, 0)
        assert np.float32(a) == np.float32(lp.a) and np.float32(b) == np.float32(lp.b)
        assert np.float32(model.lrad) == np.float32(lp.lrad)
        la, lb, bb, ta, tb = P.TYPE_PARAMS[code]
        mean = model.photons_per_m * ta * e32 ** tb
        assert mean == pytest.approx(total_lightyield(part, med, (300.0, 800.0)), rel=1e-12)


def test_injector_distributions():
    rng = np.random.default_rng(3)
    inj = I.VolumeInjector(I.Cylinder(np.array([0.0, 0.0, -100.0]), 200.0, 400.0), I.LogUniform(1e2, 1e6),
                           (PEMinus, PGamma), (0.25, 0.75))
    ps = inj.sample(20000, rng)
    pos = np.array([p.position for p in ps])
    d = np.array([p.direction for p in ps])
    e = np.array([p.energy for p in ps])
    r = np.hypot(pos[:, 0], pos[:, 1])
    # injectors.jl:59-70: z uniform over the height, r = sqrt(u·radius)
    assert np.all(np.abs(pos[:, 2] + 100.0) <= 100.0)
    assert stats.kstest(r ** 2 / 400.0, "uniform").pvalue > 1e-3
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert stats.kstest(d[:, 2], stats.uniform(-1, 2).cdf).pvalue > 1e-3     # isotropic: cos θ uniform
    assert stats.kstest(np.log(e), stats.uniform(math.log(1e2), math.log(1e4)).cdf).pvalue > 1e-3
    frac_gamma = np.mean([p.type is PGamma for p in ps])
    assert abs(frac_gamma - 0.75) < 5 * math.sqrt(0.75 * 0.25 / 20000)
    assert all(p.time == 0.0 for p in ps)
    box = I.Cuboid(np.array([1.0, 2.0, 3.0]), 2.0, 4.0, 6.0).sample(1000, rng)
    assert np.all(np.abs(box - [1, 2, 3]) <= [1, 2, 3])
    assert np.all(I.FixedPosition(np.array([5.0, 0, 0])).sample(3, rng) == [5.0, 0, 0])


def test_particles_to_abi_layout():
    from neutrinotelescopes_amd import _abi
    import ctypes as C
    ps = [Particle(np.array([1.0, 2, 3]), np.array([0, 0, 1.0]), 4.5, 1e3, PGamma),
          Particle(np.array([-1.0, 0, 0]), np.array([1.0, 0, 0]), 0.0, 2.0, PEPlus)]
    raw = I.particles_to_abi(ps)
    arr = (_abi.nt_particle * 2).from_buffer_copy(raw.tobytes())
    assert raw.nbytes == 2 * C.sizeof(_abi.nt_particle)
    assert list(arr[0].position) == [1.0, 2.0, 3.0] and arr[0].time == 4.5 and arr[0].energy == 1e3
    assert arr[0].type == _abi.NT_PARTICLE_GAMMA and arr[1].type == _abi.NT_PARTICLE_EPLUS


def _rec(volume=0, center=(0.0, 0.0, -100.0), height=200.0, radius=400.0, size=(0.0, 0.0, 0.0), e=(1e2, 1e6),
         cdf=(0.25, 0.25, 1.0)):
    return {"volume": volume, "center": center, "height": height, "radius": radius, "size": size,
            "e_min": e[0], "e_max": e[1], "type_cdf": cdf}


def test_injector_oracle_follows_the_reference_law():
    """The device injector's FP64 restatement draws the same law as the host
    VolumeInjector (injectors.jl:59-111): z uniform over the height,
    r² / radius uniform (r = sqrt(u·radius)), cos θ and φ uniform, ln E
    uniform, types by their probabilities."""
    from oracle import injector as J
    n = 20000
    x = J.inject(_rec(), n, 77, philox)
    r2 = x[:, 0] ** 2 + x[:, 1] ** 2
    assert np.all(np.abs(x[:, 2] + 100.0) <= 100.0)
    assert stats.kstest(r2 / 400.0, "uniform").pvalue > 1e-3
    assert stats.kstest((np.arctan2(x[:, 1], x[:, 0]) + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    np.testing.assert_allclose(np.linalg.norm(x[:, 3:6], axis=1), 1.0, atol=1e-12)
    assert stats.kstest((x[:, 5] + 1) / 2, "uniform").pvalue > 1e-3
    assert stats.kstest(np.log(x[:, 7] / 1e2) / np.log(1e4), "uniform").pvalue > 1e-3
    assert abs(np.mean(x[:, 8] == 2) - 0.75) < 5 * math.sqrt(0.75 * 0.25 / n)
    assert np.all(x[:, 8] != 1) and np.all(x[:, 6] == 0)
    # the host mirror draws the same law (two-sample KS)
    inj = I.VolumeInjector(I.Cylinder(np.array([0.0, 0.0, -100.0]), 200.0, 400.0), I.LogUniform(1e2, 1e6),
                           (PEMinus, PGamma), (0.25, 0.75))
    ps = inj.sample(n, np.random.default_rng(8))
    pos = np.array([p.position for p in ps])
    assert stats.ks_2samp(pos[:, 0] ** 2 + pos[:, 1] ** 2, r2).pvalue > 1e-3
    assert stats.ks_2samp(pos[:, 2], x[:, 2]).pvalue > 1e-3
    assert stats.ks_2samp(np.log([p.energy for p in ps]), np.log(x[:, 7])).pvalue > 1e-3
    # cuboid and fixed position
    xc = J.inject(_rec(volume=1, center=(1.0, 2.0, 3.0), size=(2.0, 4.0, 6.0)), 2000, 3, philox)
    assert np.all(np.abs(xc[:, :3] - [1, 2, 3]) <= [1, 2, 3])
    xf = J.inject(_rec(volume=2, center=(5.0, 0.0, 0.0), e=(10.0, 10.0)), 50, 3, philox)
    assert np.all(xf[:, :3] == [5.0, 0.0, 0.0]) and np.allclose(xf[:, 7], 10.0)


def test_injector_record_from_host_injector():
    from neutrinotelescopes_amd import _abi
    inj = I.VolumeInjector(I.Cylinder(np.array([1.0, 2.0, 3.0]), 50.0, 20.0), I.LogUniform(1e3, 1e5),
                           (PGamma, PEPlus), (3.0, 1.0))
    r = I.injector_to_abi(inj)
    assert r.volume == _abi.NT_VOLUME_CYLINDER and r.height == 50.0 and r.radius == 20.0
    assert list(r.center) == [1.0, 2.0, 3.0] and (r.e_min, r.e_max) == (1e3, 1e5)
    np.testing.assert_allclose(list(r.type_cdf), [0.0, 0.25, 1.0])
    with pytest.raises(ValueError):
        I.injector_to_abi(I.VolumeInjector(I.Cylinder(np.zeros(3), 1.0, 1.0), "not log-uniform"))
