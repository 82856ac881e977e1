"""End-to-end flow surrogate on the GPU: tables from the engine, training by
the project's kernels, accuracy against direct propagation (and what the
trained network gains over the untrained one), and the device likelihood
path against a host evaluation of the same terms."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(engine):
    from neutrinotelescopes_amd import make_cascadia_medium_properties, CherenkovSpectrum
    from neutrinotelescopes_amd import surrogate as S
    from neutrinotelescopes_amd.propagation import PropagationEngine
    med = make_cascadia_medium_properties(0.95)
    spec = CherenkovSpectrum(S.WL_RANGE, 30, med)
    eng = PropagationEngine(0)
    tables = S.generate_tables(eng, med, spec, n_runs=32, photons_per_run=10 ** 8, seed=3)
    sur, hist = S.train_surrogate(tables, steps=8000, lr=2e-3, seed=3)
    yield eng, med, spec, sur, hist, tables
    sur.model.close()
    eng.close()


def test_tables_and_training(trained):
    _, _, _, _, hist, tables = trained
    assert len(tables.t) > 10 ** 5
    assert set(np.unique(tables.flags)) == {1, 2}
    assert np.all(np.isfinite(tables.features)) and np.all(tables.count >= 0)
    print("holdout loss", hist["initial_holdout"], "->", hist["final_holdout"])
    assert hist["final_holdout"] < hist["initial_holdout"] - 1.0


def test_trained_network_beats_untrained_against_direct(trained):
    from neutrinotelescopes_amd import surrogate as S
    from neutrinotelescopes_amd.flow import FlowModel
    eng, med, spec, sur, _, _ = trained
    rng = np.random.default_rng(21)
    modules = S.array_detector()
    casc = S.random_cascades(24, rng)
    lam_dir, hits, _ = S.direct_propagation(eng, med, spec, casc, modules, 2 * 10 ** 8, seed=22)
    nh = np.zeros((24, len(modules)))
    np.add.at(nh, (hits[:, 0].astype(int), hits[:, 1].astype(int)), 1)
    dist = np.linalg.norm(modules[None] - casc.pos[:, None], axis=-1)
    nh = np.where(dist >= S.FALLBACK_DISTANCE, nh, 0)
    err = S.module_errors(S.surrogate_expectations(sur, casc, modules), lam_dir, nh)
    untrained = S.Surrogate(FlowModel(0, seed=77), sur.norm, sur.scale)
    err0 = S.module_errors(S.surrogate_expectations(untrained, casc, modules), lam_dir, nh)
    sel = np.argwhere(nh >= 400)[:24]
    bins = np.arange(-20.0, 1000.0, 4.0)
    lam_sur = S.surrogate_expectations(sur, casc, modules)
    kl = S.time_kl(sur, casc, modules, lam_sur, hits, [tuple(p) for p in sel], bins)
    kl0 = S.time_kl(untrained, casc, modules, lam_sur, hits, [tuple(p) for p in sel], bins)
    untrained.model.close()
    # the method without the network: log λ = the count scale's offset alone
    base = S.scale_only_surrogate(sur)
    err_b = S.module_errors(S.surrogate_expectations(base, casc, modules), lam_dir, nh)
    kl_b = S.time_kl(base, casc, modules, lam_sur, hits, [tuple(p) for p in sel], bins)
    base.model.close()
    print("pairs", err["pairs"], "count error median", err["median_abs_rel_err"], "untrained",
          err0["median_abs_rel_err"], "count scale alone", err_b["median_abs_rel_err"],
          "KL", np.median(kl), "untrained", np.median(kl0), "count scale alone", np.median(kl_b))
    assert err["pairs"] >= 10
    assert err["median_abs_rel_err"] < 0.3
    assert err["median_abs_rel_err"] < err0["median_abs_rel_err"] / 2
    assert err["median_abs_rel_err"] < err_b["median_abs_rel_err"] / 2
    assert np.median(kl) < np.median(kl0) / 4
    assert np.median(kl) < np.median(kl_b) / 4


def test_device_likelihood_matches_host_terms(trained):
    """surrogate_llh (device rows, one forward over pairs, per-hit logpdf
    from the pair outputs) against the same terms built on the host from
    pmt_features and evaluated row by row."""
    from neutrinotelescopes_amd import surrogate as S
    eng, med, spec, sur, _, _ = trained
    rng = np.random.default_rng(31)
    modules = S.array_detector()
    casc = S.random_cascades(5, rng)
    lam, hits, _ = S.direct_propagation(eng, med, spec, casc, modules, 10 ** 7, seed=32)
    obs = S.observe(lam, hits, np.random.default_rng(33))
    ev = S.to_device(casc, modules, obs, str(sur.model.dev))
    llh = S.surrogate_llh(sur, ev).cpu().numpy()
    f = S.pair_features(casc, modules).reshape(-1, 5)
    near = np.repeat(ev.near.reshape(-1), 16)
    _, lc = sur.model.forward(sur.rows(f, None, obs.counts.reshape(-1), np.full(len(f), 2, np.uint8)),
                              want_params=False)
    hp = ((obs.hits[:, 0].astype(int) * len(modules) + obs.hits[:, 1].astype(int)) * 16 + obs.hits[:, 2].astype(int))
    _, lt = sur.model.forward(sur.rows(f[hp], obs.hits[:, 3], None, np.full(len(hp), 1, np.uint8)),
                              want_params=False)
    host = np.zeros(5)
    np.add.at(host, np.repeat(np.arange(5), len(modules) * 16), np.where(near, 0, -lc.double().cpu().numpy()))
    np.add.at(host, obs.hits[:, 0].astype(int), np.where(near[hp], 0, -lt.double().cpu().numpy()))
    print("llh", llh, "host", host)
    # pair features differ between the FP32 device and FP64 host paths only by rounding
    np.testing.assert_allclose(llh, host, rtol=2e-4, atol=1.0)
    # event_llh with a fallback that has nothing to do (no pair within 5 m) is the surrogate alone
    if not ev.near.any():
        nf = S.NearFallback(eng, med, casc, modules, ev)
        assert nf.evs.size == 0
        assert np.array_equal(S.event_llh(sur, ev, nf).cpu().numpy(), llh)


def test_fallback_near_modules(trained):
    """A cascade 3 m from a module: that pair takes the direct-propagation
    terms and the event likelihood stays finite."""
    from neutrinotelescopes_amd import surrogate as S
    eng, med, spec, sur, _, _ = trained
    modules = S.array_detector()
    rng = np.random.default_rng(41)
    casc = S.random_cascades(3, rng)
    casc.pos[0] = modules[57] + np.array([3.0, 0.0, 0.0])
    casc.energy[0] = 2e4
    lam, hits, _ = S.direct_propagation(eng, med, spec, casc, modules, 10 ** 7, seed=42)
    obs = S.observe(lam, hits, np.random.default_rng(43))
    ev = S.to_device(casc, modules, obs, str(sur.model.dev))
    assert ev.near[0, 57] and ev.near.sum() >= 1
    fb = S.fallback_terms(eng, med, spec, casc, modules, obs.counts, ev.near, ev.near_hits, photons=10 ** 7,
                          seed=44)
    assert fb[0] != 0 and np.all(fb[1:] == 0) and np.isfinite(fb[0])
    llh = S.surrogate_llh(sur, ev, fb).cpu().numpy()
    assert np.all(np.isfinite(llh))


def test_near_fallback_on_device_equals_host_terms(trained):
    """The device form of the fallback (NearFallback: asynchronous launch into
    device columns, reductions on the device) gives the host terms of
    fallback_terms for the same seed, for two events with near modules and
    observed hits there (time densities included); event_llh, which runs it
    beside the surrogate on a second stream, equals the surrogate with the
    host terms added, on two calls in a row."""
    from neutrinotelescopes_amd import surrogate as S
    eng, med, spec, sur, _, _ = trained
    modules = S.array_detector()
    rng = np.random.default_rng(61)
    casc = S.random_cascades(6, rng)
    casc.pos[1] = modules[57] + np.array([3.0, 0.0, 0.0])
    casc.pos[4] = modules[120] + np.array([0.0, -2.5, 1.0])
    casc.energy[1], casc.energy[4] = 2e4, 5e4
    lam, hits, _ = S.direct_propagation(eng, med, spec, casc, modules, 10 ** 7, seed=62)
    obs = S.observe(lam, hits, np.random.default_rng(63))
    ev = S.to_device(casc, modules, obs, str(sur.model.dev))
    assert ev.near[1, 57] and ev.near[4, 120] and len(ev.near_hits) > 100
    host = S.fallback_terms(eng, med, spec, casc, modules, obs.counts, ev.near, ev.near_hits, seed=64)
    nf = S.NearFallback(eng, med, casc, modules, ev)
    import torch
    st = torch.cuda.current_stream()
    nf.launch(64, st)
    dev = nf.finish(st).cpu().numpy()
    has = ev.near.any(axis=1)
    assert np.all(dev[~has] == 0) and np.all(host[~has] == 0)
    np.testing.assert_allclose(dev[has], host[has], rtol=1e-9)
    ref = S.surrogate_llh(sur, ev, host).cpu().numpy()
    for _ in range(2):
        got = S.event_llh(sur, ev, nf, seed=64).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-9)
    # device columns too small for the hits: the launch is repeated with room for all
    small = S.NearFallback(eng, med, casc, modules, ev, capacity=64)
    small.launch(64, st)
    np.testing.assert_allclose(small.finish(st).cpu().numpy(), dev, rtol=1e-12)
    assert small.cols.capacity > 64


def test_full_size_likelihood_1024_cascades(trained):
    """The event-likelihood configuration at full size: 1024 cascades
    (log-uniform 1e3-1e6 GeV) on the 10-string x 20-module x 16-PMT array,
    3 276 800 (cascade, PMT) pairs in one pass, cascades placed within 5 m of
    modules so that fallback pairs occur: every event's likelihood is finite,
    the fallback terms are nonzero exactly for the events with near pairs,
    and a sample of events equals the host evaluation of the same terms."""
    from neutrinotelescopes_amd import surrogate as S
    eng, med, spec, sur, _, _ = trained
    rng = np.random.default_rng(51)
    modules = S.array_detector()
    n_ev = 1024
    casc = S.random_cascades(n_ev, rng)
    for j, m in enumerate((3, 57, 110, 199)):                     # four events with a near module
        casc.pos[j] = modules[m] + np.array([2.0 + 0.5 * j, 1.0, 0.0])
    counts = rng.poisson(1.0, (n_ev, len(modules), 16))
    # near pairs observe nothing: their direct-propagation terms stay moderate
    # (an observed hit on a PMT the direct run never reached scores -1e30)
    near_h = np.linalg.norm(modules[None] - casc.pos[:, None], axis=-1) < S.FALLBACK_DISTANCE
    counts[near_h] = 0
    k = np.repeat(np.arange(counts.size), counts.reshape(-1))
    hits = np.column_stack([k // (len(modules) * 16), (k // 16) % len(modules), k % 16,
                            rng.uniform(-10, 300, k.size)])
    obs = S.Observation(counts, hits)
    ev = S.to_device(casc, modules, obs, str(sur.model.dev))
    assert ev.counts.shape[0] == n_ev * len(modules) * 16 == 3_276_800
    assert ev.near[:4].any(axis=1).all()
    fb = S.fallback_terms(eng, med, spec, casc, modules, obs.counts, ev.near, ev.near_hits, photons=10 ** 7,
                          seed=52)
    has_near = ev.near.any(axis=1)
    assert np.array_equal(has_near, near_h.any(axis=1))
    assert np.all(fb[has_near] < 0) and np.all(fb[~has_near] == 0) and np.all(fb > -1e29)
    llh = S.surrogate_llh(sur, ev, fb).cpu().numpy()
    assert llh.shape == (n_ev,) and np.all(np.isfinite(llh))
    # the same likelihood with the fallback on the device, beside the surrogate
    llh_dev = S.event_llh(sur, ev, S.NearFallback(eng, med, casc, modules, ev), seed=52).cpu().numpy()
    np.testing.assert_allclose(llh_dev, llh, rtol=1e-9)
    # host terms for a sample of events (pairs masked where near, as on the device)
    pick = np.array([0, 5, 100, 517, 1023])
    sub = S.Cascades(casc.pos[pick], casc.dir[pick], casc.energy[pick])
    f = S.pair_features(sub, modules).reshape(-1, 5)
    near = np.repeat(ev.near[pick].reshape(-1), 16)
    _, lc = sur.model.forward(sur.rows(f, None, counts[pick].reshape(-1), np.full(len(f), 2, np.uint8)),
                              want_params=False)
    host = np.zeros(len(pick))
    np.add.at(host, np.repeat(np.arange(len(pick)), len(modules) * 16), np.where(near, 0, -lc.double().cpu().numpy()))
    for j, e in enumerate(pick):
        he = hits[hits[:, 0] == e]
        hp = (he[:, 1].astype(int) * 16 + he[:, 2].astype(int))
        fe = f[j * len(modules) * 16:(j + 1) * len(modules) * 16]
        _, lt = sur.model.forward(sur.rows(fe[hp], he[:, 3], None, np.full(len(hp), 1, np.uint8)),
                                  want_params=False)
        host[j] += np.sum(np.where(near[j * len(modules) * 16 + hp], 0, -lt.double().cpu().numpy()))
    np.testing.assert_allclose(llh[pick] - fb[pick], host, rtol=2e-4, atol=1.0)
