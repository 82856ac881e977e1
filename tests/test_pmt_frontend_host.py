"""PMT front end, host side (no GPU): the configuration the reference derives
(pipeline.jl:31-82, spe_templates.jl:51-72, pulse_templates.jl:27-38,115-127)
and the FP64 oracle (oracle/ntpmt_oracle.c) pinned by closed forms and by
independent library routines: the Gumbel FWHM factor of the reference's
fwhm call, the transit-time Gamma's FWHM, the SPE peak-to-valley ratio, the
low-pass response, the Kaiser resampling filter's response, the filtered
pulse's area, the oracle's pulse, filter, decimation and NNLS (scipy nnls,
KKT conditions, exact recovery), the hit-draw laws (KS tests) and the whole
chain on a noise-free pulse.  The ABI's argument checks run without a device."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import optimize, signal, special, stats

from neutrinotelescopes_amd import _abi
from neutrinotelescopes_amd import pmt_frontend as pf
import _oracle


@pytest.fixture(scope="module")
def cfg():
    return pf.std_pmt_config()


@pytest.fixture(scope="module")
def cabi(cfg):
    c, keep = cfg.to_abi()
    c._keep = keep
    return c


# ---------------------------------------------------------------- configuration
def test_gumbel_fwhm_factor_closed_form():
    # the reference's fwhm(Gumbel(0, w), w) takes half the density at z = 1:
    # crossings solve z + e^{−z} = 1 + e^{−1} + ln 2
    target = 1 + math.exp(-1) + math.log(2)
    g = lambda z: z + math.exp(-z) - target
    z0 = optimize.newton(g, -1.2, fprime=lambda z: 1 - math.exp(-z), tol=1e-15)
    z1 = optimize.newton(g, 1.9, fprime=lambda z: 1 - math.exp(-z), tol=1e-15)
    assert pf.gumbel_fwhm_width_factor() == pytest.approx(z1 - z0, rel=1e-12)
    # not the Gumbel's own FWHM (z + e^{−z} = 1 + ln 2 around the mode 0)
    h = lambda z: z + math.exp(-z) - 1 - math.log(2)
    true = optimize.brentq(h, 0, 5) - optimize.brentq(h, -5, 0)
    assert true == pytest.approx(2.4464, abs=1e-4) and abs(pf.gumbel_fwhm_width_factor() - true) > 0.5
    # the reference's straight-line fit over w = 0.5:0.01:5 is this ratio
    ws = np.arange(0.5, 5.0001, 0.01)
    fw = ws * pf.gumbel_fwhm_width_factor()
    slope, icpt = np.polyfit(fw, ws, 1)
    assert pf.gumbel_width_from_fwhm(5.0) == pytest.approx(slope * 5.0 + icpt, rel=1e-12)


def test_transit_time_gamma(cfg):
    d = cfg.tt_dist
    assert d.mean() == pytest.approx(25.0, rel=1e-14)
    # FWHM from a dense grid around the mode, independently of the root finder
    x = np.linspace(d.mode() - 2, d.mode() + 2, 400001)
    p = stats.gamma.pdf(x, d.alpha, scale=d.theta)
    above = x[p >= p.max() / 2]
    assert above[-1] - above[0] == pytest.approx(1.5, abs=2e-5)
    # large shape: close to a normal of σ = √α·θ, FWHM = 2√(2 ln 2)·σ
    assert 2 * math.sqrt(2 * math.log(2)) * math.sqrt(d.alpha) * d.theta == pytest.approx(1.5, rel=2e-3)


def test_spe_peak_to_valley(cfg):
    s = cfg.spe_template
    x = np.linspace(0, 1, 200001)
    assert s.norm_mu / s.pdf(x).min() == pytest.approx(3.1, rel=1e-8)
    # mixture density integrates to one; the truncated normal's mean by quadrature
    assert s.pdf(np.linspace(0, 60, 600001)).sum() * 1e-4 == pytest.approx(1.0, abs=2e-4)
    grid = np.linspace(0, 60, 600001)
    assert (grid * s.pdf(grid)).sum() * 1e-4 == pytest.approx(s.mean(), abs=2e-4)


def test_lowpass_response(cfg):
    lp = cfg.lp_filter
    fs = cfg.sampling_freq
    assert abs(lp.response(0.0, fs)) == pytest.approx(1.0, abs=1e-15)
    assert abs(lp.response(fs / 2, fs)) < 1e-15
    assert abs(lp.response(0.125, fs)) == pytest.approx(1 / math.sqrt(2), abs=1e-14)
    # the bilinear transform of the analog prototype 1/(1 + s/Ω) at the pre-warped Ω
    b, a = signal.bilinear([1.0], [1.0 / (2 * fs * math.tan(math.pi * 0.125 / fs)), 1.0], fs)
    np.testing.assert_allclose([lp.b0, lp.b1], b, rtol=1e-12)
    np.testing.assert_allclose([1.0, lp.a1], a, rtol=1e-12)
    x = np.random.default_rng(0).normal(size=500)
    np.testing.assert_allclose(pf.filt(lp, x), signal.lfilter([lp.b0, lp.b1], [1.0, lp.a1], x), rtol=0, atol=1e-13)


def test_kaiser_design():
    n, alpha = pf.kaiserord(0.025, 60)
    assert n == math.ceil(52.05 / (2.285 * math.pi * 0.025)) + 1 and alpha * math.pi == pytest.approx(5.65326)
    w = pf.kaiser(11, 1.7)
    k = np.arange(11)
    np.testing.assert_allclose(w, special.i0(math.pi * 1.7 * np.sqrt(1 - (2 * k / 10 - 1) ** 2))
                               / special.i0(math.pi * 1.7), rtol=1e-13)
    h = pf.resample_filter(0.125)
    assert len(h) == 9313 and h.sum() == pytest.approx(32.0, rel=1e-13)
    np.testing.assert_allclose(h, h[::-1], rtol=0, atol=1e-16)


def test_decimation_filter_response(cfg):
    taps = cfg.decim_taps
    assert len(taps) == 291 and np.argmax(taps) == 145
    np.testing.assert_allclose(taps, taps[::-1], rtol=0, atol=1e-16)
    f = np.linspace(0, 0.5, 20001)                  # cycles per input sample
    H = np.abs(np.exp(-2j * np.pi * np.outer(f, np.arange(len(taps)) - 145)) @ taps)
    # pass band up to 0.8 of the output Nyquist (1/16), stop band beyond 1.2 of it
    assert np.max(np.abs(H[f <= 0.05] - 1)) < 2e-3
    assert np.max(H[f >= 0.075]) < 10 ** (-55 / 20)


def test_filtered_pulse(cfg):
    fp = cfg.pulse_model_filt
    assert (fp.t0, fp.dt, len(fp.values)) == (-10.0, 0.5, 121)
    # the pulse is a density (area amplitude), the low-pass keeps DC: area ≈ 1
    # (the part of the exponential tail beyond 50 ns is below 1e-9)
    assert fp.values.sum() * fp.dt == pytest.approx(1.0, abs=2e-3)
    assert np.all(fp.values[fp.t0 + fp.dt * np.arange(121) < 0] == 0)
    assert cfg.noise_amp == pytest.approx(0.04, rel=1e-15)
    assert (cfg.n_wave, cfg.n_adc, cfg.n_pulse) == (601, 76, 3001)


# ---------------------------------------------------------------- oracle pieces
def test_oracle_pulse_and_filtered_pulse(cfg, cabi):
    lib = _oracle.load()
    x = np.linspace(-5, 25, 3001)
    got = np.array([lib.ntor_pmt_pulse(C.byref(cabi), v) for v in x])
    np.testing.assert_allclose(got, cfg.pulse_model.dist.pdf(x) * cfg.pulse_model.amplitude, rtol=1e-13, atol=1e-300)
    assert got.sum() * 0.01 == pytest.approx(1.0, abs=1e-4)
    s = np.linspace(-12, 52, 6401)
    got = np.array([lib.ntor_pmt_filt_pulse(C.byref(cabi), v) for v in s])
    np.testing.assert_allclose(got, cfg.pulse_model_filt(s), rtol=1e-13, atol=1e-300)


def test_oracle_iir_and_decimation(cfg, cabi):
    lib = _oracle.load()
    rng = np.random.default_rng(1)
    x = rng.normal(size=601)
    y = np.zeros(601)
    lib.ntor_pmt_iir(C.byref(cabi), 601, _oracle.dptr(x), y.ctypes.data_as(C.POINTER(C.c_double)))
    lp = cfg.lp_filter
    np.testing.assert_allclose(y, signal.lfilter([lp.b0, lp.b1], [1.0, lp.a1], x), rtol=0, atol=1e-13)
    d = np.zeros(76)
    lib.ntor_pmt_decimate(C.byref(cabi), 601, _oracle.dptr(y), 8, 76, d.ctypes.data_as(C.POINTER(C.c_double)))
    full = np.convolve(y, cfg.decim_taps)          # full[j] = Σ_m h_m y_{j−m}
    np.testing.assert_allclose(d, full[8 * np.arange(76) + 145], rtol=0, atol=1e-13)
    # a pass-band sine comes out sampled at the input instants 8k
    t = np.arange(601)
    sine = np.sin(2 * np.pi * 0.01 * t + 0.3)
    lib.ntor_pmt_decimate(C.byref(cabi), 601, _oracle.dptr(sine), 8, 76, d.ctypes.data_as(C.POINTER(C.c_double)))
    inner = slice(20, 56)                          # away from the zero padding at the ends
    np.testing.assert_allclose(d[inner], sine[8 * np.arange(76)][inner], atol=3e-3)


def test_oracle_fnnls_unique_solutions():
    """Full column rank: the NNLS solution is unique; the oracle equals
    scipy's Lawson-Hanson on random problems with active constraints."""
    rng = np.random.default_rng(3)
    for trial in range(20):
        m, n = rng.integers(8, 40), rng.integers(3, 8)
        A = rng.normal(size=(m, n))
        b = rng.normal(size=m)
        x, st = _oracle.fnnls(A, b, tol=1e-12)
        xs, _ = optimize.nnls(A, b)
        assert st == 0
        np.testing.assert_allclose(x, xs, rtol=1e-8, atol=1e-10)


def test_oracle_fnnls_kkt_and_recovery(cfg, cabi):
    """The unfolding problem itself (underdetermined): KKT conditions hold,
    and separated non-negative pulses are recovered exactly."""
    t0 = 100.0
    n_adc, n_pulse = cfg.n_adc, cfg.n_pulse
    tk = t0 + 4.0 * np.arange(n_adc)
    tau = t0 + 0.1 * np.arange(n_pulse)
    A = cfg.pulse_model_filt(tk[:, None] - tau[None, :])
    x_true = np.zeros(n_pulse)
    x_true[[600, 1400, 2200]] = [1.0, 2.5, 0.7]
    x, st = _oracle.fnnls(A, A @ x_true)
    assert st == 0
    np.testing.assert_allclose(x, x_true, atol=1e-6)
    rng = np.random.default_rng(4)
    b = A @ x_true + rng.normal(0, 0.01, n_adc)
    x, st = _oracle.fnnls(A, b)
    assert st == 0 and np.all(x >= 0)
    w = A.T @ (b - A @ x)
    assert np.max(w[x == 0]) <= 1e-8                       # no descent direction left
    assert np.max(np.abs(w[x > 0])) <= 1e-9                # stationary on the support


def test_oracle_hit_draw_laws(cfg, cabi):
    n = 40000
    rng = np.random.default_rng(5)
    t = np.full(n, 100.0, np.float32)
    src = rng.integers(0, 5, n).astype(np.uint32)
    pid = np.arange(n, dtype=np.uint64) * 104729
    tp, q, fl = _oracle.pmt_hit_draws(cabi, t, src, pid, seed=11)
    assert not fl.any()
    tt = tp - 100.0 + cfg.tt_dist.mean()
    assert stats.kstest(tt, stats.gamma(cfg.tt_dist.alpha, scale=cfg.tt_dist.theta).cdf).pvalue > 0.01
    s = cfg.spe_template
    a = (s.trunc_low - s.norm_mu) / s.norm_sigma
    cdf = lambda x: s.expon_weight * stats.expon.cdf(x, scale=s.expon_rate) + (1 - s.expon_weight) * \
        stats.truncnorm.cdf(x, a, np.inf, loc=s.norm_mu, scale=s.norm_sigma)
    assert stats.kstest(q, cdf).pvalue > 0.01
    assert np.all(q >= 0)
    # the normals of the Box-Muller pairs
    r = np.random.default_rng(6).integers(0, 2 ** 32, (20000, 4), dtype=np.uint64).astype(np.uint32)
    out = np.zeros(2)
    z = []
    for row in r:
        _oracle.load().ntor_pmt_normals((C.c_uint32 * 4)(*row), out.ctypes.data_as(C.POINTER(C.c_double)))
        z.extend(out)
    assert stats.kstest(z, "norm").pvalue > 0.01


def test_oracle_chain_stages(cfg):
    """The oracle's channel chain stage by stage against independent
    restatements: waveform = Σ q·pulse + noise (noise law: the same Box-Muller
    normals), filtered = scipy lfilter, digitised = the centred convolution at
    8k, charges = an NNLS solution (KKT) of the matrix built in numpy from the
    interpolated filtered pulse."""
    c, keep = cfg.to_abi()
    t = np.array([123.4, 130.0, 131.5, 260.0])
    q = np.array([1.0, 0.8, 2.0, 1.3])
    r = _oracle.pmt_channel(c, t, q, 5, seed=1)
    assert r["status"] == 0 and r["t0"] == pytest.approx(73.4, abs=1e-12)
    ts = r["t0"] + 0.5 * np.arange(601)
    clean = (cfg.pulse_model.dist.pdf(ts[:, None] - t[None, :]) * q).sum(1)
    noise = (r["wave"] - clean) / cfg.noise_amp
    assert abs(noise.mean()) < 0.15 and abs(noise.std() - 1) < 0.1
    lp = cfg.lp_filter
    np.testing.assert_allclose(r["filt"], signal.lfilter([lp.b0, lp.b1], [1.0, lp.a1], r["wave"]), atol=1e-13)
    full = np.convolve(r["filt"], cfg.decim_taps)
    np.testing.assert_allclose(r["digi"], full[8 * np.arange(76) + 145], atol=1e-13)
    tk = r["t0"] + 4.0 * np.arange(76)
    tau = r["t0"] + 0.1 * np.arange(3001)
    A = cfg.pulse_model_filt(tk[:, None] - tau[None, :])
    x = r["x"]
    w = A.T @ (r["digi"] - A @ x)
    assert np.all(x >= 0) and np.max(w[x == 0]) <= 1e-8 and np.max(np.abs(w[x > 0])) <= 1e-9
    xo, st = _oracle.fnnls(A, r["digi"])
    np.testing.assert_allclose(x, xo, atol=1e-9)


def test_oracle_chain_noise_free(cfg):
    """One 1-p.e. pulse, no noise: the waveform is the pulse and the unfolded
    charge sits at the pulse time.  The total is 1.14 p.e.: the template is the
    filtered pulse without the resampling low-pass, which widens the digitised
    pulse, so the reference's chain reads a single p.e. as slightly more."""
    import dataclasses
    quiet = dataclasses.replace(cfg, noise_amp=0.0)
    c, keep = quiet.to_abi()
    r = _oracle.pmt_channel(c, np.array([123.4]), np.array([1.0]), 5, seed=1)
    assert r["status"] == 0
    ts = r["t0"] + 0.5 * np.arange(601)
    np.testing.assert_allclose(r["wave"], quiet.pulse_model.dist.pdf(ts - 123.4), rtol=1e-13, atol=1e-300)
    x = r["x"]
    tau = r["t0"] + 0.1 * np.arange(3001)
    assert 1.0 < x.sum() < 1.25
    assert np.sum(x * tau) / x.sum() == pytest.approx(123.4, abs=1.0)


# ---------------------------------------------------------------- ABI (no device)
def test_abi_rejects_bad_configurations(cfg):
    lib = _abi.load_library()
    h = C.c_void_p()
    assert lib.ntpmt_create(0, None, C.byref(h)) == _abi.NT_ERR_ARG
    import dataclasses
    bad = [dataclasses.replace(cfg, adc_freq=0.3),            # sampling/adc not an integer
           dataclasses.replace(cfg, unf_pulse_res=0.3),       # (1/adc)/res not an integer
           dataclasses.replace(cfg, noise_amp=float("nan")),
           dataclasses.replace(cfg, time_range=1e6),          # grids beyond the limits
           dataclasses.replace(cfg, tt_dist=pf.GammaDist(0.5, 1.0)),
           dataclasses.replace(cfg, decim_taps=cfg.decim_taps[:-1])]   # even number of taps
    for b in bad:
        c, keep = b.to_abi()
        assert lib.ntpmt_create(0, C.byref(c), C.byref(h)) == _abi.NT_ERR_ARG
        assert lib.ntpmt_last_error()
    assert lib.ntpmt_make_reco_pulses(None, None, None, 0, 0, 0, 0, None, None, None) == _abi.NT_ERR_ARG
    assert lib.ntpmt_sizes(None, None, None, None) == _abi.NT_ERR_ARG
