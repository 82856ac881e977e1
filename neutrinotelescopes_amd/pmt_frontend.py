"""PMT front end: transit-time jitter, SPE charges, waveform digitisation and
NNLS unfolding (src/pmt_frontend/*.jl) for every (module, PMT) channel of an
event at once, on the device (libntprop.so, C ABI include/ntpmt.h).

The configuration types mirror the reference's (`PMTConfig`,
`STD_PMT_CONFIG`, `ExponTruncNormalSPE`, `PDFPulseTemplate`,
`InterpolatedPulse`, pipeline.jl:42-82, spe_templates.jl, pulse_templates.jl).
Their derived numbers — transit-time Gamma parameters, the SPE exponential
weight for a peak-to-valley ratio, the Gumbel width for a FWHM, the low-pass
coefficients, the resampling filter and the filtered pulse — are computed here
once per configuration on the host.  `make_reco_pulses` (pipeline.jl:146-160)
runs on the GPU: `PMTFrontEnd.make_reco_pulses`.
"""
from __future__ import annotations

import ctypes as C
import functools
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi

NTPMT_ST_TT_GIVEUP, NTPMT_ST_SPE_GIVEUP, NTPMT_ST_DEGENERATE, NTPMT_ST_ITER_CAP, NTPMT_ST_RANK_CAP = 1, 2, 4, 8, 16


# --------------------------------------------------------------------------
# root finding and minimisation on a bracket (host configuration only)
# --------------------------------------------------------------------------
def _brent_root(f, a, b, xtol=1e-15, maxiter=200):
    from scipy.optimize import brentq
    return brentq(f, a, b, xtol=xtol, rtol=4 * np.finfo(float).eps, maxiter=maxiter)


def _brent_min(f, a, b):
    from scipy.optimize import minimize_scalar
    r = minimize_scalar(f, bounds=(a, b), method="bounded", options={"xatol": 1e-12, "maxiter": 500})
    return r.x, r.fun


def fwhm(pdf, xmode: float, xlims=(-20.0, 20.0)) -> float:
    """utils.jl:301-307: distance between the two points where the density
    falls to half its value at `xmode`, one on each side of `xmode`."""
    half = pdf(xmode) / 2.0
    z0 = _brent_root(lambda x: pdf(x) - half, xlims[0], xmode)
    z1 = _brent_root(lambda x: pdf(x) - half, xmode, xlims[1])
    return z1 - z0


# --------------------------------------------------------------------------
# distributions
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class Gumbel:
    mu: float
    theta: float

    def pdf(self, x):
        z = (np.asarray(x, dtype=np.float64) - self.mu) / self.theta
        return np.exp(-(z + np.exp(-z))) / self.theta

    def cdf(self, x):
        return np.exp(-np.exp(-(np.asarray(x, dtype=np.float64) - self.mu) / self.theta))

    def mode(self):
        return self.mu

    def __add__(self, c: float) -> "Gumbel":
        return Gumbel(self.mu + float(c), self.theta)


@dataclass(frozen=True)
class Truncated:
    """Distributions.truncated(d, lower, upper): density d.pdf / (F(u) − F(l))
    on [l, u], 0 elsewhere."""
    untruncated: Gumbel
    lower: float
    upper: float

    def pdf(self, x):
        x = np.asarray(x, dtype=np.float64)
        z = self.untruncated.cdf(self.upper) - self.untruncated.cdf(self.lower)
        inside = (x >= self.lower) & (x <= self.upper)
        return np.where(inside, self.untruncated.pdf(x) / z, 0.0)


def truncated(d: Gumbel, lower: float, upper: float) -> Truncated:
    return Truncated(d, float(lower), float(upper))


@dataclass(frozen=True)
class GammaDist:
    """Distributions.Gamma(α, θ) (shape, scale)."""
    alpha: float
    theta: float

    def mean(self):
        return self.alpha * self.theta

    def mode(self):
        return max(self.alpha - 1.0, 0.0) * self.theta

    def pdf(self, x):
        from scipy.stats import gamma
        return gamma.pdf(x, self.alpha, scale=self.theta)


def calc_gamma_shape_mean_fwhm(mean: float, target_fwhm: float) -> float:
    """pipeline.jl:31-39: the scale θ of the Gamma(mean/θ, θ) whose FWHM
    (utils.jl:301, at its mode, within (0, 100)) is `target_fwhm`; the root
    is bracketed by [0.1, 10]·fwhm²/mean as in the reference."""
    def excess(theta):
        d = GammaDist(mean / theta, theta)
        return fwhm(d.pdf, d.mode(), xlims=(0.0, 100.0)) - target_fwhm
    return _brent_root(excess, 0.1 * target_fwhm ** 2 / mean, 10.0 * target_fwhm ** 2 / mean)


# --------------------------------------------------------------------------
# SPE charge (spe_templates.jl)
# --------------------------------------------------------------------------
def _truncnorm_pdf(x, mu, sigma, low):
    from scipy.stats import norm
    return np.where(np.asarray(x) >= low, norm.pdf(x, mu, sigma) / norm.sf(low, mu, sigma), 0.0)


@dataclass(frozen=True)
class ExponTruncNormalSPE:
    """spe_templates.jl:21-27,74-99: mixture of Exponential(expon_rate) (the
    scale, as Distributions.Exponential(θ)) with weight expon_weight and
    Normal(norm_mu, norm_sigma) truncated to [trunc_low, ∞)."""
    expon_rate: float
    norm_sigma: float
    norm_mu: float
    trunc_low: float
    expon_weight: float

    @classmethod
    def make(cls, *, expon_rate, norm_sigma, norm_mu, trunc_low, expon_weight=None, peak_to_valley=None):
        if expon_weight is None and peak_to_valley is None:
            raise ValueError("Set one of `expon_weight` and `peak_to_valley")
        if expon_weight is None:
            expon_weight = get_expon_weight_for_pv(peak_to_valley, expon_rate, norm_sigma, norm_mu, trunc_low)
        return cls(float(expon_rate), float(norm_sigma), float(norm_mu), float(trunc_low), float(expon_weight))

    def pdf(self, x):
        x = np.asarray(x, dtype=np.float64)
        ex = np.where(x >= 0, np.exp(-x / self.expon_rate) / self.expon_rate, 0.0)
        return self.expon_weight * ex + (1 - self.expon_weight) * _truncnorm_pdf(x, self.norm_mu, self.norm_sigma,
                                                                              self.trunc_low)

    def mean(self):
        from scipy.stats import norm
        a = (self.trunc_low - self.norm_mu) / self.norm_sigma
        tn = self.norm_mu + self.norm_sigma * norm.pdf(a) / norm.sf(a)
        return self.expon_weight * self.expon_rate + (1 - self.expon_weight) * tn


def spe_peak(d: ExponTruncNormalSPE) -> float:
    return d.norm_mu


def make_spe_dist(d: ExponTruncNormalSPE) -> ExponTruncNormalSPE:
    return d


def peak_to_valley(spe: ExponTruncNormalSPE) -> float:
    """spe_templates.jl:51-55: spe_peak (norm_mu, not the density there)
    divided by the smallest density on [0, 1]."""
    _, fmin = _brent_min(lambda x: float(spe.pdf(x)), 0.0, 1.0)
    return spe_peak(spe) / fmin


def get_expon_weight_for_pv(ptov, expon_rate, norm_sigma, norm_mu, trunc_low) -> float:
    """spe_templates.jl:58-72: the exponential weight in [0.01, 0.99] whose
    peak-to-valley ratio is `ptov`.  The reference minimises (pv − ptov)² with
    Brent's method (x tolerance ≈ 1.5e-8); here the root of pv − ptov is
    bracketed from the lower end, where pv falls from 66 (w = 0.01)."""
    pv = lambda w: peak_to_valley(ExponTruncNormalSPE(expon_rate, norm_sigma, norm_mu, trunc_low, w)) - ptov
    grid = np.linspace(0.01, 0.99, 99)
    vals = [pv(w) for w in grid]
    for a, b, fa, fb in zip(grid[:-1], grid[1:], vals[:-1], vals[1:]):
        if fa == 0.0:
            return float(a)
        if fa * fb < 0:
            return float(_brent_root(pv, a, b))
    raise ValueError(f"no exponential weight in [0.01, 0.99] gives peak-to-valley {ptov}")


# --------------------------------------------------------------------------
# pulse templates (pulse_templates.jl)
# --------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def gumbel_fwhm_width_factor() -> float:
    """fwhm(Gumbel(0, w), w) / w as the reference evaluates it
    (pulse_templates.jl:27-36): utils.jl's fwhm is called with xmode = w,
    where the Gumbel's mode is 0, so the half level is half the density at
    z = 1 and the factor is 3.0875 (the Gumbel's own FWHM is 2.4464·w).  It
    does not depend on w, so the reference's straight-line fit over
    w = 0.5:0.01:5 is width = fwhm / factor."""
    g = Gumbel(0.0, 1.0)
    return fwhm(g.pdf, 1.0)


def gumbel_width_from_fwhm(f: float) -> float:
    return float(f) / gumbel_fwhm_width_factor()


@dataclass(frozen=True)
class PDFPulseTemplate:
    dist: Truncated
    amplitude: float


@dataclass(frozen=True)
class InterpolatedPulse:
    """Linear interpolation of `values` at t0 + i·dt, 0 outside (the
    reference's linear_interpolation with extrapolation_bc = 0)."""
    t0: float
    dt: float
    values: np.ndarray
    amplitude: float = 1.0

    def __call__(self, s):
        s = np.asarray(s, dtype=np.float64)
        n = len(self.values)
        u = (s - self.t0) / self.dt
        i = np.clip(np.floor(u).astype(np.int64), 0, n - 2)
        f = u - i
        v = self.values[i] * (1 - f) + self.values[i + 1] * f
        v = np.where(s == self.t0 + (n - 1) * self.dt, self.values[-1], v)
        return np.where((u >= 0) & (u <= n - 1), v, 0.0) * self.amplitude


def get_template_mode(p: PDFPulseTemplate) -> float:
    """pulse_templates.jl:56-58: the untruncated distribution's mode."""
    return p.dist.untruncated.mode()


def evaluate_pulse_template(p, pulse_time: float, timestamps):
    """pulse_templates.jl:79-97 (host; configuration use)."""
    s = np.asarray(timestamps, dtype=np.float64) - pulse_time
    if isinstance(p, PDFPulseTemplate):
        return p.dist.pdf(s) * p.amplitude
    return p(s)


@dataclass(frozen=True)
class LowpassFilter:
    """First-order digital low-pass y_n = b0·x_n + b1·x_{n−1} − a1·y_{n−1}
    (zero initial state): digitalfilter(Lowpass(cutoff; fs), Butterworth(1))
    of DSP.jl, i.e. the bilinear transform of 1/(1 + s/Ω) with the cutoff
    pre-warped: Ω = tan(π·cutoff/fs), b0 = b1 = Ω/(1 + Ω), a1 = (Ω − 1)/(Ω + 1).
    DC gain 1, a zero at the Nyquist frequency, |H| = 1/√2 at the cutoff."""
    b0: float
    b1: float
    a1: float

    def response(self, f, fs):
        z = np.exp(-2j * np.pi * np.asarray(f) / fs)
        return (self.b0 + self.b1 * z) / (1 + self.a1 * z)


def butterworth1_lowpass(cutoff: float, fs: float) -> LowpassFilter:
    om = math.tan(math.pi * cutoff / fs)
    k = om / (1.0 + om)
    return LowpassFilter(k, k, (om - 1.0) / (om + 1.0))


def filt(f: LowpassFilter, x) -> np.ndarray:
    """DSP.filt with zero initial state (host; configuration use)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.empty_like(x)
    xp = yp = 0.0
    for i, xi in enumerate(x):
        yp = f.b0 * xi + f.b1 * xp - f.a1 * yp
        xp = xi
        y[i] = yp
    return y


def make_filtered_pulse(orig_pulse: PDFPulseTemplate, sampling_freq: float, eval_range, lp: LowpassFilter
                        ) -> InterpolatedPulse:
    """pulse_templates.jl:115-127: the pulse sampled on
    eval_range[0]:1/sampling_freq:eval_range[1], low-pass filtered, linear
    in between, 0 outside."""
    dt = 1.0 / sampling_freq
    n = int(math.floor((eval_range[1] - eval_range[0]) / dt + 1e-9)) + 1
    ts = eval_range[0] + dt * np.arange(n)
    return InterpolatedPulse(float(eval_range[0]), dt, filt(lp, evaluate_pulse_template(orig_pulse, 0.0, ts)), 1.0)


# --------------------------------------------------------------------------
# resampling filter (DSP.jl resample_filter, FIR design by the window method)
# --------------------------------------------------------------------------
def kaiserord(transition_width: float, attenuation: float = 60.0):
    """Kaiser's length and shape estimate for a transition width in units of
    the Nyquist frequency: n = ⌈(A − 7.95)/(2.285·π·Δ)⌉ + 1 and α = β/π with
    β = 0.1102(A − 8.7) (A > 50), 0.5842(A − 21)^0.4 + 0.07886(A − 21)
    (21 ≤ A ≤ 50), 0 otherwise."""
    n = math.ceil((attenuation - 7.95) / (2.285 * math.pi * transition_width)) + 1
    if attenuation > 50:
        beta = 0.1102 * (attenuation - 8.7)
    elif attenuation >= 21:
        beta = 0.5842 * (attenuation - 21) ** 0.4 + 0.07886 * (attenuation - 21)
    else:
        beta = 0.0
    return n, beta / math.pi


def kaiser(n: int, alpha: float) -> np.ndarray:
    """I0(πα·√(1 − (2k/(n−1) − 1)²)) / I0(πα), k = 0 … n−1."""
    return np.kaiser(n, math.pi * alpha)


def resample_filter(rate: float, n_phi: int = 32, rel_bw: float = 1.0, attenuation: float = 60.0) -> np.ndarray:
    """DSP.jl's filter for resampling by a floating-point `rate` with an
    `n_phi`-phase filter bank: low-pass at f = (rate < 1 ? rate : 1)/n_phi of
    the bank's Nyquist frequency, transition 0.2·f, Kaiser window of the
    length kaiserord gives rounded up to a multiple of n_phi and made odd,
    taps w·sinc(w·(k − (n−1)/2))·window normalised to unit sum, times n_phi."""
    f_nyq = (rate if rate < 1.0 else 1.0) / n_phi
    cutoff = f_nyq * rel_bw
    n, alpha = kaiserord(cutoff * 0.2, attenuation)
    n = n_phi * math.ceil(n / n_phi)
    if n % 2 == 0:
        n += 1
    k = np.arange(n) - (n - 1) / 2.0
    h = cutoff * np.sinc(cutoff * k) * kaiser(n, alpha)
    return h / h.sum() * n_phi


def decimation_taps(rate: float, n_phi: int = 32) -> np.ndarray:
    """The taps DSP.resample(x, rate) applies for rate = 1/R (R integer):
    every output sample sits at the input sample R·k exactly, so the bank
    always uses the one phase whose taps are centred there — taps
    h[c mod n_phi :: n_phi] with c = (len(h) − 1)/2, an odd, symmetric filter
    centred on its middle tap."""
    h = resample_filter(rate, n_phi)
    c = (len(h) - 1) // 2
    taps = h[c % n_phi::n_phi]
    assert len(taps) % 2 == 1 and (c - c % n_phi) // n_phi == (len(taps) - 1) // 2
    return taps.copy()


# --------------------------------------------------------------------------
# PMTConfig
# --------------------------------------------------------------------------
@dataclass
class PMTConfig:
    """pipeline.jl:42-69."""
    spe_template: ExponTruncNormalSPE
    pulse_model: PDFPulseTemplate
    pulse_model_filt: InterpolatedPulse
    noise_amp: float
    sampling_freq: float        # GHz
    unf_pulse_res: float        # ns
    adc_freq: float             # GHz
    tt_dist: GammaDist
    lp_filter: LowpassFilter
    decim_taps: np.ndarray = field(repr=False, default=None)
    # make_waveform / unfold_waveform arguments of make_reco_pulses
    time_range: float = 300.0
    lead_time: float = 50.0
    min_charge: float = 0.2
    nnls_tol: float = 1e-8
    # unfolding factor rows kept in LDS on the device (0: the library's choice);
    # a tuning knob that does not change results
    lds_factor_rows: int = 0

    @classmethod
    def make(cls, st: ExponTruncNormalSPE, pm: PDFPulseTemplate, snr_db: float, sampling_freq: float,
             unf_pulse_res: float, adc_freq: float, lp_cutoff: float, tt_mean: float, tt_fwhm: float) -> "PMTConfig":
        """The reference's positional constructor (pipeline.jl:55-69)."""
        mode = get_template_mode(pm)
        lp = butterworth1_lowpass(lp_cutoff, sampling_freq)
        filtered = make_filtered_pulse(pm, sampling_freq, (-10.0, 50.0), lp)
        tt_theta = calc_gamma_shape_mean_fwhm(tt_mean, tt_fwhm)
        tt = GammaDist(tt_mean / tt_theta, tt_theta)
        return cls(st, pm, filtered, mode / 10 ** (snr_db / 10), float(sampling_freq), float(unf_pulse_res),
                   float(adc_freq), tt, lp, decimation_taps(adc_freq / sampling_freq))

    # sizes the configuration implies (waveforms.jl:35-53, 56-86, 128-147)
    @property
    def n_wave(self) -> int:
        return int(math.floor(self.time_range * self.sampling_freq + 1e-9)) + 1

    @property
    def n_adc(self) -> int:
        return int(math.floor(self.time_range * self.adc_freq + 1e-9)) + 1

    @property
    def n_pulse(self) -> int:
        return int(math.floor(self.time_range / self.unf_pulse_res + 1e-9)) + 1

    def to_abi(self):
        """The ntpmt_config struct; the returned tuple keeps its arrays alive."""
        taps = np.ascontiguousarray(self.decim_taps, dtype=np.float64)
        fp = np.ascontiguousarray(self.pulse_model_filt.values * self.pulse_model_filt.amplitude, dtype=np.float64)
        pm = self.pulse_model
        c = _abi.ntpmt_config()
        c.tt_shape, c.tt_scale, c.tt_subtract = self.tt_dist.alpha, self.tt_dist.theta, self.tt_dist.mean()
        s = self.spe_template
        c.spe_expon_weight, c.spe_expon_scale = s.expon_weight, s.expon_rate
        c.spe_norm_mu, c.spe_norm_sigma, c.spe_trunc_low = s.norm_mu, s.norm_sigma, s.trunc_low
        c.pulse_mu, c.pulse_beta = pm.dist.untruncated.mu, pm.dist.untruncated.theta
        c.pulse_lo, c.pulse_hi, c.pulse_amplitude = pm.dist.lower, pm.dist.upper, pm.amplitude
        c.noise_amp = self.noise_amp
        c.sampling_freq, c.adc_freq, c.unf_pulse_res = self.sampling_freq, self.adc_freq, self.unf_pulse_res
        c.time_range, c.lead_time, c.min_charge = self.time_range, self.lead_time, self.min_charge
        c.iir_b0, c.iir_b1, c.iir_a1 = self.lp_filter.b0, self.lp_filter.b1, self.lp_filter.a1
        c.nnls_tol, c.nnls_max_iter = self.nnls_tol, 30 * self.n_pulse
        c.decim_taps = taps.ctypes.data_as(C.POINTER(C.c_double))
        c.n_decim_taps = len(taps)
        c.filt_pulse = fp.ctypes.data_as(C.POINTER(C.c_double))
        c.n_filt_pulse = len(fp)
        c.filt_pulse_t0, c.filt_pulse_dt = self.pulse_model_filt.t0, self.pulse_model_filt.dt
        c.lds_factor_rows = int(self.lds_factor_rows)
        return c, (taps, fp)


@functools.lru_cache(maxsize=None)
def std_pmt_config() -> PMTConfig:
    """STD_PMT_CONFIG (pipeline.jl:71-82)."""
    spe = ExponTruncNormalSPE.make(expon_rate=1.0, norm_sigma=0.3, norm_mu=1.0, trunc_low=0.0, peak_to_valley=3.1)
    pulse = PDFPulseTemplate(dist=truncated(Gumbel(0.0, gumbel_width_from_fwhm(5.0)) + 4, 0.0, 20.0),
                             amplitude=1.0)
    return PMTConfig.make(spe, pulse, 20, 2.0, 0.1, 0.25, 0.125, 25, 1.5)


def __getattr__(name):
    if name == "STD_PMT_CONFIG":
        return std_pmt_config()
    raise AttributeError(name)


# --------------------------------------------------------------------------
# the device front end
# --------------------------------------------------------------------------
@dataclass
class RecoPulses:
    """Unfolded pulses of every channel (per pulse: module_id, pmt_id, time,
    charge, in channel then time order) and the per-channel table
    (channel_module_id, channel_pmt_id, n_hits, n_pulses, t0, status and,
    when asked for, the digitised waveform [channels][n_adc])."""
    module_id: np.ndarray
    pmt_id: np.ndarray
    time: np.ndarray
    charge: np.ndarray
    channel_module_id: np.ndarray
    channel_pmt_id: np.ndarray
    n_hits: np.ndarray
    n_pulses: np.ndarray
    t0: np.ndarray
    status: np.ndarray
    waveform: np.ndarray | None = None

    def __len__(self):
        return len(self.time)

    def channel(self, module_id: int, pmt_id: int):
        """(times, charges) of one channel (empty arrays if it has no hits)."""
        m = (self.module_id == module_id) & (self.pmt_id == pmt_id)
        return self.time[m], self.charge[m]


class PMTFrontEnd:
    """A device context holding one PMTConfig (ntpmt_create)."""

    def __init__(self, config: PMTConfig | None = None, device: int = 0):
        self.lib = _abi.load_library()
        self.config = config or std_pmt_config()
        self.device = device
        cfg, keep = self.config.to_abi()
        h = C.c_void_p()
        _abi.check(self.lib.ntpmt_create(device, C.byref(cfg), C.byref(h)), "ntpmt_create")
        self.handle = h
        nw, na, npl = C.c_int32(), C.c_int32(), C.c_int32()
        _abi.check(self.lib.ntpmt_sizes(self.handle, C.byref(nw), C.byref(na), C.byref(npl)), "ntpmt_sizes")
        self.n_wave, self.n_adc, self.n_pulse = nw.value, na.value, npl.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.ntpmt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def make_reco_pulses_device(self, cols: dict, n_hits: int, seed: int, select=None, time_column: str = "time",
                                capacity: int | None = None, waveforms: bool = False, stream=None):
        """Device form: `cols` holds torch device tensors (time or tres,
        module_id, pmt_id, source_index, photon_id); `select` an optional
        uint8 device tensor.  Returns a dict of device output tensors and the
        device channel count; nothing is synchronised."""
        import torch
        dev = torch.device(f"cuda:{self.device}")
        cap = int(capacity if capacity is not None else max(1, min(n_hits, 65536 * 17)))
        out = {
            "module_id": torch.zeros(cap, dtype=torch.int16, device=dev),
            "pmt_id": torch.zeros(cap, dtype=torch.uint8, device=dev),
            "n_hits": torch.zeros(cap, dtype=torch.int32, device=dev),
            "n_pulses": torch.zeros(cap, dtype=torch.int32, device=dev),
            "status": torch.zeros(cap, dtype=torch.int32, device=dev),
            "t0": torch.zeros(cap, dtype=torch.float64, device=dev),
            "time": torch.zeros(cap * self.n_adc, dtype=torch.float64, device=dev),
            "charge": torch.zeros(cap * self.n_adc, dtype=torch.float64, device=dev),
        }
        if waveforms:
            out["waveform"] = torch.zeros(cap * self.n_adc, dtype=torch.float64, device=dev)
        n_ch = torch.zeros(1, dtype=torch.int64, device=dev)
        h = _abi.nt_hits()
        for k, v in cols.items():
            setattr(h, k, v.data_ptr())
        h.capacity = n_hits
        p = _abi.ntpmt_pulses()
        for k, v in out.items():
            setattr(p, k, v.data_ptr())
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        tc = {"time": 0, "tres": 1}[time_column]
        _abi.check(self.lib.ntpmt_make_reco_pulses(
            self.handle, C.byref(h), C.c_void_p(select.data_ptr() if select is not None else None), int(n_hits), tc,
            C.c_uint64(seed & (2 ** 64 - 1)), cap, C.byref(p), C.c_void_p(n_ch.data_ptr()),
            C.c_void_p(s.cuda_stream)), "ntpmt_make_reco_pulses")
        return out, n_ch, cap

    def make_reco_pulses(self, hits, seed: int = 0, select=None, time_column: str = "time",
                         waveforms: bool = False) -> RecoPulses:
        """make_reco_pulses (pipeline.jl:146-160) for the host hits `hits`
        (a PhotonHits with module_id, pmt_id, source_index, photon_id and the
        time column), restricted to `select` (bool per hit, e.g. the detected
        flags of PropagationEngine.resample; None: all hits).  Every
        (module_id, pmt_id) present is a channel, pmt_id 0 (hits between the
        PMTs) included: leave those out with `select`, as the reference's
        per-PMT loops over PMTs 1-16 do."""
        import torch
        dev = torch.device(f"cuda:{self.device}")
        n = len(hits.time)
        up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt, copy=False)).to(dev)
        tcol = hits.time if time_column == "time" else hits.tres
        cols = {time_column: up(tcol, np.float32),
                "module_id": up(hits.module_id, np.uint16).view(torch.int16),
                "pmt_id": up(hits.pmt_id, np.uint8),
                "source_index": up(hits.source_index, np.uint32).view(torch.int32),
                "photon_id": up(hits.photon_id, np.uint64).view(torch.int64)}
        sel = None
        n_sel = n
        if select is not None:
            sm = np.asarray(select, dtype=bool)
            n_sel = int(sm.sum())
            sel = up(sm.astype(np.uint8), np.uint8)
        out, n_ch, cap = self.make_reco_pulses_device(cols, max(n, 0), seed, sel, time_column,
                                                      capacity=max(1, min(n_sel, 65536 * 17)), waveforms=waveforms)
        torch.cuda.synchronize(dev)
        nc = int(n_ch.item())
        assert nc <= cap
        h = {k: v.cpu().numpy() for k, v in out.items()}
        na = self.n_adc
        npl = h["n_pulses"][:nc].astype(np.int64)
        sel_idx = np.concatenate([c * na + np.arange(k) for c, k in enumerate(npl)]) if nc else np.zeros(0, np.int64)
        chm = h["module_id"][:nc].view(np.uint16)
        chp = h["pmt_id"][:nc]
        return RecoPulses(
            module_id=np.repeat(chm, npl), pmt_id=np.repeat(chp, npl),
            time=h["time"][sel_idx], charge=h["charge"][sel_idx],
            channel_module_id=chm, channel_pmt_id=chp, n_hits=h["n_hits"][:nc], n_pulses=h["n_pulses"][:nc],
            t0=h["t0"][:nc], status=h["status"][:nc],
            waveform=h["waveform"].reshape(cap, na)[:nc] if waveforms else None)


_FRONTENDS: dict = {}


def make_reco_pulses(hits, pmt_config: PMTConfig | None = None, *, seed: int = 0, select=None,
                     time_column: str = "time", device: int = 0, waveforms: bool = False) -> RecoPulses:
    """pipeline.jl:146-160 for every (module, PMT) channel of `hits` at once
    (the reference calls it once per channel's hits)."""
    cfg = pmt_config or std_pmt_config()
    key = (id(cfg), device)
    if key not in _FRONTENDS:
        _FRONTENDS[key] = (cfg, PMTFrontEnd(cfg, device))
    return _FRONTENDS[key][1].make_reco_pulses(hits, seed, select, time_column, waveforms)
