"""rqflow.py — FP64 numpy restatement of the reference's rational-quadratic
spline flow and its training loss, the checker for the flow kernels
(libntflow).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's baseline legs,
never by the library or any product path.

Reference (paths relative to the reference repository root):
  constrain_spline_params, _normalize_bin_sizes, _normalize_knot_slopes
      src/surrogate_models/rq_spline_flow.jl:127-216
  safe_quadratic_root          rq_spline_flow.jl:215-254
  rqs_univariate               rq_spline_flow.jl:262-324
  inv_rqs_univariate           rq_spline_flow.jl:334-389
  eval_transformed_normal_logpdf  rq_spline_flow.jl:396-428
  RQNormFlow (params → last row log-expectation)  extended_cascade_model.jl:53-125
  log_likelihood_with_poisson  extended_cascade_model.jl:132-139

Generalisation: `n_layers` splines are composed (y = S_1(S_2(…S_L(x)))),
each with its own 3K+1 parameters; n_layers = 1 is exactly the reference's
eval_transformed_normal_logpdf.  The parameter vector of one row is
[spline_1 … spline_L, shift, scale logit, log λ] (the reference's order).

Every function accepts complex arrays: branches and comparisons use the real
part, so the complex-step derivative  Im f(p + i·h) / h  (h = 1e-30) gives the
exact derivative of the FP64 function, independently of the hand-written
backward of the kernels.

MLP (extended_cascade_model.jl:71-97 with ReLU): P = relu(relu(X W1 + b1) W2 + b2) W3 + b3.
`bf16=True` rounds at the points where the kernels round to bfloat16 (round to
nearest even): X, W1, W2, W3, the hidden activations, and the back-propagated
dP, dH2, dH1 before they enter a matrix product; biases and the bias gradients
stay FP32/FP64.
"""
from __future__ import annotations

import math

import numpy as np

RANGE_MIN = -20.0          # setup_time_expectation_model, extended_cascade_model.jl:166-177
RANGE_MAX = 100.0
MIN_BIN_SIZE = 1e-4        # constrain_spline_params defaults, rq_spline_flow.jl:170
MIN_KNOT_SLOPE = 1e-4
LOG_2PI = math.log(2.0 * math.pi)

ROW_TIME = 1               # row flags: time term present
ROW_COUNT = 2              #            Poisson count term present


def n_params(n_bins: int, n_layers: int) -> int:
    return n_layers * (3 * n_bins + 1) + 3


# ---------------------------------------------------------------------------
# spline parameters
# ---------------------------------------------------------------------------
def _softmax(u):
    m = np.max(u.real, axis=-1, keepdims=True)
    e = np.exp(u - m)
    return e / np.sum(e, axis=-1, keepdims=True)


def _softplus(x):
    # log(1 + e^x) without overflow (real part decides the branch)
    big = x.real > 30.0
    safe = np.where(big, 0.0, x)
    return np.where(big, x, np.log1p(np.exp(safe)))


def normalize_bin_sizes(u, total_size, min_bin_size=MIN_BIN_SIZE):
    """rq_spline_flow.jl:127-137."""
    k = u.shape[-1]
    assert k * min_bin_size <= total_size
    return _softmax(u) * (total_size - k * min_bin_size) + min_bin_size


def normalize_knot_slopes(u, min_knot_slope=MIN_KNOT_SLOPE):
    """rq_spline_flow.jl:146-155."""
    offset = math.log(math.exp(1.0 - min_knot_slope) - 1.0)
    return _softplus(u + offset) + min_knot_slope


def constrain_spline_params(p, range_min=RANGE_MIN, range_max=RANGE_MAX):
    """rq_spline_flow.jl:170-216.  p: [..., 3K+1] → x_pos, y_pos, slopes [..., K+1]."""
    k = (p.shape[-1] - 1) // 3
    size = range_max - range_min
    bw = normalize_bin_sizes(p[..., :k], size)
    bh = normalize_bin_sizes(p[..., k:2 * k], size)
    lead = p.shape[:-1]
    lo = np.full(lead + (1,), range_min, dtype=p.dtype)
    hi = np.full(lead + (1,), range_max, dtype=p.dtype)
    xp = np.concatenate([lo, range_min + np.cumsum(bw[..., :-1], axis=-1), hi], axis=-1)
    yp = np.concatenate([lo, range_min + np.cumsum(bh[..., :-1], axis=-1), hi], axis=-1)
    return xp, yp, normalize_knot_slopes(p[..., 2 * k:])


def _bin(knots, v):
    """Index of the bin [knots_j, knots_j+1) holding v; the first bin when none
    does (rq_spline_flow.jl:271-281)."""
    k = knots.shape[-1] - 1
    r = knots.real
    inb = (v.real[..., None] >= r[..., :-1]) & (v.real[..., None] < r[..., 1:])
    idx = np.argmax(inb, axis=-1)
    return np.where(inb.any(axis=-1), idx, 0).clip(0, k - 1)


def _take(a, idx):
    return np.take_along_axis(a, idx[..., None], axis=-1)[..., 0]


def _clamp01(z):
    return np.where(z.real < 0.0, 0.0 * z, np.where(z.real > 1.0, 0.0 * z + 1.0, z))


def safe_quadratic_root(a, b, c):
    """rq_spline_flow.jl:215-254."""
    disc = b * b - 4.0 * a * c
    pos = disc.real > 0.0
    s = np.sqrt(np.where(pos, disc, 1.0))
    s = np.where(pos, s, 0.0 * s)
    bpos = b.real >= 0.0
    num = np.where(bpos, 2.0 * c, -b + s)
    den = np.where(bpos, -b - s, 2.0 * a)
    return num / den


def rqs_forward(xp, yp, d, x):
    """rqs_univariate, rq_spline_flow.jl:262-324: y, log|dy/dx|."""
    j = _bin(xp, x)
    x0, x1 = _take(xp, j), _take(xp, j + 1)
    y0, y1 = _take(yp, j), _take(yp, j + 1)
    d0, d1 = _take(d, j), _take(d, j + 1)
    w, h = x1 - x0, y1 - y0
    s = h / w
    z = _clamp01((x - x0) / w)
    sq_z, z1mz, sq_1mz = z * z, z - z * z, (1.0 - z) ** 2
    st = d1 + d0 - 2.0 * s
    num = h * (s * sq_z + d0 * z1mz)
    den = s + st * z1mz
    y = y0 + num / den
    ld = 2.0 * np.log(s) + np.log(d1 * sq_z + 2.0 * s * z1mz + d0 * sq_1mz) - 2.0 * np.log(den)
    below = x.real <= xp[..., 0].real
    above = x.real >= xp[..., -1].real
    y = np.where(below, (x - xp[..., 0]) * d[..., 0] + yp[..., 0], y)
    y = np.where(above, (x - xp[..., -1]) * d[..., -1] + yp[..., -1], y)
    ld = np.where(below, np.log(d[..., 0]), ld)
    ld = np.where(above, np.log(d[..., -1]), ld)
    return y, ld


def rqs_inverse(xp, yp, d, y):
    """inv_rqs_univariate, rq_spline_flow.jl:334-389: x, log|dx/dy|."""
    j = _bin(yp, y)
    x0, x1 = _take(xp, j), _take(xp, j + 1)
    y0, y1 = _take(yp, j), _take(yp, j + 1)
    d0, d1 = _take(d, j), _take(d, j + 1)
    w, h = x1 - x0, y1 - y0
    s = h / w
    wv = _clamp01((y - y0) / h)
    st = d1 + d0 - 2.0 * s
    c = -s * wv
    b = d0 - st * wv
    a = s - b
    z = _clamp01(safe_quadratic_root(a, b, c))
    x = w * z + x0
    sq_z, z1mz, sq_1mz = z * z, z - z * z, (1.0 - z) ** 2
    den = s + st * z1mz
    ld = -2.0 * np.log(s) - np.log(d1 * sq_z + 2.0 * s * z1mz + d0 * sq_1mz) + 2.0 * np.log(den)
    below = y.real <= yp[..., 0].real
    above = y.real >= yp[..., -1].real
    x = np.where(below, (y - yp[..., 0]) / d[..., 0] + xp[..., 0], x)
    x = np.where(above, (y - yp[..., -1]) / d[..., -1] + xp[..., -1], x)
    ld = np.where(below, -np.log(d[..., 0]), ld)
    ld = np.where(above, -np.log(d[..., -1]), ld)
    return x, ld


def flow_logpdf(y, p, n_bins, n_layers, range_min=RANGE_MIN, range_max=RANGE_MAX):
    """eval_transformed_normal_logpdf (rq_spline_flow.jl:396-428) for a stack of
    n_layers splines: z_0 = y, z_l = S_l^{-1}(z_{l-1}), x = z_L,
    logpdf = log N(x; shift, 100·σ(scale logit)) + Σ log|dz_l/dz_{l-1}|.
    p: [..., n_params] (the trailing log λ is ignored)."""
    m = 3 * n_bins + 1
    z = y
    ld = 0.0 * y
    for layer in range(n_layers):
        xp, yp, d = constrain_spline_params(p[..., layer * m:(layer + 1) * m], range_min, range_max)
        z, l = rqs_inverse(xp, yp, d, z)
        ld = ld + l
    shift = p[..., n_layers * m]
    scale = 100.0 / (1.0 + np.exp(-p[..., n_layers * m + 1]))
    return -np.log(scale) - 0.5 * (LOG_2PI + ((z - shift) / scale) ** 2) + ld


def flow_sample_inverse_check(x, p, n_bins, n_layers, range_min=RANGE_MIN, range_max=RANGE_MAX):
    """Forward map of the stack (x → y), the inverse of the transformation
    flow_logpdf undoes; used by the round-trip checks."""
    m = 3 * n_bins + 1
    z = x
    ld = 0.0 * x
    for layer in reversed(range(n_layers)):
        xp, yp, d = constrain_spline_params(p[..., layer * m:(layer + 1) * m], range_min, range_max)
        z, l = rqs_forward(xp, yp, d, z)
        ld = ld + l
    return z, ld


def _lgamma(n):
    return np.vectorize(math.lgamma, otypes=[float])(np.asarray(n, dtype=float) + 1.0)


def row_loss(p, t, n, flags, n_bins, n_layers, offset=0.0, weight=1.0):
    """Per-row negative log-likelihood: −[time·logpdf(t) + count·(n log λ − λ −
    lgamma(n+1))] (extended_cascade_model.jl:132-139 with the terms switched per
    row by `flags`), with log λ = p_λ + offset, times the row weight."""
    lp = flow_logpdf(t, p, n_bins, n_layers)
    loglam = p[..., -1] + offset
    pois = n * loglam - np.exp(loglam) - _lgamma(n)
    ft = (flags & ROW_TIME) != 0
    fc = (flags & ROW_COUNT) != 0
    return -weight * (np.where(ft, lp, 0.0 * lp) + np.where(fc, pois, 0.0 * pois))


def row_loss_grad(p, t, n, flags, n_bins, n_layers, offset=0.0, weight=1.0, h=1e-30):
    """d row_loss / d p by complex step (exact to FP64 rounding)."""
    p = np.asarray(p, dtype=np.float64)
    g = np.empty_like(p)
    for k in range(p.shape[-1]):
        pc = p.astype(np.complex128)
        pc[..., k] += 1j * h
        g[..., k] = row_loss(pc, t, n, flags, n_bins, n_layers, offset, weight).imag / h
    return g


# ---------------------------------------------------------------------------
# MLP with the kernels' bfloat16 rounding points
# ---------------------------------------------------------------------------
def bf16(a):
    """Round FP32 values to bfloat16 (round to nearest even), returned as FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def mlp_forward(x, w, use_bf16=True):
    """x: [B, n_in]; w: dict W1 [n_in, H], b1 [H], W2 [H, H], b2, W3 [H, P], b3."""
    xq = _r(x, use_bf16)
    a1 = xq @ _r(w["W1"], use_bf16) + w["b1"]
    h1 = np.maximum(a1, 0.0)
    h1q = _r(h1, use_bf16)
    a2 = h1q @ _r(w["W2"], use_bf16) + w["b2"]
    h2 = np.maximum(a2, 0.0)
    h2q = _r(h2, use_bf16)
    p = h2q @ _r(w["W3"], use_bf16) + w["b3"]
    return {"xq": xq, "h1": h1, "h1q": h1q, "h2": h2, "h2q": h2q, "p": p}


def loss_and_grad(x, t, n, flags, w, n_bins, n_layers, use_bf16=True, offset=0.0, weight=1.0):
    """Mean row loss over the batch and its gradient for every weight (FP64
    back-propagation through the MLP, complex-step gradient of the flow)."""
    f = mlp_forward(x, w, use_bf16)
    p = f["p"]
    B = x.shape[0]
    loss = row_loss(p, t, n, flags, n_bins, n_layers, offset, weight)
    dp = row_loss_grad(p, t, n, flags, n_bins, n_layers, offset, weight) / B
    dpq = _r(dp, use_bf16)
    g = {"W3": f["h2q"].T @ dpq, "b3": dp.sum(axis=0)}
    dh2 = (dpq @ _r(w["W3"], use_bf16).T) * (f["h2"] > 0)
    dh2q = _r(dh2, use_bf16)
    g["W2"] = f["h1q"].T @ dh2q
    g["b2"] = dh2.sum(axis=0)
    dh1 = (dh2q @ _r(w["W2"], use_bf16).T) * (f["h1"] > 0)
    dh1q = _r(dh1, use_bf16)
    g["W1"] = f["xq"].T @ dh1q
    g["b1"] = dh1.sum(axis=0)
    return float(np.mean(loss)), g, p


def ref_mlp_forward(cond, w1, b1, w2, b2):
    """The embedding of the reference's saved RQNormFlow (extended_cascade_model.jl:
    71-97; with hidden_structure [768, 768] the loop at :77 adds no second hidden
    layer): Dense(24 → H, relu), Dropout (inactive at inference), Dense(H → 3K + 4),
    evaluated in FP64 as `m.embedding(Float64.(cond))` asks (:109).
    cond: [N, 24]; w1: [H, 24], w2: [3K + 4, H] (Flux's [out, in] layout)."""
    x = np.asarray(cond, dtype=np.float64)
    h = np.maximum(x @ np.asarray(w1, np.float64).T + np.asarray(b1, np.float64), 0.0)
    return h @ np.asarray(w2, np.float64).T + np.asarray(b2, np.float64)


def adam_update(w, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Flux.Optimise.Adam (the reference's optimiser, extended_cascade_model.jl:203)
    with bias correction; step counts from 1."""
    out = {}
    for k in w:
        m[k] = beta1 * m[k] + (1 - beta1) * g[k]
        v[k] = beta2 * v[k] + (1 - beta2) * g[k] ** 2
        mh = m[k] / (1 - beta1 ** step)
        vh = v[k] / (1 - beta2 ** step)
        out[k] = w[k] - lr * mh / (np.sqrt(vh) + eps)
    return out
