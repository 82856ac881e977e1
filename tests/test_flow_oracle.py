"""Pins of the FP64 flow oracle (oracle/rqflow.py) without a GPU: the
reference's own invariants (test/runtests.jl:14-44), closed forms, and
independent derivatives (finite differences against the complex step and the
hand-back-propagated MLP)."""
import math

import numpy as np
import pytest

from oracle import rqflow as rq


@pytest.fixture(autouse=True)
def _quiet():
    with np.errstate(all="ignore"):
        yield


def test_roundtrip_and_logdet_symmetry_reference_case():
    """runtests.jl:14-23: K = 10, range [-5, 5], x = -10:0.1:10, params ~ N(0,1)."""
    rng = np.random.default_rng(1)
    k = 10
    x = np.arange(-100, 101) / 10.0
    p = rng.standard_normal((x.size, 3 * k + 1))
    xp, yp, d = rq.constrain_spline_params(p, -5.0, 5.0)
    y, ld = rq.rqs_forward(xp, yp, d, x)
    xr, ldi = rq.rqs_inverse(xp, yp, d, y)
    np.testing.assert_allclose(ld, -ldi, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(xr, x, atol=1e-5)
    # knots: ends at the range, increasing, slopes >= the minimum
    assert np.all(xp[:, 0] == -5.0) and np.all(xp[:, -1] == 5.0)
    assert np.all(np.diff(xp, axis=1) >= rq.MIN_BIN_SIZE * (1 - 1e-9))
    assert np.all(np.diff(yp, axis=1) >= rq.MIN_BIN_SIZE * (1 - 1e-9))
    assert np.all(d >= rq.MIN_KNOT_SLOPE)


def test_logdet_is_log_derivative():
    rng = np.random.default_rng(2)
    k = 8
    x = rng.uniform(-30, 120, 500)
    p = rng.standard_normal((x.size, 3 * k + 1)) * 2
    xp, yp, d = rq.constrain_spline_params(p)
    h = 1e-5
    y1, ld = rq.rqs_forward(xp, yp, d, x)
    yp_, _ = rq.rqs_forward(xp, yp, d, x + h)
    ym_, _ = rq.rqs_forward(xp, yp, d, x - h)
    inside = (np.abs(x[:, None] - xp) > 1e-3).all(axis=1)     # away from knots
    # central difference: truncation O(h²), cancellation ≈ ulp(y)/h ≈ 2e-9
    np.testing.assert_allclose(((yp_ - ym_) / (2 * h))[inside], np.exp(ld)[inside], rtol=1e-5, atol=1e-8)


def test_zero_params_give_identity():
    """All-zero parameters: equal bins and unit slopes (the softplus offset of
    rq_spline_flow.jl:151-152), i.e. the identity map with zero log-det."""
    k = 8
    x = np.linspace(-50, 150, 401)
    p = np.zeros((x.size, 3 * k + 1))
    xp, yp, d = rq.constrain_spline_params(p)
    np.testing.assert_allclose(d, 1.0, rtol=1e-12)
    y, ld = rq.rqs_forward(xp, yp, d, x)
    np.testing.assert_allclose(y, x, atol=1e-9)
    np.testing.assert_allclose(ld, 0.0, atol=1e-9)
    # the flow then reduces to the base normal N(shift, 100·σ(a))
    pf = np.zeros((x.size, rq.n_params(k, 3)))
    pf[:, -3] = 7.0
    pf[:, -2] = -1.0
    s = 100.0 / (1 + math.exp(1.0))
    lp = rq.flow_logpdf(x, pf, k, 3)
    ref = -math.log(s) - 0.5 * (math.log(2 * math.pi) + ((x - 7.0) / s) ** 2)
    np.testing.assert_allclose(lp, ref, atol=1e-9)


@pytest.mark.parametrize("layers", [1, 3])
def test_density_integrates_to_one(layers):
    """∫ exp(logpdf) dy = 1.  (runtests.jl:35-43 integrates over [-10, 10]
    only; with the 100·σ scale of rq_spline_flow.jl:406 the mass lies far
    outside that window, so the whole line is integrated here.)"""
    rng = np.random.default_rng(3 + layers)
    k = 8
    for _ in range(3):
        p = rng.standard_normal(rq.n_params(k, layers))
        p[-2] = -2.5                          # base scale ≈ 7.6
        # Quadrature nodes adapted to the map: composed splines have slopes
        # down to ~1e-5, i.e. density spikes far narrower than any fixed grid,
        # so the nodes are the images of a dense uniform grid of the base
        # variable.  The integrand itself is the oracle's logpdf at those nodes.
        x = np.linspace(-150, 150, 300001)
        pb = np.broadcast_to(p, (x.size, p.size))
        y, _ = rq.flow_sample_inverse_check(x, pb, k, layers)
        assert np.all(np.diff(y) > 0)
        lp = rq.flow_logpdf(y, pb, k, layers)
        integral = np.trapezoid(np.exp(lp), y)
        assert abs(integral - 1.0) < 1e-6, integral


def test_forward_inverse_roundtrip_stack():
    rng = np.random.default_rng(5)
    k, layers = 8, 3
    x = rng.normal(10, 40, 2000)
    p = rng.standard_normal((x.size, rq.n_params(k, layers)))
    y, ldf = rq.flow_sample_inverse_check(x, p, k, layers)
    z = y
    ldi = 0 * y
    m = 3 * k + 1
    for layer in range(layers):
        xp, yp, d = rq.constrain_spline_params(p[:, layer * m:(layer + 1) * m])
        z, l = rq.rqs_inverse(xp, yp, d, z)
        ldi = ldi + l
    np.testing.assert_allclose(z, x, atol=1e-6)
    np.testing.assert_allclose(ldi, -ldf, atol=1e-7)


def test_complex_step_matches_finite_differences():
    rng = np.random.default_rng(6)
    k, layers = 8, 3
    B = 64
    p = rng.standard_normal((B, rq.n_params(k, layers)))
    t = rng.uniform(-25, 150, B)
    n = rng.integers(0, 20, B).astype(float)
    flags = rng.integers(1, 4, B)
    g = rq.row_loss_grad(p, t, n, flags, k, layers)
    h = 1e-6
    for j in range(p.shape[1]):
        pp, pm = p.copy(), p.copy()
        pp[:, j] += h
        pm[:, j] -= h
        fd = (rq.row_loss(pp, t, n, flags, k, layers) - rq.row_loss(pm, t, n, flags, k, layers)) / (2 * h)
        np.testing.assert_allclose(g[:, j], fd, rtol=1e-4, atol=1e-5)


def test_poisson_term_gradient_closed_form():
    k, layers = 8, 1
    p = np.zeros((5, rq.n_params(k, layers)))
    p[:, -1] = np.log([0.5, 1, 2, 5, 10])
    n = np.array([0, 1, 2, 3, 4.5])
    g = rq.row_loss_grad(p, np.zeros(5), n, np.full(5, rq.ROW_COUNT), k, layers)
    np.testing.assert_allclose(g[:, -1], np.exp(p[:, -1]) - n, rtol=1e-12)
    assert np.all(g[:, :-1] == 0)


def test_mlp_backprop_matches_finite_differences():
    rng = np.random.default_rng(7)
    k, layers = 8, 3
    n_in, H, P = 5, 16, rq.n_params(k, layers)
    w = {"W1": rng.standard_normal((n_in, H)) * 0.5, "b1": rng.standard_normal(H) * 0.1,
         "W2": rng.standard_normal((H, H)) * 0.3, "b2": rng.standard_normal(H) * 0.1,
         "W3": rng.standard_normal((H, P)) * 0.3, "b3": rng.standard_normal(P) * 0.1}
    B = 32
    x = rng.standard_normal((B, n_in))
    t = rng.uniform(-10, 80, B)
    n = rng.integers(0, 6, B).astype(float)
    flags = rng.integers(1, 4, B)
    loss, g, _ = rq.loss_and_grad(x, t, n, flags, w, k, layers, use_bf16=False)
    h = 1e-6
    for name in ("W1", "b2", "W3"):
        flat = w[name].reshape(-1)
        for idx in rng.choice(flat.size, 6, replace=False):
            wp = {kk: vv.copy() for kk, vv in w.items()}
            wm = {kk: vv.copy() for kk, vv in w.items()}
            wp[name].reshape(-1)[idx] += h
            wm[name].reshape(-1)[idx] -= h
            lp_, _, _ = rq.loss_and_grad(x, t, n, flags, wp, k, layers, use_bf16=False)
            lm_, _, _ = rq.loss_and_grad(x, t, n, flags, wm, k, layers, use_bf16=False)
            fd = (lp_ - lm_) / (2 * h)
            assert abs(g[name].reshape(-1)[idx] - fd) <= 1e-5 + 1e-4 * abs(fd), (name, idx)


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    out = rq.bf16(vals)
    # bf16 ulp at 1 is 2^-7: 1 + 2^-8 is a tie → even (1); 1 + 2^-9 rounds
    # down; 1 + 3·2^-9 is above the midpoint and rounds up
    assert out[0] == 1.0 and out[3] == 1.0
    assert out[1] == 1.0
    assert out[2] == 1.0 + 2 ** -7
    assert out[4] == -2.5
