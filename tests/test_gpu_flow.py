"""Flow kernels (libntflow) against the FP64 oracle (oracle/rqflow.py):
forward outputs, per-row losses, the per-row flow gradient, the full weight
gradient through the bfloat16 MFMA products, and training."""
import math

import numpy as np
import pytest

from oracle import rqflow as rq

pytestmark = pytest.mark.gpu

K, L = 8, 3


@pytest.fixture(scope="module")
def flow(engine):
    from neutrinotelescopes_amd.flow import FlowModel
    m = FlowModel(0, seed=1)
    yield m
    m.close()


def _weights(seed, scale=1.0):
    from neutrinotelescopes_amd import flow as F
    rng = np.random.default_rng(seed)
    flat = F.init_weights(seed)
    w = F.unflatten(flat.copy())
    w["b1"][:] = rng.normal(0, 0.1, w["b1"].shape)
    w["b2"][:] = rng.normal(0, 0.1, w["b2"].shape)
    w["b3"][:F.N_PARAMS] = rng.normal(0, 0.5, F.N_PARAMS)
    w["W3"] *= scale
    return F.flatten(w)


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 5)).astype(np.float32)
    t = rng.uniform(-30, 150, n).astype(np.float32)
    cnt = rng.poisson(3.0, n).astype(np.float32)
    flags = rng.integers(1, 4, n).astype(np.uint8)
    return x, t, cnt, flags


def _oracle_w(flat):
    from neutrinotelescopes_amd import flow as F
    w = F.unflatten(flat.astype(np.float64))
    return {"W1": w["W1"][:5], "b1": w["b1"], "W2": w["W2"], "b2": w["b2"],
            "W3": w["W3"][:, :F.N_PARAMS], "b3": w["b3"][:F.N_PARAMS]}


def test_forward_matches_oracle(flow):
    from neutrinotelescopes_amd.flow import FlowRows
    flat = _weights(3)
    flow.set_weights(flat)
    x, t, cnt, flags = _rows(1000, 4)          # not a multiple of the 64-row tile
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    p, loss = flow.forward(rows)
    p = p.cpu().numpy().astype(np.float64)
    ref = rq.mlp_forward(x, _oracle_w(flat))["p"]
    assert np.all(p[:, 78:] == 0.0)
    err = np.abs(p[:, :78] - ref) / (1.0 + np.abs(ref))
    print("forward max rel err", err.max(), "median", np.median(err))
    assert np.median(err) < 1e-6
    assert err.max() < 1e-4          # a bf16 rounding flip of one hidden unit stays below this
    # per-row loss of the kernel's own outputs against the FP64 flow
    with np.errstate(all="ignore"):
        lref = rq.row_loss(p[:, :78], t.astype(np.float64), cnt.astype(np.float64), flags, K, L)
    le = np.abs(loss.cpu().numpy() - lref) / (1.0 + np.abs(lref))
    print("row loss max rel err", le.max(), "median", np.median(le))
    assert np.median(le) < 1e-6
    assert np.quantile(le, 0.999) < 1e-4


def test_row_gradient_matches_complex_step(flow):
    """d loss / d p of single rows (read back as the b3 gradient of a one-row
    batch) against the oracle's complex-step derivative at the same p."""
    from neutrinotelescopes_amd.flow import FlowRows
    flat = _weights(5)
    flow.set_weights(flat)
    x, t, cnt, flags = _rows(64, 6)
    flags[:] = 3
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    p, _ = flow.forward(rows)
    p = p.cpu().numpy().astype(np.float64)[:, :78]
    with np.errstate(all="ignore"):
        gref = rq.row_loss_grad(p, t.astype(np.float64), cnt.astype(np.float64), flags, K, L)
    rel = []
    for r in range(64):
        _, g = flow.loss_grad(rows.slice(slice(r, r + 1)), 1.0)
        gb3 = g[-80:][:78].astype(np.float64)
        rel.append(np.linalg.norm(gb3 - gref[r]) / max(np.linalg.norm(gref[r]), 1e-6))
    rel = np.array(rel)
    print("row gradient rel err: median", np.median(rel), "max", rel.max())
    assert np.median(rel) < 1e-5
    assert np.quantile(rel, 0.95) < 1e-3


def test_offset_and_weight_rows(flow):
    """Rows with (log λ offset, weight): loss and per-row gradient against the
    oracle with the same offset and weight."""
    from neutrinotelescopes_amd.flow import FlowRows
    flat = _weights(9)
    flow.set_weights(flat)
    x, t, cnt, flags = _rows(48, 10)
    rng = np.random.default_rng(11)
    off = rng.uniform(-8, 12, 48).astype(np.float32)
    wgt = np.exp(-off).astype(np.float32)
    cnt = rng.poisson(np.exp(off + 0.3)).astype(np.float32)
    rows = FlowRows.from_numpy(x, t, cnt, flags, np.column_stack([off, wgt]))
    p, loss = flow.forward(rows)
    p = p.cpu().numpy().astype(np.float64)[:, :78]
    args = (t.astype(np.float64), cnt.astype(np.float64), flags, K, L, off.astype(np.float64),
            wgt.astype(np.float64))
    with np.errstate(all="ignore"):
        lref = rq.row_loss(p, *args)
        gref = rq.row_loss_grad(p, *args)
    np.testing.assert_allclose(loss.cpu().numpy(), lref, rtol=1e-4, atol=1e-4)
    for r in range(0, 48, 3):
        _, g = flow.loss_grad(rows.slice(slice(r, r + 1)), 1.0)
        gb3 = g[-80:][:78].astype(np.float64)
        assert np.linalg.norm(gb3 - gref[r]) <= 1e-3 * max(np.linalg.norm(gref[r]), 1e-3), r


def test_weight_gradient_matches_oracle(flow):
    from neutrinotelescopes_amd import flow as F
    from neutrinotelescopes_amd.flow import FlowRows
    flat = _weights(7, scale=0.5)
    flow.set_weights(flat)
    x, t, cnt, flags = _rows(512, 8)
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    loss, g = flow.loss_grad(rows, 1.0 / 512)
    with np.errstate(all="ignore"):
        lref, gref, _ = rq.loss_and_grad(x.astype(np.float64), t.astype(np.float64), cnt.astype(np.float64),
                                         flags, _oracle_w(flat), K, L, use_bf16=True)
    assert abs(loss / 512 - lref) < 1e-4 * (1 + abs(lref))
    gw = F.unflatten(g.astype(np.float64))
    gw = {"W1": gw["W1"][:5], "b1": gw["b1"], "W2": gw["W2"], "b2": gw["b2"],
          "W3": gw["W3"][:, :78], "b3": gw["b3"][:78]}
    for name in gref:
        a, b = gw[name].ravel(), gref[name].ravel()
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(name, "rel", rel, "cos", cos)
        assert rel < 1e-3 and cos > 0.99999, name
    # padding of the weight vector receives no gradient
    full = F.unflatten(g)
    assert np.all(full["W1"][5:] == 0) and np.all(full["W3"][:, 78:] == 0) and np.all(full["b3"][78:] == 0)


def _synthetic(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    mu = 10 + 20 * x[:, 0]
    sd = 3 + 2 * (x[:, 1] + 1)
    t = rng.normal(mu, sd).astype(np.float32)
    lam = np.exp(1.0 + x[:, 2])
    cnt = rng.poisson(lam).astype(np.float32)
    flags = np.where(rng.random(n) < 0.5, 1, 2).astype(np.uint8)
    return x, t, cnt, flags, mu, sd, lam


def test_training_learns_synthetic_density_and_rate(flow):
    import torch
    from neutrinotelescopes_amd.flow import FlowRows, init_weights
    flow.set_weights(init_weights(11))
    x, t, cnt, flags, *_ = _synthetic(1 << 17, 12)
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    xt, tt, ct, ft, mu, sd, lam = _synthetic(4096, 13)
    test = FlowRows.from_numpy(xt, tt, ct, ft)
    _, l0 = flow.forward(test, want_params=False)
    loss0 = float(l0.mean())
    B = 16384
    g = torch.Generator(device="cpu").manual_seed(0)
    for step in range(600):
        idx = torch.randint(0, rows.n, (B,), generator=g).to(rows.t.device)
        flow.train_step(rows.slice(idx), 3e-3)
    p, l1 = flow.forward(test)
    loss1 = float(l1.mean())
    # expected loss of the true model: 0.5 per row type mix of Gaussian and Poisson NLL
    with np.errstate(all="ignore"):
        true_nll = np.where(ft == 1, 0.5 * np.log(2 * np.pi * sd ** 2) + 0.5 * ((tt - mu) / sd) ** 2,
                            lam - ct * np.log(lam) + np.vectorize(math.lgamma)(ct + 1.0))
    print("loss before", loss0, "after", loss1, "true model", true_nll.mean())
    assert loss1 < loss0 - 1.0
    assert loss1 < true_nll.mean() + 0.1
    loglam = p.cpu().numpy()[:, 77]
    assert np.median(np.abs(loglam - np.log(lam))) < 0.1


def test_pair_rows_and_time_logpdf(flow):
    """Device pair rows against the host features (surrogate.pmt_features) and
    the per-hit time logpdf against the forward pass's time-row loss."""
    import torch
    from neutrinotelescopes_amd import surrogate as S
    from neutrinotelescopes_amd.flow import FlowRows
    flow.set_weights(_weights(12))
    rng = np.random.default_rng(13)
    casc = S.random_cascades(7, rng)
    modules = S.array_detector()[::9]
    axes = S.pom_axes()
    f = S.pair_features(casc, modules).reshape(-1, 5)
    norm = S.FeatureNorm(f.mean(0), f.std(0) + 0.1)
    scale = S.CountScale(np.linspace(0, 6, 9), rng.normal(-20, 3, 9))
    src = np.zeros((7, 8), np.float32)
    src[:, :3], src[:, 3:6], src[:, 6] = casc.pos, casc.dir, np.log10(casc.energy)
    dev = flow.dev
    x, aux = flow.pair_rows(torch.from_numpy(src).to(dev), torch.from_numpy(modules.astype(np.float32)).to(dev),
                            torch.from_numpy(axes.astype(np.float32)).to(dev), norm.mean, norm.std,
                            np.column_stack([scale.log_d, scale.log_g]), float(S.log_yield(0.0)))
    x = x.cpu().numpy()
    np.testing.assert_allclose(x[:, :5], norm(f), atol=2e-4)
    assert np.all(x[:, 5:] == 0)
    a = aux.cpu().numpy()
    np.testing.assert_allclose(a[:, 0], scale(f), rtol=1e-5, atol=1e-3)
    assert np.all(a[:, 1] == 1.0)
    # time logpdf from the pair outputs == −loss of time rows with the same features
    n_pairs = x.shape[0]
    p, _ = flow.forward(FlowRows.from_numpy(x, flags=np.full(n_pairs, 2, np.uint8)))
    hp = rng.integers(0, n_pairs, 3000).astype(np.int32)
    ht = rng.uniform(-30, 400, 3000).astype(np.float32)
    lp = flow.time_logpdf(p, torch.from_numpy(hp).to(dev), torch.from_numpy(ht).to(dev)).cpu().numpy()
    _, l = flow.forward(FlowRows.from_numpy(x[hp], ht, None, np.full(3000, 1, np.uint8)), want_params=False)
    np.testing.assert_allclose(lp, -l.cpu().numpy(), rtol=1e-6, atol=1e-6)
    bad = flow.time_logpdf(p, torch.tensor([-1, n_pairs], dtype=torch.int32, device=dev),
                           torch.zeros(2, device=dev)).cpu().numpy()
    assert np.all(np.isneginf(bad))


def test_train_step_gather_equals_gathered_rows(flow):
    """ntf_train_step_gather (the kernel reads the batch rows in place) gives
    bit for bit the weights and losses of ntf_train_step on the gathered
    batch, over several Adam steps; an index outside the arrays counts as an
    absent row (the same as a row of no kind)."""
    import torch
    from neutrinotelescopes_amd.flow import FlowModel, FlowRows, init_weights
    x, t, cnt, flags = _rows(5000, 21)
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    a, b = FlowModel(0, seed=1), FlowModel(0, seed=1)
    try:
        a.set_weights(init_weights(22))
        b.set_weights(init_weights(22))
        la = torch.zeros(1, dtype=torch.float64, device=rows.t.device)
        lb = torch.zeros_like(la)
        g = torch.Generator(device="cpu").manual_seed(3)
        for step in range(4):
            idx = torch.randint(0, rows.n, (1000,), generator=g).to(rows.t.device)
            sl = rows.slice(idx)
            if step == 3:                       # one index outside the arrays
                idx[5] = rows.n
                sl.flags[5] = 0
            a.train_step(sl, 1e-3, la)
            b.train_step_gather(rows, idx, 1e-3, lb)
            assert float(la.cpu()[0]) == float(lb.cpu()[0])
        assert np.array_equal(a.get_weights(), b.get_weights())
    finally:
        a.close()
        b.close()
