"""Pins the resampling oracle (oracle/resample.py), the checker of
nt_resample: its Philox against the C oracle (and so the published vectors),
the law of weighted sampling without replacement (inclusion probabilities of
successive draws, computed exactly), the Poisson group counts, and the cap at
the number of photons with positive weight.  CPU only."""
import itertools

import numpy as np
import pytest
from scipy import stats

import _oracle as O
from oracle import resample as R


def test_vectorised_philox_equals_c_oracle():
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, (500, 4), dtype=np.uint64)
    key = [int(x) for x in rng.integers(0, 2 ** 32, 2)]
    a = R.philox_np(ctr[:, 0], ctr[:, 1], ctr[:, 2], ctr[:, 3], key[0], key[1])
    b = np.array([O.philox([int(x) for x in c], key) for c in ctr], np.uint64)
    for j in range(4):
        assert np.array_equal(a[j], b[:, j])


def exact_inclusion(w, n):
    """P(i in sample) for n successive draws without replacement ∝ w, by
    enumerating the ordered samples."""
    w = np.asarray(w, float)
    p = np.zeros(len(w))
    for seq in itertools.permutations(np.flatnonzero(w > 0), n):
        pr, left = 1.0, w.sum()
        for i in seq:
            pr *= w[i] / left
            left -= w[i]
        p[list(seq)] += pr
    return p


@pytest.mark.parametrize("n", [1, 2, 3])
def test_inclusion_probabilities_of_weighted_sampling(n):
    w = np.array([0.5, 0.25, 0.15, 0.1, 0.0, 0.3], np.float32)
    groups = 30000
    g = np.repeat(np.arange(groups), len(w))
    weight = np.tile(w, groups)
    pid = np.arange(g.size, dtype=np.uint64)
    k = R.keys(weight, np.zeros(g.size, np.uint32), pid, seed=5)
    sel, drawn = R.select_top(g, k, weight, lambda grp, wsum: n)
    assert all(v == n for v in drawn.values())
    freq = sel.reshape(groups, len(w)).mean(axis=0)
    exact = exact_inclusion(w, n)
    assert freq[4] == 0.0                                  # weight 0: never drawn
    sigma = np.sqrt(exact * (1 - exact) / groups)
    assert np.all(np.abs(freq - exact) <= 5 * sigma + 1e-12), (freq, exact)
    # a with-replacement or uniform draw would fail: they differ from the exact law by far more
    assert np.abs(exact - n * w / w.sum()).max() > 10 * sigma.max() or n == 1


def test_group_counts_are_poisson_and_capped():
    groups, per = 3000, 100
    g = np.repeat(np.arange(groups), per)
    weight = np.full(g.size, 0.2, np.float32)                # Σw = 20 per group
    k = R.keys(weight, np.zeros(g.size, np.uint32), np.arange(g.size, dtype=np.uint64), seed=9)
    sel, drawn = R.select_top(g, k, weight, lambda grp, wsum: R.P.poisson(wsum, R._philox_scalar, 9, grp,
                                                                         R.COUNT_WORD))
    n = np.array([drawn[i] for i in range(groups)])
    assert np.array_equal(sel.reshape(groups, per).sum(axis=1), n)
    assert n.mean() == pytest.approx(20, abs=4 * np.sqrt(20 / groups))
    assert n.var() == pytest.approx(20, rel=0.1)
    # binned χ² against the Poisson pmf (tails pooled below 10 and above 30)
    edges = np.arange(10, 32)
    obs = np.histogram(np.clip(n, 10, 30), bins=edges)[0]
    pmf = stats.poisson.pmf(edges[:-1], 20)
    pmf[0] += stats.poisson.cdf(9, 20)
    pmf[-1] += stats.poisson.sf(30, 20)
    assert stats.chisquare(obs, pmf * groups / pmf.sum()).pvalue > 1e-3
    # Σw = 50 but 3 photons with positive weight: n is capped at 3
    wcap = np.array([20.0, 20.0, 10.0, 0.0, 0.0], np.float32)
    g2 = np.repeat(np.arange(400), 5)
    w2 = np.tile(wcap, 400)
    k2 = R.keys(w2, np.zeros(g2.size, np.uint32), np.arange(g2.size, dtype=np.uint64), seed=3)
    sel2, d2 = R.select_top(g2, k2, w2, lambda grp, wsum: R.P.poisson(wsum, R._philox_scalar, 3, grp,
                                                                      R.COUNT_WORD))
    assert all(v == 3 for v in d2.values())
    assert not sel2.reshape(400, 5)[:, 3:].any()


def test_resample_independent_of_hit_order():
    rng = np.random.default_rng(4)
    n = 5000
    mod = rng.integers(1, 6, n).astype(np.uint16)
    pmt = rng.integers(0, 17, n).astype(np.uint8)
    w = rng.exponential(0.2, n).astype(np.float32)
    w[rng.random(n) < 0.1] = 0
    src = rng.integers(0, 3, n).astype(np.uint32)
    pid = rng.permutation(n).astype(np.uint64)
    a, _ = R.resample(mod, pmt, w, src, pid, seed=17)
    perm = rng.permutation(n)
    b, _ = R.resample(mod[perm], pmt[perm], w[perm], src[perm], pid[perm], seed=17)
    assert np.array_equal(a[perm], b)
    assert not a[w == 0].any()
