"""Pins the oracle's stats::p.adjust (every method test_de can pass,
R/test.R:127) by definitions independent of R's loops:

* the decision rule of each procedure: at any level a, {i : adj_i <= a} is
  the set the step-down / step-up / single-step rule rejects (Bonferroni,
  Holm, Hochberg, Benjamini-Hochberg, Benjamini-Yekutieli);
* closed testing by brute force over every subset for small n: Holm's
  adjusted p is max over I containing i of |I| min_I p (Bonferroni local
  tests), Hommel's is max over I containing i of the Simes p-value of I;
* scipy.stats.false_discovery_control for BH and BY;
* p.adjust's own rules: NaN stays NaN and n counts the others, n <= 1
  returns p, hommel with two p-values is hochberg.
"""
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.lib()
    return oracle


def _reject(p, a, method):
    """Indices the procedure rejects at level a (p without NaN)."""
    n = p.size
    o = np.argsort(p, kind="stable")
    ps = p[o]
    j = np.arange(1, n + 1)
    if method == "bonferroni":
        return set(np.flatnonzero(p <= a / n))
    if method == "holm":  # step-down
        k = 0
        while k < n and ps[k] <= a / (n - k):
            k += 1
        return set(o[:k])
    if method in ("hochberg", "BH", "BY"):  # step-up
        if method == "hochberg":
            crit = a / (n - j + 1)
        elif method == "BH":
            crit = j * a / n
        else:
            crit = j * a / (n * np.sum(1.0 / j))
        ok = np.flatnonzero(ps <= crit)
        return set(o[:ok.max() + 1]) if ok.size else set()
    raise ValueError(method)


@pytest.mark.parametrize("method", ["bonferroni", "holm", "hochberg", "BH", "BY"])
def test_decision_rules(orc, method):
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(2, 60))
        p = np.concatenate([rng.uniform(0, 0.05, n // 3), rng.uniform(0, 1, n - n // 3)])
        p[rng.integers(0, n, 3)] = p[0]  # ties
        adj = orc.p_adjust(p, method)
        assert np.all((adj >= p - 1e-15) & (adj <= 1.0))
        for a in rng.uniform(0, 0.3, 25):
            assert set(np.flatnonzero(adj <= a)) == _reject(p, a, method), (method, a)


def _simes(ps):
    s = np.sort(ps)
    return np.min(s.size * s / np.arange(1, s.size + 1))


@pytest.mark.parametrize("method", ["holm", "hommel"])
def test_closed_testing_brute_force(orc, method):
    rng = np.random.default_rng(4)
    for trial in range(30):
        n = int(rng.integers(2, 9))
        p = rng.uniform(0, 0.2, n)
        if trial % 3 == 0:
            p[1] = p[0]
        adj = orc.p_adjust(p, method)
        for i in range(n):
            worst = 0.0
            for r in range(1, n + 1):
                for I in itertools.combinations(range(n), r):
                    if i not in I:
                        continue
                    sub = p[list(I)]
                    loc = min(1.0, len(I) * sub.min()) if method == "holm" else _simes(sub)
                    worst = max(worst, loc)
            assert abs(adj[i] - worst) <= 1e-15 * max(worst, 1e-300) * 4, (method, p, i)


def test_scipy_fdr(orc):
    from scipy.stats import false_discovery_control
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 1, 5000) ** 3
    for m, sm in (("BH", "bh"), ("BY", "by")):
        np.testing.assert_allclose(orc.p_adjust(p, m), false_discovery_control(p, method=sm),
                                   rtol=1e-13, atol=0)


def test_p_adjust_rules(orc):
    p = np.array([0.01, np.nan, 0.04, 0.03, np.nan])
    for m in orc.P_ADJUST:
        adj = orc.p_adjust(p, m)
        assert np.isnan(adj[1]) and np.isnan(adj[4])
        # n = 3 (the non-NaN entries)
        np.testing.assert_array_equal(adj[[0, 2, 3]], orc.p_adjust(p[[0, 2, 3]], m))
        one = orc.p_adjust(np.array([0.3, np.nan]), m)
        assert one[0] == 0.3
    two = np.array([0.02, 0.04])
    np.testing.assert_array_equal(orc.p_adjust(two, "hommel"), orc.p_adjust(two, "hochberg"))
    np.testing.assert_array_equal(orc.p_adjust(two, "none"), two)
    np.testing.assert_array_equal(orc.p_adjust(two, "fdr"), orc.p_adjust(two, "BH"))
