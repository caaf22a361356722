"""Pins the oracle's special functions and test_de arithmetic (CPU only).

Fixed points: 50-digit mpmath vectors (tests/golden/special_vectors.json),
scipy.special / scipy.stats, and a brute-force restatement of the BH step-up
rule from its definition.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.special as sp
import scipy.stats as st

GOLD = os.path.join(os.path.dirname(__file__), "golden", "special_vectors.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


def test_digamma_trigamma_golden(orc, gold):
    for x, v in gold["digamma"]:
        ref = float(v)
        assert abs(orc.digamma(x) - ref) <= 1e-15 * max(1.0, abs(ref)), x
    for x, v in gold["trigamma"]:
        ref = float(v)
        assert abs(orc.trigamma(x) - ref) <= 1e-15 * abs(ref), x


def test_digamma_trigamma_scipy_grid(orc):
    xs = np.concatenate([np.logspace(-9, 7, 400), np.linspace(0.01, 25, 400)])
    for x in xs:
        d = orc.digamma(x)
        assert abs(d - sp.digamma(x)) <= 1e-15 * max(1.0, abs(d)) + 1e-15
        t = orc.trigamma(x)
        assert abs(t - sp.polygamma(1, x)) <= 2e-15 * t


def test_lbeta_golden(orc, gold):
    for a, b, v in gold["lbeta"]:
        ref = float(v)
        assert abs(orc.lbeta(a, b) - ref) <= 1e-14 * max(1.0, abs(ref)), (a, b)


def test_t_tail_golden(orc, gold):
    # log of 2*pt(t, df, lower=FALSE); the continued fraction is accurate to
    # ~1e-9 relative near its crossover at df ~ 5e5, far inside the 1e-6
    # p-value parity bar.
    for t, df, v in gold["pt2_log"]:
        ref = float(v)
        p, lp = orc.pt2_upper(t, df)
        assert abs(lp - ref) <= 2e-9 * max(1.0, abs(ref)), (t, df, lp, ref)
        assert abs(p - math.exp(ref)) <= 2e-9 * math.exp(ref) + 1e-300


def test_t_tail_scipy_and_edges(orc):
    rng = np.random.default_rng(0)
    for _ in range(300):
        df = float(rng.choice([1, 2, 5, 48, 998, 4998, 99998]))
        t = float(np.exp(rng.uniform(-6, 4)))
        p, lp = orc.pt2_upper(t, df)
        ref = math.log(2) + st.t.logsf(t, df)
        assert abs(lp - ref) <= 2e-9 * max(1.0, abs(ref)), (t, df)
    assert orc.pt2_upper(0.0, 10)[0] == 1.0
    assert orc.pt2_upper(math.inf, 10)[0] == 0.0
    assert math.isnan(orc.pt2_upper(math.nan, 10)[0])
    assert math.isnan(orc.pt2_upper(1.0, 0.0)[0])
    # symmetric in t
    assert orc.pt2_upper(-2.5, 7)[0] == orc.pt2_upper(2.5, 7)[0]
    # deep tail stays finite in log space where p underflows
    p, lp = orc.pt2_upper(200.0, 499998)
    assert p == 0.0 and lp < -19000 and math.isfinite(lp)


def bh_bruteforce(p):
    """Benjamini-Hochberg from its step-up definition, O(n^2):
    adj_i = min(1, min_{j : p_j >= p_i} n * p_j / rank_j), rank over non-NaN."""
    p = np.asarray(p, float)
    ok = ~np.isnan(p)
    v = p[ok]
    n = v.size
    ranks = np.array([np.sum(v <= x) for x in v])  # rank with ties at max
    adj = np.array([min(1.0, min(n * v[j] / ranks[j] for j in range(n) if v[j] >= x))
                    for x in v])
    out = np.full(p.shape, np.nan)
    out[ok] = adj
    return out


def test_bh_against_definition(orc):
    rng = np.random.default_rng(1)
    for trial in range(20):
        n = int(rng.integers(1, 60))
        p = rng.uniform(0, 1, n) ** 3
        if trial % 3 == 0:
            p[rng.integers(0, n, 3)] = np.nan
        if trial % 4 == 0:
            p[: n // 2] = np.round(p[: n // 2], 1)  # ties
        got = orc.p_adjust_bh(p)
        ref = bh_bruteforce(p)
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=0, equal_nan=True)


def test_test_de_lfc_clipping_and_bh_order(orc):
    # the reference test-test.R:32-62 case: beta rows (0,.5),(0,-1),(0,4),(0,-5)
    beta = np.array([[0, 0.5], [0, -1], [0, 4], [0, -5.0]])
    nsamples = 6
    bread = np.stack([np.eye(2) / nsamples] * 4)
    r = orc.de_test(beta, bread, None, nsamples, [0, 1], max_lfc=2)
    expected = np.clip(beta @ np.array([0, 1.0]) / np.log(2), -2, 2)
    np.testing.assert_allclose(r["lfc"], expected, rtol=0, atol=0)
    # p-values: 2*pt(|l|/sqrt(1/6), df=4)
    tt = np.abs(beta[:, 1]) / np.sqrt(1 / 6)
    np.testing.assert_allclose(r["pval"], 2 * st.t.sf(tt, 4), rtol=1e-12)
    assert np.all(r["adj_pval"] >= r["pval"])


def test_test_de_clustered_takes_max(orc):
    # test-test.R:88-118: a 5x larger clustered sandwich gives larger p-values
    rng = np.random.default_rng(4)
    G, p, n = 6, 2, 12
    beta = rng.normal(size=(G, p))
    b1 = np.stack([np.eye(p) / n] * G)
    r0 = orc.de_test(beta, b1, None, n, [0, 1])
    r1 = orc.de_test(beta, b1, 5 * b1, n, [0, 1])
    assert np.all(r1["pval"] >= r0["pval"])
    r2 = orc.de_test(beta, 5 * b1, b1, n, [0, 1])  # max() is symmetric
    np.testing.assert_array_equal(r1["pval"], r2["pval"])


def test_size_factors_definitions():
    """calculate_sf (R/size_factors.R:37-127) by its definitions: column sums
    over their geometric mean; mean minus minimum of log1p over genes, over
    its geometric mean; scale-free and gene-permutation-free."""
    import oracle
    rng = np.random.default_rng(8)
    Y = rng.poisson(3.0, (40, 25)).astype(float) + (rng.random((40, 25)) < 0.05)
    Y[:, Y.sum(axis=0) == 0] = 1.0
    cs = Y.sum(axis=0)
    np.testing.assert_allclose(oracle.size_factors(Y), cs / np.exp(np.mean(np.log(cs))),
                               rtol=1e-14)
    lg = np.log1p(Y)
    ps = lg.mean(axis=0) - lg.min(axis=0)
    np.testing.assert_allclose(oracle.size_factors(Y, "psinorm"),
                               ps / np.exp(np.mean(np.log(ps))), rtol=1e-13)
    np.testing.assert_allclose(oracle.size_factors(7.0 * Y), oracle.size_factors(Y), rtol=1e-13)
    perm = rng.permutation(40)
    np.testing.assert_allclose(oracle.size_factors(Y[perm], "psinorm"),
                               oracle.size_factors(Y, "psinorm"), rtol=1e-13)
    assert np.all(oracle.size_factors(Y[:1]) == 1.0)
    Y0 = Y.copy()
    Y0[:, 3] = 0
    with pytest.raises(ValueError):
        oracle.size_factors(Y0)


def test_group_data_definition():
    """group_data (R/group_data.R:38-69): codes by first appearance, a stable
    order, and every array permuted alike."""
    import oracle
    Y = np.arange(2 * 7).reshape(2, 7)
    X = np.arange(7)[:, None] * 1.0
    cl = ["b", "a", "b", "c", "a", "b", "c"]
    Yr, Xr, cr, order = oracle.group_data(Y, X, cl)
    np.testing.assert_array_equal(order, [0, 2, 5, 1, 4, 3, 6])
    np.testing.assert_array_equal(cr, [1, 1, 1, 2, 2, 3, 3])
    np.testing.assert_array_equal(Yr, Y[:, order])
    np.testing.assert_array_equal(Xr[:, 0], order)
