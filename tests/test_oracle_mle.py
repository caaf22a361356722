"""Pins the oracle's restatement of fit_dispersion's stats::nlminb call
(R/overdispersion.R:25-115) on the CPU.

What fixes it without R:
* the log-likelihood in its cancellation-free form equals the reference's
  lgamma-sum form (src/overdispersion.cpp:53-88) where that one is well
  conditioned, and lgamma(a+y) - lgamma(a) - y log a equals a 50-digit mpmath
  evaluation;
* the score is its derivative in log(theta) (central differences);
* with rel.tol -> 0 nlminb's iteration converges to the root of the score
  (the root driver's answer);
* at rel.tol = 1e-3, where every step is a Newton step inside the trust
  region, nlminb's par is fixed by its documented stopping rule: Newton steps
  x_{k+1} = x_k - g_k/h_k from log(start) until the first k whose predicted
  reduction g_k^2 / (2 h_k) is at most rel.tol |f(x_k)|, and par = x_{k+1};
* the reference's own scenarios (test-overdispersion.R:47-152) with numpy
  draws of the same distributions (R's set.seed streams cannot be
  regenerated here).
"""
import math

import numpy as np
import pytest


def nb_sample(rng, mu, alpha, size=None):
    lam = rng.gamma(1.0 / alpha, alpha * np.asarray(mu, dtype=np.float64), size=size)
    return rng.poisson(lam).astype(np.float64)


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.lib()
    return oracle


def test_lgamma_diff_mpmath(orc):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    for a in [1e-3, 0.3, 2.0, 9.5, 10.0, 37.0, 1e3, 1e6, 1e9]:
        for y in [1, 2, 5, 16, 17, 40, 1000, 2.5]:
            ref = mp.loggamma(mp.mpf(a) + y) - mp.loggamma(a) - y * mp.log(a)
            got = orc.lgamma_diff(a, y)
            scale = max(abs(float(ref)), y * 1e-3, 1e-300)
            assert abs(got - float(ref)) <= 1e-13 * scale, (a, y, got, float(ref))


def test_loglik_acc_matches_reference_form(orc):
    rng = np.random.default_rng(3)
    n = 300
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    beta = np.array([1.0, 0.3])
    mu = np.exp(X @ beta)
    y = nb_sample(rng, mu, 0.4)
    table = orc.make_table(y, n // 2)
    for lt in [-3.0, -1.0, 0.0, 1.5]:
        ref = orc.nb_loglik(y, mu, lt, X, True, table)
        got = orc.nb_loglik_acc(y, mu, lt, X, True)
        assert abs(got - ref) <= 1e-10 * abs(ref), (lt, got, ref)
        # without the Cox-Reid term too
        ref0 = orc.nb_loglik(y, mu, lt, X, False, table)
        assert abs(orc.nb_loglik_acc(y, mu, lt, X, False) - ref0) <= 1e-10 * abs(ref0)


def test_score_is_loglik_derivative(orc):
    rng = np.random.default_rng(4)
    n = 500
    X = np.column_stack([np.ones(n), rng.normal(size=n), rng.integers(0, 2, n)])
    beta = np.array([0.2, 0.4, -0.3])
    mu = np.exp(X @ beta)
    y = nb_sample(rng, mu, 0.7)
    for lt in [-6.0, -2.0, -0.4, 1.0]:
        h = 1e-4
        fd = (orc.nb_loglik_acc(y, mu, lt + h, X) - orc.nb_loglik_acc(y, mu, lt - h, X)) / (2 * h)
        s = orc.nb_score_acc(y, mu, lt, X)
        assert abs(fd - s) <= 1e-6 * max(1.0, abs(s)), (lt, fd, s)


def test_port_converges_to_root_at_tiny_rel_tol(orc):
    rng = np.random.default_rng(5)
    for alpha, n, mu0 in [(0.5, 2000, 3.0), (0.05, 3000, 8.0), (2.0, 1500, 0.5)]:
        X = np.column_stack([np.ones(n), rng.normal(size=n)])
        beta = np.array([math.log(mu0), 0.2])
        y = nb_sample(rng, np.exp(X @ beta), alpha)
        off = np.zeros(n)
        root, _ = orc.fit_dispersion(beta, X, y, off, 1e-3, 200)
        th, it, code, nf = orc.fit_dispersion_port(beta, X, y, off, 1e-12, 200)
        assert code in (3, 4, 5), code
        assert abs(th - root) <= 1e-6 * root, (alpha, th, root)


def test_port_stopping_rule_newton_path(orc):
    """Genes whose nlminb run takes only Newton steps inside the trust region:
    par follows from the stopping rule alone."""
    rng = np.random.default_rng(6)
    checked = 0
    for trial in range(40):
        n = int(rng.integers(500, 4000))
        X = np.column_stack([np.ones(n), rng.integers(0, 2, n), rng.normal(size=n)])
        beta = np.array([rng.normal(0.5, 1.0), rng.normal(0, 0.5), rng.normal(0, 0.2)])
        alpha = float(np.exp(rng.normal(-1.0, 1.0)))
        mu = np.exp(X @ beta)
        y = nb_sample(rng, mu, alpha)
        off = np.zeros(n)
        ybar, v = y.mean(), y.var(ddof=1)
        start = (v - ybar) / ybar ** 2
        if not (start > 0):
            start = 0.5
        if orc.nb_score_acc(y, mu, math.log(1e-8), X) < 0:
            continue
        x = math.log(start)
        f = -orc.nb_loglik_acc(y, mu, x, X)
        newton_only = True
        for k in range(50):
            g = -orc.nb_score_acc(y, mu, x, X)
            h = -orc.nb_dscore_acc(y, mu, x, X)
            if not h > 0 or abs(g / h) > (1.0 if k == 0 else 1e300):
                newton_only = False
                break
            xn = x - g / h
            fn = -orc.nb_loglik_acc(y, mu, xn, X)
            pred = 0.5 * g * g / h
            if not (f - fn > 0.1 * pred and f - fn < 0.75 * abs(g * (xn - x))):
                newton_only = False  # not the plain accept-with-unchanged-radius path
                break
            if pred <= 1e-3 * abs(f):
                break
            x, f = xn, fn
        if not newton_only:
            continue
        th, it, code, nf = orc.fit_dispersion_port(beta, X, y, off, 1e-3, 200)
        assert code == 4 and it == k + 1, (code, it, k)
        # mu differs in the last bit between numpy and the oracle loop
        assert abs(math.log(th) - xn) <= 1e-10 * max(1.0, abs(xn))
        checked += 1
    assert checked >= 10


def test_reference_scenarios(orc):
    """test-overdispersion.R:47-152 (tolerance 1e-8, max_iter 50 / 100)."""
    # all-zero counts -> exactly 0
    n = 20
    th, it, code, nf = orc.fit_dispersion_port([0.0], np.ones((n, 1)), np.zeros(n), np.zeros(n),
                                                1e-8, 50)
    assert th == 0.0
    rng = np.random.default_rng(10)
    # near-Poisson: small dispersion
    n = 200
    y = rng.poisson(10, n).astype(float)
    X = np.ones((n, 1))
    th, _, code, _ = orc.fit_dispersion_port([math.log(10)], X, y, np.zeros(n), 1e-8, 100)
    assert np.isfinite(th) and th >= 0 and th < 0.2
    # NB data has the larger dispersion, near the true 0.5 (relative 0.3)
    rng = np.random.default_rng(11)
    n = 400
    y_pois = rng.poisson(10, n).astype(float)
    y_nb = nb_sample(rng, np.full(n, 10.0), 0.5)
    X = np.ones((n, 1))
    d_p = orc.fit_dispersion_port([math.log(10)], X, y_pois, np.zeros(n), 1e-8, 100)[0]
    d_nb = orc.fit_dispersion_port([math.log(10)], X, y_nb, np.zeros(n), 1e-8, 100)[0]
    assert np.isfinite(d_nb) and d_nb > d_p
    assert abs(d_nb - 0.5) <= 0.3 * 0.5
    # the Cox-Reid flag changes theta
    rng = np.random.default_rng(12)
    n = 200
    y = nb_sample(rng, np.full(n, 5.0), 0.8)
    X = np.ones((n, 1))
    d_cr = orc.fit_dispersion_port([math.log(5)], X, y, np.zeros(n), 1e-8, 100, cr=True)[0]
    d_ncr = orc.fit_dispersion_port([math.log(5)], X, y, np.zeros(n), 1e-8, 100, cr=False)[0]
    assert np.isfinite(d_cr) and np.isfinite(d_ncr) and d_cr != d_ncr
    # the same scenarios with the root driver
    d_cr_root = orc.fit_dispersion([math.log(5)], X, y, np.zeros(n), 1e-8, 100, cr=True)[0]
    d_ncr_root = orc.fit_dispersion([math.log(5)], X, y, np.zeros(n), 1e-8, 100, cr=False)[0]
    assert d_cr_root != d_ncr_root
    assert abs(d_cr - d_cr_root) <= 1e-4 * d_cr_root


def test_cpu_fit_mle_drivers(orc):
    """cpu_fit runs the nlminb restatement by default; its passes are the
    far-left check plus the objective evaluations."""
    rng = np.random.default_rng(13)
    G, n = 12, 1500
    X = np.column_stack([np.ones(n), rng.integers(0, 2, n)])
    beta = np.column_stack([rng.normal(0.5, 1, G), rng.normal(0, 0.5, G)])
    Y = np.vstack([nb_sample(rng, np.exp(X @ b), 0.3) for b in beta])
    off = np.zeros(n)
    a = orc.cpu_fit(Y, X, off, od_mode=orc.OD_MLE)
    r = orc.cpu_fit(Y, X, off, od_mode=orc.OD_MLE, mle_driver="root")
    np.testing.assert_array_equal(a["beta"], r["beta"])
    for g in range(G):
        th, it, code, nf = orc.fit_dispersion_port(a["beta"][g], X, Y[g], off, 1e-3, 200)
        assert a["theta"][g] == th
        assert a["theta_iters"][g] == (nf + 1 if th > 0 or nf > 0 else 0)
        assert r["theta"][g] == orc.fit_dispersion(a["beta"][g], X, Y[g], off, 1e-3, 200)[0]
