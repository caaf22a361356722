"""Pins the oracle's fit functions on the CPU against what the mathematics fixes.

* beta_fit: for intercept-only and two-group designs with a constant offset the
  reference update (src/beta.cpp:38-42) reduces to the scalar recursion
  eta <- eta + 1 - (exp(eta + off) + k) / (k + ybar) per group, so both the
  early-stopped iterate and the iteration count are known in closed form; for
  general designs the fixed point solves the NB score equation
  sum_i x_i (y_i - mu_i) / (1 + mu_i / k) = 0.
* bread: inverse of the negative NB log-likelihood Hessian in beta (central
  second differences); Poisson case (X' diag(mu) X)^-1.
* clustered meat: cluster score sums are gradients of the per-cluster NB
  log-likelihood (finite differences); one cell per cluster gives
  n/(n-1) * HC0.
* NB score / dscore: central differences of the log-likelihood in log(theta).
* init_beta: numpy least squares (test-beta.R:41-77), intercept-only mean
  (test-beta.R:19-39), column scaling (test-beta.R:79-100).
* estimate_dispersion: test-overdispersion.R:1-45.
"""
import math

import numpy as np
import pytest


def nb_sample(rng, mu, alpha, size=None):
    """y ~ NB(mean mu, dispersion alpha) as Poisson(Gamma(1/alpha, alpha*mu))."""
    lam = rng.gamma(1.0 / alpha, alpha * np.asarray(mu, dtype=np.float64), size=size)
    return rng.poisson(lam).astype(np.float64)


def kfloat(alpha0):
    """k as src/beta.cpp:25-28 holds it: float(1.0 / float(alpha0))."""
    return float(np.float32(1.0 / float(np.float32(alpha0))))


def scalar_irls(ybar_groups, off, k, eta0, max_iter, eps):
    """Closed-form per-group recursion of the reference update; returns the
    per-group eta and the iteration count (convergence on the beta scale of
    a [1, g] design: |d_A| and |d_B - d_A|)."""
    eta = np.array(eta0, dtype=np.float64)
    epsd = float(np.float32(eps))
    it = 0
    while it < max_iter:
        d = 1.0 - (np.exp(eta + off) + k) / (k + np.asarray(ybar_groups))
        eta = eta + d
        it += 1
        dbeta = [d[0]] if d.size == 1 else [d[0], d[1] - d[0]]
        if max(abs(v) for v in dbeta) < epsd:
            break
    return eta, it


@pytest.mark.parametrize("alpha0,mu,off", [(0.37, 3.0, 0.0), (1.9, 0.2, 0.4),
                                           (0.05, 40.0, -0.7)])
def test_beta_fit_intercept_recursion(orc, alpha0, mu, off):
    rng = np.random.default_rng(7)
    n = 500
    y = nb_sample(rng, mu * np.exp(off), alpha0, size=n)
    X = np.ones((n, 1))
    offv = np.full(n, off)
    k = kfloat(alpha0)
    eta0 = math.log1p(y.mean() / math.exp(off))
    b, it = orc.beta_fit(y, X, [eta0], offv, alpha0, 200, 1e-3)
    eta, it_ref = scalar_irls([y.mean()], off, k, [eta0], 200, 1e-3)
    assert it == it_ref
    assert abs(b[0] - eta[0]) <= 1e-12 * max(1, abs(eta[0]))
    # tight tolerance -> the NB MLE log(ybar) - off
    b2, _ = orc.beta_fit(y, X, [eta0], offv, alpha0, 5000, 1e-30)
    assert abs(b2[0] - (math.log(y.mean()) - off)) < 1e-12


def test_beta_fit_two_group_recursion(orc):
    rng = np.random.default_rng(8)
    n = 600
    g = (rng.uniform(size=n) < 0.4).astype(float)
    mu = np.where(g > 0, 0.3, 2.5)
    alpha0 = 0.8
    y = nb_sample(rng, mu, alpha0)
    X = np.column_stack([np.ones(n), g])
    off = 0.25
    k = kfloat(alpha0)
    yA, yB = y[g == 0].mean(), y[g == 1].mean()
    beta0 = np.array([0.1, -0.2])
    b, it = orc.beta_fit(y, X, beta0, np.full(n, off), alpha0, 200, 1e-3)
    eta, it_ref = scalar_irls([yA, yB], off, k, [beta0[0], beta0[0] + beta0[1]], 200, 1e-3)
    assert it == it_ref
    np.testing.assert_allclose(b, [eta[0], eta[1] - eta[0]], rtol=1e-11, atol=1e-12)
    b2, _ = orc.beta_fit(y, X, beta0, np.full(n, off), alpha0, 20000, 1e-30)
    np.testing.assert_allclose(b2, [math.log(yA) - off, math.log(yB / yA)], rtol=0,
                               atol=1e-11)


def test_beta_fit_float_k_semantics(orc):
    # alpha0 whose reciprocal is not a float: the recursion only matches when
    # k is float-rounded exactly as the reference stores it.
    rng = np.random.default_rng(9)
    n = 300
    y = nb_sample(rng, np.full(n, 0.05), 3.0)
    y[0] = 1.0
    X = np.ones((n, 1))
    alpha0 = 3.0000001234
    eta0 = [math.log1p(y.mean())]
    b, it = orc.beta_fit(y, X, eta0, np.zeros(n), alpha0, 200, 1e-3)
    eta_f, it_f = scalar_irls([y.mean()], 0.0, kfloat(alpha0), eta0, 200, 1e-3)
    eta_d, _ = scalar_irls([y.mean()], 0.0, 1.0 / alpha0, eta0, 200, 1e-3)
    assert it == it_f
    assert abs(b[0] - eta_f[0]) < 1e-13
    assert abs(eta_f[0] - eta_d[0]) > 1e-10  # the rounding matters at this scale


def nb_score_beta(X, y, mu, alpha):
    return X.T @ ((y - mu) / (1.0 + alpha * mu))


def test_beta_fit_fixed_point_is_score_root(orc):
    rng = np.random.default_rng(10)
    n, p = 2000, 4
    X = np.column_stack([np.ones(n), rng.uniform(size=n) < 0.5, rng.normal(size=(n, 2))])
    sf = np.exp(rng.normal(0, 0.3, n))
    beta_true = np.array([0.5, -0.4, 0.2, 0.1])
    alpha0 = 0.6
    y = nb_sample(rng, sf * np.exp(X @ beta_true), alpha0)
    b, it = orc.beta_fit(y, X, np.zeros(p), np.log(sf), alpha0, 10000, 1e-30)
    k = kfloat(alpha0)
    mu = sf * np.exp(X @ b)
    s = nb_score_beta(X, y, mu, 1.0 / k)
    assert np.max(np.abs(s)) < 1e-9 * n
    assert np.all(np.abs(b - beta_true) < 0.2)


def test_beta_fit_nan_stops(orc):
    # k -> inf (alpha0 = 0) makes m = inf/inf = NaN: the loop stops after one
    # iteration with NaN (src/beta.cpp:45).
    n = 50
    y = np.arange(n) % 3.0
    b, it = orc.beta_fit(y, np.ones((n, 1)), [0.0], np.zeros(n), 0.0, 200, 1e-3)
    assert it == 1 and math.isnan(b[0])


def nb_loglik_beta(beta, X, y, sf, alpha):
    mu = sf * np.exp(X @ beta)
    if alpha == 0:
        return np.sum(y * np.log(mu) - mu)
    return np.sum(y * np.log(mu) - (y + 1 / alpha) * np.log1p(alpha * mu))


def fd_hessian(f, x, h=1e-4):
    p = x.size
    H = np.zeros((p, p))
    for a in range(p):
        for b in range(p):
            ea, eb = np.eye(p)[a] * h, np.eye(p)[b] * h
            H[a, b] = (f(x + ea + eb) - f(x + ea - eb) - f(x - ea + eb) + f(x - ea - eb)) / (4 * h * h)
    return H


@pytest.mark.parametrize("alpha", [0.0, 0.3, 2.0])
def test_bread_is_inverse_nb_information(orc, alpha):
    rng = np.random.default_rng(11)
    n, p = 300, 3
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    sf = np.exp(rng.normal(0, 0.3, n))
    beta = np.array([1.0, 0.3, -0.2])
    y = nb_sample(rng, sf * np.exp(X @ beta), 0.5)
    bread = orc.compute_hessian(beta, alpha, y, X, sf)
    H = fd_hessian(lambda b: nb_loglik_beta(b, X, y, sf, alpha), beta)
    np.testing.assert_allclose(bread, np.linalg.inv(-H), rtol=2e-5)
    np.testing.assert_allclose(bread, bread.T, rtol=1e-13, atol=0)
    if alpha == 0.0:
        mu = sf * np.exp(X @ beta)
        np.testing.assert_allclose(bread, np.linalg.inv(X.T @ (mu[:, None] * X)), rtol=1e-12)


def fd_grad(f, x, h=1e-5):
    return np.array([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(x.size)])


def test_clustered_meat_from_cluster_scores(orc):
    rng = np.random.default_rng(12)
    n, p = 240, 3
    X = np.column_stack([np.ones(n), rng.uniform(size=n) < 0.5, rng.normal(size=n)])
    sf = np.exp(rng.normal(0, 0.3, n))
    beta = np.array([0.7, 0.2, -0.1])
    alpha = 0.4
    y = nb_sample(rng, sf * np.exp(X @ beta), alpha)
    sizes = [30, 70, 1, 40, 99]
    ends = np.cumsum(sizes)
    meat = orc.clustered_meat(X, y, beta, alpha, sf, ends)
    S = []
    start = 0
    for e in ends:
        sl = slice(start, e)
        S.append(fd_grad(lambda b: nb_loglik_beta(b, X[sl], y[sl], sf[sl], alpha), beta))
        start = e
    S = np.array(S)
    C = len(sizes)
    ref = C / (C - 1) / n * S.T @ S
    np.testing.assert_allclose(meat, ref, rtol=1e-6)
    # a single block has adj = 1
    m1 = orc.clustered_meat(X, y, beta, alpha, sf, [n])
    s1 = fd_grad(lambda b: nb_loglik_beta(b, X, y, sf, alpha), beta)
    np.testing.assert_allclose(m1, np.outer(s1, s1) / n, rtol=1e-6)
    # one cell per cluster: n/(n-1) * HC0
    m_each = orc.clustered_meat(X, y, beta, alpha, sf, np.arange(1, n + 1))
    np.testing.assert_allclose(m_each, n / (n - 1) * orc.meat_hc0(X, y, beta, alpha, sf),
                               rtol=1e-12)


def test_sandwich_clustered_vs_unclustered(orc):
    # test-variance.R:58-82 properties on our own data: symmetry of the
    # sandwich bread*meat*bread*n and PSD of the meat.
    rng = np.random.default_rng(13)
    n, p = 120, 2
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    beta = rng.normal(size=p) * 0.3
    y = rng.poisson(10, n).astype(float)
    sf = np.ones(n)
    br = orc.compute_hessian(beta, 1.0, y, X, sf)
    m = orc.clustered_meat(X, y, beta, 1.0, sf, [40, 80, 120])
    S = br @ m @ br * n
    np.testing.assert_allclose(S, S.T, rtol=1e-12)
    assert np.all(np.linalg.eigvalsh(m) >= -1e-12)


def test_mom_dispersion_consistent(orc):
    # with the true mean, (y-mu)^2 - mu has expectation alpha*mu^2: the MOM
    # estimate approaches alpha; y == mu exactly gives a negative numerator -> 0.
    rng = np.random.default_rng(14)
    n = 400000
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    sf = np.exp(rng.normal(0, 0.3, n))
    beta = np.array([1.0, 0.3])
    alpha = 0.5
    y = nb_sample(rng, sf * np.exp(X @ beta), alpha)
    th = orc.mom_dispersion(y, X, beta, sf)
    assert abs(th - alpha) < 0.03
    mu = sf * np.exp(X @ beta)
    assert orc.mom_dispersion(mu, X, beta, sf) == 0.0
    # definition on a tiny input: n/(n-p) * sum((y-mu)^2 - mu) / sum(mu^2)
    yt = np.array([0.0, 3, 1, 8, 2, 0, 5])
    Xt = np.column_stack([np.ones(7), np.arange(7) / 7.0])
    bt = np.array([0.4, 0.9])
    mt = np.ones(7) * np.exp(Xt @ bt)
    ref = 7 / 5 * np.sum((yt - mt) ** 2 - mt) / np.sum(mt * mt)
    assert abs(orc.mom_dispersion(yt, Xt, bt, np.ones(7)) - max(ref, 0)) < 1e-14


def test_nb_score_is_loglik_derivative(orc):
    rng = np.random.default_rng(15)
    n, p = 400, 3
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    mu = np.exp(X @ np.array([1.2, 0.3, -0.2]))
    y = nb_sample(rng, mu, 0.7)
    table = orc.make_table(y, n // 2)
    assert table[0].size > 0
    for cr in (True, False):
        for lt in [-3.0, -0.8, 0.0, 1.5]:
            h = 1e-5
            f = lambda t: orc.nb_loglik(y, mu, t, X, cr)
            fd = (f(lt + h) - f(lt - h)) / (2 * h)
            s = orc.nb_score(y, mu, lt, X, cr)
            assert abs(s - fd) <= 1e-6 * max(1.0, abs(fd)), (cr, lt, s, fd)
            g = lambda t: orc.nb_score(y, mu, t, X, cr)
            fd2 = (g(lt + h) - g(lt - h)) / (2 * h)
            d = orc.nb_dscore(y, mu, lt, X, cr)
            assert abs(d - fd2) <= 1e-5 * max(1.0, abs(fd2)), (cr, lt, d, fd2)
            # the frequency-table branch gives the same numbers
            assert abs(orc.nb_score(y, mu, lt, X, cr, table) - s) <= 1e-10 * max(1, abs(s))
            assert abs(orc.nb_dscore(y, mu, lt, X, cr, table) - d) <= 1e-10 * max(1, abs(d))
            assert abs(orc.nb_loglik(y, mu, lt, X, cr, table) - f(lt)) <= 1e-10 * abs(f(lt))


def test_dpsi_differences_mpmath(orc):
    import mpmath as mp
    mp.mp.dps = 50
    for a in [1e-8, 0.3, 9.9, 10.0, 1e3, 1e7, 1e12]:
        for y in [1, 3, 16, 17, 250, 2.5, 4e4]:
            r = mp.digamma(mp.mpf(a) + y) - mp.digamma(mp.mpf(a))
            r1 = mp.polygamma(1, mp.mpf(a) + y) - mp.polygamma(1, mp.mpf(a))
            assert abs(orc.dpsi_diff(a, y) - float(r)) <= 2e-14 * abs(float(r))
            assert abs(orc.dpsi1_diff(a, y) - float(r1)) <= 2e-14 * abs(float(r1))
    assert orc.dpsi_diff(5.0, 0.0) == 0.0 and orc.dpsi1_diff(5.0, 0.0) == 0.0


def nb_score_mpmath(y, mu, log_theta, X):
    """The Cox-Reid adjusted score of src/overdispersion.cpp:95-178 at 50 digits
    (no Taylor branches needed at this precision; includes the 1e-6 ridge)."""
    import mpmath as mp
    mp.mp.dps = 50
    th = mp.e ** mp.mpf(log_theta)
    a = 1 / th
    n, p = X.shape
    w = [1 / (1 / mp.mpf(m) + th) for m in mu]
    b = mp.matrix(p, p)
    db = mp.matrix(p, p)
    for i in range(n):
        for r in range(p):
            for c in range(p):
                b[r, c] += X[i, r] * X[i, c] * w[i]
                db[r, c] += -X[i, r] * X[i, c] * w[i] ** 2
    for r in range(p):
        b[r, r] += mp.mpf("1e-6")
    cr = -mp.mpf("0.5") * mp.mpf("0.99") * sum((b ** -1 * db)[k, k] for k in range(p))
    dg = sum(mp.digamma(a + yi) - mp.digamma(a) for yi in y) * a
    sy = sum(y)
    corr = sum((yi - 1) * yi for yi in y) / (2 * a) if a > 1e5 else 0
    dg = sy - corr if max(y) * 1e6 < a else min(dg, sy - corr)
    ll = sum(mp.log(1 + m * th) - m / (m + a) + yi / (m + a) for yi, m in zip(y, mu)) * a
    return float(ll - dg + cr * th)


def test_nb_score_acc(orc):
    rng = np.random.default_rng(21)
    n, p = 60, 2
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    # well-conditioned: the accurate form equals the reference form
    mu = np.exp(X @ np.array([1.0, 0.3]))
    y = nb_sample(rng, mu, 0.5)
    for lt in [-2.0, 0.0, 1.0]:
        s0 = orc.nb_score(y, mu, lt, X, True)
        s1 = orc.nb_score_acc(y, mu, lt, X, True)
        assert abs(s0 - s1) <= 1e-9 * max(1, abs(s0))
        d0 = orc.nb_dscore(y, mu, lt, X, True)
        d1 = orc.nb_dscore_acc(y, mu, lt, X, True)
        assert abs(d0 - d1) <= 1e-9 * max(1, abs(d0))
    # near-Poisson gene, tiny theta: the reference form cancels; the accurate
    # one matches 50-digit arithmetic
    mu2 = np.exp(X @ np.array([-3.0, 0.2]))
    y2 = rng.poisson(mu2).astype(float)
    y2[0] = 2.0
    for lt in [-14.0, -13.0, -12.0]:
        ref = nb_score_mpmath(y2, mu2, lt, X)
        acc = orc.nb_score_acc(y2, mu2, lt, X, True)
        # rounding floor of the remaining sum_y - sum_y cancellation: ~1e-14
        assert abs(acc - ref) <= 1e-8 * abs(ref) + 1e-13, (lt, acc, ref)


def test_make_table(orc):
    y = np.array([3, 0, 0, 7, 3, 0, 1.0])
    k, f = orc.make_table(y, 4)
    np.testing.assert_array_equal(k, [0, 1, 3, 7])
    np.testing.assert_array_equal(f, [3, 1, 2, 1])
    k, f = orc.make_table(y, 3)
    assert k.size == 0


def test_fit_dispersion_cases(orc):
    n = 20
    X = np.ones((n, 1))
    th, it = orc.fit_dispersion([0.0], X, np.zeros(n), np.zeros(n), 1e-8, 50)
    assert th == 0.0 and it == 0                      # test-overdispersion.R:47-65
    # equidispersed data: the score at theta = 1e-8 is negative -> 0
    y = np.tile([9.0, 10.0, 11.0], 30)
    th, _ = orc.fit_dispersion([math.log(10)], np.ones((90, 1)), y, np.zeros(90))
    assert th == 0.0
    # NB data: the returned theta is a root of the score and a local max of ll
    rng = np.random.default_rng(16)
    n = 3000
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    beta = np.array([1.5, 0.4])
    y = nb_sample(rng, np.exp(X @ beta), 0.5)
    th, it = orc.fit_dispersion(beta, X, y, np.zeros(n), 1e-3, 200)
    assert 0.4 < th < 0.6 and 2 <= it < 30
    mu = np.exp(X @ beta)
    table = orc.make_table(y, n // 2)
    s = orc.nb_score(y, mu, math.log(th), X, True, table)
    d = orc.nb_dscore(y, mu, math.log(th), X, True, table)
    assert abs(s) < 1e-8 * abs(d)
    ll = lambda t: orc.nb_loglik(y, mu, t, X, True, table)
    assert ll(math.log(th)) >= max(ll(math.log(th) + 1e-3), ll(math.log(th) - 1e-3))


def test_estimate_dispersion(orc):
    rng = np.random.default_rng(17)
    Y = rng.poisson(10, (5, 10)).astype(float)
    d = orc.estimate_dispersion(Y, np.zeros(10))
    v, m = Y.var(axis=1, ddof=1), Y.mean(axis=1)
    ref = (v - m) / m ** 2
    ref[(ref < 0) | np.isnan(ref)] = 0.01
    np.testing.assert_allclose(d, ref, rtol=1e-14)
    Y2 = np.vstack([np.zeros(10), Y])
    assert orc.estimate_dispersion(Y2, np.zeros(10))[0] == 0.01
    d1 = orc.estimate_dispersion(Y, np.full(10, math.log(2.0)))
    ref1 = (v - m / 2) / m ** 2
    ref1[(ref1 < 0)] = 0.01
    np.testing.assert_allclose(d1, ref1, rtol=1e-14)


def test_init_beta_qr(orc):
    rng = np.random.default_rng(18)
    G, n, p = 4, 30, 3
    Y = rng.poisson(10, (G, n)).astype(float)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    off = rng.uniform(-0.5, 0.5, n)
    b = orc.init_beta_qr(Y, X, off)
    Z = np.log1p(Y / np.exp(off))
    ref = np.linalg.lstsq(X, Z.T, rcond=None)[0].T
    np.testing.assert_allclose(b, ref, rtol=1e-10, atol=1e-12)
    # intercept-only: column means of log1p(y)
    b1 = orc.init_beta_qr(Y, np.ones((n, 1)), np.zeros(n))
    np.testing.assert_allclose(b1[:, 0], np.log1p(Y).mean(axis=1), rtol=1e-12)
    # scaling a column scales its coefficient
    Xs = X.copy()
    Xs[:, 1] *= 5
    bs = orc.init_beta_qr(Y, Xs, off)
    np.testing.assert_allclose(bs[:, 0], b[:, 0], rtol=1e-10)
    np.testing.assert_allclose(bs[:, 1] * 5, b[:, 1], rtol=1e-10)
    with pytest.raises(ValueError):
        orc.init_beta_qr(Y, np.column_stack([X, X[:, 1] * 2]), off)


def test_init_beta_univariate_and_rough(orc):
    rng = np.random.default_rng(19)
    G, n = 3, 50
    Y = rng.poisson(4, (G, n)).astype(float)
    x1 = rng.normal(size=n)
    X = np.column_stack([np.ones(n), x1])
    sf = np.exp(rng.normal(0, 0.2, n))
    b = orc.init_beta_univariate(Y, X, sf)
    for g in range(G):
        z = np.log((Y[g] + 0.5) / sf)
        slope = np.polyfit(x1, z, 1)[0]
        assert abs(b[g, 1] - slope) < 1e-12
        assert abs(b[g, 0] - math.log((Y[g].sum() + 1e-8) / (sf.sum() + 1e-8))) < 1e-14
    r = orc.init_beta_rough(Y, 2)
    np.testing.assert_allclose(r[:, 0], np.log1p(Y.mean(axis=1)), rtol=1e-15)
    assert np.all(r[:, 1] == 0)


def test_cpu_fit_pipeline(orc):
    rng = np.random.default_rng(20)
    G, n = 12, 400
    X = np.column_stack([np.ones(n), rng.uniform(size=n) < 0.5])
    sf = np.exp(rng.normal(0, 0.3, n))
    B = np.column_stack([rng.normal(0, 1, G), rng.normal(0, 0.5, G)])
    alpha = np.exp(rng.normal(0, 1, G))
    Y = np.vstack([nb_sample(rng, sf * np.exp(X @ B[g]), alpha[g]) for g in range(G)])
    off = np.log(sf)
    ends = np.array([100, 250, 400], np.int32)
    out = orc.cpu_fit(Y, X, off, clusters_ends=ends, od_mode=orc.OD_MOM)
    assert out["beta"].shape == (G, 2) and out["bread"].shape == (G, 2, 2)
    assert np.all(out["theta"] >= 0) and np.all(out["theta_iters"] == 0)
    # composition of the per-gene restatements (R/main.R:413-502)
    a0 = orc.estimate_dispersion(Y, off)
    b0 = orc.init_beta_qr(Y, X, off)
    for g in range(G):
        b, it = orc.beta_fit(Y[g], X, b0[g], off, a0[g], 200, 1e-3)
        assert it == out["beta_iters"][g]
        np.testing.assert_array_equal(b, out["beta"][g])
        th = orc.mom_dispersion(Y[g], X, b, sf)
        assert th == out["theta"][g]
        br = orc.compute_hessian(b, th, Y[g], X, sf)
        np.testing.assert_array_equal(br, out["bread"][g])
        m = orc.clustered_meat(X, Y[g], b, th, sf, ends)
        np.testing.assert_allclose(out["sandwich"][g], br @ m @ br * n, rtol=1e-14)
    # Poisson mode: theta == 0 (test-main.R:1-55)
    o0 = orc.cpu_fit(Y, X, off, od_mode=orc.OD_NONE)
    assert np.all(o0["theta"] == 0) and o0["sandwich"] is None
    # a gene subset reproduces the same rows
    sub = np.array([5, 0, 9])
    o2 = orc.cpu_fit(Y, X, off, clusters_ends=ends, genes=sub)
    np.testing.assert_array_equal(o2["beta"], out["beta"][sub])
    np.testing.assert_array_equal(o2["beta_iters"], out["beta_iters"][sub])
