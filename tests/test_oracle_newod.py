"""Pins the oracle's restatement of the "new" / "I" overdispersion
(fit_overdispersion_cppp, src/new_overdispersion.cpp:195-481) on the CPU.

What fixes it without R:
* the 16-point Gauss-Hermite rule equals numpy's hermgauss and the
  eigendecomposition of the Jacobi matrix (Golub-Welsch, as the reference);
* H_log_gh_pmhalf approximates log of int_0^inf x^p exp(s (x log x - lgamma x)
  - t x) dx (p = +-1/2): against adaptive quadrature (scipy.integrate.quad);
* a converged gene's kappa is a fixed point of F_k up to the stopping rule
  |k_new - k| < eps (1 + k): F at 1/kappa equals 1/kappa to that tolerance;
* the reference's conditioning: its Laplace mode is found with the
  overdispersion tolerance as Newton tolerance (:284), so kappa moves by far
  more than a tiny change of beta -- the bar of the GPU comparison;
* the reference's own scenario (test-main.R:216-252) with numpy draws of the
  same distributions: finite theta, kappa_iters >= 0.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.lib()
    return oracle


def test_gauss_hermite_16(orc):
    x, w = orc.gh16()
    hx, hw = np.polynomial.hermite.hermgauss(16)
    np.testing.assert_allclose(x, hx, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, hw, rtol=1e-13)
    J = np.diag(np.sqrt(0.5 * np.arange(1, 16)), 1)
    J = J + J.T
    ev, V = np.linalg.eigh(J)
    np.testing.assert_allclose(x, ev, atol=1e-13)
    np.testing.assert_allclose(w, math.sqrt(math.pi) * V[0] ** 2, rtol=1e-12)


def test_h_log_gh_against_quadrature(orc):
    integrate = pytest.importorskip("scipy.integrate")
    from scipy.special import gammaln
    for s, t in [(20.0, 30.0), (50.0, 60.0), (200.0, 230.0)]:
        lp, lm = orc.h_log_gh_pmhalf(s, t, 1e-10)
        for p, got in [(0.5, lp), (-0.5, lm)]:
            h = lambda x: p * np.log(x) + s * (x * np.log(x) - gammaln(x)) - t * x
            xs = np.linspace(1e-3, 50, 20001)
            m = np.max(h(xs))
            val, _ = integrate.quad(lambda x: np.exp(h(x) - m), 0, np.inf, limit=500,
                                    epsabs=0, epsrel=1e-11)
            ref = m + math.log(val)
            assert abs(got - ref) <= 1e-5 * max(1.0, abs(ref)), (s, t, p, got, ref)
    assert orc.h_log_gh_pmhalf(10.0, 5.0) is None  # t <= s: the reference stops


def test_kappa_is_fixed_point(orc):
    rng = np.random.default_rng(7)
    checked = 0
    for trial in range(30):
        n = int(rng.integers(400, 3000))
        X = np.column_stack([np.ones(n), rng.integers(0, 2, n)])
        beta = np.array([rng.normal(1.0, 0.8), rng.normal(0, 0.5)])
        alpha = float(np.exp(rng.normal(-1.0, 0.8)))
        mu = np.exp(X @ beta)
        y = rng.poisson(rng.gamma(1 / alpha, alpha * mu)).astype(float)
        r = orc.fit_overdispersion_new(y, X, beta, np.zeros(n), 1.0, 200, 1e-6)
        if r["err"] or not r["kappa_converged"]:
            continue
        k = 1.0 / r["kappa"]
        # F_k(k) = k at the fixed point: f = log F - log k ~ 0; its size is
        # bounded by the step that stopped the iteration
        r2 = orc.fit_overdispersion_new(y, X, beta, np.zeros(n), 1.0, 200, 1e-10)
        assert r2["err"] == 0
        assert abs(1.0 / r2["kappa"] - k) <= 1e-5 * (1.0 + k) * 10
        checked += 1
    assert checked >= 10


def test_reference_scenario(orc):
    """test-main.R:216-252: 3 genes x 8 samples of Poisson(6), intercept and a
    normal covariate; every theta finite, theta_iters >= 0."""
    rng = np.random.default_rng(14)
    G, n = 3, 8
    Y = rng.poisson(6, (G, n)).astype(float)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    out = orc.cpu_fit(Y, X, np.zeros(n), od_mode=orc.OD_NEW, max_iter=50, tolerance=1e-3)
    assert out["theta"].shape == (G,)
    assert np.all(np.isfinite(out["theta"]))
    assert np.all(out["theta_iters"] >= 0)


def test_kappa_conditioning(orc):
    """A 1e-12 relative change of beta moves the reference's kappa by more
    than 1e-6 for some genes at unchanged iteration and evaluation counts: the
    Laplace mode's Newton stops at eps_theta (src/new_overdispersion.cpp:284),
    so F_k is piecewise smooth with jumps of that order.  Results of two
    correct implementations that sum in different orders differ by as much."""
    from devil_amd import synth
    w = synth.make_workload("C1", G=96, n=3000, seed=5)
    Y = synth.counts_host(w).astype(np.float64)
    ref = orc.cpu_fit(Y, w.X, w.offset, od_mode=orc.OD_NEW)
    worst = 0.0
    for g in range(0, 96, 3):
        b = ref["beta"][g]
        r0 = orc.fit_overdispersion_new(Y[g], w.X, b, w.offset, 1.0, 200, 1e-3)
        r1 = orc.fit_overdispersion_new(Y[g], w.X, b * (1 + 1e-12), w.offset, 1.0, 200, 1e-3)
        if r0["err"] or r1["kappa_iters"] != r0["kappa_iters"] or r1["evals"] != r0["evals"]:
            continue
        worst = max(worst, abs(r1["kappa"] - r0["kappa"]) / r0["kappa"])
    assert 1e-6 < worst < 1e-3
