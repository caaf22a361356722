"""GPU parity of the count-matrix preparation passes through the C ABI:
size factors (R/size_factors.R:37-127) against the oracle, group_data
(R/group_data.R:38-69) and the CSC densify (sparse assays, R/main.R:1-9)
exact, and fit_devil taking them end to end."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import devil_amd
    devil_amd.lib()
    return devil_amd


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.lib()
    return oracle


def _counts(rng, G, n, zero_frac=0.6):
    mu = np.exp(rng.normal(0.5, 1.2, (G, 1))) * rng.lognormal(0, 0.3, (1, n))
    Y = rng.poisson(mu).astype(np.int32)
    Y[rng.random((G, n)) < zero_frac] = 0
    Y[:, Y.sum(axis=0) == 0] = 1  # no empty cells
    return Y


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.float64])
@pytest.mark.parametrize("method", ["normed_sum", "psinorm"])
def test_size_factors_match_oracle(dv, orc, dtype, method):
    rng = np.random.default_rng(1)
    for G, n in [(300, 5000), (7, 1037), (2000, 20011)]:
        say(f"[case] size factors {method} {np.dtype(dtype).name} G={G} n={n}")
        Y = _counts(rng, G, n).astype(dtype)
        sf = dv.calculate_sf(torch.as_tensor(Y, device="cuda"), method).cpu().numpy()
        ref = orc.size_factors(Y.astype(np.float64), method)
        np.testing.assert_allclose(sf, ref, rtol=1e-12, atol=0)
        # a cell whose counts are all equal has psinorm factor 0 before the
        # scaling, and log(0) = -inf then makes every factor inf / NaN, as in R
        if np.all(np.isfinite(ref)):
            assert abs(np.exp(np.mean(np.log(sf))) - 1.0) < 1e-12
        else:
            assert method == "psinorm"


def test_size_factors_edge_cases(dv):
    say("[case] size factors edge cases")
    Y = torch.as_tensor(np.arange(12, dtype=np.int32).reshape(1, 12) + 1, device="cuda")
    assert torch.all(dv.calculate_sf(Y, "normed_sum") == 1.0)  # one row: unit factors
    Yz = np.ones((5, 9), np.int32)
    Yz[:, 4] = 0
    with pytest.raises(dv.DevilError) as ei:
        dv.calculate_sf(torch.as_tensor(Yz, device="cuda"), "normed_sum")
    assert ei.value.code == 1
    assert "all zeros" in str(ei.value)
    with pytest.raises(dv.DevilError) as ei:
        dv.calculate_sf(torch.as_tensor(np.ones((5, 9), np.int32), device="cuda"), "edgeR")
    assert ei.value.code == 7


def test_group_data_matches_oracle(dv, orc):
    rng = np.random.default_rng(2)
    G, n, p = 50, 3001, 3
    Y = _counts(rng, G, n)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    clusters = rng.choice(np.array(["p7", "p2", "p9", "p1"]), size=n)
    say("[case] group_data")
    out = dv.group_data(torch.as_tensor(Y, device="cuda"), X, clusters)
    Yr, Xr, cr, _ = orc.group_data(Y, X, clusters)
    np.testing.assert_array_equal(out["input_matrix"].cpu().numpy(), Yr)
    np.testing.assert_array_equal(out["design_matrix"], Xr)
    np.testing.assert_array_equal(out["clusters"], cr)
    # host counts come back as a host array; the grouped clusters are sorted
    out2 = dv.group_data(Y.astype(np.float64), X, clusters)
    np.testing.assert_array_equal(out2["input_matrix"], Yr.astype(np.float64))
    assert np.all(np.diff(out2["clusters"]) >= 0)
    # a bad permutation is rejected before any read
    ctx = dv.context(0)
    Yd = torch.as_tensor(Y, device="cuda")
    bad = torch.as_tensor(np.r_[np.arange(n - 1), n + 5].astype(np.int64), device="cuda")
    outb = torch.empty_like(Yd)
    torch.cuda.synchronize()  # torch's stream before the library's
    rc = dv.lib().devil_group_cells(ctx.h, Yd.data_ptr(), 0, G, n, n, bad.data_ptr(),
                                    outb.data_ptr(), n)
    assert rc == 1


def test_csc_densify_exact(dv):
    sp = pytest.importorskip("scipy.sparse")
    rng = np.random.default_rng(3)
    for G, n, dens in [(200, 3000, 0.1), (33, 1, 0.5), (64, 777, 0.0)]:
        say(f"[case] csc densify G={G} n={n} density={dens}")
        A = sp.random(G, n, density=dens, format="csc", random_state=4,
                      data_rvs=lambda k: rng.integers(1, 50, k).astype(float))
        D = dv.densify(A).cpu().numpy()
        assert D.dtype == np.int32
        np.testing.assert_array_equal(D, A.toarray().astype(np.int32))
    A = sp.random(20, 40, density=0.3, format="csc", random_state=5)
    D = dv.densify(A).cpu().numpy()
    assert D.dtype == np.float64
    np.testing.assert_array_equal(D, A.toarray())
    # malformed structure: a row index beyond G is rejected, nothing written
    ctx = dv.context(0)
    colptr = torch.as_tensor(np.array([0, 1, 2], np.int64), device="cuda")
    rowind = torch.as_tensor(np.array([0, 9], np.int32), device="cuda")
    vals = torch.ones(2, dtype=torch.float64, device="cuda")
    out = torch.full((3, 2), 5, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # torch's stream before the library's
    rc = dv.lib().devil_csc_to_dense(ctx.h, colptr.data_ptr(), rowind.data_ptr(),
                                     vals.data_ptr(), 3, 2, 0, out.data_ptr(), 2)
    assert rc == 1 and torch.all(out == 5)


def test_fit_with_estimated_sf_and_sparse_input(dv):
    sp = pytest.importorskip("scipy.sparse")
    from devil_amd import synth
    w = synth.make_workload("C1", G=100, n=3000, seed=9)
    Y = synth.counts_host(w)
    say("[case] fit_devil size_factors='normed_sum' and a sparse assay")
    fit_s = dv.fit_devil(Y, w.X, size_factors="normed_sum")
    sf = dv.calculate_sf(torch.as_tensor(Y, device="cuda"), "normed_sum")
    fit_n = dv.fit_devil(torch.as_tensor(Y, device="cuda"), w.X, size_factors=sf)
    np.testing.assert_array_equal(fit_s["beta"].cpu().numpy(), fit_n["beta"].cpu().numpy())
    fit_d = dv.fit_devil(Y, w.X, size_factors=w.sf)
    fit_sp = dv.fit_devil(sp.csc_matrix(Y.astype(np.float64)), w.X, size_factors=w.sf)
    np.testing.assert_array_equal(np.asarray(fit_d["beta"]),
                                  fit_sp["beta"].cpu().numpy())


def test_sandwich_hc0_matches_oracle(dv, orc):
    """compute_sandwich without clusters: bread (HC0 meat) bread n against the
    oracle's compute_hessian and compute_meat (src/variance.cpp:157-178); with
    clusters it equals the fit's clustered sandwich."""
    from devil_amd import synth
    say("[case] compute_sandwich HC0 and clustered")
    for cfg, G, n, seed in [("C1", 40, 3001, 3), ("C3", 24, 5000, 4), ("C4", 20, 4000, 5)]:
        w = synth.make_workload(cfg, G=G, n=n, seed=seed)
        Y = synth.counts_host(w)
        fit = dv.fit_devil(Y, w.X, size_factors=w.sf, clusters=w.clusters)
        beta = np.asarray(fit["beta"])
        th = np.asarray(fit["overdispersion"])
        sw = dv.compute_sandwich(w.X, Y, beta, th, w.sf).cpu().numpy()
        for g in range(G):
            b = orc.compute_hessian(beta[g], th[g], Y[g].astype(float), w.X, w.sf)
            m = orc.meat_hc0(w.X, Y[g].astype(float), beta[g], th[g], w.sf)
            ref = b @ m @ b * n
            np.testing.assert_allclose(sw[g], ref, rtol=1e-9,
                                       atol=1e-12 * np.abs(ref).max())
        if w.clusters is not None:
            swc = dv.compute_sandwich(w.X, Y, beta, th, w.sf, clusters=w.clusters).cpu().numpy()
            np.testing.assert_array_equal(swc, np.asarray(fit["beta_sandwiches"]))
