"""GPU parity: the HIP fit (through the C ABI) against the FP64 oracle on the
same seeded inputs.

Bars (FP64 path): beta within 1e-8 relative (floor 1e-6), identical per-gene
IRLS iteration counts, theta / bread / sandwich within 1e-8 relative, and
p-values within 1e-6 relative.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e[both_nan] = 0.0
    return e


def assert_rel(a, b, tol, floor, what):
    e = relerr(a, b, floor)
    worst = np.nanmax(e) if e.size else 0.0
    assert np.all(np.isfinite(e)) and worst <= tol, f"{what}: max rel err {worst:.3e} > {tol}"


@pytest.fixture(scope="module")
def dv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import devil_amd
    devil_amd.lib()
    return devil_amd


def run_both(dv, orc, w, Y, *, od, clusters=None, init="qr", max_iter=200, tol=1e-3,
             disp=None):
    say(f"[case] {w.name} G={Y.shape[0]} n={Y.shape[1]} p={w.p} od={od} init={init} "
        f"dtype={Y.dtype} clusters={'yes' if clusters is not None else 'no'}")
    fit = dv.fit_devil(Y, w.X, clusters=clusters, overdispersion=od, size_factors=w.sf,
                       max_iter=max_iter, tolerance=tol, init_mode=init,
                       init_overdispersion=disp)
    ends = None if clusters is None else dv.cluster_blocks(clusters, Y.shape[1])
    odm = {"MOM": orc.OD_MOM, "MLE": orc.OD_MLE, "new": orc.OD_NEW, False: orc.OD_NONE}[od]
    im = {"qr": orc.INIT_QR, "rough": orc.INIT_ROUGH, "univariate": orc.INIT_UNIVARIATE}[init]
    ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset, clusters_ends=ends, od_mode=odm,
                      init_mode=im, max_iter=max_iter, tolerance=tol,
                      disp_init=None if disp is None else np.full(Y.shape[0], disp))
    return fit, ref


def check_fit(fit, ref, *, theta_tol=1e-8):
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"], ref["beta_iters"])
    assert_rel(fit["beta"], ref["beta"], 1e-8, 1e-6, "beta")
    assert_rel(fit["overdispersion"], ref["theta"], theta_tol, 1e-8, "theta")
    assert_rel(fit["beta_sandwiches_null"], ref["bread"], 1e-8, 1e-12, "bread")
    if ref["sandwich"] is not None:
        assert_rel(fit["beta_sandwiches"], ref["sandwich"], 1e-8, 1e-12, "sandwich")


def test_fit_c1(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C1", G=400, n=5000)
    Y = synth.counts_host(w)
    fit, ref = run_both(dv, orc, w, Y, od="MOM")
    check_fit(fit, ref)
    st = fit["status"]
    hit = fit["iterations"]["beta_iters"] == 200
    np.testing.assert_array_equal((st & 1) != 0, hit)


def test_fit_clustered_c4(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C4", G=120, n=6000)
    Y = synth.counts_host(w)
    fit, ref = run_both(dv, orc, w, Y, od="MOM", clusters=w.clusters)
    check_fit(fit, ref)


def test_fit_mle_c2(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C2", G=80, n=4000)
    Y = synth.counts_host(w)
    fit, ref = run_both(dv, orc, w, Y, od="MLE")
    check_fit(fit, ref, theta_tol=1e-8)
    same = np.mean(fit["theta_passes"] == ref["theta_iters"])
    assert same >= 0.95
    # cpu_fit reports theta_iters as NA in MLE mode (R/main.R:431-468)
    assert np.all(np.isnan(fit["iterations"]["theta_iters"]))


@pytest.mark.parametrize("init", ["rough", "univariate"])
def test_fit_init_modes(dv, orc, init):
    from devil_amd import synth
    w = synth.make_workload("C1", G=150, n=3000, seed=11)
    Y = synth.counts_host(w)
    fit, ref = run_both(dv, orc, w, Y, od="MOM", init=init)
    check_fit(fit, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fit_count_dtypes(dv, orc, dtype):
    from devil_amd import synth
    w = synth.make_workload("C1", G=100, n=2500, seed=12)
    Y = synth.counts_host(w, dtype=dtype)
    fit, ref = run_both(dv, orc, w, Y, od="MOM")
    check_fit(fit, ref)


@pytest.mark.parametrize("p,n", [(1, 3000), (3, 2999), (10, 3001)])
def test_fit_design_widths(dv, orc, p, n):
    """p from 1 to 10; n not a multiple of 4 or 16 exercises the unaligned and
    partial-block count loads."""
    from devil_amd import synth
    rng = np.random.default_rng(100 + p)
    G = 60
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(p - 1)])
    w = synth.Workload("wide", G, n, p, X, np.exp(rng.normal(0, 0.3, n)),
                       np.column_stack([rng.normal(0, 1, G)] +
                                       [rng.normal(0, 0.1, G) for _ in range(p - 1)]),
                       np.exp(rng.normal(0, 1, G)), seed=99 + p)
    Y = synth.counts_host(w)
    clusters = np.repeat([1, 2, 3], [n // 3, n // 3, n - 2 * (n // 3)])
    fit, ref = run_both(dv, orc, w, Y, od="MOM", clusters=clusters)
    check_fit(fit, ref)


def test_fit_edge_cases(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C1", G=50, n=2000, seed=13)
    Y = synth.counts_host(w)
    Y[3] = 0                       # all-zero gene
    Y[7] = 0
    Y[7, 5] = 40                   # a single non-zero count
    fit, ref = run_both(dv, orc, w, Y, od="MOM", max_iter=5)  # many genes hit max_iter
    check_fit(fit, ref)
    fit2, ref2 = run_both(dv, orc, w, Y, od="MLE", max_iter=200)
    check_fit(fit2, ref2)
    # alpha0 = 0 -> k = inf -> NaN step after one iteration (src/beta.cpp:45)
    fit3, ref3 = run_both(dv, orc, w, Y, od=False, disp=0.0)
    assert np.all(fit3["iterations"]["beta_iters"] == 1)
    assert np.all(np.isnan(fit3["beta"]))
    np.testing.assert_array_equal(fit3["iterations"]["beta_iters"], ref3["beta_iters"])


def test_fit_device_resident_and_deterministic(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C1", G=300, n=4000, seed=14)
    Y = synth.counts_host(w)
    say("[case] device-resident counts, two runs")
    host = dv.fit_devil(Y, w.X, size_factors=w.sf)
    Yd = torch.as_tensor(Y, device="cuda")
    d1 = dv.fit_devil(Yd, w.X, size_factors=w.sf)
    d2 = dv.fit_devil(Yd, w.X, size_factors=w.sf)
    # the design matrix on the GPU (C- and column-ordered) is read as is
    Xc = torch.as_tensor(np.ascontiguousarray(w.X), device="cuda")
    Xf = torch.as_tensor(np.ascontiguousarray(w.X.T), device="cuda").t()
    d3 = dv.fit_devil(Yd, Xc, size_factors=w.sf)
    d4 = dv.fit_devil(Yd, Xf, size_factors=w.sf)
    for key in ("beta", "overdispersion", "beta_sandwiches_null"):
        a = d1[key].cpu().numpy()
        np.testing.assert_array_equal(a, d2[key].cpu().numpy())
        np.testing.assert_array_equal(a, host[key])
        np.testing.assert_array_equal(a, d3[key].cpu().numpy())
        np.testing.assert_array_equal(a, d4[key].cpu().numpy())
    # size factors on the GPU: the offset log(sf) is taken there (the two
    # logarithms may differ in the last bit)
    d5 = dv.fit_devil(Yd, Xf, size_factors=torch.as_tensor(w.sf, device="cuda"))
    np.testing.assert_array_equal(d5["iterations"]["beta_iters"].cpu().numpy(),
                                  host["iterations"]["beta_iters"])
    assert_rel(d5["beta"].cpu().numpy(), host["beta"], 1e-10, 1e-3, "beta (GPU sf)")
    r1 = dv.test_de(d1, contrast=[0, 1])
    r5 = dv.test_de(host, contrast=[0, 1])
    np.testing.assert_array_equal(r1["pval"], r5["pval"])


def test_test_de_parity(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C4", G=200, n=5000, seed=15)
    Y = synth.counts_host(w)
    say("[case] test_de on a clustered fit")
    fit = dv.fit_devil(Y, w.X, clusters=w.clusters, size_factors=w.sf)
    res = dv.test_de(fit, contrast=[0, 1, 0, 0, 0, 0], max_lfc=2)
    ref = orc.de_test(fit["beta"], fit["beta_sandwiches_null"], fit["beta_sandwiches"], w.n,
                      [0, 1, 0, 0, 0, 0], max_lfc=2)
    assert_rel(res["log_pval"], ref["log_pval"], 1e-10, 1e-3, "log p")
    assert_rel(res["pval"], ref["pval"], 1e-6, 1e-300, "p")
    assert_rel(res["adj_pval"], ref["adj_pval"], 1e-6, 1e-300, "adj p")
    np.testing.assert_allclose(res["lfc"], ref["lfc"], rtol=1e-15, atol=0)
    # unclustered fit: p from the bread only
    fit0 = dv.fit_devil(Y, w.X, size_factors=w.sf)
    r0 = dv.test_de(fit0, contrast=[0, 1, 0, 0, 0, 0])
    f0 = orc.de_test(fit0["beta"], fit0["beta_sandwiches_null"], None, w.n, [0, 1, 0, 0, 0, 0])
    assert_rel(r0["pval"], f0["pval"], 1e-6, 1e-300, "p (bread)")
    np.testing.assert_array_equal(np.isnan(r0["adj_pval"]), np.isnan(f0["adj_pval"]))


def test_init_logarithm_accuracy(dv):
    """The init pass's logarithm (table-driven: log u = e log 2 + log F +
    log1p(r), |r| <= 2^-8) against numpy on isolated values.  An intercept-only
    design on 16 cells gives Q = 1/4 and R = 4 exactly, so beta0 of a gene is
    exactly (1/16) sum_i log1p(y_i / s_i); each gene has one nonzero count, so
    16 beta0 is a single value.  Counts 1 .. 2^31 - 1 and size factors 1e-3 ..
    1e3 cover every table slot region and binade in between: within 5 ulp
    (numpy's and the device's exp(-offset) may differ in the last bit)."""
    from devil_amd import _lib as L
    ctx = dv.context(0)
    lib = dv.lib()
    n = 16
    ys = [1, 2, 3, 5, 7, 10, 31, 100, 255, 256, 257, 1000, 4095, 12345, 65535, 99991,
          2 ** 20 + 1, 2 ** 24 + 3, 2 ** 30 - 1, 2 ** 31 - 1]
    sf = np.array([1.0, 1e-3, 0.37, 0.5, 0.9999, 1.0001, 1.9, 2.0, 2.1, 3.3, 7.0, 15.5, 100.0,
                   999.0, 1e3, 0.0123])
    G = len(ys) * n
    Y = np.zeros((G, n), np.int32)
    for a, y in enumerate(ys):
        for c in range(n):
            Y[a * n + c, c] = y
    off = np.log(sf)
    say("[case] init logarithm, 320 isolated values")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                    device="cuda")
    Yd, Xd, offd = t(Y, torch.int32), t(np.ones(n)), t(off)
    b0 = torch.empty(G, dtype=torch.float64, device="cuda")
    ctx.check(lib.devil_init_beta(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n, Xd.data_ptr(), 1,
                                  offd.data_ptr(), L.INIT_QR, b0.data_ptr()))
    z = 16.0 * b0.cpu().numpy()
    ref = np.log1p(Y.max(axis=1) * np.exp(-off[np.arange(G) % n]))
    e = np.abs(z - ref) / ref
    say(f"  max rel err {e.max():.2e}")
    assert e.max() <= 5 * 2.0 ** -53


def test_bh_matches_oracle(dv, orc):
    """The device BH (sort, suffix minimum) equals the oracle's step-up
    definition bit for bit on the device's own p-values, with ties (repeated
    genes) and NaNs (a NaN bread), and at G = 50000."""
    rng = np.random.default_rng(21)
    for G in (1, 7, 1000, 50000):
        say(f"[case] BH G={G}")
        p = 3
        beta = rng.normal(0, 0.05, (G, p))
        beta[: G // 3] = beta[0]                  # ties
        bread = np.tile(np.eye(p) * 1e-3, (G, 1, 1))
        if G > 5:
            bread[5] = np.nan                     # NaN p-value
        fit = {"beta": beta, "beta_sandwiches_null": bread, "beta_sandwiches": None,
               "design_matrix": np.zeros((500, p))}
        res = dv.test_de(fit, contrast=[0, 1, 0])
        ref = orc.p_adjust_bh(res["pval"])
        np.testing.assert_array_equal(res["adj_pval"], ref)
    # every p-value NaN (no non-NaN p-values: m = 0)
    fit = {"beta": np.zeros((4, 2)), "beta_sandwiches_null": np.full((4, 2, 2), np.nan),
           "beta_sandwiches": None, "design_matrix": np.zeros((50, 2))}
    res = dv.test_de(fit, contrast=[0, 1])
    assert np.all(np.isnan(res["pval"])) and np.all(np.isnan(res["adj_pval"]))


def test_stage_entry_points(dv, orc):
    """Each C-ABI stage against the oracle function it restates."""
    import ctypes as C
    from devil_amd import _lib as L
    from devil_amd import synth
    say("[case] stage entry points")
    w = synth.make_workload("C2", G=64, n=3000, seed=16)
    Y = synth.counts_host(w)
    ctx = dv.context(0)
    lib = dv.lib()
    dev = "cuda"
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    Yd = t(Y, torch.int32)
    Xd = t(w.X.T.reshape(-1))
    off = t(w.offset)
    G, n, p = Y.shape[0], Y.shape[1], w.p
    Yf = Y.astype(np.float64)
    a0 = torch.empty(G, dtype=torch.float64, device=dev)
    ctx.check(lib.devil_estimate_dispersion(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n,
                                            off.data_ptr(), a0.data_ptr()))
    # (v - ybar/mean(sf)) / ybar^2 cancels; alpha0 is used rounded to float
    assert_rel(a0.cpu().numpy(), orc.estimate_dispersion(Yf, w.offset), 1e-10, 1e-12, "alpha0")
    b0 = torch.empty(G * p, dtype=torch.float64, device=dev)
    ctx.check(lib.devil_init_beta(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n, Xd.data_ptr(), p,
                                  off.data_ptr(), L.INIT_QR, b0.data_ptr()))
    b0n = b0.cpu().numpy().reshape(p, G).T
    assert_rel(b0n, orc.init_beta_qr(Yf, w.X, w.offset), 1e-10, 1e-8, "beta0")
    beta = b0.clone()
    torch.cuda.synchronize()  # the clone runs on torch's stream, the calls on the library's
    it = torch.empty(G, dtype=torch.int32, device=dev)
    ctx.check(lib.devil_beta_fit(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n, Xd.data_ptr(), p,
                                 beta.data_ptr(), off.data_ptr(), a0.data_ptr(), 200, 1e-3,
                                 L.PREC_FP64, it.data_ptr(), None))
    bn = beta.cpu().numpy().reshape(p, G).T
    a0n = a0.cpu().numpy()
    for g in range(G):
        rb, ri = orc.beta_fit(Yf[g], w.X, b0n[g], w.offset, a0n[g], 200, 1e-3)
        assert ri == int(it[g])
        assert_rel(bn[g], rb, 1e-8, 1e-6, f"beta gene {g}")
    th = torch.empty(G, dtype=torch.float64, device=dev)
    ctx.check(lib.devil_mom_dispersion(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n,
                                       Xd.data_ptr(), p, beta.data_ptr(), off.data_ptr(),
                                       th.data_ptr()))
    thn = th.cpu().numpy()
    ref_th = [orc.mom_dispersion(Yf[g], w.X, bn[g], w.sf) for g in range(G)]
    assert_rel(thn, ref_th, 1e-9, 1e-9, "theta MOM")
    thm = torch.empty(G, dtype=torch.float64, device=dev)
    itm = torch.empty(G, dtype=torch.int32, device=dev)
    # default driver: nlminb's iteration and stopping rules restated
    ctx.check(lib.devil_fit_dispersion_mle(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n,
                                           Xd.data_ptr(), p, beta.data_ptr(), off.data_ptr(),
                                           200, 1e-3, thm.data_ptr(), itm.data_ptr()))
    ref_mle = [orc.fit_dispersion_port(bn[g], w.X, Yf[g], w.offset)[0] for g in range(G)]
    assert_rel(thm.cpu().numpy(), ref_mle, 1e-8, 1e-8, "theta MLE (nlminb)")
    # the root of the score
    ctx.set_mle_driver("root")
    try:
        ctx.check(lib.devil_fit_dispersion_mle(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n,
                                               Xd.data_ptr(), p, beta.data_ptr(),
                                               off.data_ptr(), 200, 1e-3, thm.data_ptr(),
                                               itm.data_ptr()))
    finally:
        ctx.set_mle_driver("nlminb")
    ref_mle = [orc.fit_dispersion(bn[g], w.X, Yf[g], w.offset)[0] for g in range(G)]
    assert_rel(thm.cpu().numpy(), ref_mle, 1e-8, 1e-8, "theta MLE (root)")
    ends = t(np.array([1000, 2000, 3000], np.int32), torch.int32)
    br = torch.empty(G * p * p, dtype=torch.float64, device=dev)
    sw = torch.empty(G * p * p, dtype=torch.float64, device=dev)
    st = torch.zeros(G, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx.check(lib.devil_bread_meat(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n, Xd.data_ptr(),
                                   p, beta.data_ptr(), th.data_ptr(), off.data_ptr(),
                                   ends.data_ptr(), 3, br.data_ptr(), sw.data_ptr(),
                                   st.data_ptr()))
    brn = br.cpu().numpy().reshape(G, p, p).transpose(0, 2, 1)
    swn = sw.cpu().numpy().reshape(G, p, p).transpose(0, 2, 1)
    for g in range(G):
        rb = orc.compute_hessian(bn[g], thn[g], Yf[g], w.X, w.sf)
        assert_rel(brn[g], rb, 1e-9, 1e-12, "bread")
        m = orc.clustered_meat(w.X, Yf[g], bn[g], thn[g], w.sf, [1000, 2000, 3000])
        assert_rel(swn[g], rb @ m @ rb * n, 1e-8, 1e-12, "sandwich")


def test_abi_errors_on_gpu(dv):
    """Argument errors come back as codes, never as a crash."""
    import ctypes as C
    from devil_amd import _lib as L
    ctx = dv.context(0)
    lib = dv.lib()
    rc = lib.devil_estimate_dispersion(ctx.h, None, 0, 5, 10, 10, None, None)
    assert rc == 1
    prob = L.Problem(G=2, n=10, p=11, mem=0)
    outs = L.Outputs()
    assert lib.devil_fit(ctx.h, C.byref(prob), C.byref(outs), 1) == 1
    assert lib.devil_fit(ctx.h, None, None, 2) == 1
    with pytest.raises(ValueError):
        dv.cluster_blocks([1, 1, 2, 1], 4)
    X = np.column_stack([np.ones(8), np.ones(8)])
    with pytest.raises(dv.DevilError) as ei:
        dv.fit_devil(np.ones((3, 8), np.int32), X)
    assert ei.value.code == 4


def test_estimate_dispersion_large_counts(dv, orc):
    """Counts whose squares sum past 2^63 (sum y^2 is kept in 128 bits): alpha0
    equals the oracle's FP64 two-pass variance."""
    from devil_amd import _lib as L
    say("[case] estimate_dispersion with counts near 2^30")
    rng = np.random.default_rng(7)
    G, n = 4, 2000
    Y = rng.poisson(3.0, size=(G, n)).astype(np.int32)
    Y[0, :16] = 1 << 30                 # 16 * 2^60 = 2^64 > int64
    Y[1, ::50] = 2_000_000_000          # 40 * 4e18
    Y[2, :] = 1 << 30                   # every cell: 2000 * 2^60
    off = np.log(rng.lognormal(0.0, 0.3, n))
    ctx = dv.context(0)
    Yd = torch.as_tensor(Y, device="cuda")
    offd = torch.as_tensor(off, device="cuda")
    a0 = torch.empty(G, dtype=torch.float64, device="cuda")
    ctx.check(dv.lib().devil_estimate_dispersion(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n,
                                                 offd.data_ptr(), a0.data_ptr()))
    ref = orc.estimate_dispersion(Y.astype(np.float64), off)
    assert_rel(a0.cpu().numpy(), ref, 1e-10, 1e-12, "alpha0")


def _mle_one(dv, y, X, beta, tol, max_iter):
    """devil_fit_dispersion_mle on one gene (device pointers)."""
    from devil_amd import _lib as L
    ctx = dv.context(0)
    n, p = X.shape
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                    device="cuda")
    Yd = t(np.asarray(y)[None, :], torch.float64)
    Xd = t(np.asarray(X, np.float64).T.reshape(-1))
    bd = t(np.asarray(beta, np.float64).reshape(1, p).T.reshape(-1))
    off = t(np.zeros(n))
    th = torch.empty(1, dtype=torch.float64, device="cuda")
    it = torch.empty(1, dtype=torch.int32, device="cuda")
    ctx.check(dv.lib().devil_fit_dispersion_mle(ctx.h, Yd.data_ptr(), L.COUNTS_F64, 1, n, n,
                                                Xd.data_ptr(), p, bd.data_ptr(),
                                                off.data_ptr(), max_iter, tol, th.data_ptr(),
                                                it.data_ptr()))
    return float(th.item())


def test_mle_reference_scenarios(dv, orc):
    """test-overdispersion.R:47-152 through the C ABI (both drivers), against
    the oracle; numpy draws of the same distributions."""
    import math
    say("[case] MLE reference scenarios")
    ctx = dv.context(0)
    rng = np.random.default_rng(10)
    n = 200
    y = rng.poisson(10, n).astype(float)
    X = np.ones((n, 1))
    b = [math.log(10)]
    for drv in ("nlminb", "root"):
        ctx.set_mle_driver(drv)
        try:
            th = _mle_one(dv, np.zeros(20), np.ones((20, 1)), [0.0], 1e-8, 50)
            assert th == 0.0
            th = _mle_one(dv, y, X, b, 1e-8, 100)
            assert np.isfinite(th) and 0 <= th < 0.2
            ref = (orc.fit_dispersion_port(b, X, y, np.zeros(n), 1e-8, 100)[0] if drv == "nlminb"
                   else orc.fit_dispersion(b, X, y, np.zeros(n), 1e-8, 100)[0])
            assert abs(th - ref) <= 1e-8 * max(ref, 1e-8)
            r2 = np.random.default_rng(11)
            yp = r2.poisson(10, 400).astype(float)
            ynb = r2.poisson(r2.gamma(2.0, 5.0, 400)).astype(float)
            X4 = np.ones((400, 1))
            dp = _mle_one(dv, yp, X4, b, 1e-8, 100)
            dnb = _mle_one(dv, ynb, X4, b, 1e-8, 100)
            assert np.isfinite(dnb) and dnb > dp and abs(dnb - 0.5) <= 0.15
        finally:
            ctx.set_mle_driver("nlminb")


def test_fit_mle_root_driver(dv, orc):
    """The root-of-the-score driver through fit_devil against the oracle's."""
    from devil_amd import synth
    w = synth.make_workload("C2", G=64, n=3000, seed=21)
    Y = synth.counts_host(w)
    ctx = dv.context(0)
    ctx.set_mle_driver("root")
    try:
        fit = dv.fit_devil(Y, w.X, overdispersion="MLE", size_factors=w.sf)
    finally:
        ctx.set_mle_driver("nlminb")
    ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset, od_mode=orc.OD_MLE,
                      mle_driver="root")
    check_fit(fit, ref, theta_tol=1e-8)


@pytest.mark.parametrize("wl,G,n", [("C1", 96, 3000), ("C3", 64, 6000)])
def test_fit_new_overdispersion(dv, orc, wl, G, n):
    """overdispersion="new" (fit_overdispersion_cppp, src/new_overdispersion.cpp:
    195-481) against the oracle's restatement.

    The reference finds the Laplace mode with the overdispersion tolerance as
    its Newton tolerance (:284, eps_theta = 1e-3), so its F_k jumps where the
    Newton count changes: a 1e-12 relative change of beta moves theta by up to
    ~1e-4 in the oracle itself (test_oracle_newod.py pins this).  The bar is
    therefore: median relative theta error <= 1e-8, 90% of genes <= 1e-6,
    every gene <= 1e-3 (the tolerance); identical kappa iteration counts for
    95% of genes; failed genes (C1 <= n, where the reference stops) flagged
    with NaN theta exactly where the oracle fails."""
    from devil_amd import synth
    w = synth.make_workload(wl, G=G, n=n, seed=5)
    Y = synth.counts_host(w)
    fit, ref = run_both(dv, orc, w, Y, od="new")
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"], ref["beta_iters"])
    assert_rel(fit["beta"], ref["beta"], 1e-8, 1e-6, "beta")
    th, rth = np.asarray(fit["overdispersion"]), ref["theta"]
    np.testing.assert_array_equal(np.isnan(th), np.isnan(rth))
    from devil_amd import _lib as L
    failed = (np.asarray(fit["status"]) & L.GENE_OD_FAILED) != 0
    np.testing.assert_array_equal(failed, np.isnan(rth))
    ok = ~np.isnan(rth)
    assert ok.sum() >= G // 2
    e = relerr(th[ok], rth[ok], 1e-8)
    say(f"theta rel err: median {np.median(e):.2e} p90 {np.quantile(e, 0.9):.2e} "
        f"max {e.max():.2e}")
    assert np.median(e) <= 1e-8 and np.quantile(e, 0.9) <= 1e-6 and e.max() <= 1e-3
    same = np.asarray(fit["iterations"]["theta_iters"])[ok] == ref["theta_iters"][ok]
    assert same.mean() >= 0.95
    # the bread follows theta: where theta agrees to 1e-8, each gene's bread
    # agrees normwise (max |difference| over max |entry|) to 1e-7
    tight = np.flatnonzero(ok)[e <= 1e-8]
    A = np.asarray(fit["beta_sandwiches_null"])[tight]
    B = ref["bread"][tight]
    nerr = np.abs(A - B).max(axis=(1, 2)) / np.abs(B).max(axis=(1, 2))
    assert nerr.max() <= 1e-7, f"bread: normwise rel err {nerr.max():.3e}"


def test_new_overdispersion_reference_scenario(dv, orc):
    """test-main.R:216-252 through fit_devil: 3 genes x 8 samples of
    Poisson(6), intercept and a normal covariate, max_iter 50, tolerance 1e-3;
    numpy draws of the same distributions."""
    say("[case] new overdispersion reference scenario")
    rng = np.random.default_rng(14)
    Y = rng.poisson(6, (3, 8)).astype(np.float64)
    X = np.column_stack([np.ones(8), rng.normal(size=8)])
    fit = dv.fit_devil(Y, X, overdispersion="I", size_factors=None, max_iter=50,
                       tolerance=1e-3)
    ref = orc.cpu_fit(Y, X, np.zeros(8), od_mode=orc.OD_NEW, max_iter=50, tolerance=1e-3)
    th = np.asarray(fit["overdispersion"])
    assert th.shape == (3,) and np.all(np.isfinite(th))
    assert np.all(np.asarray(fit["iterations"]["theta_iters"]) >= 0)
    assert_rel(th, ref["theta"], 1e-3, 1e-8, "theta")


@pytest.mark.parametrize("G", [1, 2, 3, 4097, 20000])
def test_p_adjust_methods(dv, orc, G):
    """devil_p_adjust (stats::p.adjust, R/test.R:127) against the oracle,
    bit for bit, with ties and NaN p-values."""
    from devil_amd import _lib as L
    say(f"[case] p_adjust G={G}")
    rng = np.random.default_rng(G)
    p = rng.uniform(0, 1, G) ** 4
    if G > 3:
        p[rng.integers(0, G, G // 10)] = p[0]
        p[rng.integers(0, G, G // 20)] = np.nan
    ctx = dv.context(0)
    pd = torch.as_tensor(p, device="cuda")
    for name, code in L.P_ADJUST.items():
        adj = torch.empty_like(pd)
        torch.cuda.synchronize()
        ctx.check(dv.lib().devil_p_adjust(ctx.h, pd.data_ptr(), G, code, adj.data_ptr()))
        np.testing.assert_array_equal(adj.cpu().numpy(), orc.p_adjust(p, name), err_msg=name)
    inplace = pd.clone()
    torch.cuda.synchronize()  # the clone runs on torch's stream, the call on its own
    ctx.check(dv.lib().devil_p_adjust(ctx.h, inplace.data_ptr(), G, L.P_ADJUST["holm"],
                                      inplace.data_ptr()))
    np.testing.assert_array_equal(inplace.cpu().numpy(), orc.p_adjust(p, "holm"))


def test_test_de_adjust_methods(dv, orc):
    """test_de(pval_adjust_method=...) applies p.adjust to its p-values."""
    from devil_amd import synth
    say("[case] test_de adjust methods")
    w = synth.make_workload("C1", G=300, n=2000)
    Y = synth.counts_host(w)
    fit = dv.fit_devil(Y, w.X, overdispersion="MOM", size_factors=w.sf)
    for m in ("holm", "hochberg", "hommel", "bonferroni", "BY", "none", "fdr"):
        r = dv.test_de(fit, [0, 1], pval_adjust_method=m)
        np.testing.assert_array_equal(r["adj_pval"], orc.p_adjust(r["pval"], m), err_msg=m)
    with pytest.raises(ValueError):
        dv.test_de(fit, [0, 1], pval_adjust_method="sidak")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_tail_matches_launched_rounds(dv, orc, precision):
    """With DEVIL_TUNE_GRAPH the IRLS rounds whose active genes fit one column
    of workgroups replay a captured two-round hipGraph: the same kernels and
    round shapes as rounds launched one by one, so the same results."""
    from devil_amd import _lib as L
    from devil_amd import synth
    w = synth.make_workload("C3", G=600, n=6000, seed=33)
    Y = synth.counts_host(w)
    say(f"[case] graph tail vs launched rounds, {precision}")
    ctx = dv.context(0)
    fits = {}
    try:
        for gv in (1, 0):
            ctx.set_tuning(L.TUNE_GRAPH, gv)
            fits[gv] = dv.fit_devil(Y, w.X, size_factors=w.sf, precision=precision)
    finally:
        ctx.set_tuning(L.TUNE_GRAPH, 0)
    np.testing.assert_array_equal(fits[1]["iterations"]["beta_iters"],
                                  fits[0]["iterations"]["beta_iters"])
    tol = 1e-10 if precision == "fp64" else 1e-4
    floor = 1e-6 if precision == "fp64" else 1e-3
    assert_rel(fits[1]["beta"], fits[0]["beta"], tol, floor, "beta")
    if precision == "fp64":
        ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset)
        np.testing.assert_array_equal(fits[1]["iterations"]["beta_iters"], ref["beta_iters"])
        assert_rel(fits[1]["beta"], ref["beta"], 1e-8, 1e-6, "beta vs oracle")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dtype", ["int32", "float64"])
def test_host_counts_in_gene_batches(dv, precision, dtype):
    """Host-resident counts (R's path) are uploaded and fitted in gene
    batches; with the batch capped at 100 genes (5 batches, the last one
    partial) and rows that need padding (n = 6001), the per-gene outputs
    equal the single-batch fit of the same counts on the device up to the
    order of the FP64 chunk sums (rounds of fewer genes get shorter chunks):
    identical iteration counts, beta / theta / bread / sandwich within 1e-10
    of each gene's largest entry, clustered sandwich included.  In the mixed
    path a last-bit change of beta moves the next rounds' FP32 roundings, so
    there the bar is the mixed path's own noise level, 1e-6."""
    from devil_amd import _lib as L, synth
    w = synth.make_workload("C4", G=437, n=6001, seed=785)
    Y = synth.counts_host(w).astype(dtype)  # float64: an R double matrix
    say(f"[case] host {dtype} counts in batches of 100 genes, {precision}")
    ctx = dv.context(0)
    if precision == "mixed":
        ctx.set_mixed_kernel("bf16_ws")
        ctx.set_tuning(L.TUNE_CHUNK_CELLS, 2048)
    try:
        ctx.set_tuning(L.TUNE_BATCH_GENES, 100)
        a = dv.fit_devil(Y, w.X, clusters=w.clusters, size_factors=w.sf, precision=precision)
        ctx.set_tuning(L.TUNE_BATCH_GENES, 0)
        b = dv.fit_devil(torch.as_tensor(Y, device="cuda"), w.X, clusters=w.clusters,
                         size_factors=w.sf, precision=precision)
    finally:
        ctx.set_tuning(L.TUNE_BATCH_GENES, 0)
        ctx.set_tuning(L.TUNE_CHUNK_CELLS, 0)
        ctx.set_mixed_kernel("auto")
    np.testing.assert_array_equal(np.asarray(a["iterations"]["beta_iters"]),
                                  b["iterations"]["beta_iters"].cpu().numpy())
    for key in ("beta", "overdispersion", "beta_sandwiches_null", "beta_sandwiches"):
        x = np.asarray(a[key]).reshape(w.G, -1)
        y = b[key].cpu().numpy().reshape(w.G, -1)
        scale = np.maximum(np.abs(y).max(axis=1, keepdims=True), 1e-300)
        e = np.nanmax(np.abs(x - y) / scale)
        say(f"  {key}: max difference / gene's largest entry {e:.2e}")
        bar = 1e-10 if precision == "fp64" else 1e-6
        assert np.array_equal(np.isnan(x), np.isnan(y)) and e <= bar, key
