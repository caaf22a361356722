"""GPU parity of the mixed-precision IRLS (FP32 MFMA contractions, FP64
partial sums, FP64 solve and stopping decisions) against the FP64 oracle.

Problems with fewer than 4096 cells run the FP64 kernels (the test sizes
here are above that).

Stated tolerance: beta within 1e-4 relative with a floor of 1e-3 on
|reference| (beta is a natural-log effect: the floor is a 0.1% fold change,
far below any coefficient's standard error), log10 p within 1e-4 relative
(floor 1 in log10 units: absolute 1e-4 for p > 0.1), theta within 1e-4
relative (floor 1e-8), identical BH-adjusted p < 0.05 calls, and identical
per-gene IRLS iteration counts (a gene whose step comes within 1e-3 * eps of
the threshold is refit from its start in FP64, so decisions that close are
FP64 decisions; a gene with a coefficient below 5e-3 in magnitude has its last
step redone in FP64).

Why the beta floor: FP32 element arithmetic leaves ~1e-7 relative noise per
cell in the working residuals, which averages to an absolute beta error of
order 1e-8 at n = 2e4 cells (less at larger n); a coefficient of 1e-5 then
has a relative error near 1e-3 although its absolute error is 1e-8.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-4
FLOOR = 1e-6
BETA_FLOOR = 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relmax(a, b, floor=FLOOR):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e[both_nan] = 0.0
    assert np.all(np.isfinite(e)), "NaN in one result only"
    if e.size and e.max() > TOL:
        i = np.unravel_index(int(np.argmax(e)), e.shape)
        say(f"  worst at {i}: {a[i]!r} vs {b[i]!r}")
    return float(e.max()) if e.size else 0.0


# per-variant stated tolerance on beta and log10 p (f32 is the default kernel;
# the bf16 kernel rounds the pair-column products to bf16 and is looser)
KERNEL_TOL = {"f32": TOL, "bf16": TOL, "bf16_ws": TOL, "bf16_ws2": TOL}


@pytest.fixture(scope="module", params=["f32", "bf16", "bf16_ws", "bf16_ws2"])
def dv(request):
    """The package with the context's mixed kernel set to each variant."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import devil_amd
    devil_amd.lib()
    ctx = devil_amd.context(0)
    say(f"[mixed kernel] {request.param}")
    ctx.set_mixed_kernel(request.param)
    devil_amd.mixed_tol = KERNEL_TOL[request.param]
    yield devil_amd
    ctx.set_mixed_kernel("auto")  # library default


def fit_and_oracle(dv, orc, w, Y, *, od="MOM", clusters=None, max_iter=200):
    say(f"[case] mixed {w.name} G={Y.shape[0]} n={Y.shape[1]} p={w.p} od={od}")
    fit = dv.fit_devil(Y, w.X, clusters=clusters, overdispersion=od, size_factors=w.sf,
                       max_iter=max_iter, precision="mixed")
    ends = None if clusters is None else dv.cluster_blocks(clusters, Y.shape[1])
    odm = {"MOM": orc.OD_MOM, "MLE": orc.OD_MLE, False: orc.OD_NONE}[od]
    ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset, clusters_ends=ends, od_mode=odm,
                      max_iter=max_iter)
    return fit, ref


def check_mixed(dv, orc, w, fit, ref, contrast):
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"], ref["beta_iters"])
    eb = relmax(fit["beta"], ref["beta"], BETA_FLOOR)
    et = relmax(fit["overdispersion"], ref["theta"], 1e-8)
    res = dv.test_de(fit, contrast=contrast)
    rr = orc.de_test(ref["beta"], ref["bread"], ref["sandwich"], w.n, contrast)
    lp = res["log_pval"] / np.log(10.0)
    lr = rr["log_pval"] / np.log(10.0)
    ep = relmax(lp, lr, 1.0)
    # the same errors at the 1e-6 floor, recorded, and the absolute beta error
    # that the 1e-3 floor stands for (FP32 element noise, ~4e-7 on C3)
    eb6 = relmax(fit["beta"], ref["beta"], FLOOR)
    ep6 = relmax(lp, lr, FLOOR)
    b = fit["beta"].cpu().numpy() if hasattr(fit["beta"], "cpu") else np.asarray(fit["beta"])
    ea = float(np.nanmax(np.abs(b - ref["beta"])))
    say(f"  beta {eb:.2e}  theta {et:.2e}  log10 p {ep:.2e}  |  floor 1e-6: beta {eb6:.2e} "
        f"log10 p {ep6:.2e}  |  beta absolute {ea:.2e}")
    tol = dv.mixed_tol
    assert eb <= tol and ep <= tol and et <= TOL
    assert ea <= 5e-6
    calls = res["adj_pval"] < 0.05
    calls_ref = rr["adj_pval"] < 0.05
    np.testing.assert_array_equal(calls, calls_ref)


def test_mixed_c1(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C1", G=400, n=5000)
    Y = synth.counts_host(w)
    fit, ref = fit_and_oracle(dv, orc, w, Y)
    check_mixed(dv, orc, w, fit, ref, [0, 1])


def test_mixed_c3_design(dv, orc):
    """The C3 design (intercept, 3-level one-hot, five covariates; p = 8)."""
    from devil_amd import synth
    w = synth.make_workload("C3", G=160, n=20000)
    Y = synth.counts_host(w)
    fit, ref = fit_and_oracle(dv, orc, w, Y)
    check_mixed(dv, orc, w, fit, ref, [0, 1, 0, 0, 0, 0, 0, 0])


def test_mixed_c2_mle(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C2", G=80, n=6000)
    Y = synth.counts_host(w)
    fit, ref = fit_and_oracle(dv, orc, w, Y, od="MLE")
    check_mixed(dv, orc, w, fit, ref, [0, 1, 0, 0])


def test_mixed_clustered_c4(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C4", G=120, n=6000)
    Y = synth.counts_host(w)
    fit, ref = fit_and_oracle(dv, orc, w, Y, clusters=w.clusters)
    check_mixed(dv, orc, w, fit, ref, [0, 1, 0, 0, 0, 0])
    # the clustered sandwich of the mixed path (bread on the bf16 tables, FP64
    # per-patient score sums): normwise per gene against the FP64 oracle
    for key, rk in (("beta_sandwiches", "sandwich"), ("beta_sandwiches_null", "bread")):
        a = np.asarray(fit[key])
        b = ref[rk]
        e = np.abs(a - b).reshape(len(b), -1).max(axis=1) / np.abs(b).reshape(len(b), -1).max(axis=1)
        say(f"  {key} normwise rel err {e.max():.2e}")
        assert e.max() <= 1e-4, key


@pytest.mark.parametrize("p,n", [(1, 4099), (5, 5001), (10, 6007)])
def test_mixed_design_widths(dv, orc, p, n):
    from devil_amd import synth
    rng = np.random.default_rng(300 + p)
    G = 64
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(p - 1)])
    w = synth.Workload("wide", G, n, p, X, np.exp(rng.normal(0, 0.3, n)),
                       np.column_stack([rng.normal(0, 1, G)] +
                                       [rng.normal(0, 0.1, G) for _ in range(p - 1)]),
                       np.exp(rng.normal(0, 1, G)), seed=299 + p)
    Y = synth.counts_host(w)
    fit, ref = fit_and_oracle(dv, orc, w, Y)
    check_mixed(dv, orc, w, fit, ref, [0] + [1] * (p > 1) + [0] * max(0, p - 2))


def test_mixed_edge_cases(dv, orc):
    from devil_amd import synth
    w = synth.make_workload("C1", G=50, n=4500, seed=13)
    Y = synth.counts_host(w)
    Y[3] = 0
    Y[7] = 0
    Y[7, 5] = 40
    Y[9] *= 1000                   # large counts: mu far from 1 in FP32
    fit, ref = fit_and_oracle(dv, orc, w, Y, max_iter=5)
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"], ref["beta_iters"])
    assert relmax(fit["beta"], ref["beta"], BETA_FLOOR) <= dv.mixed_tol


def test_mixed_matches_fp64_gpu_at_scale(dv):
    """Mixed against the FP64 GPU path on a larger problem (no oracle run)."""
    from devil_amd import synth
    w = synth.make_workload("C3", G=2048, n=60000)
    say("[case] mixed vs fp64 GPU, C3 design G=2048 n=60000")
    Yd = synth.counts_device(w, "cuda")
    f64 = dv.fit_devil(Yd, w.X, size_factors=w.sf, precision="fp64")
    mix = dv.fit_devil(Yd, w.X, size_factors=w.sf, precision="mixed")
    it64 = f64["iterations"]["beta_iters"].cpu().numpy()
    itm = mix["iterations"]["beta_iters"].cpu().numpy()
    same = float(np.mean(it64 == itm))
    eb = relmax(mix["beta"].cpu().numpy(), f64["beta"].cpu().numpy(), BETA_FLOOR)
    say(f"  iterations equal {same:.4f}  beta {eb:.2e}")
    assert same == 1.0 and eb <= dv.mixed_tol


def test_mixed_stage_entry_point(dv, orc):
    """devil_beta_fit with DEVIL_PREC_MIXED against the oracle's beta_fit."""
    from devil_amd import _lib as L
    from devil_amd import synth
    say("[case] devil_beta_fit mixed")
    w = synth.make_workload("C2", G=48, n=5000, seed=17)
    Y = synth.counts_host(w)
    ctx = dv.context(0)
    lib = dv.lib()
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                    device="cuda")
    G, n, p = Y.shape[0], Y.shape[1], w.p
    Yf = Y.astype(np.float64)
    b0 = orc.init_beta_qr(Yf, w.X, w.offset)
    a0 = orc.estimate_dispersion(Yf, w.offset)
    # device buffers held by name for the duration of the call
    Yd, Xd, offd, a0d = t(Y, torch.int32), t(w.X.T.reshape(-1)), t(w.offset), t(a0)
    beta = t(b0.T.reshape(-1))
    it = torch.empty(G, dtype=torch.int32, device="cuda")
    ctx.check(lib.devil_beta_fit(ctx.h, Yd.data_ptr(), L.COUNTS_I32, G, n, n, Xd.data_ptr(), p,
                                 beta.data_ptr(), offd.data_ptr(), a0d.data_ptr(), 200, 1e-3,
                                 L.PREC_MIXED, it.data_ptr(), None))
    bn = beta.cpu().numpy().reshape(p, G).T
    for g in range(G):
        rb, ri = orc.beta_fit(Yf[g], w.X, b0[g], w.offset, a0[g], 200, 1e-3)
        assert ri == int(it[g])
        assert relmax(bn[g], rb, BETA_FLOOR) <= dv.mixed_tol


def test_mixed_small_coefficients_c5(dv):
    """Four problems of the C5 set (n = 6386 .. 39023, p = 3 .. 7) whose mixed fits
    left 1.0-1.5e-4 relative error (floor 1e-3) on coefficients below 1e-3
    before genes with small coefficients had their last step redone in FP64:
    within the tolerance against the FP64 path, iteration counts equal."""
    from devil_amd import synth
    probs = synth.make_problems("C5")
    ctx = dv.context(0)
    for j in (10, 22, 23, 39):
        w = probs[j]
        say(f"[case] C5 problem {j}: G={w.G} n={w.n} p={w.p}")
        Yd = synth.counts_device(w, "cuda")
        f64 = dv.fit_devil(Yd, w.X, size_factors=w.sf, precision="fp64")
        mix = dv.fit_devil(Yd, w.X, size_factors=w.sf, precision="mixed")
        assert ctx.fit_info()["polished"] > 0
        np.testing.assert_array_equal(mix["iterations"]["beta_iters"].cpu().numpy(),
                                      f64["iterations"]["beta_iters"].cpu().numpy())
        eb = relmax(mix["beta"].cpu().numpy(), f64["beta"].cpu().numpy(), BETA_FLOOR)
        say(f"  beta {eb:.2e}")
        assert eb <= TOL


def test_mixed_misaligned_device_rows(dv):
    """Device counts whose rows are not 16-byte aligned (n odd) run the
    LDS-staged kernel on the library's aligned copy: results equal the same
    counts fitted from host memory (padded rows) bit for bit."""
    from devil_amd import synth
    w = synth.make_workload("C3", G=256, n=20011, seed=31)
    Y = synth.counts_host(w)
    say("[case] mixed, misaligned device rows, n=20011")
    ctx = dv.context(0)
    ctx.set_mixed_kernel("bf16_ws")
    try:
        fh = dv.fit_devil(Y, w.X, size_factors=w.sf, precision="mixed")
        ih = ctx.fit_info()
        fd = dv.fit_devil(torch.as_tensor(Y, device="cuda"), w.X, size_factors=w.sf,
                          precision="mixed")
        idv = ctx.fit_info()
    finally:
        ctx.set_mixed_kernel("auto")
    assert ih["mixed_kernel"] == "bf16_ws" and idv["mixed_kernel"] == "bf16_ws"
    assert ih["aligned_copy"] == 0 and idv["aligned_copy"] == 1
    np.testing.assert_array_equal(np.asarray(fh["beta"]), fd["beta"].cpu().numpy())
    np.testing.assert_array_equal(np.asarray(fh["iterations"]["beta_iters"]),
                                  fd["iterations"]["beta_iters"].cpu().numpy())
