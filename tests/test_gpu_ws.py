"""GPU: the wave-specialised mixed IRLS kernel (operands staged in LDS by a
loader wavefront and per-wave count rings) computes exactly what the bf16
register kernel computes -- same operands, same products, same summation
order -- so the two fits must agree bit for bit.  The shapes cover a chunk
that holds only a row's partial last step (n = 5128 with 1024-cell chunks),
rows without a partial step (n = 6144), a partial step inside a longer chunk
(n = 5000), 16-byte-misaligned rows (n = 5001: read in place by 4-byte
direct-to-LDS loads, or from the padded copy a mixed fit makes by default),
workgroups with idle tiles (G not a multiple of the tiles per
workgroup; 17 genes leave the second tile of a two-tile wave idle), float
counts, and p = 4 and 8, for the kernel with one and with two 16-gene tiles per
compute wave.  The same kernel computes the
mixed path's bread where there are no clusters (test_mixed_bread_matches_fp64)
and, with patients, the bread and each patient's score sums in one pass
(misaligned rows, patients ending inside a step, patients shorter than a
step, a single patient).
"""
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
    yield devil_amd
    ctx = devil_amd.context(0)
    ctx.set_mixed_kernel("auto")
    ctx.set_tuning(devil_amd._lib.TUNE_CHUNK_CELLS, 0)


def fit_with(dv, kernel, Y, w, pad=1):
    ctx = dv.context(0)
    ctx.set_mixed_kernel(kernel)
    ctx.set_tuning(dv._lib.TUNE_CHUNK_CELLS, 1024)
    ctx.set_tuning(dv._lib.TUNE_PAD_COPY, pad)
    try:
        fit = dv.fit_devil(Y, w.X, size_factors=w.sf, precision="mixed")
    finally:
        ctx.set_tuning(dv._lib.TUNE_PAD_COPY, 1)
    info = ctx.fit_info()
    return fit, info


@pytest.mark.parametrize("kernel", ["bf16_ws", "bf16_ws2"])
@pytest.mark.parametrize("config,G,n,dtype,pad", [
    ("C3", 487, 5128, "int32", 1),
    ("C3", 487, 6144, "int32", 1),
    ("C2", 300, 5000, "int32", 1),
    ("C3", 250, 5001, "int32", 1),
    ("C3", 250, 5001, "int32", 0),
    ("C2", 301, 4097, "float32", 0),
    ("C2", 487, 5128, "float32", 1),
    ("C3", 17, 6144, "int32", 1),
])
def test_ws_equals_bf16(dv, config, G, n, dtype, pad, kernel):
    from devil_amd import synth
    w = synth.make_workload(config, G=G, n=n, seed=4242 + n)
    Y = synth.counts_device(w, device="cuda")
    if dtype == "float32":
        Y = Y.float()
    say(f"[case] {kernel} {config} G={G} n={n} p={w.p} {dtype} pad copy {pad}")
    a, ia = fit_with(dv, "bf16", Y, w, pad)
    b, ib = fit_with(dv, kernel, Y, w, pad)
    assert ib["aligned_copy"] == (1 if pad and n % 4 else 0)
    say(f"  kernels {ia['mixed_kernel']} / {ib['mixed_kernel']}, replayed {ia['replayed']} / "
        f"{ib['replayed']}")
    assert ia["mixed_kernel"] == "bf16" and ib["mixed_kernel"] == kernel
    for key in ("beta", "overdispersion"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        same = (x == y) | (np.isnan(x) & np.isnan(y))
        assert same.all(), f"{key}: {int((~same).sum())} entries differ"
    assert np.array_equal(a["iterations"]["beta_iters"].cpu().numpy(),
                          b["iterations"]["beta_iters"].cpu().numpy())
    assert ia["replayed"] == ib["replayed"]


def test_mixed_bread_matches_fp64(dv):
    """Without clusters the mixed path computes the bread on the bf16 tables
    (FP32 element math, bf16x3 pair products, FP64 sums and inverse): within
    1e-5 of the FP64 bread relative to each gene's largest entry."""
    from devil_amd import synth
    w = synth.make_workload("C3", G=487, n=6144, seed=777)
    Y = synth.counts_device(w, device="cuda")
    ctx = dv.context(0)
    ctx.set_tuning(dv._lib.TUNE_CHUNK_CELLS, 0)
    ref = dv.fit_devil(Y, w.X, size_factors=w.sf, precision="fp64")
    ctx.set_mixed_kernel("bf16_ws")
    fit = dv.fit_devil(Y, w.X, size_factors=w.sf, precision="mixed")
    x = fit["beta_sandwiches_null"].cpu().numpy()
    y = ref["beta_sandwiches_null"].cpu().numpy()
    scale = np.abs(y).reshape(len(y), -1).max(axis=1)
    e = np.abs(x - y).reshape(len(y), -1).max(axis=1) / scale
    say(f"  bread rel err max {e.max():.2e} median {np.median(e):.2e}")
    assert e.max() < 1e-5
    st = fit["status"].cpu().numpy() & 4  # DEVIL_GENE_SINGULAR
    assert np.array_equal(st, ref["status"].cpu().numpy() & 4)


def test_mixed_bread_misaligned_rows(dv):
    """The mixed bread read from misaligned rows in place (4-byte
    direct-to-LDS loads) equals the one read from the padded copy."""
    from devil_amd import synth
    w = synth.make_workload("C3", G=300, n=6143, seed=778)
    Y = synth.counts_device(w, device="cuda")
    say("[case] mixed bread, n = 6143, in place vs padded copy")
    a, ia = fit_with(dv, "bf16_ws", Y, w, 0)
    b, ib = fit_with(dv, "bf16_ws", Y, w, 1)
    assert ia["aligned_copy"] == 0 and ib["aligned_copy"] == 1
    for key in ("beta", "overdispersion", "beta_sandwiches_null"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        same = (x == y) | (np.isnan(x) & np.isnan(y))
        assert same.all(), f"{key}: {int((~same).sum())} entries differ"


def fit_clustered(dv, Y, w, clusters, precision, pad=1, kernel="bf16_ws", od="MOM"):
    ctx = dv.context(0)
    ctx.set_mixed_kernel(kernel)
    ctx.set_tuning(dv._lib.TUNE_PAD_COPY, pad)
    try:
        fit = dv.fit_devil(Y, w.X, clusters=clusters, size_factors=w.sf, precision=precision,
                           overdispersion=od)
    finally:
        ctx.set_tuning(dv._lib.TUNE_PAD_COPY, 1)
    return fit, ctx.fit_info()


def normwise(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    a = a.reshape(len(a), -1)
    b = b.reshape(len(b), -1)
    return np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)


def test_mixed_clustered_misaligned_rows(dv):
    """The mixed bread + per-patient score pass read from misaligned rows in
    place (4-byte direct-to-LDS loads) equals the one read from the padded
    copy, bread and sandwich bit for bit."""
    from devil_amd import synth
    w = synth.make_workload("C4", G=200, n=6001, seed=779)
    Y = synth.counts_device(w, device="cuda")
    say("[case] mixed clustered, n = 6001, in place vs padded copy")
    a, ia = fit_clustered(dv, Y, w, w.clusters, "mixed", pad=0)
    b, ib = fit_clustered(dv, Y, w, w.clusters, "mixed", pad=1)
    assert ia["aligned_copy"] == 0 and ib["aligned_copy"] == 1
    for key in ("beta", "overdispersion", "beta_sandwiches_null", "beta_sandwiches"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        same = (x == y) | (np.isnan(x) & np.isnan(y))
        assert same.all(), f"{key}: {int((~same).sum())} entries differ"


@pytest.mark.parametrize("sizes", ["one patient", "tiny patients", "ragged patients"])
def test_mixed_clustered_patient_shapes(dv, sizes):
    """Patient blocks that end inside a 32-cell step, patients of 1-40 cells
    (chunks of less than one step), and a single patient (adj = 1): the mixed
    sandwich within 1e-4 normwise of the FP64 path's, which the parity tests
    hold to the oracle.  A single patient's score sum is the whole gradient
    at the fitted beta, zero up to the IRLS tolerance, so its sandwich is
    that residual's outer product: there the bar is 1e-4 of the bread's
    largest entry (the scale of a sandwich from more than one patient)."""
    from devil_amd import synth
    w = synth.make_workload("C4", G=96, n=5003, seed=780)
    n = w.n  # the synthetic patient sizes round the cell count
    Y = synth.counts_device(w, device="cuda")
    rng = np.random.default_rng(781)
    if sizes == "one patient":
        lab = np.zeros(n, np.int64)
    elif sizes == "tiny patients":
        cuts = np.cumsum(rng.integers(1, 41, size=n))
        lab = np.searchsorted(cuts, np.arange(n), side="right")
    else:
        cuts = np.sort(rng.choice(np.arange(1, n), size=24, replace=False))
        lab = np.searchsorted(cuts, np.arange(n), side="right")
    C = int(np.unique(lab).shape[0])
    say(f"[case] mixed clustered, {sizes}: {C} patients over {n} cells")
    ref, _ = fit_clustered(dv, Y, w, lab, "fp64")
    fit, info = fit_clustered(dv, Y, w, lab, "mixed")
    assert info["mixed_kernel"] == "bf16_ws"
    for key in ("beta_sandwiches_null", "beta_sandwiches"):
        e = normwise(fit[key], ref[key])
        say(f"  {key} normwise rel err max {e.max():.2e} median {np.median(e):.2e}")
        if key == "beta_sandwiches" and C == 1:
            a = fit[key].cpu().numpy().reshape(w.G, -1)
            b = ref[key].cpu().numpy().reshape(w.G, -1)
            scale = np.abs(ref["beta_sandwiches_null"].cpu().numpy()).reshape(w.G, -1).max(axis=1)
            e = np.abs(a - b).max(axis=1) / scale
            say(f"  {key} error / largest bread entry: max {e.max():.2e}")
        assert e.max() <= 1e-4, key


@pytest.mark.parametrize("od,dtype", [("MLE", "int32"), ("MOM", "float32")])
def test_mixed_clustered_mle_and_float_counts(dv, od, dtype):
    """The mixed bread + per-patient score pass after the MLE overdispersion
    (theta from the FP64 nlminb driver) and on float32 counts: the same
    iteration counts as the FP64 path, beta within 1e-4 (floor 1e-3), theta
    within 1e-4 and the bread and sandwich within 1e-4 normwise of the FP64
    path's."""
    from devil_amd import synth
    w = synth.make_workload("C4", G=128, n=6000, seed=782)
    Y = synth.counts_device(w, device="cuda")
    if dtype == "float32":
        Y = Y.float()
    say(f"[case] mixed clustered, {od}, {dtype} counts, G={w.G} n={w.n}")
    ref, _ = fit_clustered(dv, Y, w, w.clusters, "fp64", od=od)
    fit, info = fit_clustered(dv, Y, w, w.clusters, "mixed", od=od)
    assert info["mixed_kernel"] == "bf16_ws"
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"].cpu().numpy(),
                                  ref["iterations"]["beta_iters"].cpu().numpy())
    b, rb = fit["beta"].cpu().numpy(), ref["beta"].cpu().numpy()
    eb = float(np.max(np.abs(b - rb) / np.maximum(np.abs(rb), 1e-3)))
    t, rt = fit["overdispersion"].cpu().numpy(), ref["overdispersion"].cpu().numpy()
    et = float(np.max(np.abs(t - rt) / np.maximum(np.abs(rt), 1e-8)))
    say(f"  beta {eb:.2e}  theta {et:.2e}")
    assert eb <= 1e-4 and et <= 1e-4
    for key in ("beta_sandwiches_null", "beta_sandwiches"):
        e = normwise(fit[key], ref[key])
        say(f"  {key} normwise rel err max {e.max():.2e}")
        assert e.max() <= 1e-4, key


def test_mixed_clustered_deterministic(dv):
    """Two mixed clustered fits of the same problem agree bit for bit (the
    partial sums have a fixed order; no floating-point atomics)."""
    from devil_amd import synth
    w = synth.make_workload("C4", G=100, n=6001, seed=783)
    Y = synth.counts_device(w, device="cuda")
    say("[case] mixed clustered, twice")
    a, _ = fit_clustered(dv, Y, w, w.clusters, "mixed")
    b, _ = fit_clustered(dv, Y, w, w.clusters, "mixed")
    for key in ("beta", "overdispersion", "beta_sandwiches_null", "beta_sandwiches"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        same = (x == y) | (np.isnan(x) & np.isnan(y))
        assert same.all(), f"{key}: {int((~same).sum())} entries differ"


def test_mixed_new_overdispersion(dv):
    """overdispersion="new" after the mixed IRLS, against the FP64 path (which
    the parity tests hold to the oracle): identical beta iteration counts,
    beta within 1e-4 (floor 1e-3), the same genes failed, and theta within the
    method's own tolerance (1e-3, src/new_overdispersion.cpp:284): its Laplace
    mode is found to that tolerance, so its F_k, and theta, jump with the
    Newton count when beta moves by FP32 rounding."""
    from devil_amd import synth, _lib as L
    w = synth.make_workload("C3", G=256, n=6144, seed=784)
    Y = synth.counts_device(w, device="cuda")
    say("[case] mixed, new overdispersion")
    ctx = dv.context(0)
    ctx.set_mixed_kernel("bf16_ws")
    ref = dv.fit_devil(Y, w.X, size_factors=w.sf, overdispersion="new", precision="fp64")
    fit = dv.fit_devil(Y, w.X, size_factors=w.sf, overdispersion="new", precision="mixed")
    np.testing.assert_array_equal(fit["iterations"]["beta_iters"].cpu().numpy(),
                                  ref["iterations"]["beta_iters"].cpu().numpy())
    b, rb = fit["beta"].cpu().numpy(), ref["beta"].cpu().numpy()
    eb = float(np.max(np.abs(b - rb) / np.maximum(np.abs(rb), 1e-3)))
    failed = (fit["status"].cpu().numpy() & L.GENE_OD_FAILED) != 0
    rfailed = (ref["status"].cpu().numpy() & L.GENE_OD_FAILED) != 0
    np.testing.assert_array_equal(failed, rfailed)
    t, rt = fit["overdispersion"].cpu().numpy(), ref["overdispersion"].cpu().numpy()
    ok = ~rfailed
    et = np.abs(t[ok] - rt[ok]) / np.maximum(np.abs(rt[ok]), 1e-8)
    say(f"  beta {eb:.2e}  theta rel err median {np.median(et):.2e} p90 "
        f"{np.quantile(et, 0.9):.2e} max {et.max():.2e} ({ok.sum()} genes)")
    assert eb <= 1e-4
    assert et.max() <= 1e-3
