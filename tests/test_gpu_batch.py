"""Many independent problems in one devil_fit call (n_problems > 1): every
problem's result equals its own single-problem fit (FP64: bit for bit when
each problem runs its own rounds; within 1e-10 with the joint rounds, whose
round shapes can order the FP64 chunk sums differently), the mixed path stays
within its stated tolerance, and errors name the failing problem."""
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


def small_problems(k=7, G=96):
    from devil_amd import synth
    return synth.make_problems(G=G, n_problems=k, n_min=300, n_max=6000, seed=77)


def job(w, Y, **kw):
    return dict(x=Y, design_matrix=w.X, size_factors=w.sf, **kw)


@pytest.fixture(params=[0, 1], ids=["own_rounds", "joint"])
def joint(dv, request):
    from devil_amd import _lib as L
    ctx = dv.context(0)
    ctx.set_tuning(L.TUNE_JOINT, request.param)
    yield request.param
    ctx.set_tuning(L.TUNE_JOINT, 0)


def same(a, b, tol, floor, what):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    if tol == 0:
        np.testing.assert_array_equal(a, b, err_msg=what)
        return
    both = np.isnan(a) & np.isnan(b)
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e[both] = 0.0
    assert np.all(np.isfinite(e)) and e.max() <= tol, f"{what}: {np.nanmax(e):.3e}"


def test_many_equals_single_fp64(dv, orc, joint):
    from devil_amd import synth
    ws = small_problems()
    say(f"[case] {len(ws)} problems, n = {[w.n for w in ws]}, p = {[w.p for w in ws]}")
    Ys = [synth.counts_host(w) for w in ws]
    jobs = [job(w, Y) for w, Y in zip(ws, Ys)]
    # device counts for half of them, a clustered and an MLE problem
    for i in (1, 3, 5):
        jobs[i]["x"] = torch.as_tensor(Ys[i], device="cuda")
    n2 = ws[2].n
    jobs[2]["clusters"] = np.repeat([1, 2, 3], [n2 // 3, n2 // 3, n2 - 2 * (n2 // 3)])
    jobs[4]["overdispersion"] = "MLE"
    many = dv.fit_devil_many(jobs)
    tol = 1e-10 if joint else 0.0
    for i, (j, m) in enumerate(zip(jobs, many)):
        one = dv.fit_devil(**j)
        same(m["iterations"]["beta_iters"], one["iterations"]["beta_iters"], 0.0, 1.0,
             f"problem {i} beta_iters")
        for key, floor in (("beta", 1e-6), ("overdispersion", 1e-8),
                           ("beta_sandwiches_null", 1e-12)):
            same(m[key], one[key], tol, floor, f"problem {i} {key}")
        if j.get("clusters") is not None:
            same(m["beta_sandwiches"], one["beta_sandwiches"], tol, 1e-12,
                 f"problem {i} beta_sandwiches")
    # and against the oracle for one problem
    w, Y = ws[0], Ys[0]
    ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset)
    np.testing.assert_array_equal(many[0]["iterations"]["beta_iters"], ref["beta_iters"])
    e = np.abs(many[0]["beta"] - ref["beta"]) / np.maximum(np.abs(ref["beta"]), 1e-6)
    assert e.max() <= 1e-8


def test_many_mixed_autotuned(dv, orc, joint):
    from devil_amd import _lib as L
    from devil_amd import synth
    ws = small_problems(k=5, G=80)
    Ys = [synth.counts_host(w) for w in ws]
    ctx = dv.context(0)
    ctx.set_mixed_kernel("auto")
    ctx.set_tuning(L.TUNE_CHUNK_CELLS, 0)
    try:
        say("[case] mixed, autotuned kernel and chunk size")
        many = dv.fit_devil_many([job(w, Y, precision="mixed") for w, Y in zip(ws, Ys)])
    finally:  # library defaults
        ctx.set_mixed_kernel("auto")
        ctx.set_tuning(L.TUNE_CHUNK_CELLS, 0)
    for w, Y, m in zip(ws, Ys, many):
        ref = orc.cpu_fit(Y.astype(np.float64), w.X, w.offset)
        np.testing.assert_array_equal(m["iterations"]["beta_iters"], ref["beta_iters"])
        e = np.abs(m["beta"] - ref["beta"]) / np.maximum(np.abs(ref["beta"]), 1e-3)
        assert e.max() <= 1e-4, (w.name, e.max())


def test_many_error_names_problem(dv):
    from devil_amd import synth
    ws = small_problems(k=3, G=32)
    Ys = [synth.counts_host(w) for w in ws]
    jobs = [job(w, Y) for w, Y in zip(ws, Ys)]
    Ybad = Ys[1].copy()
    Ybad[:] = 1
    X = np.column_stack([np.ones(ws[1].n), np.ones(ws[1].n)])  # rank deficient
    jobs[1] = dict(x=Ybad, design_matrix=X)
    with pytest.raises(dv.DevilError) as ei:
        dv.fit_devil_many(jobs)
    assert ei.value.code == 4
    assert "problem 1" in str(ei.value)


def test_joint_matches_own_rounds_mixed(dv):
    """Joint rounds over problems of several design widths, kernels (the
    LDS-staged mixed kernel above 4096 cells, the FP64 kernel below) and row
    alignments against the same problems with their own rounds."""
    from devil_amd import _lib as L
    from devil_amd import synth
    ws = synth.make_problems(G=300, n_problems=9, n_min=1500, n_max=20000, seed=91)
    say(f"[case] joint vs own rounds, n = {[w.n for w in ws]}, p = {[w.p for w in ws]}")
    Ys = [synth.counts_device(w, device="cuda") for w in ws]
    jobs = [dict(x=Y, design_matrix=w.X, size_factors=w.sf, precision="mixed")
            for w, Y in zip(ws, Ys)]
    ctx = dv.context(0)
    out = {}
    try:
        for jv in (0, 1):
            ctx.set_tuning(L.TUNE_JOINT, jv)
            out[jv] = dv.fit_devil_many(jobs)
            out[(jv, "info")] = ctx.fit_info()
    finally:
        ctx.set_tuning(L.TUNE_JOINT, 0)
    for i, (a, b) in enumerate(zip(out[1], out[0])):
        same(a["iterations"]["beta_iters"], b["iterations"]["beta_iters"], 0.0, 1.0,
             f"problem {i} beta_iters")
        same(a["beta"], b["beta"], 1e-4, 1e-3, f"problem {i} beta")
        same(a["overdispersion"], b["overdispersion"], 1e-4, 1e-3, f"problem {i} theta")
        # the own rounds may run another mixed kernel (the autotuner's choice):
        # bread normwise per gene, as the mixed path's bread is stated
        x = a["beta_sandwiches_null"].cpu().numpy()
        y = b["beta_sandwiches_null"].cpu().numpy()
        e = np.abs(x - y).reshape(len(y), -1).max(axis=1) / np.abs(y).reshape(len(y), -1).max(axis=1)
        assert e.max() <= 1e-4, f"problem {i} bread: {e.max():.3e}"
