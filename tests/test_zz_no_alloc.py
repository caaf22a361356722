"""devil_test_de_noalloc (the no-allocation variant of devil_test_de, its
scratch in a fixed per-GPU device array) against devil_test_de through the C
ABI: bit-identical outputs on two calls in a row, the second with fewer
genes, and a clear error beyond the array's size."""
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


def _inputs(G, p, seed, clustered):
    rng = np.random.default_rng(seed)
    beta = rng.normal(0, 0.7, (G, p))
    A = rng.normal(size=(G, p, p)) * 0.1
    bread = A @ A.transpose(0, 2, 1) + np.eye(p)[None] * 0.02
    sw = bread * rng.uniform(0.5, 2.0, (G, 1, 1)) if clustered else None
    # ties and NaNs in the p-values exercise the BH sort
    beta[: G // 10, 1] = beta[G // 10: 2 * (G // 10), 1]
    bread[G - 3:, 1, 1] = np.nan
    return beta, bread, sw


def _run(dv, fn, beta, bread, sw, n, contrast):
    G, p = beta.shape
    ctx = dv.context(0)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    bcol = t(beta.T.reshape(-1))
    br = t(bread.transpose(0, 2, 1).reshape(-1))
    swd = None if sw is None else t(sw.transpose(0, 2, 1).reshape(-1))
    c = t(contrast)
    out = [torch.full((G,), -7.0, dtype=torch.float64, device="cuda") for _ in range(4)]
    ptr = lambda a: None if a is None else a.data_ptr()
    torch.cuda.synchronize()  # the fills above run on torch's stream, the call on its own
    rc = getattr(dv.lib(), fn)(ctx.h, bcol.data_ptr(), br.data_ptr(), ptr(swd), G, p, n,
                               c.data_ptr(), 10.0, out[0].data_ptr(), out[1].data_ptr(),
                               out[2].data_ptr(), out[3].data_ptr())
    return rc, [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("clustered", [False, True])
def test_noalloc_matches_test_de(dv, clustered):
    p, n = 6, 100000
    contrast = np.zeros(p)
    contrast[1] = 1.0
    for G, seed in [(50000, 1), (3001, 2)]:
        say(f"[case] test_de_noalloc G={G} p={p} clustered={clustered}")
        beta, bread, sw = _inputs(G, p, seed, clustered)
        rc0, ref = _run(dv, "devil_test_de", beta, bread, sw, n, contrast)
        rc1, got = _run(dv, "devil_test_de_noalloc", beta, bread, sw, n, contrast)
        assert rc0 == 0 and rc1 == 0
        for a, b, name in zip(got, ref, ["pval", "log_pval", "adj_pval", "lfc"]):
            np.testing.assert_array_equal(a, b, err_msg=name)


def test_noalloc_rejects_too_many_genes(dv):
    say("[case] test_de_noalloc beyond its scratch")
    p = 10
    G = 400000  # 8 G p bytes alone exceed the 16 MiB array
    ctx = dv.context(0)
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    ptr = dummy.data_ptr()
    rc = dv.lib().devil_test_de_noalloc(ctx.h, ptr, ptr, None, G, p, 1000, ptr, 10.0, ptr, ptr,
                                        ptr, ptr)
    assert rc == 1
    assert b"scratch" in dv.lib().devil_last_error(ctx.h)
