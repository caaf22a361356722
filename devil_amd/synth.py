"""Synthetic single-cell count workloads (the benchmark configurations).

y_gi ~ NB(mean s_i exp(x_i' beta_g), size theta_g), drawn as
Poisson(Gamma(theta_g, mean/theta_g)); theta_g ~ LogNormal(0, 1) is the NB
size, s_i ~ LogNormal(0, 0.3) the size factors.  Host generation uses
numpy's Philox bit generator; device generation uses torch's Philox-based
generator on the GPU in gene chunks.  Seeds are 20260000 + k for config k.

Configurations (genes x cells, design width p):
  C1  2000 x 5000,     p=2 [1, Bern(.5)]; beta0~N(0,1), beta1~N(0,.5^2); MOM
  C2  20000 x 100000,  p=4 [1, Bern, z1, z2]; beta0~N(-2,1.5^2) (~90% zeros),
                            beta1~N(0,.5^2), beta2,3~N(0,.25^2); MLE
  C3  20000 x 500000,  p=8 [1, 1{lvl2}, 1{lvl3}, z1..z5]; beta0~N(0,1),
                            level effects N(0,.5^2), covariates N(0,.1^2); MOM
  C4  15000 x ~1.2M,   p=6 [1, Bern, z1..z4]; 40 patients of U{2000..60000}
                            cells, patient effect N(0,.3^2) per gene; MOM, clustered
  C5  64 independent problems over 15000 shared genes: problem j has
      n_j = round(exp(U(log 500, log 2e5))) cells (sum ~2M) and design
      [1, Bern, z1..z(p_j - 2)] with p_j = 2 + (j mod 9); beta per (problem,
      gene) as in C4; MOM (make_problems)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    G: int
    n: int
    p: int
    X: np.ndarray              # n x p float64
    sf: np.ndarray             # n size factors
    beta: np.ndarray           # G x p true coefficients
    size: np.ndarray           # G NB size (theta_g)
    od: str = "MOM"
    cluster_ends: np.ndarray | None = None
    patient_sd: float = 0.0
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def offset(self) -> np.ndarray:
        return np.log(self.sf)

    @property
    def clusters(self) -> np.ndarray | None:
        if self.cluster_ends is None:
            return None
        sizes = np.diff(np.concatenate([[0], self.cluster_ends]))
        return np.repeat(np.arange(1, len(sizes) + 1), sizes)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def make_workload(config: str, G: int | None = None, n: int | None = None,
                  seed: int | None = None) -> Workload:
    """Design, size factors and true parameters of a configuration (counts are
    drawn separately by ``counts_host`` / ``counts_device``).  ``G`` and
    ``n`` shrink a configuration for tests."""
    k = {"C1": 1, "C2": 2, "C3": 3, "C4": 4}[config]
    seed = 20260000 + k if seed is None else seed
    r = _rng(seed)
    if config == "C1":
        G = G or 2000
        n = n or 5000
        X = np.column_stack([np.ones(n), r.uniform(size=n) < 0.5])
        B = np.column_stack([r.normal(0, 1, G), r.normal(0, 0.5, G)])
        od, ends, psd = "MOM", None, 0.0
    elif config == "C2":
        G = G or 20000
        n = n or 100000
        X = np.column_stack([np.ones(n), r.uniform(size=n) < 0.5, r.normal(size=(n, 2))])
        B = np.column_stack([r.normal(-2, 1.5, G), r.normal(0, 0.5, G),
                             r.normal(0, 0.25, (G, 2))])
        od, ends, psd = "MLE", None, 0.0
    elif config == "C3":
        G = G or 20000
        n = n or 500000
        lvl = r.integers(0, 3, n)
        X = np.column_stack([np.ones(n), lvl == 1, lvl == 2, r.normal(size=(n, 5))])
        B = np.column_stack([r.normal(0, 1, G), r.normal(0, 0.5, (G, 2)),
                             r.normal(0, 0.1, (G, 5))])
        od, ends, psd = "MOM", None, 0.0
    else:
        G = G or 15000
        n_pat = 40
        if n is None:
            sizes = r.integers(2000, 60001, n_pat)
        else:
            sizes = np.maximum(1, np.round(r.uniform(0.2, 1.8, n_pat) * n / n_pat)).astype(int)
            sizes[-1] = max(1, n - sizes[:-1].sum())
        n = int(sizes.sum())
        X = np.column_stack([np.ones(n), r.uniform(size=n) < 0.5, r.normal(size=(n, 4))])
        B = np.column_stack([r.normal(0, 1, G), r.normal(0, 0.5, G), r.normal(0, 0.1, (G, 4))])
        od, ends, psd = "MOM", np.cumsum(sizes).astype(np.int32), 0.3
    X = X.astype(np.float64)
    sf = np.exp(r.normal(0, 0.3, n))
    size = np.exp(r.normal(0, 1, G))
    return Workload(config, G, n, X.shape[1], X, sf, B, size, od, ends, psd, seed)


def make_problems(config: str = "C5", G: int | None = None, n_problems: int = 64,
                  n_min: int = 500, n_max: int = 200000, seed: int | None = None) -> list:
    """The independent problems of the many-problems configuration: one
    Workload per problem (shared genes, own cells, design width and
    coefficients).  ``G``, ``n_problems``, ``n_min`` and ``n_max`` shrink it
    for tests."""
    if config != "C5":
        raise ValueError("make_problems builds the C5 configuration")
    seed = 20260005 if seed is None else seed
    r = _rng(seed)
    G = G or 15000
    ns = np.round(np.exp(r.uniform(np.log(n_min), np.log(n_max), n_problems))).astype(int)
    out = []
    for j in range(n_problems):
        p = 2 + (j % 9)
        n = int(ns[j])
        X = np.column_stack([np.ones(n), r.uniform(size=n) < 0.5] +
                            [r.normal(size=n) for _ in range(p - 2)]).astype(np.float64)
        B = np.column_stack([r.normal(0, 1, G), r.normal(0, 0.5, G)] +
                            [r.normal(0, 0.1, G) for _ in range(p - 2)])
        sf = np.exp(r.normal(0, 0.3, n))
        size = np.exp(r.normal(0, 1, G))
        out.append(Workload(f"C5.{j}", G, n, p, X, sf, B, size, "MOM", None, 0.0,
                            seed + 1000 * (j + 1)))
    return out


def counts_host(w: Workload, genes: np.ndarray | None = None, dtype=np.int32) -> np.ndarray:
    """Draw the G x n counts on the host (small configurations)."""
    r = _rng(w.seed + 1000)
    lin = w.X @ w.beta.T  # n x G
    mu = w.sf[:, None] * np.exp(lin)
    if w.cluster_ends is not None and w.patient_sd > 0:
        C = len(w.cluster_ends)
        u = r.normal(0, w.patient_sd, (C, w.G))
        sizes = np.diff(np.concatenate([[0], w.cluster_ends]))
        mu = mu * np.exp(np.repeat(u, sizes, axis=0))
    lam = r.gamma(w.size[None, :], mu / w.size[None, :])
    Y = r.poisson(lam).T.astype(dtype)
    if genes is not None:
        Y = Y[genes]
    return np.ascontiguousarray(Y)


def counts_device(w: Workload, device="cuda", chunk_genes: int | None = None,
                  dtype=None):
    """Draw the G x n counts on the GPU with torch (int32, gene-major), in
    gene chunks so temporaries stay small."""
    import torch
    dtype = dtype or torch.int32
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(w.seed + 2000)
    Y = torch.empty((w.G, w.n), dtype=dtype, device=dev)
    X = torch.as_tensor(w.X, dtype=torch.float32, device=dev)
    logsf = torch.as_tensor(np.log(w.sf), dtype=torch.float32, device=dev)
    B = torch.as_tensor(w.beta, dtype=torch.float32, device=dev)
    size = torch.as_tensor(w.size, dtype=torch.float32, device=dev)
    if w.cluster_ends is not None and w.patient_sd > 0:
        sizes = np.diff(np.concatenate([[0], w.cluster_ends]))
        pid = torch.as_tensor(np.repeat(np.arange(len(sizes)), sizes), device=dev)
        U = torch.randn((w.G, len(sizes)), generator=gen, device=dev) * w.patient_sd
    else:
        pid = U = None
    cg = chunk_genes or max(1, int(2e8 // w.n))
    for g0 in range(0, w.G, cg):
        g1 = min(w.G, g0 + cg)
        lin = B[g0:g1] @ X.T + logsf[None, :]
        if U is not None:
            lin = lin + U[g0:g1][:, pid]
        mu = torch.exp(lin)
        shape = size[g0:g1, None].expand(-1, w.n).contiguous()
        lam = torch._standard_gamma(shape, generator=gen) * (mu / size[g0:g1, None])
        Y[g0:g1] = torch.poisson(lam, generator=gen).to(dtype)
        del lin, mu, shape, lam
    return Y
