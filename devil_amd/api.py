"""Python mirror of devil's user API for the fitting path.

``fit_devil`` and ``test_de`` keep the argument names, meanings and error
behaviour of the R functions (R/main.R:164-273, R/test.R:85-134), and run the
fit through libdevil_hip.so on an MI355X.  Counts may be a numpy array (host;
the library streams them to HBM in gene batches) or a torch tensor already on
the GPU (nothing is copied); with GPU counts the design matrix and size factors
may be GPU tensors too.  There is no CPU fitting path.
"""
from __future__ import annotations

import ctypes as C
from typing import Any

import numpy as np

from . import _lib as L

_OD_MODES = {"MOM": L.OD_MOM, "MLE": L.OD_MLE, "old": L.OD_MLE, "new": L.OD_NEW,
             "I": L.OD_NEW}


def _inputs_ready(device: int) -> None:
    """The library's stream does not wait for other streams: finish the work
    torch has queued on its current stream (device inputs just made) before a
    library call reads those tensors.  Outputs are complete when a call
    returns."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _counts_dtype(dt) -> int:
    s = str(dt).replace("torch.", "")
    if s == "int32":
        return L.COUNTS_I32
    if s == "float32":
        return L.COUNTS_F32
    if s == "float64":
        return L.COUNTS_F64
    raise TypeError(f"counts dtype {dt} not supported (int32, float32, float64)")


def cluster_blocks(clusters, n: int) -> np.ndarray:
    """R/main.R:190-196: codes by first appearance, must be sorted, then the
    cumulative run-length ends (1-based)."""
    lab = np.asarray(clusters.cpu() if _is_torch(clusters) else clusters)
    if lab.shape[0] != n:
        raise ValueError("clusters must have one entry per cell")
    if n == 0:
        return np.zeros(0, np.int32)
    # codes by first appearance never decrease exactly when every label's
    # cells form one run: the runs' labels are then all distinct
    change = np.flatnonzero(lab[1:] != lab[:-1])
    run_labels = lab[np.concatenate([[0], change + 1])]
    if np.unique(run_labels).shape[0] != run_labels.shape[0]:
        raise ValueError("Input data should be grouped in blocks of patients. "
                         "Use the `group_data` function to group them.")
    ends = np.concatenate([change + 1, [n]]).astype(np.int32)
    return ends


def _is_sparse(x) -> bool:
    return type(x).__module__.startswith("scipy.sparse")


def densify(x, device: int = 0):
    """A scipy.sparse G x n count matrix (R's dgCMatrix: CSC, column = cell)
    as a dense gene-major torch tensor on the GPU, built by the library's
    scatter kernel (devil_csc_to_dense).  int32 when every value is integral
    and fits, else float64."""
    import torch
    csc = x.tocsc()
    G, n = csc.shape
    dev = torch.device("cuda", device)
    data = np.asarray(csc.data, dtype=np.float64)
    integral = bool(np.all(np.floor(data) == data)) and (
        data.size == 0 or (data.min() >= -2**31 and data.max() < 2**31))
    out_dtype = L.COUNTS_I32 if integral else L.COUNTS_F64
    out = torch.empty((G, n), dtype=torch.int32 if integral else torch.float64, device=dev)
    colptr = torch.as_tensor(np.asarray(csc.indptr, dtype=np.int64), device=dev)
    rowind = torch.as_tensor(np.asarray(csc.indices, dtype=np.int32), device=dev)
    vals = torch.as_tensor(data, device=dev)
    ctx = L.context(device)
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_csc_to_dense(ctx.h, colptr.data_ptr(), rowind.data_ptr(),
                                         vals.data_ptr(), G, n, out_dtype, out.data_ptr(), n))
    return out


def calculate_sf(Y, method: str = "normed_sum", verbose: bool = False, device: int = 0):
    """R/size_factors.R:37-127 on the GPU (devil_size_factors): "normed_sum"
    or "psinorm"; "edgeR" raises as the reference does without edgeR.
    ``Y`` is a genes x cells torch CUDA tensor (int32, float32 or float64);
    returns a float64 tensor of n size factors."""
    import torch
    if method not in L.SF_METHODS:
        raise ValueError(f"'arg' should be one of \"normed_sum\", \"psinorm\", \"edgeR\"")
    if verbose:
        print(f"Calculating size factors using method: {method}")
    if not Y.is_contiguous() and Y.stride(1) != 1:
        Y = Y.contiguous()
    G, n = (int(v) for v in Y.shape)
    sf = torch.empty(n, dtype=torch.float64, device=Y.device)
    ctx = L.context(Y.device.index or 0)
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_size_factors(ctx.h, Y.data_ptr(), _counts_dtype(Y.dtype), G, n,
                                         Y.stride(0), L.SF_METHODS[method], sf.data_ptr()))
    return sf


def group_data(input_matrix, design_matrix, clusters, device: int = 0):
    """R/group_data.R:38-69: cells reordered so that each cluster is one
    contiguous block (clusters coded by first appearance, stable order).  The
    count matrix is permuted on the GPU (devil_group_cells); returns
    dict(input_matrix, design_matrix, clusters) like the reference's list."""
    import torch
    n = int(input_matrix.shape[1])
    if int(design_matrix.shape[0]) != n:
        raise ValueError("Dimension mismatch: ncol(input_matrix) must equal nrow(design_matrix)")
    lab = np.asarray(clusters.cpu() if _is_torch(clusters) else clusters)
    if lab.shape[0] != n:
        raise ValueError("Dimension mismatch: length(clusters) must match the number of cells")
    _, first, inv = np.unique(lab, return_index=True, return_inverse=True)
    rank = np.empty(len(first), np.int64)
    rank[np.argsort(first)] = np.arange(len(first))
    codes = rank[inv] + 1
    order = np.argsort(codes, kind="stable")
    Y = input_matrix
    on_host = not (_is_torch(Y) and Y.is_cuda)
    dev = torch.device("cuda", device) if on_host else Y.device
    if on_host:
        Y = torch.as_tensor(np.ascontiguousarray(Y), device=dev)
    if not Y.is_contiguous() and Y.stride(1) != 1:
        Y = Y.contiguous()
    G = int(Y.shape[0])
    out = torch.empty((G, n), dtype=Y.dtype, device=dev)
    ordd = torch.as_tensor(order, device=dev)
    ctx = L.context(dev.index or 0)
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_group_cells(ctx.h, Y.data_ptr(), _counts_dtype(Y.dtype), G, n,
                                        Y.stride(0), ordd.data_ptr(), out.data_ptr(), n))
    if _is_torch(design_matrix):
        X = design_matrix[torch.as_tensor(order, device=design_matrix.device)]
    else:
        X = np.asarray(design_matrix)[order]
    return {"input_matrix": out.cpu().numpy() if on_host else out, "design_matrix": X,
            "clusters": codes[order].astype(np.float64)}


def compute_sandwich(design_matrix, y, beta, overdispersion, size_factors, clusters=None,
                     device: int = 0):
    """R/variance.R:16-26 for every gene at once on the GPU: bread meat bread n
    with the clustered meat (contiguous patient blocks, as fit_devil requires)
    or, without clusters, the HC0 meat (1/n) sum_i s_i s_i' of compute_meat.
    ``y`` is genes x cells, ``beta`` genes x p, ``overdispersion`` per gene.
    The bread uses theta as fit_devil's does.  Returns a genes x p x p tensor."""
    import torch
    dev = torch.device("cuda", device)
    Y = y if (_is_torch(y) and y.is_cuda) else torch.as_tensor(np.ascontiguousarray(y),
                                                               device=dev)
    if not Y.is_contiguous() and Y.stride(1) != 1:
        Y = Y.contiguous()
    G, n = (int(v) for v in Y.shape)
    f64 = lambda a: (a.to(device=dev, dtype=torch.float64) if _is_torch(a)
                     else torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev))
    X = f64(design_matrix)
    p = int(X.shape[1])
    Xd = X.t().contiguous().reshape(-1)
    bcol = f64(beta).reshape(G, p).t().contiguous().reshape(-1)
    th = f64(overdispersion).reshape(-1).contiguous()
    off = torch.log(f64(size_factors).reshape(-1)).contiguous()
    bread = torch.empty(G * p * p, dtype=torch.float64, device=dev)
    sw = torch.empty(G * p * p, dtype=torch.float64, device=dev)
    ctx = L.context(device)
    lib = L.lib()
    args = (ctx.h, Y.data_ptr(), _counts_dtype(Y.dtype), G, n, Y.stride(0), Xd.data_ptr(), p,
            bcol.data_ptr(), th.data_ptr(), off.data_ptr())
    if clusters is None:
        _inputs_ready(ctx.device)
        ctx.check(lib.devil_sandwich_hc0(*args, bread.data_ptr(), sw.data_ptr(), None))
    else:
        ends = torch.as_tensor(cluster_blocks(clusters, n), dtype=torch.int32, device=dev)
        _inputs_ready(ctx.device)
        ctx.check(lib.devil_bread_meat(*args, ends.data_ptr(), int(ends.numel()),
                                       bread.data_ptr(), sw.data_ptr(), None))
    return sw.view(G, p, p).transpose(1, 2)


def compute_offset_vector(off, n: int, size_factors) -> np.ndarray:
    """R/size_factors.R:172-184: rep(off, n) + log(size_factors)."""
    v = np.full(n, float(off))
    if size_factors is not None:
        v = v + np.log(np.asarray(size_factors, dtype=np.float64))
    return v


class _Job:
    """One problem's C-ABI structs, the buffers they point to and the
    translation of its outputs to fit_devil's return value."""

    def __init__(self, x, design_matrix, clusters=None, overdispersion: Any = "MOM",
                 init_overdispersion=None, init_beta_rough: bool = False, offset: float = 0,
                 size_factors=None, max_iter: int = 200, tolerance: float = 1e-3,
                 precision: str = "fp64", init_mode: str | None = None, device: int = 0):
        # a pandas DataFrame carries the gene names as its index (R's rownames)
        self.gene_names = None
        if hasattr(x, "index") and hasattr(x, "to_numpy"):
            self.gene_names = list(x.index)
            x = x.to_numpy()
        # sparse assays (R/main.R:1-9): densified on the GPU, gene-major
        if _is_sparse(x):
            x = densify(x, device)
        # size factors estimated from the counts (R/main.R:215-224) need the
        # counts on the GPU; the fit then reads that copy
        if isinstance(size_factors, str):
            if not (_is_torch(x) and x.is_cuda):
                import torch
                x = torch.as_tensor(np.ascontiguousarray(x), device=torch.device("cuda", device))
            size_factors = calculate_sf(x, size_factors)
        self.on_device = _is_torch(x) and x.is_cuda
        x_dev = _is_torch(design_matrix) and design_matrix.is_cuda and self.on_device
        sf_dev = _is_torch(size_factors) and size_factors.is_cuda and self.on_device
        if x_dev:
            Xa = None
            if design_matrix.dim() != 2:
                raise ValueError("design_matrix must be a matrix")
            n, p = (int(v) for v in design_matrix.shape)
        else:
            Xa = np.asarray(design_matrix.cpu() if _is_torch(design_matrix) else design_matrix,
                            dtype=np.float64)
            if Xa.ndim != 2:
                raise ValueError("design_matrix must be a matrix")
            n, p = Xa.shape
        G = int(x.shape[0])
        if int(x.shape[1]) != n:
            raise ValueError("ncol(x) must equal nrow(design_matrix)")
        if overdispersion is False:
            od = L.OD_NONE
        elif overdispersion in _OD_MODES:
            od = _OD_MODES[overdispersion]
        else:
            raise ValueError(f"Unknown overdispersion mode: {overdispersion}")
        if sf_dev:
            import torch
            sf = size_factors.to(torch.float64).reshape(-1)
            if sf.numel() != n:
                raise ValueError("size_factors must have one entry per cell")
            offset_vector = float(offset) + torch.log(sf)
        else:
            sf = np.ones(n) if size_factors is None else np.asarray(
                size_factors.cpu() if _is_torch(size_factors) else size_factors,
                dtype=np.float64)
            offset_vector = compute_offset_vector(offset, n, sf)
        ends = None if clusters is None else cluster_blocks(clusters, n)
        if init_mode is None:
            init = L.INIT_ROUGH if init_beta_rough else L.INIT_QR
        else:
            init = {"qr": L.INIT_QR, "rough": L.INIT_ROUGH,
                    "univariate": L.INIT_UNIVARIATE}[init_mode]
        prec = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}[precision]
        disp = None
        if init_overdispersion is not None:
            disp = np.broadcast_to(np.asarray(init_overdispersion, dtype=np.float64),
                                   (G,)).copy()
        self.G, self.n, self.p, self.od = G, n, p, od
        self.meta = dict(size_factors=sf, offset_vector=offset_vector,
                         design_matrix=design_matrix, input_matrix=x,
                         input_parameters={"max_iter": max_iter, "tolerance": tolerance})
        common = dict(G=G, n=n, p=p, n_clusters=0 if ends is None else len(ends),
                      beta_init=None, init_mode=init, od_mode=od, max_iter=int(max_iter),
                      tolerance=float(tolerance), precision=prec)
        if self.on_device:
            import torch
            dev = x.device
            if not x.is_contiguous() and x.stride(1) != 1:
                x = x.contiguous()
            t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=dev)
            # X column-major on the device: a C-ordered X is uploaded as is and
            # transposed there (no host copies of the n x p matrix); a GPU X is
            # used in place when it is column-major already
            if x_dev:
                Xd = design_matrix.to(device=dev, dtype=torch.float64).t().contiguous().reshape(-1)
            elif Xa.flags.f_contiguous:
                Xd = t(Xa.ravel(order="F"))
            else:
                Xd = t(np.ascontiguousarray(Xa)).t().contiguous().reshape(-1)
            offd = (offset_vector.to(dev).contiguous() if sf_dev else t(offset_vector))
            self.hold = dict(x=x, Xd=Xd, offd=offd,
                             endsd=None if ends is None else t(ends, torch.int32),
                             dispd=None if disp is None else t(disp))
            e = lambda *shape, dt=torch.float64: torch.empty(*shape, dtype=dt, device=dev)
            self.out = dict(beta=e(G * p), theta=e(G), bit=e(G, dt=torch.int32),
                            tit=e(G, dt=torch.int32), bread=e(G * p * p),
                            sw=e(G * p * p) if ends is not None else None,
                            st=e(G, dt=torch.int32))
            ptr = lambda a: None if a is None else a.data_ptr()
            h = self.hold
            self.prob = L.Problem(mem=L.MEM_DEVICE, counts=x.data_ptr(),
                                  counts_dtype=_counts_dtype(x.dtype), counts_ld=x.stride(0),
                                  X=ptr(h["Xd"]), offset=ptr(h["offd"]),
                                  cluster_ends=ptr(h["endsd"]), disp_init=ptr(h["dispd"]),
                                  **common)
        else:
            X = np.asfortranarray(Xa)
            Y = np.asarray(x)
            if not Y.flags.c_contiguous:
                Y = np.ascontiguousarray(Y)
            if Y.dtype not in (np.int32, np.float32, np.float64):
                Y = Y.astype(np.float64)
            self.hold = dict(Y=Y, X=X, off=offset_vector, ends=ends, disp=disp)
            self.out = dict(beta=np.zeros((G, p), order="F"), theta=np.zeros(G),
                            bit=np.zeros(G, np.int32), tit=np.zeros(G, np.int32),
                            bread=np.zeros((G, p, p)),
                            sw=np.zeros((G, p, p)) if ends is not None else None,
                            st=np.zeros(G, np.int32))
            ptr = lambda a: None if a is None else a.ctypes.data
            self.prob = L.Problem(mem=L.MEM_HOST, counts=ptr(Y),
                                  counts_dtype=_counts_dtype(Y.dtype), counts_ld=n, X=ptr(X),
                                  offset=ptr(offset_vector), cluster_ends=ptr(ends),
                                  disp_init=ptr(disp), **common)
        o = self.out
        self.outs = L.Outputs(beta=ptr(o["beta"]), theta=ptr(o["theta"]),
                              beta_iters=ptr(o["bit"]), theta_iters=ptr(o["tit"]),
                              bread=ptr(o["bread"]), sandwich=ptr(o["sw"]), status=ptr(o["st"]))

    def result(self):
        G, p, o = self.G, self.p, self.out
        if self.on_device:
            beta = o["beta"].view(p, G).T
            bread = o["bread"].view(G, p, p).transpose(1, 2)
            sw = None if o["sw"] is None else o["sw"].view(G, p, p).transpose(1, 2)
        else:
            # per-gene blocks are column-major: [g, a, b] = M[a + b p]
            beta = o["beta"]
            bread = o["bread"].transpose(0, 2, 1)
            sw = None if o["sw"] is None else o["sw"].transpose(0, 2, 1)
        # theta_iters: 0 for MOM / no overdispersion, NA for MLE (cpu_fit,
        # R/main.R:431-468); the MLE score passes are kept as theta_passes
        if self.od == L.OD_MLE:
            if self.on_device:
                import torch
                tit = torch.full((G,), float("nan"), dtype=torch.float64, device=o["tit"].device)
            else:
                tit = np.full(G, np.nan)
        else:
            tit = o["tit"]
        res = {
            "beta": beta,
            "beta_sandwiches_null": bread,
            "beta_sandwiches": sw,
            "overdispersion": o["theta"],
            "iterations": {"beta_iters": o["bit"], "theta_iters": tit},
            "theta_passes": o["tit"],
            "status": o["st"],
            "gene_names": self.gene_names,
        }
        res.update(self.meta)
        return res


def fit_devil(x, design_matrix, clusters=None, overdispersion: Any = "MOM",
              init_overdispersion=None, init_beta_rough: bool = False, offset: float = 0,
              size_factors=None, verbose: bool = False, max_iter: int = 200,
              tolerance: float = 1e-3, CUDA: bool = True, batch_size: int = 1024,
              precision: str = "fp64", device: int = 0, init_mode: str | None = None):
    """Fit the per-gene NB GLM (R/main.R:164-273 with cpu_fit semantics).

    ``x`` is genes x cells (numpy on the host or a torch CUDA tensor).
    ``overdispersion`` is "MOM", "MLE"/"old" or False.  ``batch_size`` is
    accepted for signature compatibility; the library sizes its own gene
    batches.  ``precision`` is "fp64" or "mixed" (FP32 element arithmetic and
    FP32/bf16 MFMA contractions with FP64 solves and stopping decisions).
    """
    if not CUDA:
        raise NotImplementedError("devil_amd only implements the GPU fit (CUDA=True)")
    ctx = L.context(device)
    job = _Job(x, design_matrix, clusters=clusters, overdispersion=overdispersion,
               init_overdispersion=init_overdispersion, init_beta_rough=init_beta_rough,
               offset=offset, size_factors=size_factors, max_iter=max_iter,
               tolerance=tolerance, precision=precision, init_mode=init_mode, device=device)
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_fit(ctx.h, C.byref(job.prob), C.byref(job.outs), 1))
    return job.result()


def fit_devil_many(problems, device: int = 0, **common):
    """Fit several independent problems in one library call (devil_fit with
    n_problems > 1: the problems run concurrently on the context's streams).

    ``problems`` is a list of dicts of fit_devil arguments (at least ``x``
    and ``design_matrix``); ``common`` supplies defaults for all of them.
    Returns the list of fit_devil results in the same order.
    """
    ctx = L.context(device)
    jobs = [_Job(**{**common, **pr, "device": device}) for pr in problems]
    if not jobs:
        return []
    probs = (L.Problem * len(jobs))(*[j.prob for j in jobs])
    outs = (L.Outputs * len(jobs))(*[j.outs for j in jobs])
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_fit(ctx.h, probs, outs, len(jobs)))
    return [j.result() for j in jobs]


def test_de(devil_fit, contrast, clusters=None, pval_adjust_method: str = "BH",
            max_lfc: float = 10, device: int = 0):
    """R/test.R:85-134 on the GPU: returns dict(name, pval, adj_pval, lfc,
    log_pval).  ``clusters`` is accepted and unused, as in the reference."""
    import torch
    if "beta_sandwiches_null" not in devil_fit:
        raise ValueError("The fit object was created with an older version of devil and is "
                         "no longer supported. Please re-fit your model with the current "
                         "version of fit_devil().")
    if pval_adjust_method not in L.P_ADJUST:
        raise ValueError(f"'arg' should be one of {', '.join(map(repr, L.P_ADJUST))}")
    ctx = L.context(device)
    dev = torch.device("cuda", device)

    def f64(a):
        if _is_torch(a):
            return a.to(device=dev, dtype=torch.float64)
        return torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)

    beta = f64(devil_fit["beta"])
    G, p = beta.shape
    n = int(devil_fit["design_matrix"].shape[0]) if hasattr(devil_fit["design_matrix"], "shape") \
        else int(np.asarray(devil_fit["design_matrix"]).shape[0])
    bcol = beta.T.contiguous()  # column-major G x p
    bread = f64(devil_fit["beta_sandwiches_null"]).transpose(1, 2).contiguous()
    sw = devil_fit.get("beta_sandwiches")
    swd = None if sw is None else f64(sw).transpose(1, 2).contiguous()
    c = f64(np.asarray(contrast, dtype=np.float64).reshape(-1))
    if c.numel() != p:
        raise ValueError("contrast length must match the number of coefficients")
    pval, logp, adj, lfc = (torch.empty(G, dtype=torch.float64, device=dev) for _ in range(4))
    ptr = lambda a: None if a is None else a.data_ptr()
    _inputs_ready(ctx.device)
    ctx.check(L.lib().devil_test_de(ctx.h, ptr(bcol), ptr(bread), ptr(swd), G, p, n, ptr(c),
                                    float(max_lfc), ptr(pval), ptr(logp), ptr(adj), ptr(lfc)))
    method = L.P_ADJUST[pval_adjust_method]
    if method != L.P_ADJUST["BH"]:  # devil_test_de applies BH
        _inputs_ready(ctx.device)
        ctx.check(L.lib().devil_p_adjust(ctx.h, ptr(pval), G, method, ptr(adj)))
    name = devil_fit.get("gene_names")
    return {"name": name, "pval": pval.cpu().numpy(), "adj_pval": adj.cpu().numpy(),
            "lfc": lfc.cpu().numpy(), "log_pval": logp.cpu().numpy()}
