"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker / timed CPU baseline.
The product package ``devil_amd`` never imports this module.

Array conventions follow ``devil_oracle.h``: counts gene-major (G x n rows),
X column-major n x p (R layout), beta column-major G x p, per-gene p x p
matrices column-major and contiguous.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)

OD_NONE, OD_MOM, OD_MLE, OD_NEW = 0, 1, 2, 3
INIT_QR, INIT_ROUGH, INIT_UNIVARIATE = 0, 1, 2


class FitArgs(C.Structure):
    _fields_ = [
        ("G", C.c_int64), ("n", C.c_int64), ("ld", C.c_int64), ("p", C.c_int32),
        ("Y", _dp), ("X", _dp), ("offset", _dp), ("cluster_ends", _ip),
        ("n_clusters", C.c_int32), ("disp_init", _dp), ("init_mode", C.c_int32),
        ("od_mode", C.c_int32), ("max_iter", C.c_int32), ("tolerance", C.c_double),
        ("n_threads", C.c_int32), ("genes", _lp), ("n_genes_subset", C.c_int64),
        ("mle_driver", C.c_int32),
    ]


class FitOut(C.Structure):
    _fields_ = [
        ("beta", _dp), ("theta", _dp), ("beta_iters", _ip), ("theta_iters", _ip),
        ("bread", _dp), ("sandwich", _dp), ("alpha0", _dp), ("beta0", _dp),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d, i32, i64 = C.c_double, C.c_int32, C.c_int64
        sig = {
            "orc_digamma": (d, [d]), "orc_trigamma": (d, [d]),
            "orc_lbeta": (d, [d, d]), "orc_log_ibeta": (d, [d, d, d, d]),
            "orc_pt2_upper": (d, [d, d, _dp]),
            "orc_inverse": (C.c_int, [_dp, C.c_int, _dp]),
            "orc_beta_fit": (C.c_int, [_dp, i64, _dp, C.c_int, _dp, _dp, C.c_float,
                                       C.c_int, C.c_float]),
            "orc_mom_dispersion": (d, [_dp, i64, _dp, C.c_int, _dp, _dp]),
            "orc_compute_hessian": (C.c_int, [_dp, d, _dp, _dp, i64, C.c_int, _dp, _dp]),
            "orc_clustered_meat": (None, [_dp, _dp, _dp, d, _dp, i64, C.c_int, _ip,
                                          C.c_int, _dp]),
            "orc_meat_hc0": (None, [_dp, _dp, _dp, d, _dp, i64, C.c_int, _dp]),
            "orc_make_table": (i64, [_dp, i64, i64, _dp, _dp]),
            "orc_nb_loglik": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int, _dp, _dp, i64]),
            "orc_nb_score": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int, _dp, _dp, i64]),
            "orc_nb_dscore": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int, _dp, _dp, i64]),
            "orc_dpsi_diff": (d, [d, d]), "orc_dpsi1_diff": (d, [d, d]),
            "orc_nb_score_acc": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int]),
            "orc_nb_dscore_acc": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int]),
            "orc_fit_dispersion": (d, [_dp, _dp, _dp, _dp, i64, C.c_int, d, C.c_int,
                                       C.c_int, C.POINTER(C.c_int)]),
            "orc_fit_dispersion_port": (d, [_dp, _dp, _dp, _dp, i64, C.c_int, d, C.c_int,
                                            C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int)]),
            "orc_lgamma_diff": (d, [d, d]),
            "orc_nb_loglik_acc": (d, [_dp, _dp, i64, d, _dp, C.c_int, C.c_int]),
            "orc_estimate_dispersion": (None, [_dp, i64, i64, i64, _dp, _dp]),
            "orc_init_beta_qr": (C.c_int, [_dp, i64, i64, i64, _dp, C.c_int, _dp, _dp]),
            "orc_init_beta_rough": (None, [_dp, i64, i64, i64, C.c_int, _dp]),
            "orc_init_beta_univariate": (None, [_dp, i64, i64, i64, _dp, C.c_int, _dp, _dp]),
            "orc_cpu_fit": (C.c_int, [C.POINTER(FitArgs), C.POINTER(FitOut)]),
            "orc_test_de": (None, [_dp, _dp, _dp, i64, C.c_int, i64, _dp, d, _dp, _dp,
                                   _dp, _dp]),
            "orc_p_adjust_bh": (None, [_dp, i64, _dp]),
            "orc_p_adjust": (None, [_dp, i64, i32, _dp]),
            "orc_size_factors": (C.c_int, [_dp, i64, i64, i64, C.c_int, _dp]),
            "orc_fit_overdispersion_new": (d, [_dp, _dp, _dp, _dp, i64, C.c_int, d, C.c_int, d,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
            "orc_h_log_gh_pmhalf": (C.c_int, [d, d, d, _dp, _dp]),
            "orc_gh16": (None, [_dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def digamma(x):
    return lib().orc_digamma(float(x))


def trigamma(x):
    return lib().orc_trigamma(float(x))


def lbeta(a, b):
    return lib().orc_lbeta(float(a), float(b))


def pt2_upper(t, df):
    lp = C.c_double()
    v = lib().orc_pt2_upper(float(t), float(df), C.byref(lp))
    return v, lp.value


def inverse(A):
    A = np.asfortranarray(A, dtype=np.float64)
    p = A.shape[0]
    out = np.zeros((p, p), order="F")
    lib().orc_inverse(A.ctypes.data_as(_dp), p, out.ctypes.data_as(_dp))
    return out


def beta_fit(y, X, beta0, off, k, max_iter, eps):
    """src/beta.cpp:25-51; returns (beta, iter)."""
    y, off = _f64(y), _f64(off)
    X = np.asfortranarray(X, dtype=np.float64)
    b = _f64(beta0).copy()
    it = lib().orc_beta_fit(_d(y), y.size, X.ctypes.data_as(_dp), X.shape[1], _d(b),
                            _d(off), float(k), int(max_iter), float(eps))
    return b, it


def mom_dispersion(y, X, beta, sf):
    y, beta, sf = _f64(y), _f64(beta), _f64(sf)
    X = np.asfortranarray(X, dtype=np.float64)
    return lib().orc_mom_dispersion(_d(y), y.size, X.ctypes.data_as(_dp), X.shape[1],
                                    _d(beta), _d(sf))


def compute_hessian(beta, alpha, y, X, sf):
    y, beta, sf = _f64(y), _f64(beta), _f64(sf)
    X = np.asfortranarray(X, dtype=np.float64)
    p = X.shape[1]
    out = np.zeros((p, p), order="F")
    lib().orc_compute_hessian(_d(beta), float(alpha), _d(y), X.ctypes.data_as(_dp),
                              y.size, p, _d(sf), out.ctypes.data_as(_dp))
    return out


def clustered_meat(X, y, beta, alpha, sf, ends):
    y, beta, sf = _f64(y), _f64(beta), _f64(sf)
    X = np.asfortranarray(X, dtype=np.float64)
    ends = np.ascontiguousarray(ends, dtype=np.int32)
    p = X.shape[1]
    out = np.zeros((p, p), order="F")
    lib().orc_clustered_meat(X.ctypes.data_as(_dp), _d(y), _d(beta), float(alpha),
                             _d(sf), y.size, p, ends.ctypes.data_as(_ip), ends.size,
                             out.ctypes.data_as(_dp))
    return out


def meat_hc0(X, y, beta, alpha, sf):
    y, beta, sf = _f64(y), _f64(beta), _f64(sf)
    X = np.asfortranarray(X, dtype=np.float64)
    p = X.shape[1]
    out = np.zeros((p, p), order="F")
    lib().orc_meat_hc0(X.ctypes.data_as(_dp), _d(y), _d(beta), float(alpha), _d(sf),
                       y.size, p, out.ctypes.data_as(_dp))
    return out


def make_table(y, stop_if_larger):
    y = _f64(y)
    keys = np.zeros(y.size + 1)
    freqs = np.zeros(y.size + 1)
    nu = lib().orc_make_table(_d(y), y.size, int(stop_if_larger), _d(keys), _d(freqs))
    if nu < 0:
        return np.zeros(0), np.zeros(0)
    return keys[:nu].copy(), freqs[:nu].copy()


def _nbfun(name, y, mu, log_theta, X, cr, table):
    y, mu = _f64(y), _f64(mu)
    X = np.asfortranarray(X, dtype=np.float64)
    if table is None or len(table[0]) == 0:
        u = f = None
        nu = 0
    else:
        u, f = _f64(table[0]), _f64(table[1])
        nu = u.size
    return getattr(lib(), name)(_d(y), _d(mu), y.size, float(log_theta),
                                X.ctypes.data_as(_dp), X.shape[1], int(bool(cr)),
                                _d(u), _d(f), nu)


def nb_loglik(y, mu, log_theta, X, cr=True, table=None):
    return _nbfun("orc_nb_loglik", y, mu, log_theta, X, cr, table)


def nb_score(y, mu, log_theta, X, cr=True, table=None):
    return _nbfun("orc_nb_score", y, mu, log_theta, X, cr, table)


def nb_dscore(y, mu, log_theta, X, cr=True, table=None):
    return _nbfun("orc_nb_dscore", y, mu, log_theta, X, cr, table)


def dpsi_diff(a, y):
    return lib().orc_dpsi_diff(float(a), float(y))


def dpsi1_diff(a, y):
    return lib().orc_dpsi1_diff(float(a), float(y))


def nb_score_acc(y, mu, log_theta, X, cr=True):
    y, mu = _f64(y), _f64(mu)
    X = np.asfortranarray(X, dtype=np.float64)
    return lib().orc_nb_score_acc(_d(y), _d(mu), y.size, float(log_theta),
                                  X.ctypes.data_as(_dp), X.shape[1], int(bool(cr)))


def nb_dscore_acc(y, mu, log_theta, X, cr=True):
    y, mu = _f64(y), _f64(mu)
    X = np.asfortranarray(X, dtype=np.float64)
    return lib().orc_nb_dscore_acc(_d(y), _d(mu), y.size, float(log_theta),
                                   X.ctypes.data_as(_dp), X.shape[1], int(bool(cr)))


def fit_dispersion(beta, X, y, off, tol=1e-3, max_iter=200, cr=True):
    y, off, beta = _f64(y), _f64(off), _f64(beta)
    X = np.asfortranarray(X, dtype=np.float64)
    it = C.c_int()
    th = lib().orc_fit_dispersion(_d(beta), X.ctypes.data_as(_dp), _d(y), _d(off),
                                  y.size, X.shape[1], float(tol), int(max_iter),
                                  int(bool(cr)), C.byref(it))
    return th, it.value


def fit_dispersion_port(beta, X, y, off, tol=1e-3, max_iter=200, cr=True):
    """fit_dispersion with nlminb's PORT driver restated: (theta, iterations,
    PORT return code, objective evaluations)."""
    X = np.asfortranarray(X, dtype=np.float64)
    y = _f64(y)
    it, code, nf = C.c_int(), C.c_int(), C.c_int()
    th = lib().orc_fit_dispersion_port(_d(_f64(beta)), X.ctypes.data_as(_dp), _d(y),
                                       _d(_f64(off)), y.size, X.shape[1], float(tol),
                                       int(max_iter), int(cr), C.byref(it), C.byref(code),
                                       C.byref(nf))
    return th, it.value, code.value, nf.value


def lgamma_diff(a, y):
    return lib().orc_lgamma_diff(float(a), float(y))


def nb_loglik_acc(y, mu, log_theta, X, cr=True):
    X = np.asfortranarray(X, dtype=np.float64)
    y = _f64(y)
    return lib().orc_nb_loglik_acc(_d(y), _d(_f64(mu)), y.size, float(log_theta),
                                   X.ctypes.data_as(_dp), X.shape[1], int(cr))


def estimate_dispersion(Y, off):
    Y, off = _f64(Y), _f64(off)
    G, n = Y.shape
    out = np.zeros(G)
    lib().orc_estimate_dispersion(_d(Y), G, n, n, _d(off), _d(out))
    return out


def init_beta_qr(Y, X, off):
    Y, off = _f64(Y), _f64(off)
    X = np.asfortranarray(X, dtype=np.float64)
    G, n = Y.shape
    p = X.shape[1]
    out = np.zeros((G, p), order="F")
    rc = lib().orc_init_beta_qr(_d(Y), G, n, n, X.ctypes.data_as(_dp), p, _d(off),
                                out.ctypes.data_as(_dp))
    if rc:
        raise ValueError("rank-deficient design")
    return out


def init_beta_rough(Y, p):
    Y = _f64(Y)
    G, n = Y.shape
    out = np.zeros((G, p), order="F")
    lib().orc_init_beta_rough(_d(Y), G, n, n, p, out.ctypes.data_as(_dp))
    return out


def init_beta_univariate(Y, X, sf):
    Y, sf = _f64(Y), _f64(sf)
    X = np.asfortranarray(X, dtype=np.float64)
    G, n = Y.shape
    p = X.shape[1]
    out = np.zeros((G, p), order="F")
    lib().orc_init_beta_univariate(_d(Y), G, n, n, X.ctypes.data_as(_dp), p, _d(sf),
                                   out.ctypes.data_as(_dp))
    return out


def cpu_fit(Y, X, offset, *, clusters_ends=None, disp_init=None, init_mode=INIT_QR,
            od_mode=OD_MOM, max_iter=200, tolerance=1e-3, n_threads=0, genes=None,
            mle_driver="nlminb"):
    """R/main.R:375-523 cpu_fit.  Returns a dict of numpy arrays.

    ``Y`` is G x n (any numeric dtype; converted to float64 rows).  ``genes``
    optionally restricts the fit to a subset of rows (outputs follow that
    order)."""
    Y = _f64(Y)
    G, n = Y.shape
    X = np.asfortranarray(X, dtype=np.float64)
    p = X.shape[1]
    offset = _f64(offset)
    ends = None if clusters_ends is None else np.ascontiguousarray(clusters_ends, np.int32)
    di = None if disp_init is None else _f64(disp_init)
    gidx = None if genes is None else np.ascontiguousarray(genes, dtype=np.int64)
    Gs = G if gidx is None else gidx.size
    a = FitArgs(G=G, n=n, ld=n, p=p, Y=_d(Y), X=X.ctypes.data_as(_dp), offset=_d(offset),
                cluster_ends=ends.ctypes.data_as(_ip) if ends is not None else None,
                n_clusters=0 if ends is None else ends.size, disp_init=_d(di),
                init_mode=init_mode, od_mode=od_mode, max_iter=max_iter,
                tolerance=tolerance, n_threads=n_threads,
                genes=gidx.ctypes.data_as(_lp) if gidx is not None else None,
                n_genes_subset=Gs, mle_driver={"nlminb": 0, "root": 1}[mle_driver])
    out = dict(beta=np.zeros((Gs, p), order="F"), theta=np.zeros(Gs),
               beta_iters=np.zeros(Gs, np.int32), theta_iters=np.zeros(Gs, np.int32),
               bread=np.zeros((Gs, p, p)), alpha0=np.zeros(Gs),
               beta0=np.zeros((Gs, p), order="F"))
    out["sandwich"] = np.zeros((Gs, p, p)) if ends is not None else None
    o = FitOut(beta=out["beta"].ctypes.data_as(_dp), theta=_d(out["theta"]),
               beta_iters=out["beta_iters"].ctypes.data_as(_ip),
               theta_iters=out["theta_iters"].ctypes.data_as(_ip),
               bread=_d(out["bread"]), sandwich=_d(out["sandwich"]),
               alpha0=_d(out["alpha0"]), beta0=out["beta0"].ctypes.data_as(_dp))
    rc = lib().orc_cpu_fit(C.byref(a), C.byref(o))
    if rc:
        raise ValueError("rank-deficient design")
    # per-gene p x p blocks are column-major: expose them as [g, a, b] = M[a + b*p]
    out["bread"] = out["bread"].transpose(0, 2, 1).copy()
    if out["sandwich"] is not None:
        out["sandwich"] = out["sandwich"].transpose(0, 2, 1).copy()
    return out


P_ADJUST = {"BH": 0, "fdr": 0, "BY": 1, "holm": 2, "hochberg": 3, "hommel": 4,
            "bonferroni": 5, "none": 6}


def p_adjust(p, method="BH"):
    """stats::p.adjust(p, method); NaN entries stay NaN."""
    p = _f64(p)
    out = np.zeros_like(p)
    lib().orc_p_adjust(_d(p), p.size, P_ADJUST[method], _d(out))
    return out


def p_adjust_bh(p):
    p = _f64(p)
    out = np.zeros_like(p)
    lib().orc_p_adjust_bh(_d(p), p.size, _d(out))
    return out


def de_test(beta, bread, sandwich, n, contrast, max_lfc=10.0):
    """R/test.R:85-134.  bread/sandwich given as [G, p, p] arrays."""
    beta = np.asfortranarray(beta, dtype=np.float64)
    G, p = beta.shape
    br = np.ascontiguousarray(np.asarray(bread, np.float64).transpose(0, 2, 1))
    sw = None if sandwich is None else np.ascontiguousarray(
        np.asarray(sandwich, np.float64).transpose(0, 2, 1))
    c = _f64(contrast)
    pval, lpv, adj, lfc = (np.zeros(G) for _ in range(4))
    lib().orc_test_de(beta.ctypes.data_as(_dp), _d(br), _d(sw), G, p, int(n), _d(c),
                      float(max_lfc), _d(pval), _d(lpv), _d(adj), _d(lfc))
    return dict(pval=pval, log_pval=lpv, adj_pval=adj, lfc=lfc)


def size_factors(Y, method="normed_sum"):
    """R/size_factors.R:37-127 calculate_sf (normed_sum or psinorm)."""
    Y = _f64(Y)
    G, n = Y.shape
    sf = np.zeros(n)
    rc = lib().orc_size_factors(_d(Y), G, n, n, {"normed_sum": 0, "psinorm": 1}[method],
                                _d(sf))
    if rc:
        raise ValueError("At least one column (sample) contains all zeros")
    return sf


def group_data(Y, X, clusters):
    """R/group_data.R:38-69: cells ordered by cluster code (codes by first
    appearance, stable), returned as (Y[:, ord], X[ord], codes[ord], ord)."""
    lab = np.asarray(clusters)
    _, first, inv = np.unique(lab, return_index=True, return_inverse=True)
    rank = np.empty(len(first), np.int64)
    rank[np.argsort(first)] = np.arange(len(first))
    codes = rank[inv] + 1
    ord_ = np.argsort(codes, kind="stable")
    return np.asarray(Y)[:, ord_], np.asarray(X)[ord_], codes[ord_].astype(float), ord_


def fit_overdispersion_new(y, X, beta, off, kappa=1.0, max_iter=200, eps_theta=1e-3):
    """src/new_overdispersion.cpp:195-481 for one gene: dict(kappa, kappa_iters,
    kappa_converged, err, evals); err != 0 where the reference stops."""
    X = np.asfortranarray(X, dtype=np.float64)
    y = _f64(y)
    it, conv, err, ev = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
    k = lib().orc_fit_overdispersion_new(_d(y), X.ctypes.data_as(_dp), _d(_f64(beta)),
                                         _d(_f64(off)), y.size, X.shape[1], float(kappa),
                                         int(max_iter), float(eps_theta), C.byref(it),
                                         C.byref(conv), C.byref(err), C.byref(ev))
    return dict(kappa=k, kappa_iters=it.value, kappa_converged=bool(conv.value),
                err=err.value, evals=ev.value)


def h_log_gh_pmhalf(s, t, tol=1e-3):
    lp, lm = C.c_double(), C.c_double()
    ok = lib().orc_h_log_gh_pmhalf(float(s), float(t), float(tol), C.byref(lp), C.byref(lm))
    return (lp.value, lm.value) if ok else None


def gh16():
    x, w = np.zeros(16), np.zeros(16)
    lib().orc_gh16(_d(x), _d(w))
    return x, w
