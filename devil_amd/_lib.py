"""ctypes binding of libdevil_hip.so (include/devil_hip.h).

There is no CPU fallback: if the library is missing, or no GPU is present,
calls raise instead of computing anything on the host.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdevil_hip.so")

DEVIL_OK = 0
MEM_HOST, MEM_DEVICE = 0, 1
COUNTS_I32, COUNTS_F32, COUNTS_F64 = 0, 1, 2
INIT_QR, INIT_ROUGH, INIT_UNIVARIATE = 0, 1, 2
OD_NONE, OD_MOM, OD_MLE, OD_NEW = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
PMAX = 10

GENE_MAXITER, GENE_NAN, GENE_SINGULAR, GENE_MOM_DEN0, GENE_MLE_ZERO = 1, 2, 4, 8, 16
GENE_MLE_NOCONV, GENE_OD_FAILED = 32, 64

_vp = C.c_void_p


class Problem(C.Structure):
    _fields_ = [
        ("G", C.c_int64), ("n", C.c_int64), ("p", C.c_int32), ("mem", C.c_int32),
        ("counts", _vp), ("counts_dtype", C.c_int32), ("counts_ld", C.c_int64),
        ("X", _vp), ("offset", _vp), ("cluster_ends", _vp), ("n_clusters", C.c_int32),
        ("disp_init", _vp), ("beta_init", _vp), ("init_mode", C.c_int32),
        ("od_mode", C.c_int32), ("max_iter", C.c_int32), ("tolerance", C.c_double),
        ("precision", C.c_int32),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("beta", _vp), ("theta", _vp), ("beta_iters", _vp), ("theta_iters", _vp),
        ("bread", _vp), ("sandwich", _vp), ("status", _vp),
    ]


class DevilError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"devil_hip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libdevil_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m devil_amd.build` "
            "(devil_amd has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    i32, i64, d = C.c_int32, C.c_int64, C.c_double
    sigs = {
        "devil_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
        "devil_destroy": (None, [_vp]),
        "devil_strerror": (C.c_char_p, [C.c_int]),
        "devil_last_error": (C.c_char_p, [_vp]),
        "devil_version": (C.c_char_p, []),
        "devil_fit": (C.c_int, [_vp, C.POINTER(Problem), C.POINTER(Outputs), C.c_int32]),
        "devil_estimate_dispersion": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, _vp]),
        "devil_init_beta": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp, i32, _vp]),
        "devil_beta_fit": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp, _vp, _vp,
                                     i32, d, i32, _vp, _vp]),
        "devil_mom_dispersion": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp, _vp,
                                           _vp]),
        "devil_fit_dispersion_mle": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp,
                                               _vp, i32, d, _vp, _vp]),
        "devil_bread_meat": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp, _vp, _vp,
                                       _vp, i32, _vp, _vp, _vp]),
        "devil_test_de": (C.c_int, [_vp, _vp, _vp, _vp, i64, i32, i64, _vp, d, _vp, _vp, _vp,
                                    _vp]),
        "devil_test_de_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, i64, i32, i64, _vp, d, _vp,
                                            _vp, _vp, _vp]),
        "devil_p_adjust": (C.c_int, [_vp, _vp, i64, i32, _vp]),
        "devil_sandwich_hc0": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, i32, _vp, _vp,
                                         _vp, _vp, _vp, _vp]),
        "devil_size_factors": (C.c_int, [_vp, _vp, i32, i64, i64, i64, i32, _vp]),
        "devil_group_cells": (C.c_int, [_vp, _vp, i32, i64, i64, i64, _vp, _vp, i64]),
        "devil_csc_to_dense": (C.c_int, [_vp, _vp, _vp, _vp, i64, i64, i32, _vp, i64]),
        "devil_set_timing": (C.c_int, [_vp, C.c_int]),
        "devil_stage_times": (C.c_int, [_vp, C.POINTER(C.c_double)]),
        "devil_set_tuning": (C.c_int, [_vp, C.c_int32, C.c_int64]),
        "devil_fit_info": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int64)]),
        "devil_kernel_stats": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ["devil_create", "devil_destroy", "devil_strerror", "devil_last_error",
            "devil_version", "devil_fit", "devil_estimate_dispersion", "devil_init_beta",
            "devil_beta_fit", "devil_mom_dispersion", "devil_fit_dispersion_mle",
            "devil_bread_meat", "devil_test_de", "devil_test_de_noalloc", "devil_p_adjust",
            "devil_sandwich_hc0", "devil_size_factors", "devil_group_cells",
            "devil_csc_to_dense",
            "devil_set_timing", "devil_stage_times",
            "devil_kernel_stats", "devil_set_tuning", "devil_fit_info"]

STAGES = ["init", "irls", "dispersion", "bread_meat", "test_de"]
KERNELS = ["irls_fp64", "irls_mixed", "mle"]
TUNE_MIXED_KERNEL, TUNE_CHUNK_CELLS, TUNE_CONCURRENCY, TUNE_MLE_DRIVER, TUNE_PAD_COPY = 0, 1, 2, 3, 4
TUNE_JOINT, TUNE_GRAPH, TUNE_BATCH_GENES = 5, 6, 7
MLE_DRIVERS = {"nlminb": 0, "root": 1}
SF_METHODS = {"normed_sum": 0, "psinorm": 1, "edgeR": 2}
# stats::p.adjust.methods
P_ADJUST = {"holm": 2, "hochberg": 3, "hommel": 4, "bonferroni": 5, "BH": 0, "BY": 1,
            "fdr": 0, "none": 6}
MIXED_KERNELS = {"f32": 0, "bf16": 1, "auto": 2, "bf16_ws": 3, "bf16_ws2": 4}


class Context:
    """Owns one devil_ctx (a HIP stream on one device)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = _vp()
        rc = L.devil_create(C.byref(h), int(device))
        if rc != DEVIL_OK:
            raise DevilError(rc, L.devil_strerror(rc).decode())
        self.h = h
        self.device = device

    def check(self, rc: int):
        if rc != DEVIL_OK:
            L = lib()
            raise DevilError(rc, f"{L.devil_strerror(rc).decode()}: "
                                 f"{L.devil_last_error(self.h).decode()}")

    def set_timing(self, enable: bool = True):
        self.check(lib().devil_set_timing(self.h, int(bool(enable))))

    def stage_times(self) -> dict:
        """Device ms per stage of the last fit / test_de (HIP events)."""
        arr = (C.c_double * len(STAGES))()
        self.check(lib().devil_stage_times(self.h, arr))
        return dict(zip(STAGES, list(arr)))

    def set_mixed_kernel(self, name: str):
        """Kernel of the mixed-precision IRLS: "f32", "bf16" or "auto"."""
        self.check(lib().devil_set_tuning(self.h, TUNE_MIXED_KERNEL, MIXED_KERNELS[name]))

    def set_mle_driver(self, name: str):
        """MLE overdispersion driver: "nlminb" (the reference's nlminb call,
        stopping rules included; default) or "root" (root of the score)."""
        self.check(lib().devil_set_tuning(self.h, TUNE_MLE_DRIVER, MLE_DRIVERS[name]))

    def set_tuning(self, key: int, value: int):
        self.check(lib().devil_set_tuning(self.h, int(key), int(value)))

    def fit_info(self) -> dict:
        """Genes refit in FP64 by the last mixed fit, and its kernel and chunk
        size (the autotuner's choice when tuning was left open)."""
        out = {}
        for k, name in enumerate(("replayed", "mixed_kernel", "chunk_cells", "polished",
                                  "aligned_copy", "joined")):
            v = C.c_int64()
            self.check(lib().devil_fit_info(self.h, k, C.byref(v)))
            out[name] = int(v.value)
        inv = {v: k for k, v in MIXED_KERNELS.items()}
        out["mixed_kernel"] = inv.get(out["mixed_kernel"], "fp64")
        return out

    def kernel_stats(self) -> dict:
        """{kernel: (launches, device ms, algorithmic bytes)} of the IRLS
        accumulate kernels in the last devil_fit (timing enabled)."""
        out = {}
        for k, name in enumerate(KERNELS):
            nl, ms, by = C.c_int64(), C.c_double(), C.c_double()
            self.check(lib().devil_kernel_stats(self.h, k, C.byref(nl), C.byref(ms),
                                                C.byref(by)))
            out[name] = (int(nl.value), float(ms.value), float(by.value))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().devil_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_cache: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    ctx = _ctx_cache.get(device)
    if ctx is None:
        ctx = Context(device)
        _ctx_cache[device] = ctx
    return ctx
