"""devil_amd: devil's per-gene negative-binomial GLM fit on AMD MI355X (gfx950).

The fit (IRLS beta, MOM / Newton-MLE overdispersion, bread and patient-
clustered sandwich) and the Wald test run as hand-written HIP kernels in
libdevil_hip.so behind a C ABI (include/devil_hip.h).  This package is the
Python host mirror of devil's R API for that path.
"""
from .api import (cluster_blocks, compute_offset_vector, fit_devil, fit_devil_many,  # noqa: F401
                  test_de)
from ._lib import Context, DevilError, context, lib  # noqa: F401

__all__ = ["fit_devil", "fit_devil_many", "test_de", "cluster_blocks", "compute_offset_vector", "Context",
           "DevilError", "context", "lib"]
