"""devil_amd: devil's per-gene negative-binomial GLM fit on AMD MI355X (gfx950).

The fit (IRLS beta, MOM / MLE overdispersion, bread and patient-clustered
sandwich), the Wald test and the count-matrix preparation (size factors,
cell grouping, sparse assays) run as hand-written HIP kernels in
libdevil_hip.so behind a C ABI (include/devil_hip.h).  This package is the
Python host mirror of devil's R API for that path.
"""
from .api import (calculate_sf, cluster_blocks, compute_offset_vector,  # noqa: F401
                  compute_sandwich, densify, fit_devil, fit_devil_many, group_data, test_de)
from ._lib import Context, DevilError, context, lib  # noqa: F401

__all__ = ["fit_devil", "fit_devil_many", "test_de", "calculate_sf", "group_data", "densify",
           "compute_sandwich",
           "cluster_blocks", "compute_offset_vector", "Context", "DevilError", "context", "lib"]
