"""grid2grid_amd: MI355X-native grid2grid::transform.

The compute path is libgrid2grid_amd.so (HIP kernels for gfx950 behind the C
ABI in include/grid2grid_amd.h). This package is the Python binding used by
the tests and the benchmark; it has no CPU fallback.
"""
from ._lib import (FLOAT32, FLOAT64, COMPLEX64, COMPLEX128, DTYPE_SIZE, DTYPE_NAMES, G2GError,
                   LIB_PATH, load)
from .layout import Layout, scalapack_layout, arbitrary_layout, line_split, random_splits
from .api import (Plan, Batch, Server, SERVE_BATCHED, SERVE_PLAIN, set_tuning, fill_layout,
                  verify_layout, comm_volume, optimal_reordering, stream_probe, transform_layouts,
                  plan_cache_stats, plan_cache_clear, plan_cache_set_capacity, scale_layout,
                  transform_plans_noalloc)

__all__ = [
    "FLOAT32", "FLOAT64", "COMPLEX64", "COMPLEX128", "DTYPE_SIZE", "DTYPE_NAMES", "G2GError",
    "LIB_PATH", "load", "Layout", "scalapack_layout", "arbitrary_layout", "line_split",
    "random_splits", "Plan", "Batch", "Server", "SERVE_BATCHED", "SERVE_PLAIN", "set_tuning",
    "fill_layout", "verify_layout", "comm_volume", "optimal_reordering", "stream_probe",
    "transform_layouts", "plan_cache_stats", "plan_cache_clear", "plan_cache_set_capacity", "scale_layout",
    "transform_plans_noalloc",
]
