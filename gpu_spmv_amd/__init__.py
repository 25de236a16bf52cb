"""gpu_spmv_amd -- SpMV in the CSR, K1, AXC and AXT formats on one AMD
Instinct MI355X (gfx950), through the C ABI of libaxspmv.so.

The numerical work happens in the library's HIP kernels; this package is the
Python side used by the tests and the benchmark.
"""
from . import _lib
from ._lib import (AXT_COM, AXT_UNC, BUILD_AX, BUILD_PLANES, F32, F64, FMT_AXC, FMT_AXT, FMT_CSR, FMT_RCT,
                   FMT_K1, KER_DEFAULT, KER_NAIVE, KER_STREAM, PLAN_TRANSPOSE, PLAN_RCT_SPLIT,
                   PLAN_AX_LAYOUT, PLAN_STAGE_X, PLAN_DETERMINISTIC, PLAN_RCT_ATOMIC, PLAN_REORDER,
                   PLAN_RCT_WIDE_KEY, AxspmvError)
from .core import (VEC_STREAM, Batch, CsrMatrix, Graph, DeviceBuffer, Plan, autotune, autotune_batch,
                   autotune_defaults, banded,
                   build_axc, build_axt, build_k1, dev_build_axc, dev_build_axt,
                   device_count, fill_values_normal, laplace2d, normal, opts, powerlaw, random_sparse,
                   rng_u64, spmv_host, stencil27, uniform, x_refresh)

__all__ = [
    "AXT_COM", "AXT_UNC", "BUILD_AX", "BUILD_PLANES", "F32", "F64", "FMT_AXC", "FMT_AXT", "FMT_CSR", "FMT_RCT",
    "FMT_K1", "KER_DEFAULT", "KER_NAIVE", "KER_STREAM", "PLAN_TRANSPOSE", "PLAN_RCT_SPLIT", "PLAN_AX_LAYOUT", "PLAN_STAGE_X", "PLAN_DETERMINISTIC", "PLAN_RCT_ATOMIC", "PLAN_REORDER", "PLAN_RCT_WIDE_KEY", "AxspmvError", "VEC_STREAM",
    "Batch", "Graph", "autotune", "autotune_batch", "autotune_defaults", "CsrMatrix", "DeviceBuffer", "Plan", "banded", "build_axc", "build_axt", "build_k1", "dev_build_axc", "dev_build_axt", "device_count",
    "fill_values_normal", "laplace2d", "normal", "opts", "powerlaw", "random_sparse", "rng_u64",
    "spmv_host", "stencil27", "uniform", "x_refresh", "_lib",
]
