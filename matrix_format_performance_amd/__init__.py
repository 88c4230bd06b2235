"""MI355X-native SpMV engine (gfx950) with the matrix_format_performance API.

All SpMV and layout work runs in libmfp.so (hand-written HIP kernels behind a
C ABI, include/mfp.h).  The first call loads that library and raises
ImportError if it has not been built
(``python -m matrix_format_performance_amd.build``): there is no CPU path.
"""
from . import _lib
from ._lib import MfpBudgetError, MfpError

from .engine import (FORMATS, LABELS, Batch, autotune, DeviceArray, DeviceMatrix, GenSpec, Graph,  # noqa: E402
                     Plan, Timer, device_info, flush_cache, generate, generate_vector, init,
                     residual, synchronize, workload)
from .reference_api import (compare_results, coo_matrix_class, csr_matrix_class,  # noqa: E402
                            ell_matrix_class, fill_row_blocks, get_rows_statistics,
                            gpu_coo_spmv, gpu_csr_adaptive_spmv, gpu_csr_spmv,
                            gpu_csr_vector_spmv, gpu_ell_spmv, gpu_hybrid_atomic_spmv, gpu_hybrid_cpu_coo_spmv,
                            gpu_hybrid_spmv, gpu_scoo_spmv, hybrid_matrix_class, matrix_class,
                            matrix_meta, matrix_rows_statistic, measurement_class,
                            resizable_gpu_memory, scoo_geometry, scoo_matrix_class)


def __getattr__(name):
    # the selector module loads on first use, so that
    # ``python -m matrix_format_performance_amd.selector`` runs it only once
    if name == "selector":
        import importlib
        return importlib.import_module(".selector", __name__)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


__all__ = [
    "MfpError", "MfpBudgetError", "FORMATS", "LABELS", "Batch", "autotune", "DeviceArray", "DeviceMatrix",
    "GenSpec", "Graph", "Plan", "Timer", "device_info", "flush_cache", "generate", "generate_vector",
    "init", "residual", "synchronize", "workload", "compare_results", "coo_matrix_class", "csr_matrix_class",
    "ell_matrix_class", "fill_row_blocks", "get_rows_statistics", "gpu_coo_spmv",
    "gpu_csr_adaptive_spmv", "gpu_csr_spmv", "gpu_csr_vector_spmv", "gpu_ell_spmv",
    "gpu_hybrid_atomic_spmv", "gpu_hybrid_cpu_coo_spmv", "gpu_hybrid_spmv", "gpu_scoo_spmv", "hybrid_matrix_class",
    "matrix_class", "matrix_meta", "matrix_rows_statistic", "measurement_class",
    "resizable_gpu_memory", "scoo_geometry", "scoo_matrix_class", "selector",
]
