"""ctypes binding of libmfp.so (the C ABI declared in include/mfp.h).

The library is loaded from this package directory only.  There is no
fallback: if libmfp.so is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfp.so")

# status / enums (include/mfp.h)
MFP_OK, MFP_ERR_HIP, MFP_ERR_ARG, MFP_ERR_BUDGET, MFP_ERR_STORAGE, MFP_ERR_NODEV, MFP_ERR_STATE = range(7)
MFP_F32, MFP_F64, MFP_BF16, MFP_F16 = range(4)
(MFP_CSR_SCALAR, MFP_CSR_VECTOR, MFP_CSR_ADAPTIVE, MFP_ELL, MFP_COO, MFP_SCOO, MFP_HYB,
 MFP_HYB_ATOMIC) = range(8)
MFP_FORMAT_COUNT = 8
(MFP_ARR_ELL_VALUES, MFP_ARR_ELL_COLS, MFP_ARR_COO_VALUES, MFP_ARR_COO_COLS, MFP_ARR_COO_ROWS,
 MFP_ARR_SCOO_SLICES_PTR, MFP_ARR_ROW_BLOCKS) = range(7)
MFP_GEN_STENCIL2D, MFP_GEN_STENCIL3D, MFP_GEN_POWERLAW, MFP_GEN_UNIFORM = range(4)
MFP_VAL_LAPLACE, MFP_VAL_UNIFORM, MFP_VAL_NORMAL = range(3)

u32, u64, i32, dbl, vp, sz = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_void_p, C.c_size_t


class MfpError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: status {status} ({detail})")
        self.status = status


class MfpBudgetError(MfpError):
    pass


class device_info(C.Structure):
    _fields_ = [("device", i32), ("cu_count", i32), ("xcd_count", i32), ("lds_per_block", u64),
                ("free_mem", u64), ("total_mem", u64), ("name", C.c_char * 256), ("arch", C.c_char * 64)]


class matrix_view(C.Structure):
    _fields_ = [("dtype", i32), ("rows", u32), ("cols", u32), ("nnz", u64), ("row_ptr", vp),
                ("columns", vp), ("values", vp)]


class row_stats(C.Structure):
    _fields_ = [("min_elements_in_rows", u32), ("max_elements_in_rows", u32),
                ("avg_elements_in_rows", u32), ("elements_in_rows_std_deviation", dbl)]


class tuning(C.Structure):
    _fields_ = [("lanes", u32), ("nnz_per_wg", u32), ("long_row", u32), ("chunk", u32),
                ("ell_width", u32), ("hyb_percent", dbl), ("hyb_width", u32), ("scoo_slices", u32),
                ("scoo_slice_size", u32), ("scoo_lanes", u32), ("scoo_threads", u32),
                ("budget_bytes", u64), ("flags", u32), ("scoo_pieces", u32)]


MFP_TUNE_NO_OFFSET16 = 1
MFP_TUNE_IS_DEFAULT = 2
MFP_TUNE_SCOO_COLSORT = 4
MFP_TUNE_STATIC_PARTIALS = 8
MFP_TUNE_SCOO_WHOLE_SLICES = 16
MFP_TUNE_ELL_PACKED = 32
MFP_BATCH_STREAM = 3
MFP_NOALLOC_SCRATCH_BYTES = 32 << 20


class candidate(C.Structure):
    _fields_ = [("format", i32), ("tuning", tuning), ("ms", dbl), ("status", i32)]


class error_report(C.Structure):
    _fields_ = [("rel_l2", dbl), ("max_row_err", dbl), ("max_abs_err", dbl)]


class plan_info(C.Structure):
    _fields_ = [("format", i32), ("dtype", i32), ("rows", u32), ("cols", u32), ("nnz", u64),
                ("device_bytes", u64), ("stored_elements", u64), ("ell_width", u32),
                ("coo_elements", u64), ("blocks", u32), ("chunks", u32), ("slices_count", u32),
                ("slice_size", u32), ("lane_size", u32), ("lanes", u32), ("nnz_per_wg", u32),
                ("column_bytes", u32), ("scoo_colsort", u32), ("scoo_packed", u32), ("scoo_long_rows", u32)]


class measurement(C.Structure):
    _fields_ = [("format", C.c_char * 64), ("elapsed", dbl), ("load_store_bytes", dbl),
                ("operations_count", dbl), ("effective_bandwidth", dbl),
                ("computational_throughput", dbl)]


class host_csr(C.Structure):
    _fields_ = [("rows", u32), ("cols", u32), ("nnz", u64), ("row_ptr", vp), ("columns", vp),
                ("data", vp)]


class host_ell(C.Structure):
    _fields_ = [("rows", u32), ("cols", u32), ("nnz", u64), ("elements_in_rows", u32),
                ("columns", vp), ("data", vp)]


class host_coo(C.Structure):
    _fields_ = [("rows", u32), ("cols", u32), ("nnz", u64), ("elements", u64), ("cols_", vp),
                ("rows_", vp), ("data", vp)]


class host_scoo(C.Structure):
    _fields_ = [("rows", u32), ("cols", u32), ("nnz", u64), ("slices_count", u32),
                ("slice_size", u32), ("lane_size", u32), ("values", vp), ("c_index", vp),
                ("r_index", vp), ("slices_ptr", vp)]


class gen_params(C.Structure):
    _fields_ = [("family", i32), ("values", i32), ("seed", u64), ("n", u32), ("rows", u32),
                ("nnz_lo", u32), ("nnz_hi", u32), ("alpha", dbl), ("bandwidth", dbl)]


P = C.POINTER

# name -> (restype, argtypes); every symbol include/mfp.h declares
SIGNATURES = {
    "mfp_init": (i32, [i32]),
    "mfp_device_query": (i32, [P(device_info)]),
    "mfp_status_string": (C.c_char_p, [i32]),
    "mfp_last_error": (C.c_char_p, []),
    "mfp_synchronize": (i32, []),
    "mfp_malloc": (i32, [P(vp), sz]),
    "mfp_free": (i32, [vp]),
    "mfp_memcpy_h2d": (i32, [vp, vp, sz]),
    "mfp_memcpy_d2h": (i32, [vp, vp, sz]),
    "mfp_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "mfp_fill": (i32, [i32, vp, u64, dbl, vp]),
    "mfp_flush_cache": (i32, [u64, vp]),
    "mfp_stream_copy": (i32, [vp, vp, u64, vp]),
    "mfp_timer_create": (i32, [P(vp)]),
    "mfp_timer_start": (i32, [vp, vp]),
    "mfp_timer_stop": (i32, [vp, vp]),
    "mfp_timer_elapsed_ms": (i32, [vp, P(C.c_float)]),
    "mfp_timer_destroy": (i32, [vp]),
    "mfp_buffer_create": (i32, [P(vp)]),
    "mfp_buffer_resize": (i32, [vp, sz]),
    "mfp_buffer_clear": (i32, [vp]),
    "mfp_buffer_ptr": (vp, [vp]),
    "mfp_buffer_size": (sz, [vp]),
    "mfp_buffer_destroy": (i32, [vp]),
    "mfp_matrix_from_host_csr": (i32, [P(vp), i32, u32, u32, u64, vp, vp, vp]),
    "mfp_matrix_from_triplets": (i32, [P(vp), i32, u32, u32, u64, vp, vp, vp, i32, vp]),
    "mfp_matrix_view_get": (i32, [vp, P(matrix_view)]),
    "mfp_matrix_download": (i32, [vp, vp, vp, vp]),
    "mfp_matrix_destroy": (i32, [vp]),
    "mfp_matrix_row_stats": (i32, [vp, P(row_stats)]),
    "mfp_matrix_convert": (i32, [vp, i32, P(vp)]),
    "mfp_spmv_f64_oracle": (i32, [vp, i32, vp, vp, vp, vp]),
    "mfp_residual": (i32, [u32, i32, vp, vp, vp, P(error_report)]),
    "mfp_tuning_default": (i32, [vp, i32, P(tuning)]),
    "mfp_plan_create": (i32, [P(vp), vp, i32, P(tuning), vp]),
    "mfp_plan_spmv": (i32, [vp, vp, vp, vp]),
    "mfp_plan_spmv_noalloc": (i32, [vp, vp, vp, vp]),
    "mfp_plan_kernel_times": (i32, [vp, vp, vp, vp, P(C.c_float), i32, P(i32)]),
    "mfp_plan_spmm": (i32, [vp, u32, vp, vp, vp]),
    "mfp_batch_create": (i32, [P(vp), P(vp), u32, u32, P(tuning), vp]),
    "mfp_batch_info": (i32, [vp, P(u32), P(u32), P(u64), vp, vp]),
    "mfp_batch_matrix": (i32, [vp, P(vp)]),
    "mfp_batch_spmm": (i32, [vp, u32, vp, vp, vp]),
    "mfp_batch_graph_create": (i32, [vp, u32, vp, vp, P(vp)]),
    "mfp_plan_graph_create": (i32, [vp, vp, vp, P(vp)]),
    "mfp_graph_launch": (i32, [vp, vp]),
    "mfp_graph_destroy": (i32, [vp]),
    "mfp_batch_destroy": (i32, [vp]),
    "mfp_plan_get_info": (i32, [vp, P(plan_info)]),
    "mfp_plan_array": (i32, [vp, i32, P(vp), P(u64)]),
    "mfp_plan_destroy": (i32, [vp]),
    "mfp_autotune_candidates": (i32, [vp, u32, vp, u32, P(u32)]),
    "mfp_autotune_run": (i32, [vp, vp, vp, vp, u32, u32, i32]),
    "mfp_matrix_features": (i32, [vp, vp]),
    "mfp_mlp_create": (i32, [P(vp), u64]),
    "mfp_mlp_destroy": (i32, [vp]),
    "mfp_mlp_param_count": (u32, []),
    "mfp_mlp_get_params": (i32, [vp, vp]),
    "mfp_mlp_set_params": (i32, [vp, vp]),
    "mfp_mlp_loss_grad": (i32, [vp, vp, vp, u32, P(dbl), vp]),
    "mfp_mlp_train_step": (i32, [vp, vp, vp, u32, C.c_float, P(dbl)]),
    "mfp_mlp_predict": (i32, [vp, vp, u32, i32, vp]),
    "mfp_fill_row_blocks": (i32, [u32, vp, u32, vp, P(u32)]),
    "mfp_scoo_geometry": (i32, [i32, u32, u32, u64, P(u32), P(u32), P(u32)]),
    "mfp_compare_results": (i32, [i32, u32, vp, vp]),
    "mfp_gpu_csr_spmv": (i32, [i32, P(host_csr), vp, vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_csr_vector_spmv": (i32, [i32, P(host_csr), vp, vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_csr_adaptive_spmv": (i32, [i32, P(host_csr), vp, vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_ell_spmv": (i32, [i32, P(host_ell), vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_coo_spmv": (i32, [i32, P(host_coo), vp, vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_scoo_spmv": (i32, [i32, i32, P(host_scoo), vp, vp, vp, vp, vp, vp, vp, P(measurement)]),
    "mfp_gpu_hybrid_spmv": (i32, [i32, P(host_ell), P(host_coo), vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  P(measurement)]),
    "mfp_gpu_hybrid_atomic_spmv": (i32, [i32, P(host_ell), P(host_coo), vp, vp, vp, vp, vp, vp, vp,
                                         vp, vp, P(measurement)]),
    "mfp_gpu_hybrid_cpu_coo_spmv": (i32, [i32, P(host_ell), P(host_coo), vp, vp, vp, vp, vp, vp, vp, vp,
                                          P(measurement)]),
    "mfp_generate": (i32, [P(gen_params), P(u32), P(u32), P(u64), vp, vp, vp, i32]),
    "mfp_generate_vector": (i32, [u64, u64, vp, i32]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmfp.so from this package (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m matrix_format_performance_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, where: str) -> None:
    if status == MFP_OK:
        return
    lib = load()
    detail = lib.mfp_last_error().decode(errors="replace") or lib.mfp_status_string(status).decode()
    if status == MFP_ERR_BUDGET:
        raise MfpBudgetError(status, where, detail)
    raise MfpError(status, where, detail)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
