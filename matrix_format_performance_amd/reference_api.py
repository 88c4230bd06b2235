"""The reference's host interface for the SpMV path, over libmfp.

Same names, argument meaning and error behaviour as the reference's C++:
  matrix_converter.h        csr/ell/coo/hybrid_matrix_class, get_rows_statistics
  scoo_matrix_class.h       scoo_matrix_class
  measurement_class.h       measurement_class, compare_results
  resizable_gpu_memory.h    resizable_gpu_memory
  gpu_matrix_multiplier.h   gpu_csr_spmv, gpu_csr_vector_spmv, gpu_ell_spmv,
                            gpu_coo_spmv, gpu_hybrid_spmv, gpu_hybrid_atomic_spmv,
                            gpu_hybrid_cpu_coo_spmv
  csr_adaptive_spmv.h       gpu_csr_adaptive_spmv (+ fill_row_blocks)
  scoo_spmv.h               gpu_scoo_spmv
The layouts are built by libmfp's device converters and the SpMVs run its HIP
kernels.  Value types are numpy float32 / float64 (the reference's two
instantiations).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import engine as E

GENERAL = "general"


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- MatrixMarket-style triplets (matrix_market::matrix_class) ---------------

@dataclass
class matrix_meta:
    rows_count: int
    cols_count: int
    non_zero_count: int
    matrix_storage_scheme: str = GENERAL


class matrix_class:
    """Triplet container with the reader API the converter uses
    (matrix_converter.cpp:29-33): meta, get_row_ids, get_col_ids, get_dbl_data.
    Indices are 0-based."""

    def __init__(self, rows, cols, row_ids, col_ids, values, storage=GENERAL):
        self._r = np.ascontiguousarray(row_ids, np.uint32)
        self._c = np.ascontiguousarray(col_ids, np.uint32)
        self._v = np.ascontiguousarray(values, np.float64)
        self.meta = matrix_meta(int(rows), int(cols), int(self._r.size), storage)

    def get_row_ids(self):
        return self._r

    def get_col_ids(self):
        return self._c

    def get_dbl_data(self):
        return self._v


def _require_general(meta):
    if meta.matrix_storage_scheme != GENERAL:
        raise RuntimeError("Only general matrices are supported")


# ---- measurement_class (common/measurement_class.{h,cpp}) ---------------------

def _cdiv(a: float, b: float) -> float:
    if b == 0.0:
        return math.nan if a == 0.0 or math.isnan(a) else math.copysign(math.inf, a)
    return a / b


class measurement_class:
    giga = 1e9

    def __init__(self, format: str = "", elapsed: float = 0.0, load_store_bytes: float | None = None,
                 operations_count: float | None = None):
        self._format = format
        self._count = 0
        if load_store_bytes is None:  # default-constructed record
            self._elapsed = 0.0
            self._bw = 0.0
            self._flops = 0.0
        else:
            self._elapsed = float(elapsed)
            self._bw = _cdiv(float(load_store_bytes), self._elapsed * self.giga)
            self._flops = _cdiv(float(operations_count), self._elapsed * self.giga)

    def get_elapsed(self):
        return self._elapsed

    def get_effective_bandwidth(self):
        return self._bw

    def get_computational_throughput(self):
        return self._flops

    def get_format(self):
        return self._format

    def __iadd__(self, rhs: "measurement_class"):
        self._elapsed += rhs.get_elapsed()
        self._bw += rhs.get_effective_bandwidth()
        self._flops += rhs.get_computational_throughput()
        self._format = rhs.get_format()
        self._count += 1
        return self

    def finalize(self):
        if self._count:
            self._elapsed /= self._count
            self._bw /= self._count
            self._flops /= self._count

    @classmethod
    def _from_c(cls, m: L.measurement) -> "measurement_class":
        r = cls(m.format.decode(), m.elapsed, m.load_store_bytes, m.operations_count)
        r.load_store_bytes = m.load_store_bytes
        r.operations_count = m.operations_count
        return r


def compare_results(y_size: int, a, b) -> None:
    """measurement_class.h:60-89 (prints to stderr, never raises)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    L.call("mfp_compare_results", E.dtype_code(a.dtype), int(y_size), _ptr(a), _ptr(b))


# ---- resizable_gpu_memory (gpu/resizable_gpu_memory.h) -------------------------

class resizable_gpu_memory:
    def __init__(self, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        h = C.c_void_p()
        L.call("mfp_buffer_create", C.byref(h))
        self.h = h

    def resize(self, new_size: int) -> None:
        L.call("mfp_buffer_resize", self.h, int(new_size) * self.dtype.itemsize)

    def clear(self) -> None:
        L.call("mfp_buffer_clear", self.h)

    def get(self) -> int:
        return L.load().mfp_buffer_ptr(self.h)

    def size_bytes(self) -> int:
        return L.load().mfp_buffer_size(self.h)

    def __del__(self):
        try:
            L.load().mfp_buffer_destroy(self.h)
        except Exception:
            pass


# ---- converters (common/matrix_converter.{h,cpp}) -----------------------------

@dataclass
class matrix_rows_statistic:
    min_elements_in_rows: int
    max_elements_in_rows: int
    avg_elements_in_rows: int
    elements_in_rows_std_deviation: float


class csr_matrix_class:
    """matrix_converter.cpp:15-87 -- built by the device converter."""

    def __init__(self, matrix: matrix_class, dtype=np.float32, row_ptr_only: bool = False):
        _require_general(matrix.meta)
        self.meta = matrix.meta
        self.dtype = np.dtype(dtype)
        dm = E.DeviceMatrix.from_triplets(matrix.meta.rows_count, matrix.meta.cols_count,
                                          matrix.get_row_ids(), matrix.get_col_ids(),
                                          matrix.get_dbl_data(), self.dtype)
        self.row_ptr, cols, vals = dm.download()
        dm.destroy()
        self.columns = None if row_ptr_only else cols
        self.data = None if row_ptr_only else vals

    @classmethod
    def from_arrays(cls, row_ptr, columns, data, cols=None):
        self = cls.__new__(cls)
        self.row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        self.columns = np.ascontiguousarray(columns, np.uint32)
        self.data = np.ascontiguousarray(data)
        self.dtype = self.data.dtype
        rows = self.row_ptr.size - 1
        self.meta = matrix_meta(rows, rows if cols is None else cols, int(self.columns.size))
        return self

    def get_matrix_size(self) -> int:
        return self.meta.non_zero_count

    def _device(self) -> E.DeviceMatrix:
        return E.DeviceMatrix.from_csr(self.row_ptr, self.columns, self.data, self.meta.cols_count)


def get_rows_statistics(meta: matrix_meta, row_ptr) -> matrix_rows_statistic:
    """matrix_converter.cpp:95-127, reduced on the device."""
    rows = int(meta.rows_count)
    row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
    nnz = int(row_ptr[rows])
    dm = E.DeviceMatrix.from_csr(row_ptr, np.zeros(nnz, np.uint32), np.zeros(nnz, np.float32))
    mn, mx, avg, sd = dm.row_stats()
    dm.destroy()
    return matrix_rows_statistic(mn, mx, avg, sd)


def _plan_for(csr: csr_matrix_class, fmt: str, **tuning):
    dm = csr._device()
    try:
        plan = E.Plan(dm, fmt, **tuning)
    except Exception:
        dm.destroy()
        raise
    return dm, plan


class ell_matrix_class:
    """matrix_converter.cpp:139-203: width = longest row, or the given width
    (longer rows truncated)."""

    def __init__(self, matrix: csr_matrix_class, elements_in_row_arg: int | None = None):
        self.meta = matrix.meta
        if elements_in_row_arg is None:
            _require_general(matrix.meta)
            dm, plan = _plan_for(matrix, "ell")
        else:
            dm, plan = _plan_for(matrix, "hyb", hyb_width=int(elements_in_row_arg))
        self.elements_in_rows = plan.info()["ell_width"]
        self.data = plan.array(L.MFP_ARR_ELL_VALUES, matrix.dtype)
        self.columns = plan.array(L.MFP_ARR_ELL_COLS, np.uint32)
        plan.destroy()
        dm.destroy()

    @staticmethod
    def estimate_size(matrix: csr_matrix_class) -> int:
        st = get_rows_statistics(matrix.meta, matrix.row_ptr)
        return st.max_elements_in_rows * matrix.meta.rows_count

    def get_matrix_size(self) -> int:
        return self.meta.rows_count * self.elements_in_rows


class coo_matrix_class:
    """matrix_converter.cpp:211-280 (element_start > 0: only in-row offsets
    >= element_start)."""

    def __init__(self, matrix: csr_matrix_class, element_start: int = 0):
        self.meta = matrix.meta
        if element_start == 0:
            _require_general(matrix.meta)
            dm, plan = _plan_for(matrix, "coo")
        else:
            dm, plan = _plan_for(matrix, "hyb", hyb_width=int(element_start))
        self.data = plan.array(L.MFP_ARR_COO_VALUES, matrix.dtype)
        self.cols = plan.array(L.MFP_ARR_COO_COLS, np.uint32)
        self.rows = plan.array(L.MFP_ARR_COO_ROWS, np.uint32)
        self._elements = self.data.size
        plan.destroy()
        dm.destroy()

    def get_matrix_size(self) -> int:
        return self._elements


class hybrid_matrix_class:
    """matrix_converter.cpp:286-310."""

    def __init__(self, matrix: csr_matrix_class):
        _require_general(matrix.meta)
        self.meta = matrix.meta
        self.ell_matrix = None
        self.coo_matrix = None

    def allocate(self, matrix: csr_matrix_class, percent: float) -> None:
        dm, plan = _plan_for(matrix, "hyb", hyb_percent=float(percent))
        w = plan.info()["ell_width"]
        ell = ell_matrix_class.__new__(ell_matrix_class)
        ell.meta, ell.elements_in_rows = matrix.meta, w
        ell.data = plan.array(L.MFP_ARR_ELL_VALUES, matrix.dtype)
        ell.columns = plan.array(L.MFP_ARR_ELL_COLS, np.uint32)
        coo = coo_matrix_class.__new__(coo_matrix_class)
        coo.meta = matrix.meta
        coo.data = plan.array(L.MFP_ARR_COO_VALUES, matrix.dtype)
        coo.cols = plan.array(L.MFP_ARR_COO_COLS, np.uint32)
        coo.rows = plan.array(L.MFP_ARR_COO_ROWS, np.uint32)
        coo._elements = coo.data.size
        plan.destroy()
        dm.destroy()
        self.ell_matrix, self.coo_matrix = ell, coo


def scoo_geometry(dtype, rows: int, sm_count: int, shared_mem_size: int):
    """scoo_matrix_class.cpp:30-62 -> (slices_count, slice_size, lane_size)."""
    s, ss, l = C.c_uint32(), C.c_uint32(), C.c_uint32()
    L.call("mfp_scoo_geometry", E.dtype_code(dtype), int(rows), int(sm_count), int(shared_mem_size),
           C.byref(s), C.byref(ss), C.byref(l))
    return s.value, ss.value, l.value


class scoo_matrix_class:
    """common/scoo_matrix_class.cpp:58-139 (sort_cols = false)."""

    def __init__(self, sm_count: int, shared_mem_size: int, matrix: csr_matrix_class):
        self.meta = matrix.meta
        self.slices_count, self.slice_size, self.lane_size = scoo_geometry(
            matrix.dtype, matrix.meta.rows_count, sm_count, shared_mem_size)
        if self.meta.rows_count == 0:
            raise ValueError("SCOO needs at least one row")
        dm, plan = _plan_for(matrix, "scoo", scoo_slices=self.slices_count,
                             scoo_slice_size=self.slice_size, scoo_lanes=self.lane_size)
        self.values = plan.array(L.MFP_ARR_COO_VALUES, matrix.dtype)
        self.c_index = plan.array(L.MFP_ARR_COO_COLS, np.uint32)
        self.r_index = plan.array(L.MFP_ARR_COO_ROWS, np.uint32)
        self.slices_ptr = plan.array(L.MFP_ARR_SCOO_SLICES_PTR, np.uint32)
        plan.destroy()
        dm.destroy()


def fill_row_blocks(rows_count: int, row_ptr, nnz_per_wg: int = 64) -> np.ndarray:
    """gpu/csr_adaptive_spmv.cu:180-240 (computed on the device)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
    n = C.c_uint32()
    L.call("mfp_fill_row_blocks", int(rows_count), _ptr(row_ptr), int(nnz_per_wg), None, C.byref(n))
    out = np.empty(n.value + 1, np.uint32)
    L.call("mfp_fill_row_blocks", int(rows_count), _ptr(row_ptr), int(nnz_per_wg), _ptr(out), C.byref(n))
    return out


# ---- GPU entry points ------------------------------------------------------------

def _meas():
    return L.measurement()


def _host_csr(m: csr_matrix_class):
    return L.host_csr(m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count,
                      _ptr(m.row_ptr), _ptr(m.columns), _ptr(m.data))


def _host_ell(m: ell_matrix_class):
    return L.host_ell(m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count,
                      m.elements_in_rows, _ptr(m.columns), _ptr(m.data))


def _host_coo(m: coo_matrix_class):
    return L.host_coo(m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count,
                      m.get_matrix_size(), _ptr(m.cols), _ptr(m.rows), _ptr(m.data))


def _dt(m) -> int:
    return E.dtype_code(m.data.dtype if hasattr(m, "data") else m.values.dtype)


def _check_vec(v, n, dtype, name):
    if v is None:
        return None
    if not (isinstance(v, np.ndarray) and v.dtype == dtype and v.flags.c_contiguous and v.size >= n):
        raise ValueError(f"{name} must be a contiguous {np.dtype(dtype)} array of >= {n} entries")
    return _ptr(v)


def _csr_call(fn, matrix, A, col_ids, row_ptr, x, y, reusable_vector, reference_y):
    hc = _host_csr(matrix)
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, matrix.data.dtype, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, matrix.data.dtype, "reference_y")
    out = _meas()
    L.call(fn, _dt(matrix), C.byref(hc), A.h, col_ids.h, row_ptr.h, x.h, y.h, rv, ry, C.byref(out))
    return measurement_class._from_c(out)


def gpu_csr_spmv(matrix, A, col_ids, row_ptr, x, y, reusable_vector, reference_y):
    return _csr_call("mfp_gpu_csr_spmv", matrix, A, col_ids, row_ptr, x, y, reusable_vector, reference_y)


def gpu_csr_vector_spmv(matrix, A, col_ids, row_ptr, x, y, reusable_vector, reference_y):
    return _csr_call("mfp_gpu_csr_vector_spmv", matrix, A, col_ids, row_ptr, x, y, reusable_vector,
                     reference_y)


def gpu_csr_adaptive_spmv(matrix, A, col_ids, row_ptr, x, y, reusable_vector, reference_y):
    return _csr_call("mfp_gpu_csr_adaptive_spmv", matrix, A, col_ids, row_ptr, x, y, reusable_vector,
                     reference_y)


def gpu_ell_spmv(matrix: ell_matrix_class, A, col_ids, x, y, reusable_vector, reference_y):
    he = _host_ell(matrix)
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, matrix.data.dtype, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, matrix.data.dtype, "reference_y")
    out = _meas()
    L.call("mfp_gpu_ell_spmv", _dt(matrix), C.byref(he), A.h, col_ids.h, x.h, y.h, rv, ry, C.byref(out))
    return measurement_class._from_c(out)


def gpu_coo_spmv(matrix: coo_matrix_class, A, col_ids, row_ids, x, y, reusable_vector, reference_y):
    hc = _host_coo(matrix)
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, matrix.data.dtype, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, matrix.data.dtype, "reference_y")
    out = _meas()
    L.call("mfp_gpu_coo_spmv", _dt(matrix), C.byref(hc), A.h, col_ids.h, row_ids.h, x.h, y.h, rv, ry,
           C.byref(out))
    return measurement_class._from_c(out)


def gpu_scoo_spmv(print_diff: bool, matrix: scoo_matrix_class, A, r_index, c_index, x, y,
                  reusable_vector, reference_y):
    hs = L.host_scoo(matrix.meta.rows_count, matrix.meta.cols_count, matrix.meta.non_zero_count,
                     matrix.slices_count, matrix.slice_size, matrix.lane_size, _ptr(matrix.values),
                     _ptr(matrix.c_index), _ptr(matrix.r_index), _ptr(matrix.slices_ptr))
    dt = matrix.values.dtype
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, dt, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, dt, "reference_y")
    out = _meas()
    L.call("mfp_gpu_scoo_spmv", E.dtype_code(dt), int(bool(print_diff)), C.byref(hs), A.h, r_index.h,
           c_index.h, x.h, y.h, rv, ry, C.byref(out))
    return measurement_class._from_c(out)


def _hyb_call(fn, matrix: hybrid_matrix_class, A_ell, A_coo, ell_col_ids, coo_col_ids, coo_row_ids,
              x, y, reusable_vector, reference_y):
    he, hc = _host_ell(matrix.ell_matrix), _host_coo(matrix.coo_matrix)
    dt = matrix.ell_matrix.data.dtype
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, dt, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, dt, "reference_y")
    out = _meas()
    L.call(fn, E.dtype_code(dt), C.byref(he), C.byref(hc), A_ell.h, A_coo.h, ell_col_ids.h,
           coo_col_ids.h, coo_row_ids.h, x.h, y.h, rv, ry, C.byref(out))
    return measurement_class._from_c(out)


def gpu_hybrid_spmv(matrix, A_ell, A_coo, ell_col_ids, coo_col_ids, coo_row_ids, x, y,
                    reusable_vector, reference_y):
    return _hyb_call("mfp_gpu_hybrid_spmv", matrix, A_ell, A_coo, ell_col_ids, coo_col_ids,
                     coo_row_ids, x, y, reusable_vector, reference_y)


def gpu_hybrid_atomic_spmv(matrix, A_ell, A_coo, ell_col_ids, coo_col_ids, coo_row_ids, x, y,
                           reusable_vector, reference_y):
    return _hyb_call("mfp_gpu_hybrid_atomic_spmv", matrix, A_ell, A_coo, ell_col_ids, coo_col_ids,
                     coo_row_ids, x, y, reusable_vector, reference_y)


def gpu_hybrid_cpu_coo_spmv(matrix: hybrid_matrix_class, A_ell, ell_col_ids, x, y, tmp, cpu_y,
                            reusable_vector, reference_y):
    """gpu/gpu_matrix_multiplier.cu:1111-1217: the ELL part on the GPU while
    the host sums the COO tail into ``cpu_y``; the tail goes through ``tmp``
    and is added to y."""
    he, hc = _host_ell(matrix.ell_matrix), _host_coo(matrix.coo_matrix)
    dt = matrix.ell_matrix.data.dtype
    cy = _check_vec(cpu_y, matrix.meta.rows_count, dt, "cpu_y")
    if cy is None and matrix.meta.rows_count:
        raise ValueError("cpu_y is required")
    rv = _check_vec(reusable_vector, matrix.meta.rows_count, dt, "reusable_vector")
    ry = _check_vec(reference_y, matrix.meta.rows_count, dt, "reference_y")
    out = _meas()
    L.call("mfp_gpu_hybrid_cpu_coo_spmv", E.dtype_code(dt), C.byref(he), C.byref(hc), A_ell.h,
           ell_col_ids.h, x.h, y.h, tmp.h, cy, rv, ry, C.byref(out))
    return measurement_class._from_c(out)
