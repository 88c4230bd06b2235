"""CPU oracle for the SpMV hot path -- TEST INFRASTRUCTURE ONLY.

numpy front end over ``liboracle.so`` (``mfp_oracle.c``), the plain-C
restatement of the reference's host converters, CPU SpMV loops, partition
builders and measurement arithmetic.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package, and only as the checker / CPU baseline: the product library
never loads it.

Each wrapper names the reference ``file:line`` it restates (paths relative to
the reference repository root).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_measurement.so")

F32, F64 = 0, 1

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir(os.environ.get("MFP_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    u32, u64, i32, dbl = C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "orc_csr_from_triplets": (i32, [u32, u32, u64, _u32p, _u32p, _f64p, i32, _u32p, _u32p, _vp]),
        "orc_rows_statistics": (None, [u32, _u32p, _u32p, C.POINTER(dbl)]),
        "orc_ell_from_csr": (None, [u32, _u32p, _u32p, _vp, i32, u32, _vp, _u32p]),
        "orc_coo_count": (u64, [u32, _u32p, u32]),
        "orc_coo_from_csr": (None, [u32, _u32p, _u32p, _vp, i32, u32, _vp, _u32p, _u32p]),
        "orc_hybrid_width": (u32, [u32, _u32p, dbl]),
        "orc_next_pow_2": (u32, [u32]),
        "orc_prev_pow_2": (u32, [u32]),
        "orc_scoo_slices_count": (u32, [u32, u32, u64, i32]),
        "orc_scoo_lane_size": (u32, [u32, u32, u64, i32]),
        "orc_scoo_build": (None, [u32, _u32p, _u32p, _vp, i32, u32, u32, _vp, _u32p, _u32p, _u32p]),
        "orc_fill_row_blocks": (u32, [i32, u32, _u32p, _vp, u32]),
        "orc_spmv_csr_seq": (None, [i32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmv_csr_seq_nofma": (None, [i32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmv_csr_lanes": (None, [i32, u32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmv_ell_seq": (None, [i32, u32, u32, _u32p, _vp, _vp, _vp]),
        "orc_spmv_coo_seq": (None, [i32, u64, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmm_stream_col": (i32, [i32, u32, u32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmv_csr_adaptive": (None, [i32, u32, _u32p, _u32p, _vp, _vp, _vp, u32, u32, u32, u32]),
        "orc_spmv_csr_f64": (None, [i32, u32, _u32p, _u32p, _vp, _vp, _f64p, _vp]),
        "orc_spmv_csr_multi_thread": (dbl, [i32, u32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_spmv_csr_seq_timed": (dbl, [i32, u32, u32, _u32p, _u32p, _vp, _vp, _vp]),
        "orc_compare_error": (dbl, [i32, u32, _vp, _vp]),
        "orc_measurement": (None, [dbl, dbl, dbl, _f64p]),
        "orc_reference_bytes": (dbl, [i32, i32, u32, u64, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def _dt(arr: np.ndarray) -> int:
    if arr.dtype == np.float32:
        return F32
    if arr.dtype == np.float64:
        return F64
    raise TypeError(f"oracle handles float32/float64 values, got {arr.dtype}")


def _np_dtype(code: int):
    return np.float32 if code == F32 else np.float64


def _c(a, dtype=None):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- converters (common/matrix_converter.cpp) ------------------------------

def csr_from_triplets(rows, cols, tr_rows, tr_cols, tr_vals, dtype=np.float32):
    """Triplets -> CSR, common/matrix_converter.cpp:15-87 (stable for ties)."""
    tr_rows, tr_cols = _c(tr_rows, np.uint32), _c(tr_cols, np.uint32)
    tr_vals = _c(tr_vals, np.float64)
    nnz = tr_rows.size
    row_ptr = np.zeros(rows + 1, np.uint32)
    columns = np.zeros(max(nnz, 1), np.uint32)
    data = np.zeros(max(nnz, 1), dtype)
    rc = _lib.orc_csr_from_triplets(rows, cols, nnz, tr_rows, tr_cols, tr_vals,
                                    _dt(data), row_ptr, columns, _ptr(data))
    if rc != 0:
        raise ValueError("triplet row/col out of range")
    return row_ptr, columns[:nnz], data[:nnz]


def rows_statistics(row_ptr):
    """common/matrix_converter.cpp:95-127 -> (min, max, avg, std)."""
    row_ptr = _c(row_ptr, np.uint32)
    out = np.zeros(3, np.uint32)
    sd = C.c_double()
    _lib.orc_rows_statistics(row_ptr.size - 1, row_ptr, out, C.byref(sd))
    return int(out[0]), int(out[1]), int(out[2]), sd.value


def ell_from_csr(row_ptr, columns, data, width=None):
    """common/matrix_converter.cpp:139-203 (width None -> max row length)."""
    row_ptr, columns, data = _c(row_ptr, np.uint32), _c(columns, np.uint32), _c(data)
    rows = row_ptr.size - 1
    if width is None:
        width = rows_statistics(row_ptr)[1]
    n = rows * width
    ed = np.zeros(max(n, 1), data.dtype)
    ec = np.zeros(max(n, 1), np.uint32)
    _lib.orc_ell_from_csr(rows, row_ptr, columns if columns.size else np.zeros(1, np.uint32),
                          _ptr(data), _dt(data), width, _ptr(ed), ec)
    return width, ed[:n], ec[:n]


def coo_from_csr(row_ptr, columns, data, element_start=0):
    """common/matrix_converter.cpp:211-280 -> (data, cols, rows)."""
    row_ptr, columns, data = _c(row_ptr, np.uint32), _c(columns, np.uint32), _c(data)
    rows = row_ptr.size - 1
    n = int(_lib.orc_coo_count(rows, row_ptr, element_start))
    cd = np.zeros(max(n, 1), data.dtype)
    cc = np.zeros(max(n, 1), np.uint32)
    cr = np.zeros(max(n, 1), np.uint32)
    _lib.orc_coo_from_csr(rows, row_ptr, columns if columns.size else np.zeros(1, np.uint32),
                          _ptr(data), _dt(data), element_start, _ptr(cd), cc, cr)
    return cd[:n], cc[:n], cr[:n]


def hybrid_width(row_ptr, percent=0.2):
    """common/matrix_converter.cpp:296-302."""
    row_ptr = _c(row_ptr, np.uint32)
    return int(_lib.orc_hybrid_width(row_ptr.size - 1, row_ptr, percent))


# ---- SCOO (common/scoo_matrix_class.cpp) ------------------------------------

def next_pow_2(x):
    return int(_lib.orc_next_pow_2(x))


def prev_pow_2(x):
    return int(_lib.orc_prev_pow_2(x))


def scoo_geometry(rows, sm_count, shared_mem_size, dtype=np.float32):
    """scoo_matrix_class.cpp:30-62 -> (slices_count, slice_size, lane_size)."""
    d = F32 if np.dtype(dtype) == np.float32 else F64
    slices = int(_lib.orc_scoo_slices_count(rows, sm_count, shared_mem_size, d))
    ss = (rows + slices - 1) // slices
    lane = int(_lib.orc_scoo_lane_size(slices, ss, shared_mem_size, d))
    return slices, ss, lane


def scoo_build(row_ptr, columns, data, slices_count, slice_size):
    """scoo_matrix_class.cpp:58-101 -> (values, c_index, r_index, slices_ptr)."""
    row_ptr, columns, data = _c(row_ptr, np.uint32), _c(columns, np.uint32), _c(data)
    rows = row_ptr.size - 1
    nnz = int(row_ptr[-1])
    v = np.zeros(max(nnz, 1), data.dtype)
    c = np.zeros(max(nnz, 1), np.uint32)
    r = np.zeros(max(nnz, 1), np.uint32)
    sp = np.zeros(slices_count + 1, np.uint32)
    _lib.orc_scoo_build(rows, row_ptr, columns if columns.size else np.zeros(1, np.uint32),
                        _ptr(data), _dt(data), slices_count, slice_size, _ptr(v), c, r, sp)
    return v[:nnz], c[:nnz], r[:nnz], sp


# ---- row blocks (gpu/csr_adaptive_spmv.cu:180-240) ---------------------------

def fill_row_blocks(row_ptr, nnz_per_wg=64):
    row_ptr = _c(row_ptr, np.uint32)
    rows = row_ptr.size - 1
    n = int(_lib.orc_fill_row_blocks(0, rows, row_ptr, None, nnz_per_wg))
    out = np.zeros(n + 1, np.uint32)
    _lib.orc_fill_row_blocks(1, rows, row_ptr, _ptr(out), nnz_per_wg)
    return out


# ---- SpMV ---------------------------------------------------------------------

def _spmv_args(row_ptr, columns, data, x):
    row_ptr, columns, data = _c(row_ptr, np.uint32), _c(columns, np.uint32), _c(data)
    x = _c(x, data.dtype)
    if columns.size == 0:
        columns = np.zeros(1, np.uint32)
        data = np.zeros(1, data.dtype)
    return row_ptr, columns, data, x


def spmv_seq(row_ptr, columns, data, x):
    """cpu/cpu_matrix_multiplier.cpp:20-61 (fma-contracted sequential dot)."""
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    _lib.orc_spmv_csr_seq(_dt(data), y.size, row_ptr, columns, _ptr(data), _ptr(x), _ptr(y))
    return y


def spmv_seq_nofma(row_ptr, columns, data, x):
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    _lib.orc_spmv_csr_seq_nofma(_dt(data), y.size, row_ptr, columns, _ptr(data), _ptr(x), _ptr(y))
    return y


def spmv_lanes(row_ptr, columns, data, x, lanes=32):
    """cpu/cpu_matrix_multiplier.cpp:170-227 generalised to `lanes`."""
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    _lib.orc_spmv_csr_lanes(_dt(data), lanes, y.size, row_ptr, columns, _ptr(data), _ptr(x), _ptr(y))
    return y


def spmv_ell(rows, width, ell_cols, ell_data, x):
    """gpu/gpu_matrix_multiplier.cu:450-471 order (sequential over width)."""
    ell_cols, ell_data = _c(ell_cols, np.uint32), _c(ell_data)
    x = _c(x, ell_data.dtype)
    if ell_cols.size == 0:
        ell_cols, ell_data = np.zeros(1, np.uint32), np.zeros(1, ell_data.dtype)
    y = np.zeros(rows, ell_data.dtype)
    _lib.orc_spmv_ell_seq(_dt(ell_data), rows, width, ell_cols, _ptr(ell_data), _ptr(x), _ptr(y))
    return y


def spmv_coo_seq(rows, coo_rows, coo_cols, coo_data, x, y=None):
    """gpu/gpu_matrix_multiplier.cu:1163-1168 order (entry by entry, product
    rounded before the add); accumulates into y (zeros when None)."""
    coo_rows, coo_cols, coo_data = _c(coo_rows, np.uint32), _c(coo_cols, np.uint32), _c(coo_data)
    x = _c(x, coo_data.dtype)
    y = np.zeros(rows, coo_data.dtype) if y is None else np.array(y, coo_data.dtype, copy=True)
    if coo_data.size:
        _lib.orc_spmv_coo_seq(_dt(coo_data), coo_data.size, coo_rows, coo_cols, _ptr(coo_data), _ptr(x), _ptr(y))
    return y


def spmm_stream(row_ptr, columns, data, X, k, W=512, threads=256):
    """Order of the library's CSR-stream SpMM (batch.hip k_spmm_stream) for
    X (cols x k, row-major): groups = threads / (lanes per X row), a lane
    holding 16 bytes of the row; each column is summed alike."""
    row_ptr, columns, data = _c(row_ptr, np.uint32), _c(columns, np.uint32), _c(data)
    X = np.asarray(X, data.dtype).reshape(-1, k)
    groups = threads // (k * data.dtype.itemsize // 16)
    rows = row_ptr.size - 1
    Y = np.zeros((rows, k), data.dtype)
    cols_ = columns if columns.size else np.zeros(1, np.uint32)
    vals_ = data if data.size else np.zeros(1, data.dtype)
    for j in range(k):
        x = np.ascontiguousarray(X[:, j])
        y = np.zeros(rows, data.dtype)
        if _lib.orc_spmm_stream_col(_dt(data), W, groups, rows, row_ptr, cols_, _ptr(vals_), _ptr(x), _ptr(y)):
            raise ValueError("bad W / groups")
        Y[:, j] = y
    return Y


def spmv_hybrid_host_coo(rows, width, ell_cols, ell_data, coo_rows, coo_cols, coo_data, x):
    """gpu/gpu_matrix_multiplier.cu:1111-1217: y = ELL part (slot order) +
    host COO tail (entry order), one rounding for the final add."""
    return spmv_ell(rows, width, ell_cols, ell_data, x) + spmv_coo_seq(rows, coo_rows, coo_cols, coo_data, x)


def spmv_adaptive(row_ptr, columns, data, x, nnz_per_wg, block_threads=256,
                  long_row=0xFFFFFFFF, chunk=4096):
    """Summation order of the library's CSR-Adaptive kernel (see header)."""
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    _lib.orc_spmv_csr_adaptive(_dt(data), y.size, row_ptr, columns, _ptr(data), _ptr(x), _ptr(y),
                               nnz_per_wg, block_threads, long_row, chunk)
    return y


def spmv_f64(row_ptr, columns, data, x):
    """FP64 truth and per-row sum |a*x| -> (y64, absdot)."""
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    rows = row_ptr.size - 1
    y = np.zeros(rows, np.float64)
    ad = np.zeros(rows, np.float64)
    _lib.orc_spmv_csr_f64(_dt(data), rows, row_ptr, columns, _ptr(data), _ptr(x), y, _ptr(ad))
    return y, ad


def spmv_multi_thread(row_ptr, columns, data, x, threads):
    """cpu/cpu_matrix_multiplier.cpp:229-308 -> (y, max thread seconds)."""
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    t = _lib.orc_spmv_csr_multi_thread(_dt(data), threads, y.size, row_ptr, columns,
                                       _ptr(data), _ptr(x), _ptr(y))
    return y, t


def spmv_seq_timed(row_ptr, columns, data, x, reps=1):
    row_ptr, columns, data, x = _spmv_args(row_ptr, columns, data, x)
    y = np.zeros(row_ptr.size - 1, data.dtype)
    t = _lib.orc_spmv_csr_seq_timed(_dt(data), reps, y.size, row_ptr, columns,
                                    _ptr(data), _ptr(x), _ptr(y))
    return y, t


# ---- measurement (common/measurement_class.{h,cpp}) -----------------------------

def compare_error(a, b):
    """measurement_class.h:60-72 error value, accumulated in the value type."""
    a, b = _c(a), _c(b)
    b = _c(b, a.dtype)
    return float(_lib.orc_compare_error(_dt(a), a.size, _ptr(a), _ptr(b)))


def measurement(elapsed, bytes_, ops):
    out = np.zeros(2, np.float64)
    _lib.orc_measurement(elapsed, bytes_, ops, out)
    return float(out[0]), float(out[1])


REF_BYTES_CSR, REF_BYTES_ELL, REF_BYTES_COO = 0, 1, 2


def reference_bytes(kind, dtype, rows, nnz, ell_width=0):
    d = F32 if np.dtype(dtype) == np.float32 else F64
    return float(_lib.orc_reference_bytes(kind, d, rows, nnz, ell_width))


# ---- optional: the reference's own measurement_class, built from its sources -------

def ref_library():
    """CDLL over oracle/_ref (None when it was not built)."""
    if not os.path.exists(_REF_PATH):
        return None
    lib = C.CDLL(_REF_PATH)
    lib.ref_measurement.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, _f64p]
    lib.ref_measurement.restype = None
    lib.ref_measurement_mean.argtypes = [C.c_uint, _f64p, _f64p, _f64p, _f64p]
    lib.ref_measurement_mean.restype = None
    lib.ref_compare_results.argtypes = [C.c_int, C.c_uint, _vp, _vp, C.c_char_p, C.c_size_t]
    lib.ref_compare_results.restype = C.c_size_t
    return lib


def tolerance_check(y, y_ref, absdot, rtol):
    """Per-row check |y - y_ref| <= rtol * sum_j |a_ij x_j| (plus a floor of
    rtol * tiny for empty rows).  Returns the worst normalised error."""
    y = np.asarray(y, np.float64)
    y_ref = np.asarray(y_ref, np.float64)
    scale = np.maximum(np.asarray(absdot, np.float64), np.finfo(np.float64).tiny)
    err = np.abs(y - y_ref) / scale
    return float(err.max()) if err.size else 0.0
