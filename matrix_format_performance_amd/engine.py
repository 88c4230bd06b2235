"""Device-resident SpMV engine: arrays, matrices, per-format plans, timing.

Thin Python handles over the C ABI (include/mfp.h).  Every computation runs in
libmfp's HIP kernels; numpy arrays are only host staging.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

FORMATS = {
    "csr_scalar": L.MFP_CSR_SCALAR,
    "csr_vector": L.MFP_CSR_VECTOR,
    "csr_adaptive": L.MFP_CSR_ADAPTIVE,
    "ell": L.MFP_ELL,
    "coo": L.MFP_COO,
    "scoo": L.MFP_SCOO,
    "hyb": L.MFP_HYB,
    "hyb_atomic": L.MFP_HYB_ATOMIC,
}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}

# the reference's JSON labels (main.cpp:179-245 via each entry point's record)
LABELS = {
    "csr_scalar": "GPU CSR",
    "csr_vector": "GPU CSR (vector)",
    "csr_adaptive": "GPU CSR-Adaptive",
    "ell": "GPU ELL",
    "coo": "GPU COO",
    "scoo": "GPU SCOO",
    "hyb": "GPU Hybrid (TODO)",
    "hyb_atomic": "GPU Hybrid (atomic)",
}

_DT = {np.dtype(np.float32): L.MFP_F32, np.dtype(np.float64): L.MFP_F64}

# CSR-Adaptive defaults of libmfp (kernels.h): row-block size per value type,
# long-row threshold and split chunk.
DEFAULT_NNZ_PER_WG = {np.dtype(np.float32): 2040, np.dtype(np.float64): 1020}
DEFAULT_LONG_ROW, DEFAULT_CHUNK = 4096, 2048


def dtype_code(dtype) -> int:
    d = np.dtype(dtype)
    if d not in _DT:
        raise TypeError(f"unsupported value dtype {d}")
    return _DT[d]


def lib():
    return L.load()


def init(device: int = 0) -> None:
    L.call("mfp_init", device)


def device_info() -> dict:
    info = L.device_info()
    L.call("mfp_device_query", C.byref(info))
    return {
        "device": info.device, "cu_count": info.cu_count, "xcd_count": info.xcd_count,
        "lds_per_block": info.lds_per_block, "free_mem": info.free_mem,
        "total_mem": info.total_mem, "name": info.name.decode(), "arch": info.arch.decode(),
    }


def synchronize() -> None:
    L.call("mfp_synchronize")


def flush_cache(nbytes: int = 1 << 30, stream=None) -> None:
    L.call("mfp_flush_cache", nbytes, stream)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device allocation holding n elements of dtype."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        L.call("mfp_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value
        self._owned = True

    @classmethod
    def from_host(cls, arr) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype)
        if arr.size:
            L.call("mfp_memcpy_h2d", d.ptr, _ptr(arr), arr.nbytes)
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            L.call("mfp_memcpy_d2h", _ptr(out), self.ptr, self.nbytes)
        return out

    def fill(self, value: float, stream=None) -> None:
        L.call("mfp_fill", dtype_code(self.dtype), self.ptr, self.n, float(value), stream)

    def free(self) -> None:
        if self._owned and self.ptr:
            L.load().mfp_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(ptr: int, n: int, dtype) -> np.ndarray:
    out = np.empty(int(n), np.dtype(dtype))
    if n:
        L.call("mfp_memcpy_d2h", _ptr(out), ptr, out.nbytes)
    return out


VALUE_TYPES = {"f32": L.MFP_F32, "f64": L.MFP_F64, "bf16": L.MFP_BF16, "f16": L.MFP_F16}
_VT_NAME = {v: k for k, v in VALUE_TYPES.items()}


class DeviceMatrix:
    """CSR matrix in HBM (u32 row_ptr/columns; f32, f64, bf16 or f16 values).

    ``dtype`` is the numpy type of the values (uint16 bit patterns for bf16 /
    f16) and ``value_type`` its name."""

    def __init__(self, handle, dtype=None):
        self.handle = handle
        v = L.matrix_view()
        L.call("mfp_matrix_view_get", handle, C.byref(v))
        self.value_type = _VT_NAME[v.dtype]
        self.dtype = {"f32": np.dtype(np.float32), "f64": np.dtype(np.float64)}.get(
            self.value_type, np.dtype(np.uint16))
        self.rows, self.cols, self.nnz = v.rows, v.cols, v.nnz
        self.row_ptr_ptr, self.columns_ptr, self.values_ptr = v.row_ptr, v.columns, v.values

    @property
    def vector_dtype(self):
        """x / y type: the value type, FP32 for 16-bit values."""
        return np.dtype(np.float64) if self.value_type == "f64" else np.dtype(np.float32)

    def convert(self, value_type: str) -> "DeviceMatrix":
        """Copy with values converted (16-bit types rounded from FP32)."""
        h = C.c_void_p()
        L.call("mfp_matrix_convert", self.handle, VALUE_TYPES[value_type], C.byref(h))
        return DeviceMatrix(h)

    def oracle_f64(self, x: "DeviceArray", stream=None):
        """FP64 SpMV on the GPU -> (y64, absdot) device arrays."""
        y64 = DeviceArray(self.rows, np.float64)
        ad = DeviceArray(self.rows, np.float64)
        L.call("mfp_spmv_f64_oracle", self.handle, dtype_code(x.dtype), x.ptr, y64.ptr, ad.ptr, stream)
        return y64, ad

    @classmethod
    def from_csr(cls, row_ptr, columns, values, cols=None) -> "DeviceMatrix":
        row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        columns = np.ascontiguousarray(columns, np.uint32)
        values = np.ascontiguousarray(values)
        rows = row_ptr.size - 1
        cols = rows if cols is None else cols
        h = C.c_void_p()
        L.call("mfp_matrix_from_host_csr", C.byref(h), dtype_code(values.dtype), rows, cols,
               columns.size, _ptr(row_ptr), _ptr(columns), _ptr(values))
        return cls(h, values.dtype)

    @classmethod
    def from_triplets(cls, rows, cols, tr_rows, tr_cols, tr_vals, dtype=np.float32) -> "DeviceMatrix":
        """Device converter (triplets -> CSR), matrix_converter.cpp:15-87."""
        tr_rows = np.ascontiguousarray(tr_rows, np.uint32)
        tr_cols = np.ascontiguousarray(tr_cols, np.uint32)
        tr_vals = np.ascontiguousarray(tr_vals, np.float64)
        h = C.c_void_p()
        L.call("mfp_matrix_from_triplets", C.byref(h), dtype_code(dtype), rows, cols, tr_rows.size,
               _ptr(tr_rows), _ptr(tr_cols), _ptr(tr_vals), 0, None)
        return cls(h, dtype)

    def download(self):
        rp = np.empty(self.rows + 1, np.uint32)
        ci = np.empty(self.nnz, np.uint32)
        va = np.empty(self.nnz, self.dtype)
        L.call("mfp_matrix_download", self.handle, _ptr(rp), _ptr(ci) if self.nnz else None,
               _ptr(va) if self.nnz else None)
        return rp, ci, va

    def row_stats(self):
        s = L.row_stats()
        L.call("mfp_matrix_row_stats", self.handle, C.byref(s))
        return (s.min_elements_in_rows, s.max_elements_in_rows, s.avg_elements_in_rows,
                s.elements_in_rows_std_deviation)

    def destroy(self):
        if self.handle:
            L.load().mfp_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def residual(y: "DeviceArray", y64: "DeviceArray", absdot: "DeviceArray" = None) -> dict:
    """Error of a device result against the FP64 oracle (computed on the GPU)."""
    r = L.error_report()
    L.call("mfp_residual", y.n, dtype_code(y.dtype), y.ptr, y64.ptr, absdot.ptr if absdot else None,
           C.byref(r))
    return {"rel_l2": r.rel_l2, "max_row_err": r.max_row_err, "max_abs_err": r.max_abs_err}


def default_tuning(matrix: DeviceMatrix, fmt: str) -> L.tuning:
    t = L.tuning()
    L.call("mfp_tuning_default", matrix.handle, FORMATS[fmt], C.byref(t))
    return t


class Plan:
    """One format of a DeviceMatrix, built on the device."""

    def __init__(self, matrix: DeviceMatrix, fmt: str, stream=None, **overrides):
        self.matrix = matrix
        self.fmt = fmt
        self.tuning = default_tuning(matrix, fmt)
        for k, v in overrides.items():
            if not hasattr(self.tuning, k):
                raise KeyError(f"unknown tuning field {k}")
            setattr(self.tuning, k, v)
        h = C.c_void_p()
        L.call("mfp_plan_create", C.byref(h), matrix.handle, FORMATS[fmt], C.byref(self.tuning), stream)
        self.handle = h

    def spmv(self, x: DeviceArray, y: DeviceArray, stream=None) -> None:
        L.call("mfp_plan_spmv", self.handle, x.ptr, y.ptr, stream)

    def graph(self, x: DeviceArray, y: DeviceArray) -> "Graph":
        """mfp_plan_graph_create: one SpMV of this plan captured in a HIP graph
        (replay with Graph.launch; the plan, x and y must outlive it)."""
        h = C.c_void_p()
        L.call("mfp_plan_graph_create", self.handle, x.ptr, y.ptr, C.byref(h))
        return Graph(h)

    def spmv_noalloc(self, x: DeviceArray, y: DeviceArray, stream=None) -> None:
        """mfp_plan_spmv_noalloc: partial results in the library's static
        scratch (no allocation; overlapping calls on one device race)."""
        L.call("mfp_plan_spmv_noalloc", self.handle, x.ptr, y.ptr, stream)

    def kernel_times(self, x: DeviceArray, y: DeviceArray, stream=None) -> list:
        """One SpMV with an event after each kernel it launches; returns the
        per-kernel times in ms (waits for the SpMV)."""
        ms = (C.c_float * 8)()
        n = C.c_int(0)
        L.call("mfp_plan_kernel_times", self.handle, x.ptr, y.ptr, stream, ms, 8, C.byref(n))
        return [float(ms[i]) for i in range(min(n.value, 8))]

    def spmm(self, k: int, X: DeviceArray, Y: DeviceArray, stream=None) -> None:
        """Y (rows x k) = A X (cols x k), row-major."""
        L.call("mfp_plan_spmm", self.handle, k, X.ptr, Y.ptr, stream)

    def info(self) -> dict:
        i = L.plan_info()
        L.call("mfp_plan_get_info", self.handle, C.byref(i))
        return {f: getattr(i, f) for f, _ in L.plan_info._fields_}

    def array(self, which: int, dtype) -> np.ndarray:
        p = C.c_void_p()
        n = C.c_uint64()
        L.call("mfp_plan_array", self.handle, which, C.byref(p), C.byref(n))
        return download(p.value, n.value, dtype) if p.value else np.empty(0, dtype)

    def destroy(self):
        if self.handle:
            L.load().mfp_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _tuning_dict(t: L.tuning) -> dict:
    return {f: getattr(t, f) for f, _ in L.tuning._fields_}


def autotune_candidates(matrix: DeviceMatrix, formats=None):
    """The (format, tuning) candidates the autotuner would time, untimed."""
    names = list(FORMATS) if formats is None else list(formats)
    mask = 0
    for f in names:
        mask |= 1 << FORMATS[f]
    n = C.c_uint32()
    L.call("mfp_autotune_candidates", matrix.handle, mask, None, 0, C.byref(n))
    arr = (L.candidate * n.value)()
    L.call("mfp_autotune_candidates", matrix.handle, mask, arr, n.value, C.byref(n))
    return [{"format": FORMAT_NAMES[c.format], "tuning": _tuning_dict(c.tuning)} for c in arr]


def autotune(matrix: DeviceMatrix, x: DeviceArray, y: DeviceArray, formats=None, reps: int = 3,
             flush: bool = True):
    """Time every candidate (format, tuning) the library sweeps for this
    matrix; returns a list of {format, tuning, ms, status} sorted by time."""
    names = list(FORMATS) if formats is None else list(formats)
    mask = 0
    for f in names:
        mask |= 1 << FORMATS[f]
    n = C.c_uint32()
    L.call("mfp_autotune_candidates", matrix.handle, mask, None, 0, C.byref(n))
    arr = (L.candidate * n.value)()
    L.call("mfp_autotune_candidates", matrix.handle, mask, arr, n.value, C.byref(n))
    L.call("mfp_autotune_run", matrix.handle, x.ptr, y.ptr, arr, n.value, reps, int(flush))
    out = [{"format": FORMAT_NAMES[c.format], "tuning": _tuning_dict(c.tuning), "ms": c.ms,
            "status": c.status} for c in arr]
    return sorted(out, key=lambda c: c["ms"])


class Graph:
    """An instantiated HIP graph (a batched launch or one plan's SpMV)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream=None) -> None:
        L.call("mfp_graph_launch", self.handle, stream)

    def destroy(self):
        if self.handle:
            L.load().mfp_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Many matrices assembled into one block-diagonal CSR, multiplied by one
    launch (CSR-Adaptive for k = 1, or the column-sorted SCOO when the
    tuning's flags hold MFP_TUNE_SCOO_COLSORT; for k > 1 CSR-vector SpMM with
    ``lanes`` lanes per row, lanes = 0: lanes per row from its length, lanes =
    1: the column-sorted SCOO SpMM)."""

    def __init__(self, matrices, lanes: int = 8, stream=None, **k1_tuning):
        arr = (C.c_void_p * len(matrices))(*[m.handle for m in matrices])
        t = None
        if k1_tuning:
            t = default_tuning(matrices[0], "csr_adaptive")
            for k, v in k1_tuning.items():
                setattr(t, k, v)
        h = C.c_void_p()
        L.call("mfp_batch_create", C.byref(h), arr, len(matrices), lanes, C.byref(t) if t else None, stream)
        self.handle = h
        self.count = len(matrices)
        r, c, n = C.c_uint32(), C.c_uint32(), C.c_uint64()
        self.row_offsets = np.zeros(self.count + 1, np.uint32)
        self.col_offsets = np.zeros(self.count + 1, np.uint32)
        L.call("mfp_batch_info", h, C.byref(r), C.byref(c), C.byref(n), _ptr(self.row_offsets),
               _ptr(self.col_offsets))
        self.rows, self.cols, self.nnz = r.value, c.value, n.value
        self.dtype = matrices[0].dtype

    def matrix_arrays(self):
        """Download the assembled block-diagonal CSR (row_ptr, columns, values)."""
        m = C.c_void_p()
        L.call("mfp_batch_matrix", self.handle, C.byref(m))
        rp = np.empty(self.rows + 1, np.uint32)
        ci = np.empty(self.nnz, np.uint32)
        va = np.empty(self.nnz, self.dtype)
        L.call("mfp_matrix_download", m, _ptr(rp), _ptr(ci), _ptr(va))
        return rp, ci, va

    def spmm(self, k: int, X: DeviceArray, Y: DeviceArray, stream=None) -> None:
        L.call("mfp_batch_spmm", self.handle, k, X.ptr, Y.ptr, stream)

    def graph(self, k: int, X: DeviceArray, Y: DeviceArray) -> Graph:
        g = C.c_void_p()
        L.call("mfp_batch_graph_create", self.handle, k, X.ptr, Y.ptr, C.byref(g))
        return Graph(g)

    def destroy(self):
        if self.handle:
            L.load().mfp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Timer:
    """hipEvent pair on a stream."""

    def __init__(self):
        h = C.c_void_p()
        L.call("mfp_timer_create", C.byref(h))
        self.h = h

    def start(self, stream=None):
        L.call("mfp_timer_start", self.h, stream)

    def stop(self, stream=None):
        L.call("mfp_timer_stop", self.h, stream)

    def ms(self) -> float:
        v = C.c_float()
        L.call("mfp_timer_elapsed_ms", self.h, C.byref(v))
        return float(v.value)

    def __del__(self):
        try:
            L.load().mfp_timer_destroy(self.h)
        except Exception:
            pass


# ---- synthetic workloads ----------------------------------------------------

@dataclass
class GenSpec:
    family: str            # stencil2d | stencil3d | powerlaw | uniform
    values: str = "uniform"  # laplace | uniform | normal
    seed: int = 0
    n: int = 0
    rows: int = 0
    nnz_lo: int = 1
    nnz_hi: int = 1
    alpha: float = 2.0
    bandwidth: float = 1.0

    def params(self) -> L.gen_params:
        fam = {"stencil2d": L.MFP_GEN_STENCIL2D, "stencil3d": L.MFP_GEN_STENCIL3D,
               "powerlaw": L.MFP_GEN_POWERLAW, "uniform": L.MFP_GEN_UNIFORM}[self.family]
        val = {"laplace": L.MFP_VAL_LAPLACE, "uniform": L.MFP_VAL_UNIFORM,
               "normal": L.MFP_VAL_NORMAL}[self.values]
        return L.gen_params(fam, val, self.seed, self.n, self.rows, self.nnz_lo, self.nnz_hi,
                            self.alpha, self.bandwidth)


def generate(spec: GenSpec, dtype=np.float32):
    """Host generator -> (row_ptr, columns, values), CSR with sorted columns."""
    p = spec.params()
    rows, cols, nnz = C.c_uint32(), C.c_uint32(), C.c_uint64()
    L.call("mfp_generate", C.byref(p), C.byref(rows), C.byref(cols), C.byref(nnz), None, None, None,
           dtype_code(dtype))
    rp = np.empty(rows.value + 1, np.uint32)
    ci = np.empty(max(nnz.value, 1), np.uint32)
    va = np.empty(max(nnz.value, 1), dtype)
    L.call("mfp_generate", C.byref(p), C.byref(rows), C.byref(cols), C.byref(nnz), _ptr(rp), _ptr(ci),
           _ptr(va), dtype_code(dtype))
    return rp, ci[: nnz.value], va[: nnz.value]


def generate_vector(seed: int, n: int, dtype=np.float32) -> np.ndarray:
    x = np.empty(max(n, 1), dtype)
    L.call("mfp_generate_vector", seed, n, _ptr(x), dtype_code(dtype))
    return x[:n]


# Benchmark workloads: 2D 5-point 256x256, power-law 1M rows, 3D 27-point 128^3
def batch_workload(count: int = 512, seed: int = 0) -> list:
    """The 512-matrix batch workload (host CSR triples, FP32): rows log-uniform
    in [1k, 200k], families in turn 2D 5-point and 3D 27-point stencils,
    uniform 4-64 nnz per row, and power-law rows (alpha U[1.5, 3])."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        rows = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        fam = ("stencil2d", "stencil3d", "uniform", "powerlaw")[i % 4]
        if fam == "stencil2d":
            spec = GenSpec("stencil2d", values="uniform", n=max(2, int(np.sqrt(rows))), seed=i)
        elif fam == "stencil3d":
            spec = GenSpec("stencil3d", values="normal", n=max(2, int(round(rows ** (1 / 3)))), seed=i)
        elif fam == "uniform":
            spec = GenSpec("uniform", values="uniform", rows=rows, nnz_lo=4, nnz_hi=64, seed=i)
        else:
            spec = GenSpec("powerlaw", values="uniform", rows=rows, nnz_lo=1, nnz_hi=min(rows // 2, 10000),
                           alpha=float(rng.uniform(1.5, 3.0)), seed=i)
        out.append(generate(spec, np.float32))
    return out


def workload(name: str, dtype=np.float32, seed: int = 0):
    if name == "stencil2d_256":
        return generate(GenSpec("stencil2d", values="laplace", n=256, seed=seed), dtype)
    if name == "powerlaw_1m_normal":
        # the precision study's power-law matrix: the same structure, values N(0, 1)
        return generate(GenSpec("powerlaw", values="normal", rows=1 << 20, nnz_lo=1, nnz_hi=50000,
                                alpha=2.0, seed=seed), dtype)
    if name == "powerlaw_1m":
        return generate(GenSpec("powerlaw", values="uniform", rows=1 << 20, nnz_lo=1, nnz_hi=50000,
                                alpha=2.0, seed=seed), dtype)
    if name == "stencil3d_128":
        return generate(GenSpec("stencil3d", values="normal", n=128, seed=seed), dtype)
    raise KeyError(name)
