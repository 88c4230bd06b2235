"""The benchmark-only rocSPARSE comparator slot (the counterpart of the
reference's "GPU CSR (cuSparse)" row, gpu_matrix_multiplier.cu:175-257): its
y against the oracle within the per-row tolerance.  Skipped when
librocsparse.so cannot be opened; the library's own SpMV never uses it."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from gpu_cases import announce, case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def slot(gpu):
    lib = C.CDLL(os.path.join(HERE, "matrix_format_performance_amd", "librocsparse_slot.so"))
    lib.mfp_rs_error.restype = C.c_char_p
    if not lib.mfp_rs_available():
        pytest.skip("librocsparse.so not available: " + lib.mfp_rs_error().decode())
    lib.mfp_rs_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mfp_rs_plan_spmv.argtypes = [C.c_void_p, C.c_void_p]
    lib.mfp_rs_plan_destroy.argtypes = [C.c_void_p]
    return lib


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("name", ["stencil2d_64", "powerlaw_20k", "edge_rows"])
def test_rocsparse_slot_matches_the_oracle(gpu, slot, name, dtype):
    rp, ci, va, cols = case(name, dtype)
    rows = rp.size - 1
    announce(f"rocsparse slot case={name} dtype={dtype}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    x = gpu.generate_vector(7, cols, np.dtype(dtype))
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rows, np.dtype(dtype))
    # rocsparse_spmv scales y by beta = 0 on some paths (long rows) before
    # adding, and 0 * NaN is NaN: y starts at zero here, not at NaN
    yd.fill(0.0)
    p = C.c_void_p()
    dt = gpu._lib.MFP_F32 if dtype == "float32" else gpu._lib.MFP_F64
    assert slot.mfp_rs_plan_create(C.byref(p), dt, rows, cols, ci.size, A.row_ptr_ptr, A.columns_ptr, A.values_ptr,
                                   xd.ptr, yd.ptr, None) == 0, slot.mfp_rs_error()
    assert slot.mfp_rs_plan_spmv(p, None) == 0, slot.mfp_rs_error()
    gpu.synchronize()
    slot.mfp_rs_plan_destroy(p)
    y = yd.to_host()
    _, absdot = oracle.spmv_f64(rp, ci, va, x)
    rtol = 1e-5 if dtype == "float32" else 1e-12
    assert oracle.tolerance_check(y, oracle.spmv_seq(rp, ci, va, x), absdot, rtol) <= rtol
    A.destroy()
