"""The allocation-free SpMV (mfp_plan_spmv_noalloc) against mfp_plan_spmv.

mfp_plan_spmv_noalloc keeps a plan's partial results (column-sorted SCOO
partial slots, CSR-Adaptive long-row chunk sums, chunk sums of the binned
16-bit CSR) in the library's static device scratch.  Each case runs two calls
in a row through the C ABI, the second on a smaller matrix, on plans created
with MFP_TUNE_STATIC_PARTIALS (no partial buffer of their own), and compares
with mfp_plan_spmv on ordinary plans of the same matrices: bit for bit where
the kernels fix the summation order (CSR-Adaptive, the binned 16-bit CSR,
formats without partial results), within the per-row tolerance of the
column-sorted SCOO (1e-5 FP32, 1e-12 FP64 of sum |a x|; its LDS additions
come in varying order).
"""
import numpy as np
import pytest

import oracle
from gpu_cases import announce

pytestmark = pytest.mark.gpu

COLSORT_SMALL = {"scoo_slice_size": 4096, "scoo_threads": 1024}


def _matrix(mfp, rows, hi, seed, dtype):
    rp, ci, va = mfp.generate(mfp.GenSpec("powerlaw", values="uniform", rows=rows, nnz_lo=1, nnz_hi=hi,
                                          seed=seed), dtype)
    return rp, ci, va


def _pair(mfp, rp, ci, va, fmt, dtype, value_type=None, **tuning):
    """y from mfp_plan_spmv on an ordinary plan and from mfp_plan_spmv_noalloc
    on a plan created with MFP_TUNE_STATIC_PARTIALS."""
    rows = rp.size - 1
    x = mfp.generate_vector(11, rows, np.dtype(dtype) if value_type is None else np.float32)
    A = mfp.DeviceMatrix.from_csr(rp, ci, va)
    B = A.convert(value_type) if value_type else A
    flags = tuning.pop("flags", 0)
    p_ref = mfp.Plan(B, fmt, flags=flags, **tuning)
    p_na = mfp.Plan(B, fmt, flags=flags | mfp._lib.MFP_TUNE_STATIC_PARTIALS, **tuning)
    xd = mfp.DeviceArray.from_host(x)
    y_ref = mfp.DeviceArray(rows, x.dtype)
    y_na = mfp.DeviceArray(rows, x.dtype)
    y_ref.fill(np.nan)
    y_na.fill(np.nan)
    p_ref.spmv(xd, y_ref)
    p_na.spmv_noalloc(xd, y_na)
    mfp.synchronize()
    out = (y_ref.to_host(), y_na.to_host(), x, p_na)
    p_ref.destroy()
    if B is not A:
        B.destroy()
    return out, A


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_colsort_scoo_two_calls(gpu, dtype):
    mfp = gpu
    for rows, hi, seed in ((200000, 20000, 5), (30000, 3000, 6)):
        rp, ci, va = _matrix(mfp, rows, hi, seed, np.dtype(dtype))
        announce(f"noalloc colsort dtype={dtype} rows={rows} nnz={ci.size}")
        (y_ref, y_na, x, p), A = _pair(mfp, rp, ci, va, "scoo", dtype, flags=mfp._lib.MFP_TUNE_SCOO_COLSORT,
                                       **COLSORT_SMALL)
        info = p.info()
        assert info["scoo_colsort"] == 1 and info["lane_size"] > info["slices_count"], info  # partial slots in use
        _, absdot = oracle.spmv_f64(rp, ci, va, x)
        tol = 1e-5 if dtype == "float32" else 1e-12
        assert not np.isnan(y_na).any()
        assert oracle.tolerance_check(y_na, y_ref, absdot, tol) <= tol
        # a plan without partial buffer refuses the allocating entry point
        yd = mfp.DeviceArray(rows, np.dtype(dtype))
        with pytest.raises(mfp.MfpError) as e:
            p.spmv(mfp.DeviceArray.from_host(x), yd)
        assert e.value.status == mfp._lib.MFP_ERR_STATE
        p.destroy()
        A.destroy()


def test_csr_adaptive_chunks_two_calls(gpu):
    mfp = gpu
    for rows, hi, seed in ((120000, 30000, 7), (20000, 9000, 8)):
        rp, ci, va = _matrix(mfp, rows, hi, seed, np.float32)
        announce(f"noalloc csr_adaptive rows={rows} nnz={ci.size}")
        (y_ref, y_na, x, p), A = _pair(mfp, rp, ci, va, "csr_adaptive", "float32")
        assert p.info()["chunks"] > 0
        assert np.array_equal(y_na, y_ref)
        p.destroy()
        A.destroy()


def test_binned_16bit_two_calls(gpu):
    mfp = gpu
    for rows, hi, seed in ((100000, 30000, 9), (15000, 8000, 10)):
        rp, ci, va = _matrix(mfp, rows, hi, seed, np.float32)
        announce(f"noalloc bf16 csr_adaptive rows={rows} nnz={ci.size}")
        (y_ref, y_na, x, p), A = _pair(mfp, rp, ci, va, "csr_adaptive", "float32", value_type="bf16")
        assert np.array_equal(y_na, y_ref)
        p.destroy()
        A.destroy()


def test_formats_without_partials_and_the_limit(gpu):
    mfp = gpu
    rp, ci, va = mfp.generate(mfp.GenSpec("stencil3d", values="normal", n=24, seed=3), np.float32)
    announce(f"noalloc ell stencil3d_24 rows={rp.size - 1}")
    (y_ref, y_na, x, p), A = _pair(mfp, rp, ci, va, "ell", "float32")
    assert np.array_equal(y_na, y_ref)
    p.destroy()
    A.destroy()
    # partial slots beyond the static scratch: refused before any launch
    rp, ci, va = _matrix(mfp, 2 << 20, 50000, 12, np.float32)
    announce(f"noalloc colsort over the scratch limit rows={rp.size - 1} nnz={ci.size}")
    A = mfp.DeviceMatrix.from_csr(rp, ci, va)
    p = mfp.Plan(A, "scoo", flags=mfp._lib.MFP_TUNE_SCOO_COLSORT | mfp._lib.MFP_TUNE_STATIC_PARTIALS,
                 scoo_slice_size=16384, scoo_threads=1024)
    info = p.info()
    need = info["lane_size"] * 16384 * 4  # upper bound: every item a partial slot
    assert need > mfp._lib.MFP_NOALLOC_SCRATCH_BYTES
    xd = mfp.DeviceArray.from_host(mfp.generate_vector(1, rp.size - 1, np.float32))
    yd = mfp.DeviceArray(rp.size - 1, np.float32)
    with pytest.raises(mfp.MfpError) as e:
        p.spmv_noalloc(xd, yd)
    assert e.value.status == mfp._lib.MFP_ERR_BUDGET
    assert "static scratch" in str(e.value)
    p.destroy()
    A.destroy()
