"""FP64 GPU oracle, residual report, and 16-bit value storage with
Dot2-compensated FP32 accumulation."""
import numpy as np
import pytest

import oracle
from gpu_cases import announce, case

pytestmark = pytest.mark.gpu


def bf16_round(f32):
    """FP32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    b = np.ascontiguousarray(f32, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def _stored_f32(dm, vt):
    """The stored 16-bit values as float32 (exact)."""
    from matrix_format_performance_amd import engine as E
    _, _, va = dm.download()
    return bf16_to_f32(va) if vt == "bf16" else va.view(np.float16).astype(np.float32)


@pytest.mark.parametrize("name", ["stencil3d_20", "powerlaw_20k", "edge_rows"])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_fp64_oracle_matches_cpu(gpu, name, dtype):
    rp, ci, va, cols = case(name, dtype)
    announce(f"fp64 oracle case={name} dtype={dtype}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    x = gpu.generate_vector(3, cols, np.dtype(dtype))
    xd = gpu.DeviceArray.from_host(x)
    y64, ad = A.oracle_f64(xd)
    gpu.synchronize()
    ry, rad = oracle.spmv_f64(rp, ci, va, x)
    scale = np.maximum(rad, 1e-300)
    assert np.max(np.abs(y64.to_host() - ry) / scale) < 1e-14
    assert np.allclose(ad.to_host(), rad, rtol=1e-14, atol=0)
    # residual of the CSR-scalar result against it
    yd = gpu.DeviceArray(rp.size - 1, np.dtype(dtype))
    gpu.Plan(A, "csr_scalar").spmv(xd, yd)
    rep = gpu.engine.residual(yd, y64, ad)
    y = yd.to_host().astype(np.float64)
    g64, gad = y64.to_host(), np.maximum(ad.to_host(), 1e-300)
    assert rep["rel_l2"] == pytest.approx(np.linalg.norm(y - g64) / np.linalg.norm(g64), rel=1e-6, abs=1e-30)
    assert rep["max_row_err"] == pytest.approx(np.max(np.abs(y - g64) / gad), rel=1e-6, abs=1e-30)
    assert rep["max_row_err"] <= (1e-5 if dtype == "float32" else 1e-12)


@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_value_conversion_is_round_to_nearest_even(gpu, vt):
    rp, ci, va, cols = case("stencil3d_20", "float32")
    announce(f"convert values {vt}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    B = A.convert(vt)
    _, _, bits = B.download()
    expect = bf16_round(va) if vt == "bf16" else va.astype(np.float16).view(np.uint16)
    assert np.array_equal(bits, expect)


@pytest.mark.parametrize("vt", ["bf16", "f16"])
@pytest.mark.parametrize("fmt,kw", [("csr_vector", {"lanes": 4}), ("csr_vector", {"lanes": 32}),
                                    ("ell", {}), ("ell", {"flags": 1}), ("csr_adaptive", {}),
                                    ("scoo", {}), ("scoo", {"flags": 4, "scoo_slice_size": 1024, "scoo_pieces": 3})])
@pytest.mark.parametrize("name", ["stencil3d_20", "uniform_band", "edge_rows", "stencil2d_64", "powerlaw_20k"])
def test_lp_spmv_compensated_accuracy(gpu, name, vt, fmt, kw):
    if fmt == "ell" and name == "powerlaw_20k":
        # a 12,000-wide ELL of 20,000 rows is legal but slow to check: the
        # same generator with rows capped at 300 entries stands in for it
        name = "powerlaw_narrow"
    rp, ci, va, cols = case(name, "float32")
    announce(f"lp spmv {vt} {fmt} {kw} case={name}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols).convert(vt)
    x = gpu.generate_vector(4, cols, np.float32)
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rp.size - 1, np.float32)
    yd.fill(np.nan)
    plan = gpu.Plan(A, fmt, **kw)
    info = plan.info()
    if fmt == "ell":
        assert info["column_bytes"] == (4 if kw.get("flags") else 2)
    plan.spmv(xd, yd)
    gpu.synchronize()
    y = yd.to_host()
    y64, ad = oracle.spmv_f64(rp, ci, _stored_f32(A, vt), x)
    err = oracle.tolerance_check(y, y64, ad, 1.0)
    if fmt == "scoo":
        # exact products added in double; each of a row's n partial slots is
        # rounded to FP32 once and the slots are added in FP32: at most
        # 2n - 1 roundings of 2^-24 relative to the row's sum of |a x|
        assert info["scoo_colsort"] == 1
        assert err <= 1e-6, err
    else:
        # Dot2: error of the stored-value product close to one FP32 rounding
        assert err <= 2.5e-7, err


def test_lp_compensation_survives_cancellation(gpu):
    # rows [2^24, 1, -2^24] . [1, 1, 1] = 1 exactly; a plain FP32 sum gives 0
    rows = 1000
    rp = np.arange(0, 3 * rows + 1, 3, dtype=np.uint32)
    ci = np.tile(np.array([0, 1, 2], np.uint32), rows)
    va = np.tile(np.array([2.0 ** 24, 1.0, -(2.0 ** 24)], np.float32), rows)
    announce("lp cancellation")
    x = np.ones(3, np.float32)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, 3).convert("bf16")
    xd = gpu.DeviceArray.from_host(x)
    for fmt in ("csr_vector", "ell", "csr_adaptive", "scoo"):
        yd = gpu.DeviceArray(rows, np.float32)
        gpu.Plan(A, fmt, **({"lanes": 2} if fmt == "csr_vector" else {})).spmv(xd, yd)
        gpu.synchronize()
        assert np.all(yd.to_host() == 1.0), fmt
    assert oracle.spmv_seq(rp, ci, va, x)[0] == 0.0  # the uncompensated order loses it


@pytest.mark.parametrize("vt,bound", [("bf16", 4e-3), ("f16", 6e-4)])
def test_lp_total_error_vs_fp64_of_original_values(gpu, vt, bound):
    # N(0,1) stencil values quantised to 16 bits: the error against FP64 of the
    # unquantised values is the quantisation (~2^-9/sqrt(3) bf16, 2^-11/sqrt(3) f16)
    rp, ci, va, cols = case("stencil3d_20", "float64")
    announce(f"lp total error {vt}")
    A64 = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    x = gpu.generate_vector(6, cols, np.float32)
    xd = gpu.DeviceArray.from_host(x)
    y_true, ad = A64.oracle_f64(xd)
    B = A64.convert(vt)
    yd = gpu.DeviceArray(rp.size - 1, np.float32)
    gpu.Plan(B, "ell").spmv(xd, yd)
    rep = gpu.engine.residual(yd, y_true, ad)
    assert rep["rel_l2"] < bound
    # the arithmetic part alone (against FP64 of the stored values) is ~1e-7
    y_st, ad_st = B.oracle_f64(xd)
    assert gpu.engine.residual(yd, y_st, ad_st)["max_row_err"] < 2.5e-7


def test_lp_formats_refused_elsewhere(gpu):
    rp, ci, va, cols = case("stencil2d_64", "float32")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols).convert("bf16")
    with pytest.raises(gpu.MfpError):
        gpu.Plan(A, "coo")
