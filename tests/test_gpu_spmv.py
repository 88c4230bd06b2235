"""GPU parity: every format's SpMV against the oracle.

Deterministic kernels (CSR-scalar, CSR-vector, CSR-Adaptive, ELL) must be
bit-exact against the oracle's emulator of their summation order.  Every
format must match the reference CPU answer (cpu_matrix_multiplier.cpp:20-61,
sequential) within 1e-5 (FP32) / 1e-12 (FP64) per row, relative to
sum_j |a_ij x_j|.
"""
import numpy as np
import pytest

import oracle
from gpu_cases import CASES, announce, case

pytestmark = pytest.mark.gpu

RTOL = {np.dtype(np.float32): 1e-5, np.dtype(np.float64): 1e-12}
DTYPES = ["float32", "float64"]


def _run(mfp, name, dtype, fmt, **tuning):
    rp, ci, va, cols = case(name, dtype)
    announce(f"spmv fmt={fmt} case={name} dtype={dtype} rows={rp.size - 1} nnz={ci.size} tuning={tuning}")
    x = mfp.generate_vector(7, cols, np.dtype(dtype))
    A = mfp.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = mfp.Plan(A, fmt, **tuning)
    xd = mfp.DeviceArray.from_host(x)
    yd = mfp.DeviceArray(rp.size - 1, np.dtype(dtype))
    yd.fill(np.nan)  # any row left unwritten shows as NaN
    plan.spmv(xd, yd)
    mfp.synchronize()
    y = yd.to_host()
    info = plan.info()
    plan.destroy()
    A.destroy()
    return y, (rp, ci, va, x), info


def _tol_check(y, rp, ci, va, x):
    y_ref = oracle.spmv_seq(rp, ci, va, x)
    y64, absdot = oracle.spmv_f64(rp, ci, va, x)
    assert not np.isnan(y).any(), "unwritten rows"
    rtol = RTOL[va.dtype]
    assert oracle.tolerance_check(y, y_ref, absdot, rtol) <= rtol
    assert oracle.tolerance_check(y, y64, absdot, rtol) <= rtol


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", CASES)
def test_csr_scalar_bit_exact(gpu, name, dtype):
    y, (rp, ci, va, x), _ = _run(gpu, name, dtype, "csr_scalar")
    assert np.array_equal(y, oracle.spmv_seq(rp, ci, va, x))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("lanes", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("name", ["stencil3d_20", "powerlaw_20k", "edge_rows"])
def test_csr_vector_bit_exact(gpu, name, dtype, lanes):
    y, (rp, ci, va, x), _ = _run(gpu, name, dtype, "csr_vector", lanes=lanes)
    assert np.array_equal(y, oracle.spmv_lanes(rp, ci, va, x, lanes))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", CASES)
def test_csr_adaptive_bit_exact_default(gpu, name, dtype):
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "csr_adaptive")
    E = gpu.engine
    assert info["nnz_per_wg"] == E.DEFAULT_NNZ_PER_WG[np.dtype(dtype)]
    ref = oracle.spmv_adaptive(rp, ci, va, x, info["nnz_per_wg"], 256, E.DEFAULT_LONG_ROW, E.DEFAULT_CHUNK)
    assert np.array_equal(y, ref)
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("W,long_row,chunk", [(64, 256, 256), (256, 512, 256), (4096, 8192, 4096), (2048, 1024, 512)])
@pytest.mark.parametrize("name", ["powerlaw_20k", "edge_rows", "uniform_band"])
def test_csr_adaptive_bit_exact_tunings(gpu, name, dtype, W, long_row, chunk):
    if long_row < chunk:
        pytest.skip("long_row must be >= chunk")
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "csr_adaptive", nnz_per_wg=W, long_row=long_row,
                                    chunk=chunk)
    blocks = oracle.fill_row_blocks(rp, W).astype(np.int64)
    lens = np.diff(rp.astype(np.int64))
    expect_chunks = sum(-(-int(lens[s]) // chunk) for s, e in zip(blocks[:-1], blocks[1:])
                        if e - s == 1 and lens[s] > long_row)
    assert info["blocks"] == blocks.size - 1 and info["chunks"] == expect_chunks
    assert np.array_equal(y, oracle.spmv_adaptive(rp, ci, va, x, W, 256, long_row, chunk))


def test_csr_adaptive_repeated_calls_reset_tickets(gpu):
    # the long-row ticket counters must be back at zero after each call
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    announce("adaptive repeated calls powerlaw_20k")
    x = gpu.generate_vector(5, cols, np.float32)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "csr_adaptive", nnz_per_wg=256, long_row=512, chunk=256)
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rp.size - 1, np.float32)
    ref = oracle.spmv_adaptive(rp, ci, va, x, 256, 256, 512, 256)
    for _ in range(5):
        yd.fill(np.nan)
        plan.spmv(xd, yd)
        gpu.synchronize()
        assert np.array_equal(yd.to_host(), ref)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", ["stencil2d_64", "stencil3d_20", "uniform_band", "edge_rows", "one_by_one"])
def test_ell_bit_exact(gpu, name, dtype):
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "ell")
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert info["ell_width"] == w
    assert np.array_equal(y, oracle.spmv_ell(rp.size - 1, w, ec, ed, x))


def _block_span(rp, ci):
    """Largest column span of a 64-row block (last column of its rows minus
    the smallest first column), or -1 for a matrix without entries."""
    span = -1
    for b0 in range(0, rp.size - 1, 64):
        r = np.arange(b0, min(b0 + 64, rp.size - 1))
        nz = r[rp[r + 1] > rp[r]]
        if nz.size:
            span = max(span, int(ci[rp[nz + 1] - 1].max()) - int(ci[rp[nz]].min()))
    return span


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", CASES)
def test_ell_packed_bit_exact(gpu, name, dtype):
    # MFP_TUNE_ELL_PACKED (FP32 / FP64): 64-row slices, 16-bit offsets, four
    # slots to a word; the same fma order as the plain ELL, so equal to the
    # oracle's ELL order bit for bit, or refused when a slice spans 65,536
    # columns or more
    rp, ci, va, cols = case(name, dtype)
    flags = gpu._lib.MFP_TUNE_ELL_PACKED
    if _block_span(rp, ci) >= 65536:
        A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
        with pytest.raises(gpu.MfpError, match="span"):
            gpu.Plan(A, "ell", flags=flags)
        A.destroy()
        return
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "ell", flags=flags)
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert info["ell_width"] == w and info["column_bytes"] == 2
    assert np.array_equal(y, oracle.spmv_ell(rp.size - 1, w, ec, ed, x))


def test_ell_packed_refusals(gpu):
    rp, ci, va, cols = case("stencil3d_20", "float32")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols).convert("bf16")
    with pytest.raises(gpu.MfpError, match="f32 or f64"):
        gpu.Plan(A, "ell", flags=gpu._lib.MFP_TUNE_ELL_PACKED)
    A.destroy()
    rp, ci, va, cols = case("stencil3d_20", "float32")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    with pytest.raises(gpu.MfpError, match="f32"):
        gpu.Plan(A, "hyb", flags=gpu._lib.MFP_TUNE_ELL_PACKED)
    plan = gpu.Plan(A, "ell", flags=gpu._lib.MFP_TUNE_ELL_PACKED)
    with pytest.raises(gpu.MfpError, match="own layout"):
        plan.array(gpu._lib.MFP_ARR_ELL_VALUES, np.float32)
    plan.destroy()
    A.destroy()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("fmt", ["coo", "scoo", "hyb", "hyb_atomic"])
@pytest.mark.parametrize("name", CASES)
def test_atomic_formats_within_tolerance(gpu, name, dtype, fmt):
    tuning = {}
    if fmt in ("hyb", "hyb_atomic") and name in ("powerlaw_20k", "edge_rows"):
        tuning = {"hyb_width": 8}  # the 20% rule would make the ELL part huge
    y, (rp, ci, va, x), _ = _run(gpu, name, dtype, fmt, **tuning)
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("ss,lanes,threads", [(64, 32, 256), (256, 4, 512), (1024, 8, 1024), (4096, 2, 512)])
def test_scoo_geometries(gpu, ss, lanes, threads):
    rp, ci, va, cols = case("uniform_band", "float32")
    rows = rp.size - 1
    slices = -(-rows // ss)
    y, (rp, ci, va, x), info = _run(gpu, "uniform_band", "float32", "scoo", scoo_slices=slices,
                                    scoo_slice_size=ss, scoo_lanes=lanes, scoo_threads=threads)
    assert info["slices_count"] == slices
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("R,P", [(256, 1), (1024, 3), (4096, 0)])
@pytest.mark.parametrize("name", CASES)
def test_scoo_colsort_within_tolerance(gpu, name, dtype, R, P):
    flags = gpu._lib.MFP_TUNE_SCOO_COLSORT
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "scoo", flags=flags, scoo_slice_size=R, scoo_pieces=P)
    rows = rp.size - 1
    assert info["scoo_colsort"] == 1 and info["slice_size"] == R
    # rows longer than max(256, mean entries per slice / 128) get slices of their own
    S = -(-rows // R)
    T = max(256, (ci.size // S) // 128)
    n_long = int(np.sum(np.diff(rp.astype(np.int64)) > T))
    assert info["scoo_long_rows"] == n_long
    assert info["slices_count"] == S + -(-n_long // R)
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("P", [1, 2])
def test_scoo_colsort_wide_columns_cut_at_bands(gpu, dtype, P):
    # 16384-row slices leave 18 bits for the column offset: the 300,000
    # columns make two bands of 262,144, and no item may straddle them
    rp, ci, va = gpu.generate(gpu.GenSpec("uniform", values="uniform", rows=300000, nnz_lo=1, nnz_hi=5, seed=5),
                              np.dtype(dtype))
    cols = rp.size - 1
    R = 16384
    announce(f"scoo colsort bands dtype={dtype} R={R} P={P}")
    x = gpu.generate_vector(3, cols, np.dtype(dtype))
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=R, scoo_pieces=P)
    info = plan.info()
    assert info["scoo_packed"] == 1
    S = -(-cols // R)
    assert info["lane_size"] >= 2 * S  # every slice has entries in both bands
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(cols, np.dtype(dtype))
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    _tol_check(yd.to_host(), rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("R", [256, 4096])
@pytest.mark.parametrize("name", CASES)
def test_scoo_colsort_whole_slices(gpu, name, dtype, R):
    # MFP_TUNE_SCOO_WHOLE_SLICES: no long-row slices and, as every case has
    # fewer columns than a band (2^(32 - log2 R) >= 2^20), one item per slice,
    # so no partial slots; SpMV and a 4-column SpMM within tolerance
    flags = gpu._lib.MFP_TUNE_SCOO_COLSORT | gpu._lib.MFP_TUNE_SCOO_WHOLE_SLICES
    y, (rp, ci, va, x), info = _run(gpu, name, dtype, "scoo", flags=flags, scoo_slice_size=R, scoo_pieces=3)
    rows = rp.size - 1
    S = -(-rows // R)
    assert info["scoo_long_rows"] == 0 and info["slices_count"] == S and info["lane_size"] == S
    _tol_check(y, rp, ci, va, x)
    k = 4
    cols = x.size
    X = np.random.default_rng(4).uniform(-1, 1, (cols, k)).astype(dtype)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=flags, scoo_slice_size=min(R, 4096))
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray(rows * k, np.dtype(dtype))
    Yd.fill(np.nan)
    plan.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(rows, k)
    plan.destroy()
    A.destroy()
    for c in range(k):
        _tol_check(np.ascontiguousarray(Y[:, c]), rp, ci, va, np.ascontiguousarray(X[:, c]))


@pytest.mark.parametrize("dtype", DTYPES)
def test_scoo_colsort_slice_across_a_band_edge_stays_whole(gpu, dtype):
    # a banded matrix with 300,000 columns and 16384-row slices (bands of
    # 262,144 columns): the slice holding rows near column 262,144 spans the
    # band edge but fewer than 262,144 columns, so it stays one item
    rp, ci, va = gpu.generate(gpu.GenSpec("uniform", values="uniform", rows=300000, nnz_lo=1, nnz_hi=8, seed=6,
                                          bandwidth=0.01), np.dtype(dtype))
    cols = rp.size - 1
    R = 16384
    S = -(-cols // R)
    assert any(int(ci[rp[sl * R]:rp[min((sl + 1) * R, cols)]].min()) < 262144 <=
               int(ci[rp[sl * R]:rp[min((sl + 1) * R, cols)]].max()) for sl in range(S))
    announce(f"scoo colsort band edge dtype={dtype}")
    flags = gpu._lib.MFP_TUNE_SCOO_COLSORT | gpu._lib.MFP_TUNE_SCOO_WHOLE_SLICES
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=flags, scoo_slice_size=R)
    info = plan.info()
    assert info["scoo_packed"] == 1 and info["lane_size"] == S
    x = gpu.generate_vector(3, cols, np.dtype(dtype))
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(cols, np.dtype(dtype))
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    _tol_check(yd.to_host(), rp, ci, va, x)


def _colsort_run(gpu, rp, ci, va, cols, label, **tuning):
    announce(f"scoo colsort {label} rows={rp.size - 1} nnz={ci.size} tuning={tuning}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, **tuning)
    x = gpu.generate_vector(9, cols, va.dtype)
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rp.size - 1, va.dtype)
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    return yd.to_host(), x, plan.info()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("R", [256, 16384])
def test_scoo_colsort_empty_matrix_writes_zeros(gpu, dtype, R):
    rows = 1000
    rp = np.zeros(rows + 1, np.uint32)
    y, _, info = _colsort_run(gpu, rp, np.zeros(0, np.uint32), np.zeros(0, np.dtype(dtype)), 700, "empty",
                              scoo_slice_size=R)
    assert info["scoo_colsort"] == 1
    assert np.array_equal(y, np.zeros(rows, np.dtype(dtype)))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("P", [0, 3])
def test_scoo_colsort_only_the_last_column_band(gpu, dtype, P):
    # 16384-row slices: bands of 262,144 columns; every entry sits in the
    # third band, so the first two bands of every slice are empty
    rng = np.random.default_rng(11)
    rows, cols = 40000, 600000
    lens = rng.integers(0, 9, rows)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(np.arange(524288, cols), n, replace=False)) for n in lens]).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size).astype(np.dtype(dtype))
    y, x, info = _colsort_run(gpu, rp, ci, va, cols, "last band", scoo_slice_size=16384, scoo_pieces=P)
    assert info["scoo_packed"] == 1
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("P", [1, 4])
def test_scoo_colsort_long_rows_fill_several_slices(gpu, dtype, P):
    # 600 rows of 300 entries among 3,000 rows of 2: with 256-row slices the
    # long-row threshold is 256, so the 600 rows make three long-row slices
    rng = np.random.default_rng(12)
    rows, cols = 3000, 5000
    lens = np.full(rows, 2)
    lens[rng.choice(rows, 600, replace=False)] = 300
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size).astype(np.dtype(dtype))
    y, x, info = _colsort_run(gpu, rp, ci, va, cols, "long rows", scoo_slice_size=256, scoo_pieces=P)
    assert info["scoo_long_rows"] == 600
    assert info["slices_count"] == -(-rows // 256) + 3
    _tol_check(y, rp, ci, va, x)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("R,P,flags", [(256, 1, 0), (256, 4, 0), (1024, 0, 0), (4096, 0, 16)])
def test_scoo_colsort_integer_values_exact(gpu, dtype, R, P, flags):
    # values and x small integers: every product, partial slot and row sum
    # is exact in FP32 and FP64 whatever the order, so y equals the integer
    # row sums bit for bit; a value read from the wrong slot or half of a
    # 16-byte entry word, or a lost entry, changes them
    rng = np.random.default_rng(21)
    rows, cols = 6000, 20000
    lens = rng.integers(0, 13, rows)
    lens[rng.choice(rows, 40, replace=False)] = rng.integers(300, 2000, 40)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.uint32)
    vi = rng.integers(-8, 9, ci.size)
    xi = rng.integers(-8, 9, cols)
    want = np.array([int(np.dot(vi[rp[r]:rp[r + 1]], xi[ci[rp[r]:rp[r + 1]]])) for r in range(rows)])
    dt = np.dtype(dtype)
    announce(f"scoo colsort integer values dtype={dtype} R={R} P={P} flags={flags}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, vi.astype(dt), cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT | flags, scoo_slice_size=R, scoo_pieces=P)
    xd = gpu.DeviceArray.from_host(xi.astype(dt))
    yd = gpu.DeviceArray(rows, dt)
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    info = plan.info()
    assert info["scoo_colsort"] == 1
    assert np.array_equal(yd.to_host(), want.astype(dt))
    plan.destroy()
    A.destroy()


@pytest.mark.parametrize("whole", [False, True])
@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_scoo_colsort_16bit_values_in_the_last_band(gpu, vt, whole):
    rng = np.random.default_rng(13)
    rows, cols = 20000, 600000
    lens = rng.integers(0, 9, rows)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(np.arange(524288, cols), n, replace=False)) for n in lens]).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size).astype(np.float32)
    announce(f"scoo colsort {vt} last band")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols).convert(vt)
    _, _, stored = A.download()
    vs = ((stored.astype(np.uint32) << 16).view(np.float32) if vt == "bf16"
          else stored.view(np.float16).astype(np.float32))
    flags = gpu._lib.MFP_TUNE_SCOO_COLSORT | (gpu._lib.MFP_TUNE_SCOO_WHOLE_SLICES if whole else 0)
    plan = gpu.Plan(A, "scoo", flags=flags, scoo_slice_size=16384, scoo_pieces=3)
    if whole:  # columns 524,288.. span less than a band of 262,144: one item per slice
        assert plan.info()["lane_size"] == -(-rows // 16384)
    x = gpu.generate_vector(9, cols, np.float32)
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rows, np.float32)
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    y64, ad = oracle.spmv_f64(rp, ci, vs, x)
    # exact products added in double, partial slots rounded to FP32
    assert oracle.tolerance_check(yd.to_host(), y64, ad, 1.0) <= 1e-6


class _Offset:
    """A device pointer one element into a DeviceArray (not 16-byte aligned)."""

    def __init__(self, arr, elements):
        self.ptr = arr.ptr + elements * arr.dtype.itemsize


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("R,P", [(256, 4), (16384, 2)])
def test_scoo_colsort_unaligned_y(gpu, dtype, R, P):
    # the finish stores four rows with one 16-byte write when y + row is
    # 16-byte aligned; a y one element in must take the per-row stores and
    # leave the elements on either side alone
    rng = np.random.default_rng(14)
    rows, cols = 40000, 30000
    lens = rng.integers(0, 12, rows)
    lens[rng.choice(rows, 40, replace=False)] = 2000  # long rows: ordinary slices skip them
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size).astype(np.dtype(dtype))
    announce(f"scoo colsort unaligned y dtype={dtype} R={R} P={P}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=R, scoo_pieces=P)
    assert plan.info()["scoo_long_rows"] > 0
    x = gpu.generate_vector(15, cols, np.dtype(dtype))
    xd = gpu.DeviceArray.from_host(x)
    big = gpu.DeviceArray(rows + 2, np.dtype(dtype))
    big.fill(np.nan)
    plan.spmv(xd, _Offset(big, 1))
    gpu.synchronize()
    out = big.to_host()
    assert np.isnan(out[0]) and np.isnan(out[-1])
    _tol_check(out[1:-1], rp, ci, va, x)


def test_scoo_colsort_rejects_bad_slices(gpu):
    rp, ci, va, cols = case("stencil2d_64", "float32")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    # not a power of two; accumulators (8 bytes per row) above 160 KB
    for R in (1000, 32768, 65536):
        with pytest.raises(gpu.MfpError):
            gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=R)


def test_ell_budget_refusal(gpu):
    # a width * rows layout above the budget is refused, not attempted
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    announce("ell budget refusal")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    with pytest.raises(gpu.MfpBudgetError):
        gpu.Plan(A, "ell", budget_bytes=1 << 20)
    A.destroy()


# (format, tuning, matrix, kernels launched, same summation order every run)
KERNEL_TIME_CASES = [
    ("csr_scalar", {}, "powerlaw_20k", 1, True),
    ("csr_vector", {"lanes": 8}, "powerlaw_20k", 1, True),
    ("csr_adaptive", {}, "powerlaw_20k", 2, True),                                   # long-row chunks + finish
    ("csr_adaptive", {"long_row": 65536, "chunk": 2048}, "powerlaw_20k", 1, True),  # no chunks
    ("ell", {}, "stencil3d_20", 1, True),
    ("coo", {}, "powerlaw_20k", 2, False),                                           # zero fill of y, kernel
    ("scoo", {}, "powerlaw_20k", 1, False),                                          # reference geometry
    ("scoo", {"flags": 4, "scoo_slice_size": 256, "scoo_pieces": 3}, "powerlaw_20k", 2, False),
    ("hyb", {"hyb_width": 4}, "powerlaw_20k", 2, False),                             # ELL, COO
    ("hyb_atomic", {"hyb_width": 4}, "powerlaw_20k", 2, False),                      # zero fill of y, kernel
]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("fmt,tuning,name,kernels,ordered", KERNEL_TIME_CASES)
def test_plan_kernel_times(gpu, fmt, tuning, name, kernels, ordered, dtype):
    # one SpMV through the timed launch route (an event pair per kernel): one
    # positive interval per kernel, and the same y as the plain launch route
    rp, ci, va, cols = case(name, dtype)
    announce(f"kernel times fmt={fmt} tuning={tuning} case={name} dtype={dtype}")
    x = gpu.generate_vector(7, cols, np.dtype(dtype))
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, fmt, **tuning)
    xd = gpu.DeviceArray.from_host(x)
    yt = gpu.DeviceArray(rp.size - 1, np.dtype(dtype))
    yp = gpu.DeviceArray(rp.size - 1, np.dtype(dtype))
    yt.fill(np.nan)
    yp.fill(np.nan)
    ms = plan.kernel_times(xd, yt)
    plan.spmv(xd, yp)
    gpu.synchronize()
    assert len(ms) == kernels and all(t > 0 for t in ms), ms
    y_timed, y_plain = yt.to_host(), yp.to_host()
    if ordered:
        assert np.array_equal(y_timed, y_plain)
    else:
        _, absdot = oracle.spmv_f64(rp, ci, va, x)
        rtol = RTOL[np.dtype(dtype)]
        assert oracle.tolerance_check(y_timed, y_plain, absdot, rtol) <= rtol
    _tol_check(y_timed, rp, ci, va, x)
    plan.destroy()
    A.destroy()


@pytest.mark.parametrize("fmt,tuning,name,kernels,ordered", KERNEL_TIME_CASES)
def test_plan_graph_replays(gpu, fmt, tuning, name, kernels, ordered):
    # one plan's SpMV captured in a HIP graph (zero fills are kernels, never
    # hipMemsetAsync): three replays, each after eager work on the same
    # stream that overwrites y, equal to the eager SpMV
    rp, ci, va, cols = case(name, "float32")
    announce(f"plan graph fmt={fmt} tuning={tuning} case={name}")
    x = gpu.generate_vector(9, cols, np.float32)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, fmt, **tuning)
    xd = gpu.DeviceArray.from_host(x)
    y = gpu.DeviceArray(rp.size - 1, np.float32)
    plan.spmv(xd, y)
    gpu.synchronize()
    y_eager = y.to_host()
    g = plan.graph(xd, y)
    _, absdot = oracle.spmv_f64(rp, ci, va, x)
    for _ in range(3):
        y.fill(np.nan)
        plan.spmv(xd, y)  # eager work between replays
        y.fill(np.nan)
        g.launch()
        gpu.synchronize()
        y_graph = y.to_host()
        if ordered:
            assert np.array_equal(y_graph, y_eager)
        else:
            assert oracle.tolerance_check(y_graph, y_eager, absdot, 1e-5) <= 1e-5
    g.destroy()
    plan.destroy()
    A.destroy()


def test_plan_kernel_times_bf16_binned(gpu):
    # the binned 16-bit CSR launches one kernel per non-empty length bin, the
    # long-row chunks and their finish: an event after each, at most eight
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    announce("kernel times bf16 binned csr_adaptive powerlaw_20k")
    x = gpu.generate_vector(7, cols, np.float32)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    B = A.convert("bf16")
    plan = gpu.Plan(B, "csr_adaptive")
    xd = gpu.DeviceArray.from_host(x)
    yt = gpu.DeviceArray(rp.size - 1, np.float32)
    yp = gpu.DeviceArray(rp.size - 1, np.float32)
    yt.fill(np.nan)
    ms = plan.kernel_times(xd, yt)
    plan.spmv(xd, yp)
    gpu.synchronize()
    # expected: a kernel per non-empty bin of row lengths (<= 2, 4, ..., 32,
    # <= 4096 entries) plus chunks and finish when a row exceeds 4096
    lens = np.diff(rp.astype(np.int64))
    short = lens[lens <= 4096]
    bins = {min(5, max(0, int(np.ceil(np.log2(max(n, 1)))) - 1)) for n in short}
    want = len(bins) + (2 if (lens > 4096).any() else 0)
    assert len(ms) == want and all(t > 0 for t in ms), (ms, want)
    y_timed = yt.to_host()
    assert not np.isnan(y_timed).any()
    assert np.array_equal(y_timed, yp.to_host())
    plan.destroy()
    B.destroy()
    A.destroy()


def _banded(max_len, rows=3000, seed=31):
    # rows of 1 .. max_len entries (at least one of max_len) within 150
    # columns of the diagonal: every 64-row slice spans far fewer than 65,536
    rng = np.random.default_rng(seed + max_len)
    lens = rng.integers(1, max_len + 1, rows)
    lens[rng.integers(0, rows)] = max_len
    cols_ = []
    for r, n in enumerate(lens):
        lo, hi = max(0, r - 150), min(rows, r + 150)
        cols_.append(np.sort(rng.choice(np.arange(lo, hi), n, replace=False)))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate(cols_).astype(np.uint32)
    return rp, ci, rng.standard_normal(ci.size)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("max_len", [12, 16, 17, 21, 24, 29, 32, 33])
def test_ell_packed_row_words_bit_exact(gpu, dtype, max_len):
    # rows of 3 to 9 four-slot words: those of 5 to 8 words are read as one
    # group, the others in groups of 4, 2 and 1; every grouping adds the
    # slots in the oracle's ELL order
    rp, ci, va = _banded(max_len)
    va = va.astype(np.dtype(dtype))
    rows = rp.size - 1
    announce(f"ell packed words max_len={max_len} dtype={dtype}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, rows)
    plan = gpu.Plan(A, "ell", flags=gpu._lib.MFP_TUNE_ELL_PACKED)
    x = gpu.generate_vector(5, rows, np.dtype(dtype))
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rows, np.dtype(dtype))
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert plan.info()["ell_width"] == w == max_len
    assert np.array_equal(yd.to_host(), oracle.spmv_ell(rows, w, ec, ed, x))
    plan.destroy()
    A.destroy()


@pytest.mark.parametrize("vt", ["bf16", "f16"])
@pytest.mark.parametrize("max_len", [12, 16, 17, 21, 24, 29, 32, 33])
def test_ell_16bit_row_words_compensated(gpu, vt, max_len):
    # the 16-bit sliced ELL over rows of 3 to 9 words (one group up to 8):
    # Dot2 keeps the error of the stored-value products near one FP32 rounding
    rp, ci, va = _banded(max_len)
    va = va.astype(np.float32)
    rows = rp.size - 1
    announce(f"ell 16-bit words max_len={max_len} vt={vt}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, rows).convert(vt)
    plan = gpu.Plan(A, "ell")
    assert plan.info()["column_bytes"] == 2
    x = gpu.generate_vector(6, rows, np.float32)
    xd = gpu.DeviceArray.from_host(x)
    yd = gpu.DeviceArray(rows, np.float32)
    yd.fill(np.nan)
    plan.spmv(xd, yd)
    gpu.synchronize()
    from oracle.selector import bf16_round
    stored = bf16_round(va) if vt == "bf16" else va.astype(np.float16).astype(np.float32)
    y64, ad = oracle.spmv_f64(rp, ci, stored, x)
    assert oracle.tolerance_check(yd.to_host(), y64, ad, 1.0) <= 2.5e-7
    plan.destroy()
    A.destroy()
