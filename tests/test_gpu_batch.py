"""SpMM (k right-hand sides), batched matrices in one persistent launch, and
HIP-graph replay of that launch."""
import numpy as np
import pytest

import oracle
from gpu_cases import announce, case

pytestmark = pytest.mark.gpu


def _columns_ref(rp, ci, va, X, lanes):
    return np.stack([oracle.spmv_lanes(rp, ci, va, np.ascontiguousarray(X[:, j]), lanes)
                     for j in range(X.shape[1])], axis=1)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("lanes", [4, 32])
@pytest.mark.parametrize("name", ["stencil3d_20", "powerlaw_20k"])
def test_spmm_columns_bit_exact(gpu, name, k, lanes, dtype):
    rp, ci, va, cols = case(name, dtype)
    announce(f"spmm case={name} k={k} lanes={lanes} dtype={dtype}")
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, (cols, k)).astype(dtype)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "csr_vector", lanes=lanes)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray((rp.size - 1) * k, np.dtype(dtype))
    Yd.fill(np.nan)
    plan.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(-1, k)
    assert np.array_equal(Y, _columns_ref(rp, ci, va, X, lanes))


def _batch_members(gpu, count, dtype, seed=0):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(count):
        rows = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        fam = ["stencil2d", "stencil3d", "uniform", "powerlaw"][i % 4]
        if fam == "stencil2d":
            spec = gpu.GenSpec("stencil2d", values="uniform", n=max(2, int(np.sqrt(rows))), seed=i)
        elif fam == "stencil3d":
            spec = gpu.GenSpec("stencil3d", values="normal", n=max(2, int(round(rows ** (1 / 3)))), seed=i)
        elif fam == "uniform":
            spec = gpu.GenSpec("uniform", values="uniform", rows=rows, nnz_lo=4, nnz_hi=64, bandwidth=0.1, seed=i)
        else:
            spec = gpu.GenSpec("powerlaw", values="uniform", rows=rows, nnz_lo=1, nnz_hi=rows // 2,
                               alpha=float(rng.uniform(1.5, 3.0)), seed=i)
        mats.append(gpu.generate(spec, np.dtype(dtype)))
    return mats


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("k", [1, 4, 16])
def test_batch_persistent_launch(gpu, k, dtype):
    announce(f"batch k={k} dtype={dtype}")
    members = _batch_members(gpu, 24, dtype, seed=k)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=8, long_row=512, chunk=256)
    for d in dms:
        d.destroy()  # the batch holds its own copy
    assert batch.count == 24 and batch.rows == sum(m[0].size - 1 for m in members)
    rp, ci, va = batch.matrix_arrays()
    # the assembly is the block-diagonal concatenation
    r_off, c_off = batch.row_offsets, batch.col_offsets
    for j, (mrp, mci, mva) in enumerate(members):
        e0 = rp[r_off[j]]
        assert np.array_equal(rp[r_off[j]:r_off[j + 1] + 1] - e0, mrp)
        assert np.array_equal(ci[e0:e0 + mci.size], mci + c_off[j])
        assert np.array_equal(va[e0:e0 + mva.size], mva)
    X = np.random.default_rng(3).uniform(-1, 1, (batch.cols, k)).astype(dtype)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray(batch.rows * k, np.dtype(dtype))
    Yd.fill(np.nan)
    batch.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(-1, k)
    if k == 1:
        W = gpu.engine.DEFAULT_NNZ_PER_WG[np.dtype(dtype)]
        ref = oracle.spmv_adaptive(rp, ci, va, X[:, 0].copy(), W, 256, 512, 256)[:, None]
    else:
        ref = _columns_ref(rp, ci, va, X, 8)
    assert np.array_equal(Y, ref)


def _binned_lanes(rp):
    """Lanes per row of a batch created with lanes = 0: 2 << b for rows of at
    most 2 << b entries, b <= 5."""
    lens = np.diff(rp.astype(np.int64))
    b = np.zeros(lens.size, np.int64)
    for q in range(5):
        b += lens > (2 << q)
    return 2 << b


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("k", [2, 4, 16])
def test_batch_binned_lanes_per_row(gpu, k, dtype):
    announce(f"batch binned k={k} dtype={dtype}")
    members = _batch_members(gpu, 24, dtype, seed=20 + k)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=0, long_row=512, chunk=256)
    for d in dms:
        d.destroy()
    rp, ci, va = batch.matrix_arrays()
    X = np.random.default_rng(5).uniform(-1, 1, (batch.cols, k)).astype(dtype)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray(batch.rows * k, np.dtype(dtype))
    Yd.fill(np.nan)
    batch.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(-1, k)
    G = _binned_lanes(rp)
    assert set(np.unique(G)) >= {2, 64}  # short rows and long rows both present
    ref = np.empty_like(Y)
    for g in np.unique(G):
        sel = G == g
        ref[sel] = _columns_ref(rp, ci, va, X, int(g))[sel]
    # each column in the CSR-vector order with the row's own lane count
    assert np.array_equal(Y, ref)


@pytest.mark.parametrize("dtype,k", [("float32", 4), ("float32", 8), ("float32", 16),
                                      ("float64", 2), ("float64", 4), ("float64", 16)])
def test_batch_stream_spmm_bit_exact(gpu, k, dtype):
    # CSR-stream SpMM (row blocks of 512 entries staged in LDS, groups of
    # lanes over consecutive entries, rows across groups finished through
    # LDS): every column equal to the oracle's emulation of that order,
    # including rows longer than a block and empty rows
    announce(f"batch stream k={k} dtype={dtype}")
    members = _batch_members(gpu, 24, dtype, seed=40 + k)
    members.append((np.zeros(4, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.dtype(dtype))))  # 3 empty rows
    dms = [gpu.DeviceMatrix.from_csr(*m, cols=max(m[0].size - 1, 1)) for m in members]
    batch = gpu.Batch(dms, lanes=gpu._lib.MFP_BATCH_STREAM)
    for d in dms:
        d.destroy()
    rp, ci, va = batch.matrix_arrays()
    lens = np.diff(rp.astype(np.int64))
    assert lens.max() > 512 and (lens == 0).any()
    X = np.random.default_rng(7 + k).uniform(-1, 1, (batch.cols, k)).astype(dtype)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray(batch.rows * k, np.dtype(dtype))
    Yd.fill(np.nan)
    batch.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(-1, k)
    assert np.array_equal(Y, oracle.spmm_stream(rp, ci, va, X, k))
    g = batch.graph(k, Xd, Yd)
    for _ in range(2):
        Yd.fill(np.nan)
        g.launch()
        gpu.synchronize()
        assert np.array_equal(Yd.to_host().reshape(-1, k), Y)
    g.destroy()
    batch.destroy()


@pytest.mark.parametrize("dtype,k", [("float32", 4), ("float32", 16), ("float64", 2), ("float64", 8)])
@pytest.mark.parametrize("name", ["stencil3d_20", "powerlaw_20k", "edge_rows"])
def test_plan_stream_spmm_bit_exact(gpu, name, k, dtype):
    # mfp_plan_spmm on a CSR-Adaptive plan runs the CSR-stream SpMM
    rp, ci, va, cols = case(name, dtype)
    announce(f"plan stream spmm case={name} k={k} dtype={dtype}")
    X = np.random.default_rng(k).uniform(-1, 1, (cols, k)).astype(dtype)
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "csr_adaptive")
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray((rp.size - 1) * k, np.dtype(dtype))
    Yd.fill(np.nan)
    plan.spmm(k, Xd, Yd)
    gpu.synchronize()
    assert np.array_equal(Yd.to_host().reshape(-1, k), oracle.spmm_stream(rp, ci, va, X, k))
    plan.destroy()
    A.destroy()


def test_plan_stream_spmm_x_above_4gib(gpu):
    # X of 70M rows x 16 floats (4.48 GB): past 4 GiB the CSR-stream gathers
    # take 64-bit offsets.  Most columns sit near the top of X, where a 32-bit
    # byte offset would wrap to another X row, and X[c, j] = c % 251 + j tells
    # the rows apart.  Integer values make every summation order exact.
    rows, cols, k = 2048, 70_000_000, 16
    announce(f"plan stream spmm X above 4 GiB: cols={cols} k={k}")
    rng = np.random.default_rng(12)
    lens = rng.integers(0, 40, rows)
    lens[7] = 1500  # a row longer than a block
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = rng.integers(cols - 3_000_000, cols, int(rp[-1])).astype(np.uint32)
    ci[::5] = rng.integers(0, 1000, ci[::5].size)
    va = rng.integers(-3, 4, ci.size).astype(np.float32)
    xc = (np.arange(cols, dtype=np.uint32) % 251).astype(np.float32)
    X = xc[:, None] + np.arange(k, dtype=np.float32)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    del X
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "csr_adaptive")
    Yd = gpu.DeviceArray(rows * k, np.float32)
    Yd.fill(np.nan)
    plan.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(rows, k)
    plan.destroy()
    A.destroy()
    Xd.free()
    Yd.free()
    rr = np.repeat(np.arange(rows), lens)
    base = np.bincount(rr, weights=va.astype(np.float64) * xc[ci], minlength=rows)
    vsum = np.bincount(rr, weights=va.astype(np.float64), minlength=rows)
    want = base[:, None] + vsum[:, None] * np.arange(k)
    assert np.array_equal(Y.astype(np.float64), want)


def test_batch_stream_spmm_refuses_narrow_x(gpu):
    members = _batch_members(gpu, 2, "float32", seed=3)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=gpu._lib.MFP_BATCH_STREAM)
    Xd = gpu.DeviceArray(batch.cols * 2, np.float32)
    Yd = gpu.DeviceArray(batch.rows * 2, np.float32)
    with pytest.raises(gpu.MfpError, match="16-byte"):
        batch.spmm(2, Xd, Yd)  # 2 floats = 8 bytes per X row
    for d in dms:
        d.destroy()
    batch.destroy()


def test_batch_graph_replays_match_eager(gpu):
    announce("batch graph replay")
    members = _batch_members(gpu, 16, "float32", seed=7)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=16, long_row=512, chunk=256)
    for k in (1, 4):
        Xd = gpu.DeviceArray(batch.cols * k, np.float32)
        Yg = gpu.DeviceArray(batch.rows * k, np.float32)
        Ye = gpu.DeviceArray(batch.rows * k, np.float32)
        g = batch.graph(k, Xd, Yg)
        for rep in range(4):
            X = np.random.default_rng(rep).uniform(-1, 1, batch.cols * k).astype(np.float32)
            gpu.engine.L.call("mfp_memcpy_h2d", Xd.ptr, X.ctypes.data, X.nbytes)
            Yg.fill(np.nan)
            g.launch()
            gpu.synchronize()
            batch.spmm(k, Xd, Ye)  # eager work between replays
            gpu.synchronize()
            assert np.array_equal(Yg.to_host(), Ye.to_host()), (k, rep)
        g.destroy()


def test_autotune_sweeps_and_times_candidates(gpu):
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    announce("autotune powerlaw_20k")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    x = gpu.DeviceArray.from_host(gpu.generate_vector(1, cols, np.float32))
    y = gpu.DeviceArray(rp.size - 1, np.float32)
    res = gpu.autotune(A, x, y, reps=2)
    fmts = [r["format"] for r in res]
    assert fmts.count("csr_vector") == 6 and fmts.count("csr_adaptive") == 3
    # hybrid splits: average and the 0.75 / 0.9 / 0.99 row-length quantiles
    lens = np.sort(np.diff(rp.astype(np.int64)))
    q = lambda p: int(lens[int(p * (lens.size - 1))])
    widths = sorted({-(-ci.size // (rp.size - 1)), q(0.75), q(0.9), q(0.99)} - {0})
    # plus the default candidate, whose width 0 selects the hyb_percent rule
    assert sorted(r["tuning"]["hyb_width"] for r in res if r["format"] == "hyb") == [0] + widths
    ok = [r for r in res if r["status"] == 0]
    assert ok and all(np.isfinite(r["ms"]) and r["ms"] > 0 for r in ok)
    assert res[0]["ms"] == min(r["ms"] for r in res)
    # ELL with width 12000 x 20k rows builds; a plan from the best candidate reproduces it
    best = res[0]
    plan = gpu.Plan(A, best["format"], **{k: v for k, v in best["tuning"].items()})
    plan.spmv(x, y)
    gpu.synchronize()
    y_seq = oracle.spmv_seq(rp, ci, va, x.to_host())
    _, absdot = oracle.spmv_f64(rp, ci, va, x.to_host())
    assert oracle.tolerance_check(y.to_host(), y_seq, absdot, 1e-5) <= 1e-5


def test_autotune_16bit_values_offer_column_sorted_scoo_only(gpu):
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    announce("autotune bf16 powerlaw_20k")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols).convert("bf16")
    x = gpu.DeviceArray.from_host(gpu.generate_vector(1, cols, np.float32))
    y = gpu.DeviceArray(rp.size - 1, np.float32)
    res = gpu.autotune(A, x, y, reps=1)
    assert {r["format"] for r in res} == {"csr_vector", "csr_adaptive", "ell", "scoo"}
    sc = [r for r in res if r["format"] == "scoo"]
    flag = gpu._lib.MFP_TUNE_SCOO_COLSORT
    assert sc and all(r["tuning"]["flags"] & flag for r in sc)
    assert sum(bool(r["tuning"]["flags"] & gpu._lib.MFP_TUNE_IS_DEFAULT) for r in sc) == 1
    assert all(r["status"] == 0 for r in sc)


def test_batch_column_sorted_k1(gpu):
    announce("batch column-sorted k=1")
    members = _batch_members(gpu, 16, "float32", seed=9)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=16, flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=4096, scoo_pieces=2)
    for d in dms:
        d.destroy()
    rp, ci, va = batch.matrix_arrays()
    Xd = gpu.DeviceArray(batch.cols, np.float32)
    Yg = gpu.DeviceArray(batch.rows, np.float32)
    Ye = gpu.DeviceArray(batch.rows, np.float32)
    g = batch.graph(1, Xd, Yg)
    for rep in range(3):
        x = np.random.default_rng(40 + rep).uniform(-1, 1, batch.cols).astype(np.float32)
        gpu.engine.L.call("mfp_memcpy_h2d", Xd.ptr, x.ctypes.data, x.nbytes)
        Yg.fill(np.nan)
        g.launch()
        gpu.synchronize()
        batch.spmm(1, Xd, Ye)  # eager work between replays
        gpu.synchronize()
        _, absdot = oracle.spmv_f64(rp, ci, va, x)
        y_seq = oracle.spmv_seq(rp, ci, va, x)
        for y in (Yg.to_host(), Ye.to_host()):
            assert not np.isnan(y).any()
            assert oracle.tolerance_check(y, y_seq, absdot, 1e-5) <= 1e-5
    g.destroy()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("name,P", [("stencil3d_20", 0), ("powerlaw_20k", 3), ("edge_rows", 1)])
def test_spmm_column_sorted_within_tolerance(gpu, name, P, k, dtype):
    rp, ci, va, cols = case(name, dtype)
    announce(f"spmm column-sorted {name} k={k} dtype={dtype} P={P}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=1024, scoo_pieces=P)
    X = np.random.default_rng(k).uniform(-1, 1, (cols, k)).astype(dtype)
    Xd = gpu.DeviceArray.from_host(X.ravel())
    Yd = gpu.DeviceArray((rp.size - 1) * k, np.dtype(dtype))
    Yd.fill(np.nan)
    plan.spmm(k, Xd, Yd)
    gpu.synchronize()
    Y = Yd.to_host().reshape(-1, k)
    assert not np.isnan(Y).any()
    rtol = 1e-5 if dtype == "float32" else 1e-12
    for j in range(k):
        x = np.ascontiguousarray(X[:, j])
        y64, absdot = oracle.spmv_f64(rp, ci, va, x)
        assert oracle.tolerance_check(Y[:, j], oracle.spmv_seq(rp, ci, va, x), absdot, rtol) <= rtol
        assert oracle.tolerance_check(Y[:, j], y64, absdot, rtol) <= rtol


def test_spmm_column_sorted_refuses_a_slice_too_large_for_k(gpu):
    rp, ci, va, cols = case("stencil3d_20", "float32")
    announce("spmm column-sorted refusal")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    plan = gpu.Plan(A, "scoo", flags=gpu._lib.MFP_TUNE_SCOO_COLSORT, scoo_slice_size=4096)
    X = gpu.DeviceArray(cols * 8, np.float32)
    Y = gpu.DeviceArray((rp.size - 1) * 8, np.float32)
    with pytest.raises(gpu.MfpError):
        plan.spmm(8, X, Y)  # 4096 x 8 doubles = 256 KB of LDS
    plan.spmm(4, X, Y)      # 128 KB fits
    gpu.synchronize()


@pytest.mark.parametrize("k", [2, 4, 8, 16])  # k <= 4: whole slices, one kernel; k > 4: long-row slices and a finish
def test_batch_column_sorted_spmm(gpu, k):
    announce(f"batch column-sorted spmm k={k}")
    members = _batch_members(gpu, 16, "float32", seed=30 + k)
    dms = [gpu.DeviceMatrix.from_csr(*m) for m in members]
    batch = gpu.Batch(dms, lanes=1)
    for d in dms:
        d.destroy()
    rp, ci, va = batch.matrix_arrays()
    Xd = gpu.DeviceArray(batch.cols * k, np.float32)
    Yg = gpu.DeviceArray(batch.rows * k, np.float32)
    Ye = gpu.DeviceArray(batch.rows * k, np.float32)
    g = batch.graph(k, Xd, Yg)  # builds the column-sorted plan for k before the capture
    for rep in range(2):
        X = np.random.default_rng(50 + rep).uniform(-1, 1, (batch.cols, k)).astype(np.float32)
        gpu.engine.L.call("mfp_memcpy_h2d", Xd.ptr, X.ctypes.data, X.nbytes)
        Yg.fill(np.nan)
        g.launch()
        gpu.synchronize()
        batch.spmm(k, Xd, Ye)
        gpu.synchronize()
        for Y in (Yg.to_host().reshape(-1, k), Ye.to_host().reshape(-1, k)):
            assert not np.isnan(Y).any()
            for j in range(k):
                x = np.ascontiguousarray(X[:, j])
                _, absdot = oracle.spmv_f64(rp, ci, va, x)
                assert oracle.tolerance_check(Y[:, j], oracle.spmv_seq(rp, ci, va, x), absdot, 1e-5) <= 1e-5
    g.destroy()
