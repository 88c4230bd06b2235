"""GPU parity of the device converters: bit-exact against the oracle's
restatement of the reference's host converters."""
import json
import os

import numpy as np
import pytest

import oracle
from gpu_cases import CASES, announce, case
from matrix_format_performance_amd import _lib as L

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _triplets(rp, ci, va, seed=0, dup_every=0):
    rows = np.repeat(np.arange(rp.size - 1, dtype=np.uint32), np.diff(rp))
    r, c, v = rows, ci.copy(), va.astype(np.float64)
    if dup_every:
        idx = np.arange(0, r.size, dup_every)
        r = np.concatenate([r, r[idx]])
        c = np.concatenate([c, c[idx]])
        v = np.concatenate([v, v[idx] + 0.25])
    perm = np.random.default_rng(seed).permutation(r.size)
    return r[perm], c[perm], v[perm]


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("name", CASES)
def test_csr_from_triplets_bit_exact(gpu, name, dtype):
    rp, ci, va, cols = case(name, dtype)
    r, c, v = _triplets(rp, ci, va, seed=1)
    announce(f"triplets->csr case={name} dtype={dtype} nnz={r.size}")
    A = gpu.DeviceMatrix.from_triplets(rp.size - 1, cols, r, c, v, np.dtype(dtype))
    grp, gci, gva = A.download()
    A.destroy()
    orp, oci, ova = oracle.csr_from_triplets(rp.size - 1, cols, r, c, v, np.dtype(dtype))
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci) and np.array_equal(gva, ova)


def test_csr_from_triplets_duplicates_keep_input_order(gpu):
    rp, ci, va, cols = case("uniform_band", "float64")
    r, c, v = _triplets(rp, ci, va, seed=2, dup_every=7)
    announce(f"triplets->csr duplicates nnz={r.size}")
    A = gpu.DeviceMatrix.from_triplets(rp.size - 1, cols, r, c, v, np.float64)
    got = A.download()
    A.destroy()
    ref = oracle.csr_from_triplets(rp.size - 1, cols, r, c, v, np.float64)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))


def test_csr_from_triplets_doc_example_and_errors(gpu):
    g = json.load(open(os.path.join(GOLD, "doc_examples.json")))["csr"]
    rows = np.repeat(np.arange(g["rows"]), np.diff(g["row_ptr"]))
    announce("triplets->csr doc example")
    A = gpu.DeviceMatrix.from_triplets(g["rows"], g["cols"], rows[::-1], np.array(g["columns"])[::-1],
                                       np.array(g["values"], float)[::-1], np.float32)
    rp, ci, va = A.download()
    A.destroy()
    assert rp.tolist() == g["row_ptr"] and ci.tolist() == g["columns"] and va.tolist() == g["values"]
    with pytest.raises(gpu.MfpError):
        gpu.DeviceMatrix.from_triplets(2, 2, [0, 2], [0, 1], [1.0, 2.0])


@pytest.mark.parametrize("name", CASES)
def test_row_statistics(gpu, name):
    rp, ci, va, cols = case(name, "float32")
    announce(f"row stats case={name}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    got = A.row_stats()
    A.destroy()
    assert got == oracle.rows_statistics(rp)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("name", ["stencil2d_64", "stencil3d_20", "uniform_band", "edge_rows", "one_by_one"])
def test_ell_layout(gpu, name, dtype):
    rp, ci, va, cols = case(name, dtype)
    announce(f"ell layout case={name} dtype={dtype}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    p = gpu.Plan(A, "ell")
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert p.info()["ell_width"] == w
    assert np.array_equal(p.array(L.MFP_ARR_ELL_VALUES, va.dtype), ed)
    assert np.array_equal(p.array(L.MFP_ARR_ELL_COLS, np.uint32), ec)
    p.destroy()
    A.destroy()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("name,width", [("powerlaw_20k", None), ("powerlaw_20k", 5), ("edge_rows", None),
                                        ("uniform_band", 10), ("stencil3d_20", None)])
def test_hybrid_layout(gpu, name, width, dtype):
    rp, ci, va, cols = case(name, dtype)
    announce(f"hybrid layout case={name} dtype={dtype} width={width}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    kw = {"hyb_width": width} if width else {"hyb_percent": 0.2}
    p = gpu.Plan(A, "hyb", **kw)
    w = width if width else oracle.hybrid_width(rp, 0.2)
    assert p.info()["ell_width"] == w
    _, ed, ec = oracle.ell_from_csr(rp, ci, va, width=w)
    cd, cc, cr = oracle.coo_from_csr(rp, ci, va, element_start=w)
    assert np.array_equal(p.array(L.MFP_ARR_ELL_VALUES, va.dtype), ed)
    assert np.array_equal(p.array(L.MFP_ARR_ELL_COLS, np.uint32), ec)
    assert np.array_equal(p.array(L.MFP_ARR_COO_VALUES, va.dtype), cd)
    assert np.array_equal(p.array(L.MFP_ARR_COO_COLS, np.uint32), cc)
    assert np.array_equal(p.array(L.MFP_ARR_COO_ROWS, np.uint32), cr)
    p.destroy()
    A.destroy()


@pytest.mark.parametrize("name", CASES)
def test_coo_and_scoo_layout(gpu, name):
    rp, ci, va, cols = case(name, "float32")
    rows = rp.size - 1
    announce(f"coo/scoo layout case={name}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    cd, cc, cr = oracle.coo_from_csr(rp, ci, va)
    p = gpu.Plan(A, "coo")
    assert np.array_equal(p.array(L.MFP_ARR_COO_VALUES, np.float32), cd)
    assert np.array_equal(p.array(L.MFP_ARR_COO_ROWS, np.uint32), cr)
    assert np.array_equal(p.array(L.MFP_ARR_COO_COLS, np.uint32), cc)
    p.destroy()
    # reference geometry (RTX 2080: 46 SMs, 48 KB) and MI355X geometry (256 CUs, 160 KB)
    for sm, shm in ((46, 49152), (256, 163840)):
        slices, ss, lane = oracle.scoo_geometry(rows, sm, shm, np.float32)
        assert gpu.scoo_geometry(np.float32, rows, sm, shm) == (slices, ss, lane)
        p = gpu.Plan(A, "scoo", scoo_slices=slices, scoo_slice_size=ss, scoo_lanes=lane)
        _, _, _, sp = oracle.scoo_build(rp, ci, va, slices, ss)
        assert np.array_equal(p.array(L.MFP_ARR_SCOO_SLICES_PTR, np.uint32), sp)
        p.destroy()
    A.destroy()


@pytest.mark.parametrize("k", range(4))
def test_row_blocks_kats(gpu, k):
    c = json.load(open(os.path.join(GOLD, "partition_kats.json")))["row_blocks"][k]
    lens = c["lengths"] if "lengths" in c else [c["lengths_repeat"][0]] * c["lengths_repeat"][1]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    announce(f"row blocks kat {k}")
    assert gpu.fill_row_blocks(rp.size - 1, rp, c["nnz_per_wg"]).tolist() == c["blocks"]


@pytest.mark.parametrize("W", [4, 64, 1024, 4096])
@pytest.mark.parametrize("name", CASES)
def test_row_blocks_match_oracle(gpu, name, W):
    rp, ci, va, cols = case(name, "float32")
    announce(f"row blocks case={name} W={W}")
    assert np.array_equal(gpu.fill_row_blocks(rp.size - 1, rp, W), oracle.fill_row_blocks(rp, W))
