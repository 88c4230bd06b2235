"""Pins for the CPU oracle (oracle/): golden vectors from the reference's own
write-up, hand-traced known answers, closed forms, invariants and brute force.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong index, a transposed operand, unfused vs fused arithmetic) fails.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _random_csr(rows, cols, max_len, seed, dtype=np.float32, dups=False):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, rows)
    r, c = [], []
    for i, n in enumerate(lens):
        n = min(int(n), cols)
        cc = rng.integers(0, cols, n) if dups else rng.choice(cols, n, replace=False)
        r.extend([i] * len(cc))
        c.extend(cc.tolist())
    r = np.array(r, np.uint32)
    c = np.array(c, np.uint32)
    v = rng.uniform(-1, 1, r.size)
    perm = rng.permutation(r.size)
    return r[perm], c[perm], v[perm]


# ---- triplets -> CSR (common/matrix_converter.cpp:15-87) --------------------------

def test_csr_doc_example():
    g = _gold("doc_examples.json")["csr"]
    rp = np.array(g["row_ptr"], np.uint32)
    rows = np.repeat(np.arange(g["rows"]), np.diff(rp))
    rng = np.random.default_rng(1)
    perm = rng.permutation(rows.size)  # input order must not matter
    out_rp, out_c, out_v = oracle.csr_from_triplets(
        g["rows"], g["cols"], rows[perm], np.array(g["columns"])[perm], np.array(g["values"], float)[perm])
    assert out_rp.tolist() == g["row_ptr"]
    assert out_c.tolist() == g["columns"]
    assert out_v.tolist() == g["values"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_invariants_and_multiset(dtype):
    r, c, v = _random_csr(300, 200, 40, seed=3)
    rp, ci, va = oracle.csr_from_triplets(300, 200, r, c, v, dtype)
    assert rp[0] == 0 and rp[-1] == r.size and np.all(np.diff(rp.astype(np.int64)) >= 0)
    for i in range(300):
        seg = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)  # strictly increasing
    # same (row, col, value) multiset, values rounded to dtype
    rows = np.repeat(np.arange(300), np.diff(rp))
    a = sorted(zip(rows.tolist(), ci.tolist(), va.tolist()))
    b = sorted(zip(r.tolist(), c.tolist(), np.asarray(v, dtype).tolist()))
    assert a == b


def test_csr_ties_keep_input_order():
    # duplicate (row, col): entries stay in input order (stable)
    r = np.array([1, 0, 1, 1], np.uint32)
    c = np.array([2, 0, 2, 1], np.uint32)
    v = np.array([10.0, 20.0, 30.0, 40.0])
    rp, ci, va = oracle.csr_from_triplets(2, 3, r, c, v, np.float64)
    assert rp.tolist() == [0, 1, 4]
    assert ci.tolist() == [0, 1, 2, 2]
    assert va.tolist() == [20.0, 40.0, 10.0, 30.0]


def test_csr_rejects_out_of_range():
    with pytest.raises(ValueError):
        oracle.csr_from_triplets(2, 2, np.array([2]), np.array([0]), np.array([1.0]))


# ---- row statistics (:95-127) ---------------------------------------------------------

def test_rows_statistics_closed_form():
    rp = np.array([0, 3, 3, 10, 11], np.uint32)  # lengths 3, 0, 7, 1
    mn, mx, avg, sd = oracle.rows_statistics(rp)
    assert (mn, mx, avg) == (0, 7, 2)  # integer average 11 // 4
    # std uses the integer average 2: ((1)^2 + 2^2 + 5^2 + 1^2) / 4
    assert sd == pytest.approx(np.sqrt((1 + 4 + 25 + 1) / 4), rel=0, abs=1e-15)


# ---- ELL / COO / hybrid ------------------------------------------------------------

def test_ell_doc_example():
    g = _gold("doc_examples.json")["ell"]
    t = g["triplets"]
    rp, ci, va = oracle.csr_from_triplets(g["rows"], g["cols"], t["rows"], t["cols"], t["values"], np.float32)
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert w == g["elements_in_rows"]
    assert ed.tolist() == g["data"]
    assert ec.tolist() == g["columns"]


def test_ell_truncation_and_coo_tail_partition():
    r, c, v = _random_csr(120, 90, 25, seed=5)
    rp, ci, va = oracle.csr_from_triplets(120, 90, r, c, v, np.float64)
    w = 7
    _, ed, ec = oracle.ell_from_csr(rp, ci, va, width=w)
    cd, cc, cr = oracle.coo_from_csr(rp, ci, va, element_start=w)
    # ELL(w) + COO tail(w) == CSR, as multisets of (row, col, value)
    rows = 120
    ell = [(i % rows, int(ec[i]), float(ed[i])) for i in range(ed.size)
           if (i // rows) < (rp[i % rows + 1] - rp[i % rows])]
    coo = list(zip(cr.tolist(), cc.tolist(), cd.tolist()))
    full = list(zip(np.repeat(np.arange(rows), np.diff(rp)).tolist(), ci.tolist(), va.tolist()))
    assert sorted(ell + coo) == sorted(full)
    # the tail starts at in-row offset w and keeps CSR order
    assert np.all(np.diff(cr.astype(np.int64)) >= 0)
    assert cd.size == oracle._lib.orc_coo_count(rows, rp, w)


def test_hybrid_width_rule():
    rp = np.array([0, 2, 4, 6, 106], np.uint32)  # lengths 2, 2, 2, 100 -> avg 26, max 100
    assert oracle.hybrid_width(rp, 0.2) == int(26 + (100 - 26) * 0.2)  # 40
    assert oracle.hybrid_width(rp, 0.0) == 26
    assert oracle.hybrid_width(rp, 1.0) == 100


# ---- SCOO (common/scoo_matrix_class.cpp) ---------------------------------------------

def test_pow2_helpers():
    assert [oracle.next_pow_2(x) for x in (1, 2, 3, 5, 64, 65)] == [1, 2, 4, 8, 64, 128]
    assert [oracle.prev_pow_2(x) for x in (1, 2, 3, 5, 64, 65, 0)] == [1, 2, 2, 4, 64, 64, 0]
    # exhaustively against bit arithmetic
    for x in range(1, 1 << 14):
        assert oracle.next_pow_2(x) == 1 << (x - 1).bit_length()
        assert oracle.prev_pow_2(x) == 1 << (x.bit_length() - 1)


@pytest.mark.parametrize("k", range(7))
def test_scoo_geometry_kats(k):
    g = _gold("partition_kats.json")["scoo"][k]
    got = oracle.scoo_geometry(g["rows"], g["sm_count"], g["shared_mem"], np.dtype(g["dtype"]))
    assert got == (g["slices"], g["slice_size"], g["lane_size"])


def test_scoo_layout_is_coo_order_with_slice_pointers():
    r, c, v = _random_csr(10, 10, 6, seed=8)
    rp, ci, va = oracle.csr_from_triplets(10, 10, r, c, v, np.float32)
    slices, ss, lane = oracle.scoo_geometry(10, 46, 49152, np.float32)  # 46 slices of 1 row
    vals, cidx, ridx, sp = oracle.scoo_build(rp, ci, va, slices, ss)
    cd, cc, cr = oracle.coo_from_csr(rp, ci, va)
    assert vals.tolist() == cd.tolist() and cidx.tolist() == cc.tolist() and ridx.tolist() == cr.tolist()
    # 10 real slices start at rows 0..9; entry 10 holds nnz; entries 11..46 stay 0
    assert sp[:10].tolist() == rp[:10].tolist()
    assert sp[10] == rp[-1]
    assert np.all(sp[11:] == 0) and sp.size == 47


# ---- row blocks (gpu/csr_adaptive_spmv.cu:180-240) -------------------------------------

def _kat_row_ptr(case):
    if "lengths" in case:
        lens = case["lengths"]
    else:
        val, n = case["lengths_repeat"]
        lens = [val] * n
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)


@pytest.mark.parametrize("k", range(4))
def test_row_blocks_hand_traces(k):
    case = _gold("partition_kats.json")["row_blocks"][k]
    got = oracle.fill_row_blocks(_kat_row_ptr(case), case["nnz_per_wg"])
    assert got.tolist() == case["blocks"]


@pytest.mark.parametrize("W", [4, 64, 1024])
def test_row_blocks_invariants(W):
    rng = np.random.default_rng(W)
    lens = rng.choice([0, 1, 3, 20, 70, 300, 2000], 3000, p=[.1, .3, .3, .2, .05, .04, .01])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    b = oracle.fill_row_blocks(rp, W).astype(np.int64)
    rows = lens.size
    assert b[0] == 0 and b[-1] == rows and np.all(np.diff(b) >= 0)
    for s, e in zip(b[:-1], b[1:]):
        if e - s > 1:
            assert rp[e] - rp[s] <= W and e - s <= W + 1
            # greedy: one more row would overflow nnz or the row cap
            if e < rows:
                assert rp[e + 1] - rp[s] > W or e + 1 - s > W + 1 or rp[e] - rp[s] == W


# ---- SpMV orders (cpu/cpu_matrix_multiplier.cpp) ---------------------------------------

def _one_row(vals, x):
    n = len(vals)
    return (np.array([0, n], np.uint32), np.arange(n, dtype=np.uint32),
            np.array(vals, np.float32), np.array(x, np.float32))


def test_seq_uses_fused_multiply_add():
    # (1+2^-12)^2 = 1 + 2^-11 + 2^-24 rounds to 1 + 2^-11 in float; the fused
    # form keeps the 2^-24 after adding -(1 + 2^-11) first.
    a = 1 + 2.0 ** -12
    rp, ci, va, x = _one_row([-(1 + 2.0 ** -11), a], [1.0, a])
    assert oracle.spmv_seq(rp, ci, va, x)[0] == np.float32(2.0 ** -24)
    assert oracle.spmv_seq_nofma(rp, ci, va, x)[0] == 0.0


def test_reduce_order_emulator_shape():
    # 64 entries, x = 1: a0 = 2^24, a16 = -2^24, a32 = 1, a48 = 1.
    # sequential: 2^24, 0 (at 16), 1, 2 -> 2
    # 32 lanes: lane0 = 2^24 + 1 = 2^24 (rounded), lane16 = -2^24 + 1; tree -> 1
    # 16 lanes: lane0 = ((2^24 - 2^24) + 1) + 1 -> 2
    v = np.zeros(64)
    v[0], v[16], v[32], v[48] = 2.0 ** 24, -(2.0 ** 24), 1, 1
    rp, ci, va, x = _one_row(v, np.ones(64))
    assert oracle.spmv_seq(rp, ci, va, x)[0] == 2
    assert oracle.spmv_lanes(rp, ci, va, x, lanes=32)[0] == 1
    assert oracle.spmv_lanes(rp, ci, va, x, lanes=64)[0] == 1
    assert oracle.spmv_lanes(rp, ci, va, x, lanes=16)[0] == 2


def test_adaptive_stream_sums_rounded_products():
    # block [0,3) (nnz_per_wg 4, lengths 2,1,1): row 0 is summed from rounded
    # products, so the 2^-24 of the fused form is lost.
    a = 1 + 2.0 ** -12
    rp = np.array([0, 2, 3, 4], np.uint32)
    ci = np.array([0, 1, 2, 3], np.uint32)
    va = np.array([-(1 + 2.0 ** -11), a, 1, 1], np.float32)
    x = np.array([1, a, 1, 1], np.float32)
    y = oracle.spmv_adaptive(rp, ci, va, x, nnz_per_wg=4)
    assert y.tolist() == [0.0, 1.0, 1.0]
    assert oracle.spmv_seq(rp, ci, va, x)[0] == np.float32(2.0 ** -24)


def test_adaptive_long_row_chunks_add_left_to_right():
    # one row of 3 chunks (chunk 2): partials 1+1 = 2, 2^24, 0 added left to
    # right -> 2^24 + 2 (exact).  In one workgroup the lane tree adds lane 2
    # (2^24) to lane 0 (1) first: 2^24 + 1 rounds to 2^24, then + 1 -> 2^24.
    v = [1, 1, 2.0 ** 24, 0, 0, 0]
    rp, ci, va, x = _one_row(v, np.ones(6))
    y = oracle.spmv_adaptive(rp, ci, va, x, nnz_per_wg=2, long_row=2, chunk=2)
    assert y[0] == np.float32(2.0 ** 24 + 2)
    y1 = oracle.spmv_adaptive(rp, ci, va, x, nnz_per_wg=2, long_row=100, chunk=2)
    assert y1[0] == np.float32(2.0 ** 24)


@pytest.mark.parametrize("W", [64, 300, 2040, 4096])
def test_adaptive_segmented_scan_covers_every_element(W):
    # small integers: every summation order is exact, so the emulator must
    # reproduce the exact row sums -- a dropped, repeated or misplaced element
    # (lane ranges, row boundaries, scan carries, empty rows) changes them
    rng = np.random.default_rng(W)
    lens = rng.choice([0, 1, 2, 3, 9, 40, 300, 1500, 5000], 4000,
                      p=[.15, .3, .2, .1, .1, .08, .04, .02, .01])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    nnz = int(rp[-1])
    ci = rng.integers(0, 4000, nnz).astype(np.uint32)
    va = rng.integers(-3, 4, nnz).astype(np.float32)
    x = rng.integers(-2, 3, 4000).astype(np.float32)
    exact = np.zeros(4000)
    np.add.at(exact, np.repeat(np.arange(4000), lens), va.astype(np.float64) * x[ci])
    for long_row, chunk in ((1 << 30, 256), (4096, 2048), (512, 256)):
        y = oracle.spmv_adaptive(rp, ci, va, x, W, 256, long_row, chunk)
        assert np.array_equal(y.astype(np.float64), exact)


def test_adaptive_scan_tree_order():
    # block of two rows (4 + 1 elements, nnz_per_wg 5 -> one element per lane):
    # row 0 spans lanes 0-3.  Hillis-Steele over the carried partials:
    # step 1: s1 = p0 + p1, s2 = p1 + p2; step 2: s2 = s0 + s2 = p0 + (p1 + p2);
    # lane 3 ends the row: y0 = s2 + p3.  With p = [2^24, 1, 1, 0] that is
    # 2^24 + 2 (exact); left to right would round twice to 2^24.
    v = [2.0 ** 24, 1, 1, 0, 0]
    rp = np.array([0, 4, 5], np.uint32)
    ci = np.arange(5, dtype=np.uint32)
    va = np.array(v, np.float32)
    x = np.ones(5, np.float32)
    y = oracle.spmv_adaptive(rp, ci, va, x, nnz_per_wg=5)
    assert y[0] == np.float32(2.0 ** 24 + 2)
    assert oracle.spmv_seq(rp, ci, va, x)[0] == np.float32(2.0 ** 24)


def test_stencil2d_closed_form():
    # 5-point Laplacian (4 / -1) with x = 1: y = 0 inside, 1 on edges, 2 at corners
    n = 16
    r, c, v = [], [], []
    for i in range(n):
        for j in range(n):
            row = i * n + j
            for di, dj, val in ((-1, 0, -1), (0, -1, -1), (0, 0, 4), (0, 1, -1), (1, 0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < n and 0 <= jj < n:
                    r.append(row), c.append(ii * n + jj), v.append(val)
    rp, ci, va = oracle.csr_from_triplets(n * n, n * n, r, c, v, np.float32)
    y = oracle.spmv_seq(rp, ci, va, np.ones(n * n, np.float32)).reshape(n, n)
    expect = np.zeros((n, n))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    assert np.array_equal(y, expect)


@pytest.mark.parametrize("dtype,rtol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_all_orders_agree_with_fp64_brute_force(dtype, rtol):
    r, c, v = _random_csr(400, 300, 120, seed=11)
    rp, ci, va = oracle.csr_from_triplets(400, 300, r, c, v, dtype)
    x = np.random.default_rng(2).uniform(-1, 1, 300).astype(dtype)
    dense = np.zeros((400, 300))
    dense[np.repeat(np.arange(400), np.diff(rp)), ci] = va
    y_true = dense @ x.astype(np.float64)
    y64, absdot = oracle.spmv_f64(rp, ci, va, x)
    assert np.allclose(y64, y_true, rtol=0, atol=1e-12)
    assert np.allclose(absdot, np.abs(dense) @ np.abs(x.astype(np.float64)), rtol=1e-14)
    for y in (oracle.spmv_seq(rp, ci, va, x), oracle.spmv_seq_nofma(rp, ci, va, x),
              oracle.spmv_lanes(rp, ci, va, x, 32), oracle.spmv_lanes(rp, ci, va, x, 2),
              oracle.spmv_adaptive(rp, ci, va, x, 64),
              oracle.spmv_adaptive(rp, ci, va, x, 64, long_row=64, chunk=16)):
        assert oracle.tolerance_check(y, y64, absdot, rtol) <= rtol
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    assert oracle.tolerance_check(oracle.spmv_ell(400, w, ec, ed, x), y64, absdot, rtol) <= rtol


def test_ell_order_is_sequential():
    # ELL pads with (0, col 0): the sum over the padded width equals the
    # sequential fma sum for finite x
    r, c, v = _random_csr(64, 64, 9, seed=13)
    rp, ci, va = oracle.csr_from_triplets(64, 64, r, c, v, np.float32)
    x = np.random.default_rng(4).uniform(-1, 1, 64).astype(np.float32)
    w, ed, ec = oracle.ell_from_csr(rp, ci, va)
    ye = oracle.spmv_ell(64, w, ec, ed, x)
    ys = oracle.spmv_seq(rp, ci, va, x)
    assert np.array_equal(ye, ys)


def test_host_coo_order_is_unfused_row_order():
    # the full COO (element_start 0) visits each row's entries in CSR order,
    # so the entry-order loop with rounded products is the unfused CSR loop
    r, c, v = _random_csr(300, 250, 20, seed=21)
    x = np.random.default_rng(6).uniform(-1, 1, 250)
    for dt in (np.float32, np.float64):
        rp, ci, va = oracle.csr_from_triplets(300, 250, r, c, v, dt)
        cd, cc, cr = oracle.coo_from_csr(rp, ci, va)
        yc = oracle.spmv_coo_seq(300, cr, cc, cd, x.astype(dt))
        assert np.array_equal(yc, oracle.spmv_seq_nofma(rp, ci, va, x.astype(dt)))
        assert not np.array_equal(yc, oracle.spmv_seq(rp, ci, va, x.astype(dt)))  # fused differs somewhere
    # accumulates into a given y instead of clearing it
    y0 = np.arange(3, dtype=np.float64)
    got = oracle.spmv_coo_seq(3, [2, 0], [0, 1], [2.0, 5.0], np.array([3.0, 7.0]), y=y0)
    assert got.tolist() == [35.0, 1.0, 8.0] and y0.tolist() == [0.0, 1.0, 2.0]


@pytest.mark.parametrize("w", [0, 3, 9, 40])
def test_hybrid_host_coo_integer_closed_form(w):
    # integer values and x: every partial sum is exact, so ELL(w) + tail(w)
    # equals the integer row sums whatever the split
    rng = np.random.default_rng(22)
    r, c, _ = _random_csr(200, 200, 30, seed=23)
    v = rng.integers(-8, 9, r.size).astype(np.float64)
    x = rng.integers(-4, 5, 200).astype(np.float32)
    rp, ci, va = oracle.csr_from_triplets(200, 200, r, c, v, np.float32)
    _, ed, ec = oracle.ell_from_csr(rp, ci, va, width=w)
    cd, cc, cr = oracle.coo_from_csr(rp, ci, va, element_start=w)
    y = oracle.spmv_hybrid_host_coo(200, w, ec, ed, cr, cc, cd, x)
    want = np.zeros(200, np.int64)
    np.add.at(want, np.repeat(np.arange(200), np.diff(rp)), va.astype(np.int64) * x.astype(np.int64)[ci])
    assert np.array_equal(y.astype(np.int64), want)
    if w == 40:  # wider than every row: the tail is empty, the ELL sum alone
        assert cd.size == 0 and np.array_equal(y, oracle.spmv_ell(200, w, ec, ed, x))


@pytest.mark.parametrize("k,dt", [(4, np.float32), (16, np.float32), (2, np.float64), (16, np.float64)])
def test_spmm_stream_order_against_fp64_and_integers(k, dt):
    # random rows with empty ones and rows longer than a 512-entry block
    rng = np.random.default_rng(31)
    rows, cols = 700, 900
    lens = rng.choice([0, 1, 3, 7, 8, 9, 30, 200, 511, 512, 513, 1500], rows)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = rng.integers(0, cols, int(rp[-1])).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size).astype(dt)
    X = rng.uniform(-1, 1, (cols, k)).astype(dt)
    Y = oracle.spmm_stream(rp, ci, va, X, k)
    rr = np.repeat(np.arange(rows), lens)
    prod = va.astype(np.float64)[:, None] * X.astype(np.float64)[ci]
    want = np.zeros((rows, k)); np.add.at(want, rr, prod)
    absd = np.zeros((rows, k)); np.add.at(absd, rr, np.abs(prod))
    eps = np.finfo(dt).eps
    assert np.all(np.abs(Y - want) <= 2 * lens[:, None] * eps * absd + 1e-300)
    # integer values: every order is exact
    vi = rng.integers(-3, 4, ci.size).astype(dt)
    Xi = rng.integers(-3, 4, (cols, k)).astype(dt)
    Yi = oracle.spmm_stream(rp, ci, vi, Xi, k)
    wi = np.zeros((rows, k)); np.add.at(wi, rr, vi.astype(np.float64)[:, None] * Xi.astype(np.float64)[ci])
    assert np.array_equal(Yi, wi.astype(dt))


def test_spmm_stream_cross_group_order_known_answer():
    # k = 16 FP32: 64 groups of 8 entries.  One row of 24 entries spans groups
    # 0, 1 and 2 with segment sums -2^24, 2^24 and 1: the row is
    # seg2 + seg1 + seg0 = (1 + 2^24) - 2^24 = 0 in float (1 + 2^24 rounds to
    # 2^24), where entry order would give 1
    v = np.zeros(24, np.float32)
    v[0], v[8], v[16] = -2.0 ** 24, 2.0 ** 24, 1.0
    rp = np.array([0, 24], np.uint32)
    ci = np.zeros(24, np.uint32)
    Y = oracle.spmm_stream(rp, ci, v, np.ones((1, 16), np.float32), 16)
    assert np.all(Y == 0.0)
    assert oracle.spmv_seq(rp, ci, v, np.ones(1, np.float32))[0] == 1.0


def test_spmm_stream_long_row_tree_known_answer():
    # a 576-entry row (longer than a block) at k = 16 FP32: 64 groups of 8
    # entries per 512-entry chunk, so group g adds entries 8g .. 8g + 7, then
    # 512 + 8g .. 512 + 8g + 7.  Group sums p0 = 1 (entry 0), p1 = 2^24
    # (entry 8) and p33 = -2^24 (entry 264 = 8 * 33): the halving
    # tree adds p1 + p33 = 0 first, then p0 + 0 = 1; the groups in order would
    # give (1 + 2^24) - 2^24 = 0
    v = np.zeros(576, np.float32)
    v[0], v[8], v[264] = 1.0, 2.0 ** 24, -2.0 ** 24
    rp = np.array([0, 576], np.uint32)
    ci = np.zeros(576, np.uint32)
    Y = oracle.spmm_stream(rp, ci, v, np.ones((1, 16), np.float32), 16)
    assert np.all(Y == 1.0)


def test_spmm_stream_long_row_chunk_order_known_answer():
    # within a group the chunks follow each other: group 0 of a 1,100-entry
    # row adds entry 0 (2^24), entry 512 (1) and entry 1,024 (-2^24) in that
    # order, (2^24 + 1) - 2^24 = 0 in float, where the order 0, 1,024, 512
    # would give 1; every other entry is 0
    v = np.zeros(1100, np.float32)
    v[0], v[512], v[1024] = 2.0 ** 24, 1.0, -2.0 ** 24
    rp = np.array([0, 1100], np.uint32)
    ci = np.zeros(1100, np.uint32)
    Y = oracle.spmm_stream(rp, ci, v, np.ones((1, 16), np.float32), 16)
    assert np.all(Y == 0.0)


def test_multi_thread_matches_single_thread():
    r, c, v = _random_csr(1000, 1000, 30, seed=17)
    rp, ci, va = oracle.csr_from_triplets(1000, 1000, r, c, v, np.float32)
    x = np.random.default_rng(5).uniform(-1, 1, 1000).astype(np.float32)
    y, t = oracle.spmv_multi_thread(rp, ci, va, x, threads=3)
    assert np.array_equal(y, oracle.spmv_seq(rp, ci, va, x)) and t >= 0


# ---- measurement arithmetic (common/measurement_class.{h,cpp}) ---------------------------

def test_measurement_closed_form():
    bw, fl = oracle.measurement(2e-3, 4e9, 1e9)
    assert bw == 4e9 / (2e-3 * 1e9) and fl == 1e9 / (2e-3 * 1e9)


def test_reference_byte_models():
    # CSR: nnz*s (A) + nnz*s (x per nnz) + nnz*4 + 2*rows*4 + rows*s
    assert oracle.reference_bytes(oracle.REF_BYTES_CSR, np.float32, 10, 100) == 400 + 400 + 400 + 80 + 40
    # COO: nnz*(3s + 8)
    assert oracle.reference_bytes(oracle.REF_BYTES_COO, np.float64, 10, 100) == 100 * (24 + 8)
    # ELL: n = width*rows as a 32-bit unsigned product (wraps like the reference)
    n = (70000 * 70000) % (1 << 32)
    assert oracle.reference_bytes(oracle.REF_BYTES_ELL, np.float32, 70000, 0, 70000) == n * 12 + 70000 * 4


def test_compare_error_closed_form():
    a = np.array([1.0, 2.0, 4.0])
    b = np.array([1.0, 2.0, 3.0])
    assert oracle.compare_error(a, b) == 1.0 / 14.0


def test_oracle_against_reference_measurement_sources():
    ref = oracle.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built (the reference tree is not present)")
    out = np.zeros(2)
    for el, by, op in [(1e-3, 1e9, 2e8), (3.3e-5, 12345678.0, 4.2e6), (0.25, 0.0, 0.0)]:
        ref.ref_measurement(b"x", el, by, op, out)
        assert tuple(out) == oracle.measurement(el, by, op)
