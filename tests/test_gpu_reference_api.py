"""The reference's entry points, driven the way main.cpp:82-283 drives them:
host formats built by the converters, shared grow-only device buffers,
x = 1, y copied into the host x buffer, compared with the CPU answer."""
import numpy as np
import pytest

import oracle
from gpu_cases import announce

pytestmark = pytest.mark.gpu


def _matrix(mfp, dtype):
    rp, ci, va = mfp.generate(mfp.GenSpec("uniform", values="uniform", rows=20000, nnz_lo=1, nnz_hi=90,
                                          bandwidth=0.05, seed=11), dtype)
    rows = np.repeat(np.arange(rp.size - 1, dtype=np.uint32), np.diff(rp))
    perm = np.random.default_rng(0).permutation(rows.size)
    return mfp.matrix_class(20000, 20000, rows[perm], ci[perm], va.astype(np.float64)[perm])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_main_cpp_sequence(gpu, dtype):
    mfp = gpu
    announce(f"reference API sequence dtype={np.dtype(dtype)}")
    mtx = _matrix(mfp, dtype)
    csr = mfp.csr_matrix_class(mtx, dtype)
    ell = mfp.ell_matrix_class(csr)
    coo = mfp.coo_matrix_class(csr)
    rows = csr.meta.rows_count
    ones = np.ones(csr.meta.cols_count, dtype)
    reference_answer = oracle.spmv_seq(csr.row_ptr, csr.columns, csr.data, ones)
    y64, absdot = oracle.spmv_f64(csr.row_ptr, csr.columns, csr.data, ones)
    rtol = 1e-5 if dtype == np.float32 else 1e-12

    A, x_gpu, y = (mfp.resizable_gpu_memory(dtype) for _ in range(3))
    col_ids, row_ptr = mfp.resizable_gpu_memory(np.uint32), mfp.resizable_gpu_memory(np.uint32)
    x = np.zeros(max(rows, csr.meta.cols_count), dtype)

    def check(m, label, exact=None):
        assert m.get_format() == label
        assert m.get_elapsed() > 0
        got = x[:rows]
        if exact is not None:
            assert np.array_equal(got, exact), label
        assert oracle.tolerance_check(got, reference_answer, absdot, rtol) <= rtol, label

    m = mfp.gpu_csr_spmv(csr, A, col_ids, row_ptr, x_gpu, y, x, reference_answer)
    check(m, "GPU CSR", exact=reference_answer)
    # the CSR byte model (gpu_matrix_multiplier.cu:120-124)
    assert m.load_store_bytes == oracle.reference_bytes(oracle.REF_BYTES_CSR, dtype, rows, csr.meta.non_zero_count)
    m = mfp.gpu_csr_vector_spmv(csr, A, col_ids, row_ptr, x_gpu, y, x, reference_answer)
    # the reference's warp: 32 lanes per row (gpu_matrix_multiplier.cu:331-362)
    check(m, "GPU CSR (vector)", exact=oracle.spmv_lanes(csr.row_ptr, csr.columns, csr.data, ones, 32))
    m = mfp.gpu_csr_adaptive_spmv(csr, A, col_ids, row_ptr, x_gpu, y, x, reference_answer)
    W = mfp.engine.DEFAULT_NNZ_PER_WG[np.dtype(dtype)]
    check(m, "GPU CSR-Adaptive", exact=oracle.spmv_adaptive(csr.row_ptr, csr.columns, csr.data, ones, W, 256,
                                                            mfp.engine.DEFAULT_LONG_ROW, mfp.engine.DEFAULT_CHUNK))
    assert m.get_effective_bandwidth() == 0.0  # csr_adaptive_spmv.cu:320 records 0 bytes
    m = mfp.gpu_ell_spmv(ell, A, col_ids, x_gpu, y, x, reference_answer)
    check(m, "GPU ELL", exact=reference_answer)
    m = mfp.gpu_coo_spmv(coo, A, col_ids, row_ptr, x_gpu, y, x, reference_answer)
    check(m, "GPU COO")
    info = mfp.device_info()
    scoo = mfp.scoo_matrix_class(info["cu_count"], info["lds_per_block"], csr)
    m = mfp.gpu_scoo_spmv(True, scoo, A, col_ids, row_ptr, x_gpu, y, x, reference_answer)
    check(m, "GPU SCOO")
    A_coo, col_ids_coo = mfp.resizable_gpu_memory(dtype), mfp.resizable_gpu_memory(np.uint32)
    hyb = mfp.hybrid_matrix_class(csr)
    hyb.allocate(csr, 0.2)
    m = mfp.gpu_hybrid_spmv(hyb, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x, reference_answer)
    check(m, "GPU Hybrid (TODO)")
    m = mfp.gpu_hybrid_atomic_spmv(hyb, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x,
                                   reference_answer)
    check(m, "GPU Hybrid (atomic)")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_hybrid_host_coo_tail(gpu, dtype):
    # main.cpp:247-279 (switched off there): the ELL widths of its percent
    # loop, the host-COO hybrid against ELL slot order + the host's entry
    # order, bit for bit
    mfp = gpu
    announce(f"reference API host-COO hybrid dtype={np.dtype(dtype)}")
    csr = mfp.csr_matrix_class(_matrix(mfp, dtype), dtype)
    rows = csr.meta.rows_count
    ones = np.ones(csr.meta.cols_count, dtype)
    reference_answer = oracle.spmv_seq(csr.row_ptr, csr.columns, csr.data, ones)
    A, x_gpu, y, tmp = (mfp.resizable_gpu_memory(dtype) for _ in range(4))
    col_ids = mfp.resizable_gpu_memory(np.uint32)
    x = np.zeros(max(rows, csr.meta.cols_count), dtype)
    cpu_y = np.full(rows, 7.0, dtype)  # zeroed by the call
    hyb = mfp.hybrid_matrix_class(csr)
    for percent in (0.0, 0.35, 0.7):
        hyb.allocate(csr, percent)
        e, c = hyb.ell_matrix, hyb.coo_matrix
        assert e.elements_in_rows == oracle.hybrid_width(csr.row_ptr, percent)
        m = mfp.gpu_hybrid_cpu_coo_spmv(hyb, A, col_ids, x_gpu, y, tmp, cpu_y, x, reference_answer)
        assert m.get_format() == "GPU Hybrid (CPU COO TODO)" and m.get_elapsed() > 0
        assert m.load_store_bytes == 1.0 and m.operations_count == 1.0  # gpu_matrix_multiplier.cu:1209-1213
        tail = oracle.spmv_coo_seq(rows, c.rows, c.cols, c.data, ones)
        assert np.array_equal(cpu_y, tail)
        want = oracle.spmv_ell(rows, e.elements_in_rows, e.columns, e.data, ones) + tail
        assert np.array_equal(x[:rows], want), percent
        assert tmp.size_bytes() >= rows * np.dtype(dtype).itemsize


def test_hybrid_host_coo_rejects_mismatched_parts(gpu):
    mfp = gpu
    csr = mfp.csr_matrix_class(_matrix(mfp, np.float32), np.float32)
    hyb = mfp.hybrid_matrix_class(csr)
    hyb.allocate(csr, 0.2)
    hyb.coo_matrix.rows = hyb.coo_matrix.rows.copy()
    if hyb.coo_matrix.rows.size:
        hyb.coo_matrix.rows[-1] = csr.meta.rows_count  # one past the last row
        A, x_gpu, y, tmp = (mfp.resizable_gpu_memory(np.float32) for _ in range(4))
        col_ids = mfp.resizable_gpu_memory(np.uint32)
        with pytest.raises(Exception, match="out of range"):
            mfp.gpu_hybrid_cpu_coo_spmv(hyb, A, col_ids, x_gpu, y, tmp, np.zeros(csr.meta.rows_count, np.float32),
                                        np.zeros(csr.meta.rows_count, np.float32), None)


def test_vector_prefill_sentinel_and_averaging(gpu):
    mfp = gpu
    announce("reference API averaging")
    mtx = _matrix(mfp, np.float32)
    csr = mfp.csr_matrix_class(mtx, np.float32)
    A, x_gpu, y = (mfp.resizable_gpu_memory(np.float32) for _ in range(3))
    col_ids, row_ptr = mfp.resizable_gpu_memory(np.uint32), mfp.resizable_gpu_memory(np.uint32)
    x = np.zeros(csr.meta.rows_count, np.float32)
    result = mfp.measurement_class()
    for _ in range(20):  # main.cpp:139-149
        result += mfp.gpu_csr_vector_spmv(csr, A, col_ids, row_ptr, x_gpu, y, x, None)
    result.finalize()
    assert result.get_format() == "GPU CSR (vector)" and result.get_elapsed() > 0
    assert not np.any(x == 24.0)  # every row overwrote the 24.0 pre-fill


def test_non_general_storage_is_rejected(gpu):
    mtx = gpu.matrix_class(2, 2, [0, 1], [0, 1], [1.0, 1.0], storage="symmetric")
    with pytest.raises(RuntimeError, match="Only general matrices are supported"):
        gpu.csr_matrix_class(mtx)
