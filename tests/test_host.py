"""CPU-only checks of libmfp's host side: the C ABI loads and exports every
declared symbol, the seeded generator, and the reference-interface records
(measurement_class, compare_results) against the reference's own sources."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import matrix_format_performance_amd as mfp
from matrix_format_performance_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(mfp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers the same set
    assert sorted(L.SIGNATURES) == names


def test_no_gpu_calls_fail_cleanly():
    # without a device every entry point returns a status; nothing crashes
    import ctypes
    lib = L.load()
    p = ctypes.c_void_p()
    st = lib.mfp_malloc(ctypes.byref(p), 16)
    assert st in (L.MFP_OK, L.MFP_ERR_NODEV, L.MFP_ERR_HIP)
    if st == L.MFP_OK:
        lib.mfp_free(p)


def test_kernel_times_rejects_null_arguments():
    # argument checks come before any device work
    import ctypes
    lib = L.load()
    ms = (ctypes.c_float * 8)()
    n = ctypes.c_int(-1)
    assert lib.mfp_plan_kernel_times(None, None, None, None, ms, 8, ctypes.byref(n)) == L.MFP_ERR_ARG
    assert n.value == -1
    assert b"null" in lib.mfp_last_error()


def test_package_fails_loudly_without_library(tmp_path):
    # a copy of the package without libmfp.so must raise on first use
    import shutil
    dst = tmp_path / "matrix_format_performance_amd"
    shutil.copytree(os.path.join(ROOT, "matrix_format_performance_amd"), dst,
                    ignore=shutil.ignore_patterns("*.so", "_build", "__pycache__"))
    code = ("import matrix_format_performance_amd as m\n"
            "try:\n m.generate_vector(0, 4)\nexcept ImportError as e:\n print('IMPORTERROR')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=tmp_path, capture_output=True, text=True)
    assert "IMPORTERROR" in r.stdout, r.stdout + r.stderr


# ---- generator ---------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_stencil2d_closed_form(n):
    rp, ci, va = mfp.generate(mfp.GenSpec("stencil2d", values="laplace", n=n))
    assert rp.size - 1 == n * n and ci.size == 5 * n * n - 4 * n
    x = np.ones(n * n, np.float32)
    # row sums: 4 - (#neighbours) = 0 interior, 1 edges, 2 corners
    rows = np.repeat(np.arange(n * n), np.diff(rp))
    y = np.zeros(n * n)
    np.add.at(y, rows, va.astype(np.float64) * x[ci])
    grid = y.reshape(n, n)
    if n >= 3:
        assert grid[1:-1, 1:-1].max() == 0 and grid[0, 1] == 1 and grid[0, 0] == 2


@pytest.mark.parametrize("n", [1, 3, 10])
def test_stencil3d_closed_form(n):
    rp, ci, va = mfp.generate(mfp.GenSpec("stencil3d", values="normal", n=n, seed=3))
    assert rp.size - 1 == n ** 3 and ci.size == (3 * n - 2) ** 3
    for r in range(0, n ** 3, max(1, n ** 3 // 50)):
        seg = ci[rp[r]:rp[r + 1]].astype(np.int64)
        assert np.all(np.diff(seg) > 0) and r in seg


def test_full_size_nnz_counts():
    # the benchmark configurations: 2D 256^2 and 3D 128^3 (closed forms)
    rp, ci, _ = mfp.workload("stencil2d_256")
    assert ci.size == 326_656
    rp, ci, _ = mfp.workload("stencil3d_128")
    assert ci.size == 55_742_968


def test_powerlaw_rows_distinct_sorted_and_deterministic():
    spec = mfp.GenSpec("powerlaw", values="uniform", rows=20000, nnz_lo=1, nnz_hi=5000, seed=9)
    rp, ci, va = mfp.generate(spec)
    rp2, ci2, va2 = mfp.generate(spec)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(va, va2)
    lens = np.diff(rp.astype(np.int64))
    assert lens.min() >= 1 and lens.max() <= 5000
    # P(len >= k) = 1/k for the Pareto (density x^-2) floor sampler
    for k in (2, 10, 100):
        assert abs((lens >= k).mean() - 1.0 / k) < 4 * np.sqrt(1.0 / k / lens.size) + 1e-3
    for r in range(0, 20000, 97):
        seg = ci[rp[r]:rp[r + 1]].astype(np.int64)
        assert np.all(np.diff(seg) > 0)
    assert np.all(np.abs(va) <= 1.0)


def test_uniform_family_and_bandwidth_window():
    spec = mfp.GenSpec("uniform", values="uniform", rows=5000, nnz_lo=4, nnz_hi=64, bandwidth=0.01, seed=2)
    rp, ci, _ = mfp.generate(spec)
    lens = np.diff(rp.astype(np.int64))
    assert lens.min() >= 4 and lens.max() <= 64
    for r in range(0, 5000, 37):
        seg = ci[rp[r]:rp[r + 1]].astype(np.int64)
        assert np.all(np.diff(seg) > 0)
        assert seg.max() - seg.min() < max(64, int(0.01 * 5000)) + 1


def test_vector_generator_range_and_grid():
    x = mfp.generate_vector(0, 100000)
    assert x.min() >= -1 and x.max() < 1
    assert np.all((x.astype(np.float64) * 2 ** 23) == np.round(x.astype(np.float64) * 2 ** 23))


# ---- measurement_class and compare_results vs the reference's sources ----------------------

def test_measurement_class_semantics():
    m = mfp.measurement_class("GPU CSR", 2e-3, 4e9, 1e9)
    assert m.get_effective_bandwidth() == 4e9 / (2e-3 * 1e9)
    acc = mfp.measurement_class()
    for el in (1e-3, 3e-3):
        acc += mfp.measurement_class("GPU ELL", el, 1e6, 2e6)
    acc.finalize()
    assert acc.get_format() == "GPU ELL"
    assert acc.get_elapsed() == pytest.approx(2e-3, rel=1e-15)


def test_measurement_class_matches_reference_build():
    import oracle
    ref = oracle.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    el = np.array([1e-3, 2.5e-5, 7e-2])
    by = np.array([1e9, 3e7, 5.5e10])
    op = np.array([2e8, 1e6, 4e9])
    out = np.zeros(3)
    ref.ref_measurement_mean(3, el, by, op, out)
    acc = mfp.measurement_class()
    for i in range(3):
        acc += mfp.measurement_class("x", el[i], by[i], op[i])
    acc.finalize()
    assert (acc.get_elapsed(), acc.get_effective_bandwidth(), acc.get_computational_throughput()) == tuple(out)


def _capture_fd2(fn):
    import tempfile
    sys.stderr.flush()
    saved = os.dup(2)
    with tempfile.TemporaryFile() as tf:
        os.dup2(tf.fileno(), 2)
        try:
            fn()
        finally:
            os.dup2(saved, 2)
            os.close(saved)
        tf.seek(0)
        return tf.read().decode()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_compare_results_text_matches_reference_build(dtype):
    import oracle
    ref = oracle.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, 1000).astype(dtype)
    for a in (b.copy(), b + dtype(1e-3), np.where(np.arange(1000) == 417, b + 1, b).astype(dtype)):
        a = np.ascontiguousarray(a, dtype)
        buf = C.create_string_buffer(4096)
        ref.ref_compare_results(0 if dtype == np.float32 else 1, 1000, a.ctypes.data_as(C.c_void_p),
                                b.ctypes.data_as(C.c_void_p), buf, 4096)
        mine = _capture_fd2(lambda: mfp.compare_results(1000, a, b))
        assert mine == buf.value.decode()
        # and the printed error equals the oracle's value
        if mine:
            assert float(mine.split()[1]) == pytest.approx(oracle.compare_error(a, b), rel=1e-5)


COMPAT_PROGRAM = r'''
#include "mfp_compat.hpp"
// every class and function template of the reference interface, for float and double
template <typename T> void all(const matrix_market::matrix_class &mtx)
{
  csr_matrix_class<T> csr(mtx);
  ell_matrix_class<T> ell(csr), ell_w(csr, 3u);
  coo_matrix_class<T> coo(csr), coo_tail(csr, 2u);
  hybrid_matrix_class<T> hyb(csr);
  hyb.allocate(csr, 0.2);
  scoo_matrix_class<T> scoo(256, 65536, csr);
  scoo_matrix_class scoo_deduced(256, 65536, csr);  // main.cpp:223 spells it without <T>
  (void)scoo_deduced.slice_size;
  resizable_gpu_memory<T> A, A2, x, y;
  resizable_gpu_memory<unsigned int> c, r, c2;
  T *h = nullptr;
  measurement_class m = gpu_csr_spmv<T>(csr, A, c, r, x, y, h, h);
  m += gpu_csr_vector_spmv<T>(csr, A, c, r, x, y, h, h);
  m += gpu_csr_adaptive_spmv<T>(csr, A, c, r, x, y, h, h);
  m += gpu_ell_spmv<T>(ell, A, c, x, y, h, h);
  m += gpu_coo_spmv<T>(coo, A, c, r, x, y, h, h);
  m += gpu_scoo_spmv<T>(true, scoo, A, r, c, x, y, h, h);
  m += gpu_hybrid_spmv<T>(hyb, A, A2, c, c2, r, x, y, h, h);
  m += gpu_hybrid_atomic_spmv<T>(hyb, A, A2, c, c2, r, x, y, h, h);
  m += cpu_csr_spmv_single_thread_naive(csr, h, h);
  m += cpu_csr_spmv_single_thread_naive_with_reduce_order(csr, h, h);
  m += cpu_csr_spmv_multi_thread_naive(csr, h, h);
  m.finalize();
  compare_results<T>(1u, h, h);
  (void)ell_matrix_class<T>::estimate_size(csr);
  (void)get_rows_statistics(csr.meta, csr.row_ptr.get());
  (void)fill_row_blocks(csr.meta.rows_count, csr.row_ptr.get());
  (void)m.get_format(); (void)m.get_elapsed(); (void)m.get_effective_bandwidth();
  (void)m.get_computational_throughput(); (void)ell.get_matrix_size(); (void)coo.get_matrix_size();
}
template void all<float>(const matrix_market::matrix_class &);
template void all<double>(const matrix_market::matrix_class &);
'''


def test_compat_header_compiles_for_float_and_double(tmp_path):
    src = tmp_path / "compat.cpp"
    src.write_text(COMPAT_PROGRAM)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-Wno-unused-parameter", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_matrix_market_reader_parses_without_a_device(tmp_path):
    prog = tmp_path / "mm.cpp"
    prog.write_text(r'''
#include "mfp_compat.hpp"
#include <iostream>
#include <sstream>
int main() {
  std::istringstream a("%%MatrixMarket matrix coordinate real general\n% c\n\n3 4 3\n3 4 -1.5\n1 1 2\n2 3 1e-3\n");
  matrix_market::reader ra(a);
  const auto &m = ra.matrix();
  std::cout << m.meta.rows_count << " " << m.meta.cols_count << " " << m.meta.non_zero_count;
  for (unsigned i = 0; i < 3; i++) std::cout << " " << m.get_row_ids()[i] << "," << m.get_col_ids()[i] << "," << m.get_dbl_data()[i];
  std::istringstream b("%%MatrixMarket matrix coordinate pattern symmetric\n2 2 1\n2 1\n");
  matrix_market::reader rb(b);
  std::cout << " sym=" << (rb.matrix().meta.matrix_storage_scheme == matrix_market::matrix_class::storage_scheme::symmetric)
            << " v=" << rb.matrix().get_dbl_data()[0];
  try { std::istringstream c("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n"); matrix_market::reader rc(c); }
  catch (const std::runtime_error &e) { std::cout << " err=" << e.what(); }
  std::cout << "\n";
}
''')
    exe = tmp_path / "mm"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30).stdout
    assert out.strip() == ("3 4 3 2,3,-1.5 0,0,2 1,2,0.001 sym=1 v=1 err=MatrixMarket: index out of range")


def test_bench_app_usage():
    app = os.path.join(ROOT, "matrix_format_performance_amd", "mfp_bench")
    if not os.path.exists(app):
        pytest.skip("mfp_bench not built")
    r = subprocess.run([app], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_bench_refuses_several_gpus_before_touching_one():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line for an unsupported run


def test_bench_algorithmic_bytes_closed_form():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    # 2D 5-point 256^2: 65,536 rows, 326,656 nnz, FP32
    rows, nnz = 65536, 5 * 65536 - 4 * 256
    assert nnz == 326656
    assert bench.b_alg(rows, rows, nnz, 4, 4) == nnz * 8 + (rows + 1) * 4 + rows * 4 + rows * 4 == 3399684
    # 16-bit values still count 4-byte columns
    assert bench.b_alg(rows, rows, nnz, 2, 4) == 3399684 - 2 * nnz
