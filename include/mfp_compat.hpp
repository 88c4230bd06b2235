/*
 * C++ source compatibility with the reference's host interface, built on the
 * C ABI of libmfp (mfp.h).  Code written against the reference's headers
 *   common/matrix_converter.h, common/scoo_matrix_class.h,
 *   common/measurement_class.h, gpu/resizable_gpu_memory.h,
 *   gpu/gpu_matrix_multiplier.h, gpu/csr_adaptive_spmv.h, gpu/scoo_spmv.h,
 *   cpu/cpu_matrix_multiplier.h
 * compiles against this one header with the same class and function
 * templates (float and double), the same argument meaning and the same
 * record labels.  Layouts are built by the library's device converters;
 * every gpu_* call runs the library's HIP kernels.
 *
 * Errors: converters throw std::runtime_error, as the reference's do
 * (matrix_converter.cpp:19-20); gpu_* calls also throw on a library error
 * (the reference ignores CUDA errors).  Result mismatches are printed to
 * stderr by compare_results and never throw.
 *
 * The cpu_* functions are the reference's CPU measurements
 * (cpu/cpu_matrix_multiplier.cpp:20-61, 170-227, 251-308), kept for its
 * driver's "CPU CSR" rows and reference answers; no gpu_* call uses them.
 * cpu_csr_spmv_mkl has no counterpart (MKL is not part of this stack).
 */
#ifndef MFP_COMPAT_HPP
#define MFP_COMPAT_HPP

#include "mfp.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <istream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace mfp_compat {

template <typename T> constexpr int dtype_of()
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "the reference API is instantiated for float and double");
  return std::is_same<T, float>::value ? MFP_F32 : MFP_F64;
}

inline void check(int status, const char *what)
{
  if (status != MFP_OK)
    throw std::runtime_error(std::string(what) + ": " + mfp_status_string(status) + " (" +
                             mfp_last_error() + ")");
}

inline void init_once()
{
  static const int status = mfp_init(0);
  check(status, "mfp_init");
}

} // namespace mfp_compat

/* ---- MatrixMarket triplets: the matrix_market::matrix_class / reader API
 * the reference takes from its matrix_reader submodule (used at
 * matrix_converter.cpp:29-33 and main.cpp:442-445).  That submodule is not
 * part of the reference copy, so the reader follows the published
 * MatrixMarket exchange format: "coordinate" real / integer / pattern
 * (pattern entries read as 1.0), 1-based indices made 0-based, entries kept
 * in file order, storage scheme reported from the header. */
namespace matrix_market {

class matrix_class {
public:
  enum class storage_scheme { general, symmetric, skew_symmetric, hermitian };

  struct matrix_meta {
    unsigned int rows_count = 0;
    unsigned int cols_count = 0;
    unsigned int non_zero_count = 0;
    storage_scheme matrix_storage_scheme = storage_scheme::general;
  };

  matrix_class() = default;
  matrix_class(matrix_meta m, std::vector<unsigned int> rows, std::vector<unsigned int> cols,
               std::vector<double> data)
    : meta(m), row_ids(std::move(rows)), col_ids(std::move(cols)), values(std::move(data))
  {
    meta.non_zero_count = (unsigned int)values.size();
  }

  matrix_meta meta;

  const unsigned int *get_row_ids() const { return row_ids.data(); }
  const unsigned int *get_col_ids() const { return col_ids.data(); }
  const double *get_dbl_data() const { return values.data(); }

private:
  std::vector<unsigned int> row_ids, col_ids;
  std::vector<double> values;
};

class reader {
public:
  explicit reader(std::istream &is)
  {
    std::string line;
    if (!std::getline(is, line)) throw std::runtime_error("MatrixMarket: empty stream");
    std::istringstream banner(line);
    std::string tag, object, fmt, field, symmetry;
    banner >> tag >> object >> fmt >> field >> symmetry;
    auto lower = [](std::string s) {
      for (auto &c : s) c = (char)std::tolower((unsigned char)c);
      return s;
    };
    if (tag != "%%MatrixMarket" || lower(object) != "matrix" || lower(fmt) != "coordinate")
      throw std::runtime_error("MatrixMarket: only 'matrix coordinate' files are supported");
    field = lower(field);
    symmetry = lower(symmetry);
    const bool pattern = field == "pattern";
    if (!pattern && field != "real" && field != "integer" && field != "double")
      throw std::runtime_error("MatrixMarket: unsupported field '" + field + "'");
    matrix_class::matrix_meta meta;
    if (symmetry == "general") meta.matrix_storage_scheme = matrix_class::storage_scheme::general;
    else if (symmetry == "symmetric") meta.matrix_storage_scheme = matrix_class::storage_scheme::symmetric;
    else if (symmetry == "skew-symmetric")
      meta.matrix_storage_scheme = matrix_class::storage_scheme::skew_symmetric;
    else if (symmetry == "hermitian") meta.matrix_storage_scheme = matrix_class::storage_scheme::hermitian;
    else throw std::runtime_error("MatrixMarket: unsupported symmetry '" + symmetry + "'");
    do {
      if (!std::getline(is, line)) throw std::runtime_error("MatrixMarket: missing size line");
    } while (line.empty() || line[0] == '%');
    unsigned long long r = 0, c = 0, n = 0;
    {
      std::istringstream sz(line);
      if (!(sz >> r >> c >> n)) throw std::runtime_error("MatrixMarket: bad size line");
    }
    if (r > 0xFFFFFFFFull || c > 0xFFFFFFFFull || n > 0xFFFFFFFFull)
      throw std::runtime_error("MatrixMarket: sizes exceed 32-bit indices");
    meta.rows_count = (unsigned int)r;
    meta.cols_count = (unsigned int)c;
    std::vector<unsigned int> ri, ci;
    std::vector<double> va;
    ri.reserve(n);
    ci.reserve(n);
    va.reserve(n);
    while (va.size() < n && std::getline(is, line)) {
      if (line.empty() || line[0] == '%') continue;
      std::istringstream el(line);
      unsigned long long i = 0, j = 0;
      double v = 1.0;
      if (!(el >> i >> j) || (!pattern && !(el >> v)))
        throw std::runtime_error("MatrixMarket: bad entry '" + line + "'");
      if (i < 1 || j < 1 || i > r || j > c) throw std::runtime_error("MatrixMarket: index out of range");
      ri.push_back((unsigned int)(i - 1));
      ci.push_back((unsigned int)(j - 1));
      va.push_back(v);
    }
    if (va.size() != n) throw std::runtime_error("MatrixMarket: fewer entries than declared");
    m = matrix_class(meta, std::move(ri), std::move(ci), std::move(va));
  }

  const matrix_class &matrix() const { return m; }

private:
  matrix_class m;
};

} // namespace matrix_market

/* ---- resizable_gpu_memory (gpu/resizable_gpu_memory.h:13-44) ------------ */
template <typename T>
class resizable_gpu_memory {
public:
  resizable_gpu_memory()
  {
    mfp_compat::init_once();
    mfp_compat::check(mfp_buffer_create(&buf), "mfp_buffer_create");
  }
  ~resizable_gpu_memory() { mfp_buffer_destroy(buf); }
  resizable_gpu_memory(const resizable_gpu_memory &) = delete;
  resizable_gpu_memory &operator=(const resizable_gpu_memory &) = delete;

  void clear() { mfp_compat::check(mfp_buffer_clear(buf), "mfp_buffer_clear"); }
  void resize(size_t new_size) { mfp_compat::check(mfp_buffer_resize(buf, new_size * sizeof(T)), "resize"); }
  T *get() { return static_cast<T *>(mfp_buffer_ptr(buf)); }
  const T *get() const { return static_cast<const T *>(mfp_buffer_ptr(buf)); }

  mfp_buffer handle() const { return buf; }

private:
  mfp_buffer buf = nullptr;
};

/* ---- measurement_class (common/measurement_class.h:11-58, .cpp:7-16) ---- */
class measurement_class {
  const size_t giga = 1E+9;

public:
  measurement_class() = default;
  measurement_class(std::string format, double elapsed_arg, double load_store_bytes, double operations_count)
    : elapsed(elapsed_arg), effective_bandwidth(load_store_bytes / (elapsed_arg * giga)),
      computational_throughput(operations_count / (elapsed_arg * giga)), matrix_format(std::move(format))
  {
  }
  measurement_class(const measurement_class &o)
    : elapsed(o.elapsed), effective_bandwidth(o.effective_bandwidth),
      computational_throughput(o.computational_throughput), matrix_format(o.matrix_format),
      measurements_count(o.measurements_count)
  {
  }
  measurement_class &operator=(const measurement_class &o)
  {
    elapsed = o.elapsed;
    effective_bandwidth = o.effective_bandwidth;
    computational_throughput = o.computational_throughput;
    matrix_format = o.matrix_format;
    measurements_count = o.measurements_count;
    return *this;
  }
  explicit measurement_class(const mfp_measurement &m)
    : elapsed(m.elapsed), effective_bandwidth(m.effective_bandwidth),
      computational_throughput(m.computational_throughput), matrix_format(m.format)
  {
  }

  double get_elapsed() const { return elapsed; }
  double get_effective_bandwidth() const { return effective_bandwidth; }
  double get_computational_throughput() const { return computational_throughput; }
  const std::string &get_format() const { return matrix_format; }

  measurement_class &operator+=(const measurement_class &rhs)
  {
    elapsed += rhs.get_elapsed();
    effective_bandwidth += rhs.get_effective_bandwidth();
    computational_throughput += rhs.get_computational_throughput();
    matrix_format = rhs.get_format();
    measurements_count++;
    return *this;
  }

  void finalize()
  {
    if (measurements_count) {
      elapsed /= measurements_count;
      effective_bandwidth /= measurements_count;
      computational_throughput /= measurements_count;
    }
  }

private:
  double elapsed{};
  double effective_bandwidth{};
  double computational_throughput{};
  std::string matrix_format;
  unsigned int measurements_count{};
};

/* compare_results (common/measurement_class.h:60-89) */
template <typename data_type>
void compare_results(unsigned int y_size, const data_type *a, const data_type *b)
{
  mfp_compat::check(mfp_compare_results(mfp_compat::dtype_of<data_type>(), y_size, a, b), "compare_results");
}

/* ---- converters (common/matrix_converter.h:12-96) ------------------------- */
struct matrix_rows_statistic {
  unsigned int min_elements_in_rows{};
  unsigned int max_elements_in_rows{};
  unsigned int avg_elements_in_rows{};
  double elements_in_rows_std_deviation{};
};

namespace mfp_compat {

// a device matrix uploaded for one conversion; freed on scope exit
struct device_csr {
  mfp_matrix m = nullptr;
  device_csr(int dtype, unsigned rows, unsigned cols, uint64_t nnz, const unsigned *rp, const unsigned *ci,
             const void *va)
  {
    init_once();
    check(mfp_matrix_from_host_csr(&m, dtype, rows, cols, nnz, rp, ci, va), "mfp_matrix_from_host_csr");
  }
  ~device_csr() { mfp_matrix_destroy(m); }
  device_csr(const device_csr &) = delete;
  device_csr &operator=(const device_csr &) = delete;
};

struct device_plan {
  mfp_plan p = nullptr;
  device_plan(mfp_matrix m, int format, const mfp_tuning &t)
  {
    check(mfp_plan_create(&p, m, format, &t, nullptr), "mfp_plan_create");
  }
  ~device_plan() { mfp_plan_destroy(p); }
  device_plan(const device_plan &) = delete;
  device_plan &operator=(const device_plan &) = delete;

  mfp_plan_info info() const
  {
    mfp_plan_info i;
    check(mfp_plan_get_info(p, &i), "mfp_plan_get_info");
    return i;
  }
  template <typename T> std::unique_ptr<T[]> array(int which, uint64_t *count_out = nullptr) const
  {
    const void *d = nullptr;
    uint64_t n = 0;
    check(mfp_plan_array(p, which, &d, &n), "mfp_plan_array");
    std::unique_ptr<T[]> h(new T[std::max<uint64_t>(n, 1)]);
    if (n) check(mfp_memcpy_d2h(h.get(), d, n * sizeof(T)), "mfp_memcpy_d2h");
    if (count_out) *count_out = n;
    return h;
  }
};

inline mfp_tuning default_tuning(mfp_matrix m, int format)
{
  mfp_tuning t;
  check(mfp_tuning_default(m, format, &t), "mfp_tuning_default");
  return t;
}

inline void require_general(const matrix_market::matrix_class::matrix_meta &meta)
{
  if (meta.matrix_storage_scheme != matrix_market::matrix_class::storage_scheme::general)
    throw std::runtime_error("Only general matrices are supported");
}

} // namespace mfp_compat

inline matrix_rows_statistic get_rows_statistics(const matrix_market::matrix_class::matrix_meta &meta,
                                                 const unsigned int *row_ptr)
{
  const uint64_t nnz = row_ptr[meta.rows_count];
  std::vector<unsigned int> cols(std::max<uint64_t>(nnz, 1), 0u);
  std::vector<float> vals(std::max<uint64_t>(nnz, 1), 0.f);
  mfp_compat::device_csr d(MFP_F32, meta.rows_count, std::max(1u, meta.cols_count), nnz, row_ptr, cols.data(),
                           vals.data());
  mfp_row_stats s;
  mfp_compat::check(mfp_matrix_row_stats(d.m, &s), "mfp_matrix_row_stats");
  return {s.min_elements_in_rows, s.max_elements_in_rows, s.avg_elements_in_rows,
          s.elements_in_rows_std_deviation};
}

template <typename data_type> class hybrid_matrix_class;

/* csr_matrix_class (matrix_converter.cpp:15-87), converted on the device */
template <typename data_type>
class csr_matrix_class {
public:
  csr_matrix_class() = delete;
  explicit csr_matrix_class(const matrix_market::matrix_class &matrix, bool row_ptr_only = false)
    : meta(matrix.meta)
  {
    mfp_compat::require_general(meta);
    mfp_compat::init_once();
    mfp_matrix m = nullptr;
    mfp_compat::check(mfp_matrix_from_triplets(&m, mfp_compat::dtype_of<data_type>(), meta.rows_count,
                                               meta.cols_count, meta.non_zero_count, matrix.get_row_ids(),
                                               matrix.get_col_ids(), matrix.get_dbl_data(), 0, nullptr),
                      "mfp_matrix_from_triplets");
    const size_t n = std::max<size_t>(meta.non_zero_count, 1);
    row_ptr.reset(new unsigned int[meta.rows_count + 1]);
    std::unique_ptr<unsigned int[]> c(new unsigned int[n]);
    std::unique_ptr<data_type[]> v(new data_type[n]);
    const int st = mfp_matrix_download(m, row_ptr.get(), c.get(), v.get());
    mfp_matrix_destroy(m);
    mfp_compat::check(st, "mfp_matrix_download");
    if (!row_ptr_only) {
      columns = std::move(c);
      data = std::move(v);
    }
  }

  const matrix_market::matrix_class::matrix_meta meta;

  size_t get_matrix_size() const { return meta.non_zero_count; }

  // the library's device matrix of this CSR (caller destroys it)
  mfp_compat::device_csr *to_device() const
  {
    return new mfp_compat::device_csr(mfp_compat::dtype_of<data_type>(), meta.rows_count,
                                      std::max(1u, meta.cols_count), meta.non_zero_count, row_ptr.get(),
                                      columns.get(), data.get());
  }

public:
  std::unique_ptr<data_type[]> data;
  std::unique_ptr<unsigned int[]> columns;
  std::unique_ptr<unsigned int[]> row_ptr;
};

/* ell_matrix_class (matrix_converter.cpp:129-203): column-major
 * data[row + k rows], padding value 0 with column 0; width = longest row, or
 * the given width with longer rows truncated */
template <typename data_type>
class ell_matrix_class {
public:
  ell_matrix_class() = delete;
  explicit ell_matrix_class(csr_matrix_class<data_type> &matrix) : meta(matrix.meta)
  {
    mfp_compat::require_general(meta);
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_tuning t = mfp_compat::default_tuning(d->m, MFP_ELL);
    mfp_compat::device_plan p(d->m, MFP_ELL, t);
    take(p);
  }
  ell_matrix_class(csr_matrix_class<data_type> &matrix, unsigned int elements_in_row_arg) : meta(matrix.meta)
  {
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_tuning t = mfp_compat::default_tuning(d->m, MFP_HYB);
    t.hyb_width = elements_in_row_arg;
    mfp_compat::device_plan p(d->m, MFP_HYB, t);
    take(p);
  }

  static size_t estimate_size(csr_matrix_class<data_type> &matrix)
  {
    const auto st = get_rows_statistics(matrix.meta, matrix.row_ptr.get());
    return (size_t)st.max_elements_in_rows * matrix.meta.rows_count;
  }

  const matrix_market::matrix_class::matrix_meta meta;

  size_t get_matrix_size() const { return (size_t)meta.rows_count * elements_in_rows; }

public:
  std::unique_ptr<data_type[]> data;
  std::unique_ptr<unsigned int[]> columns;
  unsigned int elements_in_rows = 0;

private:
  friend class hybrid_matrix_class<data_type>;
  ell_matrix_class(const matrix_market::matrix_class::matrix_meta &m, const mfp_compat::device_plan &p) : meta(m)
  {
    take(p);
  }
  void take(const mfp_compat::device_plan &p)
  {
    elements_in_rows = p.info().ell_width;
    data = p.array<data_type>(MFP_ARR_ELL_VALUES);
    columns = p.array<unsigned int>(MFP_ARR_ELL_COLS);
  }
};

/* coo_matrix_class (matrix_converter.cpp:211-280): row-major triplets in CSR
 * order; with element_start, only entries at in-row offset >= element_start */
template <typename data_type>
class coo_matrix_class {
public:
  coo_matrix_class() = delete;
  explicit coo_matrix_class(csr_matrix_class<data_type> &matrix) : meta(matrix.meta)
  {
    mfp_compat::require_general(meta);
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_compat::device_plan p(d->m, MFP_COO, mfp_compat::default_tuning(d->m, MFP_COO));
    take(p);
  }
  coo_matrix_class(csr_matrix_class<data_type> &matrix, unsigned int element_start) : meta(matrix.meta)
  {
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_tuning t = mfp_compat::default_tuning(d->m, MFP_HYB);
    t.hyb_width = element_start;
    if (element_start == 0) {
      mfp_compat::device_plan p(d->m, MFP_COO, mfp_compat::default_tuning(d->m, MFP_COO));
      take(p);
    } else {
      mfp_compat::device_plan p(d->m, MFP_HYB, t);
      take(p);
    }
  }

  const matrix_market::matrix_class::matrix_meta meta;

  size_t get_matrix_size() const { return elements_count; }

public:
  std::unique_ptr<data_type[]> data;
  std::unique_ptr<unsigned int[]> cols;
  std::unique_ptr<unsigned int[]> rows;

private:
  friend class hybrid_matrix_class<data_type>;
  coo_matrix_class(const matrix_market::matrix_class::matrix_meta &m, const mfp_compat::device_plan &p) : meta(m)
  {
    take(p);
  }
  void take(const mfp_compat::device_plan &p)
  {
    uint64_t n = 0;
    data = p.array<data_type>(MFP_ARR_COO_VALUES, &n);
    cols = p.array<unsigned int>(MFP_ARR_COO_COLS);
    rows = p.array<unsigned int>(MFP_ARR_COO_ROWS);
    elements_count = n;
  }
  size_t elements_count{};
};

/* hybrid_matrix_class (matrix_converter.cpp:286-310): ELL width
 * avg + (max - avg) percent, remainder in COO */
template <typename data_type>
class hybrid_matrix_class {
public:
  hybrid_matrix_class() = delete;
  explicit hybrid_matrix_class(csr_matrix_class<data_type> &matrix) : meta(matrix.meta)
  {
    mfp_compat::require_general(meta);
  }

  void allocate(csr_matrix_class<data_type> &matrix, double percent)
  {
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_tuning t = mfp_compat::default_tuning(d->m, MFP_HYB);
    t.hyb_percent = percent;
    t.hyb_width = 0;
    mfp_compat::device_plan p(d->m, MFP_HYB, t);
    ell_matrix.reset(new ell_matrix_class<data_type>(meta, p));
    coo_matrix.reset(new coo_matrix_class<data_type>(meta, p));
  }

  const matrix_market::matrix_class::matrix_meta meta;

  std::unique_ptr<ell_matrix_class<data_type>> ell_matrix;
  std::unique_ptr<coo_matrix_class<data_type>> coo_matrix;
};

/* scoo_matrix_class (common/scoo_matrix_class.cpp:30-139, sort_cols = false) */
template <typename data_type>
class scoo_matrix_class {
public:
  scoo_matrix_class() = delete;
  scoo_matrix_class(unsigned int sm_count, size_t shared_mem_size, csr_matrix_class<data_type> &matrix)
    : meta(matrix.meta)
  {
    if (meta.rows_count == 0) throw std::runtime_error("SCOO needs at least one row");
    mfp_compat::check(mfp_scoo_geometry(mfp_compat::dtype_of<data_type>(), meta.rows_count, sm_count,
                                        shared_mem_size, &slices_count, &slice_size, &lane_size),
                      "mfp_scoo_geometry");
    std::unique_ptr<mfp_compat::device_csr> d(matrix.to_device());
    mfp_tuning t = mfp_compat::default_tuning(d->m, MFP_SCOO);
    t.scoo_slices = slices_count;
    t.scoo_slice_size = slice_size;
    t.scoo_lanes = lane_size;
    mfp_compat::device_plan p(d->m, MFP_SCOO, t);
    uint64_t n = 0;
    values = p.array<data_type>(MFP_ARR_COO_VALUES, &n);
    c_index = p.array<unsigned int>(MFP_ARR_COO_COLS);
    r_index = p.array<unsigned int>(MFP_ARR_COO_ROWS);
    slices_ptr = p.array<unsigned int>(MFP_ARR_SCOO_SLICES_PTR);
    elements_count = n;
  }

  const matrix_market::matrix_class::matrix_meta meta;

public:
  unsigned int slices_count{};
  unsigned int slice_size{};
  unsigned int lane_size{};

  std::unique_ptr<data_type[]> values;
  std::unique_ptr<unsigned int[]> c_index;
  std::unique_ptr<unsigned int[]> r_index;
  std::unique_ptr<unsigned int[]> slices_ptr;

private:
  size_t elements_count{};
};

/* fill_row_blocks (gpu/csr_adaptive_spmv.cu:180-240), computed on the device;
 * returns the block boundaries (count + 1 entries) */
inline std::vector<unsigned int> fill_row_blocks(unsigned int rows_count, const unsigned int *row_ptr,
                                                 unsigned int nnz_per_wg = 64)
{
  mfp_compat::init_once();
  unsigned int n = 0;
  mfp_compat::check(mfp_fill_row_blocks(rows_count, row_ptr, nnz_per_wg, nullptr, &n), "mfp_fill_row_blocks");
  std::vector<unsigned int> out(n + 1);
  mfp_compat::check(mfp_fill_row_blocks(rows_count, row_ptr, nnz_per_wg, out.data(), &n), "mfp_fill_row_blocks");
  return out;
}

/* ---- the GPU entry points (gpu/gpu_matrix_multiplier.h:15-138,
 * gpu/csr_adaptive_spmv.h:13-23, gpu/scoo_spmv.h:12-23): upload, x = 1, one
 * timed SpMV, y into reusable_vector, compare against reference_y ---------- */
namespace mfp_compat {

template <typename T> mfp_host_csr host_csr(const csr_matrix_class<T> &m)
{
  return {m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count, m.row_ptr.get(), m.columns.get(),
          m.data.get()};
}
template <typename T> mfp_host_ell host_ell(const ell_matrix_class<T> &m)
{
  return {m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count, m.elements_in_rows, m.columns.get(),
          m.data.get()};
}
template <typename T> mfp_host_coo host_coo(const coo_matrix_class<T> &m)
{
  return {m.meta.rows_count, m.meta.cols_count, m.meta.non_zero_count, m.get_matrix_size(), m.cols.get(),
          m.rows.get(), m.data.get()};
}

} // namespace mfp_compat

#define MFP_COMPAT_CSR_ENTRY(NAME, CFUNC)                                                                     \
  template <typename data_type>                                                                                \
  measurement_class NAME(const csr_matrix_class<data_type> &matrix, resizable_gpu_memory<data_type> &A,       \
                         resizable_gpu_memory<unsigned int> &col_ids, resizable_gpu_memory<unsigned int> &row_ptr, \
                         resizable_gpu_memory<data_type> &x, resizable_gpu_memory<data_type> &y,               \
                         data_type *reusable_vector, const data_type *reference_y)                             \
  {                                                                                                            \
    const mfp_host_csr h = mfp_compat::host_csr(matrix);                                                       \
    mfp_measurement m;                                                                                         \
    mfp_compat::check(CFUNC(mfp_compat::dtype_of<data_type>(), &h, A.handle(), col_ids.handle(),               \
                            row_ptr.handle(), x.handle(), y.handle(), reusable_vector, reference_y, &m),       \
                      #NAME);                                                                                  \
    return measurement_class(m);                                                                               \
  }

MFP_COMPAT_CSR_ENTRY(gpu_csr_spmv, mfp_gpu_csr_spmv)
MFP_COMPAT_CSR_ENTRY(gpu_csr_vector_spmv, mfp_gpu_csr_vector_spmv)
MFP_COMPAT_CSR_ENTRY(gpu_csr_adaptive_spmv, mfp_gpu_csr_adaptive_spmv)
#undef MFP_COMPAT_CSR_ENTRY

template <typename data_type>
measurement_class gpu_ell_spmv(const ell_matrix_class<data_type> &matrix, resizable_gpu_memory<data_type> &A,
                               resizable_gpu_memory<unsigned int> &col_ids, resizable_gpu_memory<data_type> &x,
                               resizable_gpu_memory<data_type> &y, data_type *reusable_vector,
                               const data_type *reference_y)
{
  const mfp_host_ell h = mfp_compat::host_ell(matrix);
  mfp_measurement m;
  mfp_compat::check(mfp_gpu_ell_spmv(mfp_compat::dtype_of<data_type>(), &h, A.handle(), col_ids.handle(),
                                     x.handle(), y.handle(), reusable_vector, reference_y, &m),
                    "gpu_ell_spmv");
  return measurement_class(m);
}

template <typename data_type>
measurement_class gpu_coo_spmv(const coo_matrix_class<data_type> &matrix, resizable_gpu_memory<data_type> &A,
                               resizable_gpu_memory<unsigned int> &col_ids,
                               resizable_gpu_memory<unsigned int> &row_ids, resizable_gpu_memory<data_type> &x,
                               resizable_gpu_memory<data_type> &y, data_type *reusable_vector,
                               const data_type *reference_y)
{
  const mfp_host_coo h = mfp_compat::host_coo(matrix);
  mfp_measurement m;
  mfp_compat::check(mfp_gpu_coo_spmv(mfp_compat::dtype_of<data_type>(), &h, A.handle(), col_ids.handle(),
                                     row_ids.handle(), x.handle(), y.handle(), reusable_vector, reference_y, &m),
                    "gpu_coo_spmv");
  return measurement_class(m);
}

template <typename data_type>
measurement_class gpu_scoo_spmv(bool print_diff, const scoo_matrix_class<data_type> &matrix,
                                resizable_gpu_memory<data_type> &A, resizable_gpu_memory<unsigned int> &r_index,
                                resizable_gpu_memory<unsigned int> &c_index, resizable_gpu_memory<data_type> &x,
                                resizable_gpu_memory<data_type> &y, data_type *reusable_vector,
                                const data_type *reference_y)
{
  const mfp_host_scoo h{matrix.meta.rows_count, matrix.meta.cols_count, matrix.meta.non_zero_count,
                        matrix.slices_count,    matrix.slice_size,      matrix.lane_size,
                        matrix.values.get(),    matrix.c_index.get(),   matrix.r_index.get(),
                        matrix.slices_ptr.get()};
  mfp_measurement m;
  mfp_compat::check(mfp_gpu_scoo_spmv(mfp_compat::dtype_of<data_type>(), print_diff ? 1 : 0, &h, A.handle(),
                                      r_index.handle(), c_index.handle(), x.handle(), y.handle(), reusable_vector,
                                      reference_y, &m),
                    "gpu_scoo_spmv");
  return measurement_class(m);
}

#define MFP_COMPAT_HYB_ENTRY(NAME, CFUNC)                                                                     \
  template <typename data_type>                                                                                \
  measurement_class NAME(const hybrid_matrix_class<data_type> &matrix, resizable_gpu_memory<data_type> &A_ell, \
                         resizable_gpu_memory<data_type> &A_coo, resizable_gpu_memory<unsigned int> &ell_col_ids, \
                         resizable_gpu_memory<unsigned int> &coo_col_ids,                                      \
                         resizable_gpu_memory<unsigned int> &coo_row_ids, resizable_gpu_memory<data_type> &x,  \
                         resizable_gpu_memory<data_type> &y, data_type *reusable_vector,                       \
                         const data_type *reference_y)                                                         \
  {                                                                                                            \
    if (!matrix.ell_matrix || !matrix.coo_matrix)                                                              \
      throw std::runtime_error(#NAME ": hybrid_matrix_class::allocate was not called");                       \
    const mfp_host_ell he = mfp_compat::host_ell(*matrix.ell_matrix);                                          \
    const mfp_host_coo hc = mfp_compat::host_coo(*matrix.coo_matrix);                                          \
    mfp_measurement m;                                                                                         \
    mfp_compat::check(CFUNC(mfp_compat::dtype_of<data_type>(), &he, &hc, A_ell.handle(), A_coo.handle(),       \
                            ell_col_ids.handle(), coo_col_ids.handle(), coo_row_ids.handle(), x.handle(),      \
                            y.handle(), reusable_vector, reference_y, &m),                                     \
                      #NAME);                                                                                  \
    return measurement_class(m);                                                                               \
  }

MFP_COMPAT_HYB_ENTRY(gpu_hybrid_spmv, mfp_gpu_hybrid_spmv)
MFP_COMPAT_HYB_ENTRY(gpu_hybrid_atomic_spmv, mfp_gpu_hybrid_atomic_spmv)
#undef MFP_COMPAT_HYB_ENTRY

/* gpu_hybrid_cpu_coo_spmv (gpu/gpu_matrix_multiplier.h:141-152) */
template <typename data_type>
measurement_class gpu_hybrid_cpu_coo_spmv(const hybrid_matrix_class<data_type> &matrix,
                                          resizable_gpu_memory<data_type> &A_ell,
                                          resizable_gpu_memory<unsigned int> &ell_col_ids,
                                          resizable_gpu_memory<data_type> &x, resizable_gpu_memory<data_type> &y,
                                          resizable_gpu_memory<data_type> &tmp, data_type *cpu_y,
                                          data_type *reusable_vector, const data_type *reference_y)
{
  if (!matrix.ell_matrix || !matrix.coo_matrix)
    throw std::runtime_error("gpu_hybrid_cpu_coo_spmv: hybrid_matrix_class::allocate was not called");
  const mfp_host_ell he = mfp_compat::host_ell(*matrix.ell_matrix);
  const mfp_host_coo hc = mfp_compat::host_coo(*matrix.coo_matrix);
  mfp_measurement m;
  mfp_compat::check(mfp_gpu_hybrid_cpu_coo_spmv(mfp_compat::dtype_of<data_type>(), &he, &hc, A_ell.handle(),
                                                ell_col_ids.handle(), x.handle(), y.handle(), tmp.handle(), cpu_y,
                                                reusable_vector, reference_y, &m),
                    "gpu_hybrid_cpu_coo_spmv");
  return measurement_class(m);
}

/* ---- the reference's CPU measurements (cpu/cpu_matrix_multiplier.h) ------
 * Each sets x = 1 over cols, as the reference does. */
namespace mfp_compat {

template <typename T> double cpu_csr_bytes(const matrix_market::matrix_class::matrix_meta &meta)
{
  // cpu_matrix_multiplier.cpp:50-54
  return (double)((size_t)meta.non_zero_count * sizeof(T) * 2 + (size_t)meta.non_zero_count * 4 +
                  2 * (size_t)meta.rows_count * 4 + (size_t)meta.rows_count * sizeof(T));
}

template <typename T>
void cpu_csr_rows(const unsigned *rp, const unsigned *ci, const T *va, const T *x, T *y, unsigned r0, unsigned r1)
{
  for (unsigned row = r0; row < r1; row++) {
    T dot = 0;
    for (unsigned e = rp[row]; e < rp[row + 1]; e++) dot = std::fma(va[e], x[ci[e]], dot);
    y[row] = dot;
  }
}

} // namespace mfp_compat

/* cpu_matrix_multiplier.cpp:20-61 (products accumulated with fma) */
template <typename data_type>
measurement_class cpu_csr_spmv_single_thread_naive(const csr_matrix_class<data_type> &matrix, data_type *x,
                                                   data_type *y)
{
  std::fill_n(x, matrix.meta.cols_count, data_type(1.0));
  const auto begin = std::chrono::steady_clock::now();
  mfp_compat::cpu_csr_rows(matrix.row_ptr.get(), matrix.columns.get(), matrix.data.get(), x, y, 0u,
                           matrix.meta.rows_count);
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
  return measurement_class("CPU CSR", elapsed, mfp_compat::cpu_csr_bytes<data_type>(matrix.meta),
                           2.0 * matrix.meta.non_zero_count);
}

/* cpu_matrix_multiplier.cpp:170-227: 32 lanes, as the GPU CSR-vector entry
 * point and the reference's warp; lane l accumulates elements l, l + 32, ...
 * with fma, then lanes are added in a halving tree (lane i += lane i + offset,
 * offsets 16, 8, 4, 2, 1). */
template <typename data_type>
measurement_class cpu_csr_spmv_single_thread_naive_with_reduce_order(const csr_matrix_class<data_type> &matrix,
                                                                     data_type *x, data_type *y)
{
  std::fill_n(x, matrix.meta.cols_count, data_type(1.0));
  const unsigned rows = matrix.meta.rows_count;
  constexpr unsigned lanes = 32;
  const auto *rp = matrix.row_ptr.get();
  const auto *ci = matrix.columns.get();
  const auto *va = matrix.data.get();
  data_type dots[lanes];
  const auto begin = std::chrono::steady_clock::now();
  for (unsigned row = 0; row < rows; row++) {
    for (unsigned l = 0; l < lanes; l++) {
      data_type d = 0;
      for (unsigned e = rp[row] + l; e < rp[row + 1]; e += lanes) d = std::fma(va[e], x[ci[e]], d);
      dots[l] = d;
    }
    for (unsigned off = lanes / 2; off > 0; off /= 2)
      for (unsigned l = 0; l < off; l++) dots[l] += dots[l + off];
    y[row] = dots[0];
  }
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
  return measurement_class("CPU CSR Reduce order", elapsed, mfp_compat::cpu_csr_bytes<data_type>(matrix.meta),
                           2.0 * matrix.meta.non_zero_count);
}

/* cpu_matrix_multiplier.cpp:251-308: hardware_concurrency threads, static
 * row split, all threads released together, elapsed = the slowest thread */
template <typename data_type>
measurement_class cpu_csr_spmv_multi_thread_naive(const csr_matrix_class<data_type> &matrix, data_type *x,
                                                  data_type *y)
{
  std::fill_n(x, matrix.meta.cols_count, data_type(1.0));
  const unsigned threads_count = std::max(1u, std::thread::hardware_concurrency());
  const unsigned rows = matrix.meta.rows_count;
  std::vector<double> times(threads_count, 0.0);
  std::vector<std::thread> threads;
  std::atomic<unsigned> started{0};
  for (unsigned t = 0; t < threads_count; t++)
    threads.emplace_back([&, t] {
      const unsigned per = rows / threads_count;
      const unsigned r0 = per * t, r1 = t == threads_count - 1 ? rows : (t + 1) * per;
      started.fetch_add(1, std::memory_order_relaxed);
      while (started.load(std::memory_order_relaxed) < threads_count) std::this_thread::yield();
      const auto begin = std::chrono::steady_clock::now();
      mfp_compat::cpu_csr_rows(matrix.row_ptr.get(), matrix.columns.get(), matrix.data.get(), x, y, r0, r1);
      times[t] = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    });
  for (auto &th : threads) th.join();
  const double elapsed = *std::max_element(times.begin(), times.end());
  return measurement_class("CPU CSR Parallel", elapsed, mfp_compat::cpu_csr_bytes<data_type>(matrix.meta),
                           2.0 * matrix.meta.non_zero_count);
}

#endif // MFP_COMPAT_HPP
