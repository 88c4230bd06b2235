// mfp_bench: the benchmark driver (counterpart of the reference's main.cpp).
//
//   mfp_bench [options] /path/to/mtx_list        one MatrixMarket path per line
//   mfp_bench [options] --synthetic NAME[;NAME]  generated matrices
//
// For each matrix and each of float and double it converts the triplets to
// CSR, ELL and COO on the device, times the CPU rows and every GPU format in
// the reference's order (main.cpp:82-283), each averaged over --reps calls
// (20, main.cpp:139-149), and finally writes the reference's seven JSON files
// (main.cpp:482-501): matrices_info.json and
// {float,double}{,_effective_bandwidth,_computational_throughput}.json, each
// {matrix: {label: number}} with sorted keys.
//
// Options:
//   --out DIR        where the JSON files go (default .)
//   --reps N         calls averaged per format (default 20)
//   --meta-only      only matrices_info.json with std_deviation (the
//                    reference's PREP_MATRIX_META mode, main.cpp:468-475)
//   --ell-skip       when the ELL layout exceeds 90% of free device memory,
//                    omit only the "GPU ELL" row instead of the whole matrix
//                    (the reference skips the matrix, main.cpp:105-112)
//   --no-cpu         omit the CPU rows (the reference answer is still computed)
//   --no-rocsparse   omit the "GPU CSR (rocSPARSE)" row (left out anyway when
//                    librocsparse.so cannot be opened)
//   --hybrid-sweep   also run the block the reference keeps switched off
//                    (main.cpp:247-279): hybrid ELL widths at 0%, 35% and 70%,
//                    each timed once with the 2-kernel, atomic and host-COO
//                    hybrids; rows of equal label keep the last width's number
//
// Synthetic names: stencil2d_256 (2D 5-point, 256 x 256), powerlaw_1m
// (1,048,576 rows, power-law row lengths 1..50,000), stencil3d_128 (3D
// 27-point, 128^3), or FAMILY:key=value,... with FAMILY in stencil2d,
// stencil3d, powerlaw, uniform and keys n, rows, lo, hi, alpha, bw, seed,
// values (laplace|uniform|normal).
#include "../../include/mfp_compat.hpp"
#include "rocsparse_slot.h"

#include <charconv>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <string>
#include <vector>

namespace {

using json_object = std::map<std::string, std::map<std::string, std::string>>;

std::string num(double v)
{
  if (!std::isfinite(v)) return "null";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v);
  std::string s(buf, r.ptr);
  if (s.find_first_of(".eE") == std::string::npos && s.find('n') == std::string::npos) s += ".0";
  return s;
}

std::string num(unsigned long long v) { return std::to_string(v); }

// nlohmann::json dump(2) layout for a two-level object with sorted keys
void write_json(const std::string &path, const json_object &obj)
{
  std::ofstream os(path);
  if (!os) throw std::runtime_error("cannot write " + path);
  if (obj.empty()) {
    os << "null";
    return;
  }
  os << "{\n";
  size_t i = 0;
  for (const auto &[k, inner] : obj) {
    os << "  \"" << k << "\": {";
    if (inner.empty()) {
      os << "}";
    } else {
      os << "\n";
      size_t j = 0;
      for (const auto &[k2, v] : inner) os << "    \"" << k2 << "\": " << v << (++j < inner.size() ? ",\n" : "\n");
      os << "  }";
    }
    os << (++i < obj.size() ? ",\n" : "\n");
  }
  os << "}";
}

std::string get_filename(const std::string &path)
{
  const size_t i = path.rfind('/');
  return i == std::string::npos ? path : path.substr(i + 1);
}

struct options {
  std::string out = ".";
  unsigned reps = 20;
  bool meta_only = false, ell_skip = false, cpu = true, rocsparse = true, hybrid_sweep = false;
  std::vector<std::string> synthetic;
  std::string list;
};

// ---- synthetic matrices -------------------------------------------------------
mfp_gen_params parse_spec(const std::string &name)
{
  mfp_gen_params p{};
  p.values = MFP_VAL_UNIFORM;
  p.alpha = 2.0;
  p.bandwidth = 1.0;
  p.nnz_lo = p.nnz_hi = 1;
  if (name == "stencil2d_256") {
    p.family = MFP_GEN_STENCIL2D;
    p.values = MFP_VAL_LAPLACE;
    p.n = 256;
    return p;
  }
  if (name == "powerlaw_1m") {
    p.family = MFP_GEN_POWERLAW;
    p.rows = 1u << 20;
    p.nnz_lo = 1;
    p.nnz_hi = 50000;
    return p;
  }
  if (name == "stencil3d_128") {
    p.family = MFP_GEN_STENCIL3D;
    p.values = MFP_VAL_NORMAL;
    p.n = 128;
    return p;
  }
  const size_t colon = name.find(':');
  const std::string fam = name.substr(0, colon);
  if (fam == "stencil2d") p.family = MFP_GEN_STENCIL2D;
  else if (fam == "stencil3d") p.family = MFP_GEN_STENCIL3D;
  else if (fam == "powerlaw") p.family = MFP_GEN_POWERLAW;
  else if (fam == "uniform") p.family = MFP_GEN_UNIFORM;
  else throw std::runtime_error("unknown synthetic matrix '" + name + "'");
  if (colon != std::string::npos) {
    std::stringstream ss(name.substr(colon + 1));
    std::string kv;
    while (std::getline(ss, kv, ',')) {
      const size_t eq = kv.find('=');
      if (eq == std::string::npos) throw std::runtime_error("bad key=value '" + kv + "'");
      const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "n") p.n = (uint32_t)std::stoul(v);
      else if (k == "rows") p.rows = (uint32_t)std::stoul(v);
      else if (k == "lo") p.nnz_lo = (uint32_t)std::stoul(v);
      else if (k == "hi") p.nnz_hi = (uint32_t)std::stoul(v);
      else if (k == "alpha") p.alpha = std::stod(v);
      else if (k == "bw") p.bandwidth = std::stod(v);
      else if (k == "seed") p.seed = std::stoull(v);
      else if (k == "values")
        p.values = v == "laplace" ? MFP_VAL_LAPLACE : v == "normal" ? MFP_VAL_NORMAL : MFP_VAL_UNIFORM;
      else throw std::runtime_error("unknown key '" + k + "'");
    }
  }
  return p;
}

matrix_market::matrix_class generate(const std::string &name)
{
  const mfp_gen_params p = parse_spec(name);
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  mfp_compat::check(mfp_generate(&p, &rows, &cols, &nnz, nullptr, nullptr, nullptr, MFP_F64), "mfp_generate");
  if (nnz > 0xFFFFFFFFull) throw std::runtime_error("more than 2^32 - 1 entries");
  std::vector<uint32_t> rp(rows + 1), ci(std::max<uint64_t>(nnz, 1));
  std::vector<double> va(std::max<uint64_t>(nnz, 1));
  mfp_compat::check(mfp_generate(&p, &rows, &cols, &nnz, rp.data(), ci.data(), va.data(), MFP_F64),
                    "mfp_generate");
  std::vector<unsigned> ri(nnz);
  for (uint32_t r = 0; r < rows; r++)
    for (uint32_t e = rp[r]; e < rp[r + 1]; e++) ri[e] = r;
  ci.resize(nnz);
  va.resize(nnz);
  matrix_market::matrix_class::matrix_meta meta;
  meta.rows_count = rows;
  meta.cols_count = cols;
  return matrix_market::matrix_class(meta, std::move(ri), std::move(ci), std::move(va));
}

// ---- "GPU CSR (rocSPARSE)": the counterpart of gpu_csr_cusparse_spmv
// (gpu/gpu_matrix_multiplier.cu:175-257), benchmark-only: upload, x = 1,
// one timed rocsparse_spmv compute stage, y into reusable_vector, compare.
// The byte model is the reference's cuSPARSE one (:245-249).
template <typename T>
measurement_class gpu_csr_rocsparse_spmv(const csr_matrix_class<T> &matrix, resizable_gpu_memory<T> &A,
                                         resizable_gpu_memory<unsigned int> &col_ids,
                                         resizable_gpu_memory<unsigned int> &row_ptr, resizable_gpu_memory<T> &x,
                                         resizable_gpu_memory<T> &y, T *reusable_vector, const T *reference_y)
{
  const auto &meta = matrix.meta;
  const size_t nnz = meta.non_zero_count, rows = meta.rows_count, cols = meta.cols_count;
  A.resize(std::max<size_t>(nnz, 1));
  col_ids.resize(std::max<size_t>(nnz, 1));
  row_ptr.resize(rows + 1);
  x.resize(std::max<size_t>(cols, 1));
  y.resize(std::max<size_t>(rows, 1));
  mfp_compat::check(mfp_memcpy_h2d(A.get(), matrix.data.get(), nnz * sizeof(T)), "upload values");
  mfp_compat::check(mfp_memcpy_h2d(col_ids.get(), matrix.columns.get(), nnz * 4), "upload columns");
  mfp_compat::check(mfp_memcpy_h2d(row_ptr.get(), matrix.row_ptr.get(), (rows + 1) * 4), "upload row_ptr");
  const int dt = mfp_compat::dtype_of<T>();
  mfp_compat::check(mfp_fill(dt, x.get(), cols, 1.0, nullptr), "x = 1");
  mfp_rs_plan plan = nullptr;
  if (mfp_rs_plan_create(&plan, dt, (uint32_t)rows, (uint32_t)cols, nnz, row_ptr.get(), col_ids.get(), A.get(),
                         x.get(), y.get(), nullptr) != MFP_OK)
    throw std::runtime_error(mfp_rs_error());
  mfp_timer t = nullptr;
  mfp_compat::check(mfp_timer_create(&t), "timer");
  mfp_compat::check(mfp_synchronize(), "sync");
  mfp_compat::check(mfp_timer_start(t, nullptr), "timer start");
  const int st = mfp_rs_plan_spmv(plan, nullptr);
  mfp_compat::check(mfp_timer_stop(t, nullptr), "timer stop");
  float ms = 0;
  mfp_compat::check(mfp_timer_elapsed_ms(t, &ms), "elapsed");
  mfp_timer_destroy(t);
  mfp_rs_plan_destroy(plan);
  if (st != MFP_OK) throw std::runtime_error(mfp_rs_error());
  mfp_compat::check(mfp_memcpy_d2h(reusable_vector, y.get(), rows * sizeof(T)), "download y");
  mfp_compare_results(dt, (uint32_t)rows, reusable_vector, reference_y);
  const double bytes = nnz * sizeof(T) + nnz * sizeof(T) + nnz * 4.0 + 2.0 * rows * 4 + rows * sizeof(T);
  return measurement_class("GPU CSR (rocSPARSE)", ms / 1000.0, bytes, 2.0 * nnz);
}

// ---- one matrix in one precision (main.cpp:82-283) ----------------------------
template <typename T>
std::vector<measurement_class> perform_measurement(const matrix_market::matrix_class &mtx, const options &opt,
                                                   size_t free_memory, unsigned sm_count, size_t shared_mem_size)
{
  std::vector<measurement_class> out;
  const auto &meta = mtx.meta;
  csr_matrix_class<T> csr(mtx);

  // the reference's memory estimate (main.cpp:105-112), in its units
  const size_t csr_size = csr.get_matrix_size();
  const size_t ell_size = ell_matrix_class<T>::estimate_size(csr);
  const size_t vec_size = std::max(meta.rows_count, meta.cols_count) * sizeof(T);
  const size_t matrix_size = std::max(csr_size, ell_size) * sizeof(T);
  const size_t estimated = matrix_size + 5 * vec_size;
  const bool over = (double)(estimated * sizeof(T)) > (double)free_memory * 0.9;
  if (over && !opt.ell_skip) return out;

  std::unique_ptr<ell_matrix_class<T>> ell;
  if (!over) ell.reset(new ell_matrix_class<T>(csr));
  coo_matrix_class<T> coo(csr);

  const size_t vn = std::max(meta.rows_count, meta.cols_count);
  std::unique_ptr<T[]> reference(new T[meta.rows_count]), reduce_order(new T[meta.rows_count]);
  std::unique_ptr<T[]> cpu_y(new T[meta.rows_count]), x(new T[std::max<size_t>(vn, 1)]);

  auto print = [](const measurement_class &m) {
    std::printf("  %-22s %12.3f us  %10.2f GB/s  %9.2f GFLOP/s\n", m.get_format().c_str(), m.get_elapsed() * 1e6,
                m.get_effective_bandwidth(), m.get_computational_throughput());
    std::fflush(stdout);
  };
  auto repeat = [&](const std::function<measurement_class(bool)> &action) {
    measurement_class r;
    for (unsigned i = 0; i < opt.reps; i++) r += action(i == 0);
    r.finalize();
    out.push_back(r);
    print(r);
  };

  {
    auto m = cpu_csr_spmv_single_thread_naive(csr, x.get(), reference.get());
    if (opt.cpu) {
      out.push_back(m);
      print(m);
    }
  }
  if (opt.cpu) repeat([&](bool) { return cpu_csr_spmv_multi_thread_naive(csr, x.get(), cpu_y.get()); });

  resizable_gpu_memory<T> A, x_gpu, y;
  resizable_gpu_memory<unsigned int> col_ids, row_ptr;
  repeat([&](bool) { return gpu_csr_spmv<T>(csr, A, col_ids, row_ptr, x_gpu, y, x.get(), reference.get()); });
  cpu_csr_spmv_single_thread_naive_with_reduce_order(csr, x.get(), reduce_order.get());
  repeat([&](bool) {
    return gpu_csr_vector_spmv<T>(csr, A, col_ids, row_ptr, x_gpu, y, x.get(), reduce_order.get());
  });
  repeat([&](bool) {
    return gpu_csr_adaptive_spmv<T>(csr, A, col_ids, row_ptr, x_gpu, y, x.get(), reference.get());
  });
  // the reference's cuSPARSE row comes here (main.cpp:194-197); its
  // rocSPARSE counterpart runs when the library is present
  if (opt.rocsparse && mfp_rs_available())
    repeat([&](bool) {
      return gpu_csr_rocsparse_spmv<T>(csr, A, col_ids, row_ptr, x_gpu, y, x.get(), reference.get());
    });
  if (ell) repeat([&](bool) { return gpu_ell_spmv<T>(*ell, A, col_ids, x_gpu, y, x.get(), reference.get()); });
  repeat([&](bool) { return gpu_coo_spmv<T>(coo, A, col_ids, row_ptr, x_gpu, y, x.get(), reference.get()); });
  {
    scoo_matrix_class<T> scoo(sm_count, shared_mem_size, csr);
    std::printf("  SCOO: slice_size=%u; lane_size=%u; slices_count=%u\n", scoo.slice_size, scoo.lane_size,
                scoo.slices_count);
    repeat([&](bool print_diff) {
      return gpu_scoo_spmv<T>(print_diff, scoo, A, col_ids, row_ptr, x_gpu, y, x.get(), reference.get());
    });
  }
  {
    resizable_gpu_memory<T> A_coo;
    resizable_gpu_memory<unsigned int> col_ids_coo;
    hybrid_matrix_class<T> hybrid(csr);
    hybrid.allocate(csr, 0.2);
    repeat([&](bool) {
      return gpu_hybrid_spmv<T>(hybrid, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x.get(),
                                reference.get());
    });
    repeat([&](bool) {
      return gpu_hybrid_atomic_spmv<T>(hybrid, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x.get(),
                                       reference.get());
    });
  }
  if (opt.hybrid_sweep) {
    resizable_gpu_memory<T> A_coo, tmp;
    resizable_gpu_memory<unsigned int> col_ids_coo;
    hybrid_matrix_class<T> hybrid(csr);
    auto once = [&](measurement_class m) {
      out.push_back(m);
      print(m);
    };
    for (double percent = 0.0; percent <= 1.0; percent += 0.35) {
      hybrid.allocate(csr, percent);
      std::printf("  hybrid sweep: %d%%, ELL width %u\n", (int)(percent * 100), hybrid.ell_matrix->elements_in_rows);
      once(gpu_hybrid_spmv<T>(hybrid, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x.get(), reference.get()));
      once(gpu_hybrid_atomic_spmv<T>(hybrid, A, A_coo, col_ids, col_ids_coo, row_ptr, x_gpu, y, x.get(),
                                     reference.get()));
      once(gpu_hybrid_cpu_coo_spmv<T>(hybrid, A, col_ids, x_gpu, y, tmp, cpu_y.get(), x.get(), reference.get()));
    }
  }
  return out;
}

int run(const options &opt)
{
  mfp_compat::init_once();
  mfp_device_info dev;
  mfp_compat::check(mfp_device_query(&dev), "mfp_device_query");
  std::printf("device %s (%s): %d CUs, %llu B LDS per block, %.1f GiB free\n", dev.name, dev.arch, dev.cu_count,
              (unsigned long long)dev.lds_per_block, dev.free_mem / 1073741824.0);

  std::vector<std::pair<std::string, std::function<matrix_market::matrix_class()>>> inputs;
  for (const auto &s : opt.synthetic) inputs.emplace_back(s, [s] { return generate(s); });
  if (!opt.list.empty()) {
    std::ifstream list(opt.list);
    if (!list) throw std::runtime_error("cannot open " + opt.list);
    std::string path;
    while (std::getline(list, path)) {
      if (path.empty()) continue;
      inputs.emplace_back(get_filename(path), [path] {
        std::ifstream is(path);
        if (!is) throw std::runtime_error("cannot open " + path);
        matrix_market::reader reader(is);
        return reader.matrix();
      });
    }
  }

  std::map<std::string, json_object> measurements, bandwidth, throughput;
  json_object info;
  for (auto &[name, load] : inputs) {
    std::printf("Start loading matrix %s\n", name.c_str());
    const matrix_market::matrix_class mtx = load();
    const auto &meta = mtx.meta;
    std::printf("Complete loading (rows: %u; cols: %u; nnz: %u; nnzpr: %u)\n", meta.rows_count, meta.cols_count,
                meta.non_zero_count, meta.rows_count ? meta.non_zero_count / meta.rows_count : 0);
    if (opt.meta_only) {
      csr_matrix_class<float> csr(mtx, true);
      info[name]["std_deviation"] = num(get_rows_statistics(csr.meta, csr.row_ptr.get()).elements_in_rows_std_deviation);
    } else {
      std::map<std::string, std::vector<measurement_class>> results;
      std::printf(" float\n");
      results["float"] = perform_measurement<float>(mtx, opt, dev.free_mem, dev.cu_count, dev.lds_per_block);
      std::printf(" double\n");
      results["double"] = perform_measurement<double>(mtx, opt, dev.free_mem, dev.cu_count, dev.lds_per_block);
      if (results["float"].empty() || results["double"].empty()) {
        std::printf("skipped %s: over the device memory estimate\n", name.c_str());
        continue;
      }
      for (auto &[type, list] : results)
        for (auto &m : list) {
          measurements[type][name][m.get_format()] = num(m.get_elapsed());
          bandwidth[type][name][m.get_format()] = num(m.get_effective_bandwidth());
          throughput[type][name][m.get_format()] = num(m.get_computational_throughput());
        }
    }
    info[name]["nnz"] = num((unsigned long long)meta.non_zero_count);
    info[name]["rows"] = num((unsigned long long)meta.rows_count);
    info[name]["cols"] = num((unsigned long long)meta.cols_count);
  }

  write_json(opt.out + "/matrices_info.json", info);
  for (const char *p : {"float", "double"}) {
    const std::string base = opt.out + "/" + p;
    write_json(base + ".json", measurements[p]);
    write_json(base + "_effective_bandwidth.json", bandwidth[p]);
    write_json(base + "_computational_throughput.json", throughput[p]);
  }
  return 0;
}

} // namespace

int main(int argc, char **argv)
{
  options opt;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error(a + " needs a value");
      return argv[++i];
    };
    try {
      if (a == "--out") opt.out = next();
      else if (a == "--reps") opt.reps = (unsigned)std::stoul(next());
      else if (a == "--meta-only") opt.meta_only = true;
      else if (a == "--ell-skip") opt.ell_skip = true;
      else if (a == "--no-cpu") opt.cpu = false;
      else if (a == "--no-rocsparse") opt.rocsparse = false;
      else if (a == "--hybrid-sweep") opt.hybrid_sweep = true;
      else if (a == "--synthetic") {
        std::stringstream ss(next());
        std::string s;
        while (std::getline(ss, s, ';')) opt.synthetic.push_back(s);
      } else if (!a.empty() && a[0] == '-') throw std::runtime_error("unknown option " + a);
      else opt.list = a;
    } catch (const std::exception &e) {
      std::cerr << e.what() << "\n";
      return 1;
    }
  }
  if (opt.synthetic.empty() && opt.list.empty()) {
    std::cerr << "Usage: " << argv[0] << " [--out DIR] [--reps N] [--meta-only] [--ell-skip] [--no-cpu] [--no-rocsparse] [--hybrid-sweep] "
                 "(/path/to/mtx_list | --synthetic NAME[;NAME...])\n";
    return 1;
  }
  if (opt.reps == 0) opt.reps = 1;
  try {
    return run(opt);
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 2;
  }
}
