// The reference's GPU entry points, with its call semantics:
// grow the caller's buffers, copy the host format in, x = 1 (and y), time the
// SpMV with an event pair, copy y out, compare against reference_y (stderr
// only), return the measurement record with the reference's label and byte
// model.  The kernels are this library's.
#include "kernels.h"
#include "mfp_internal.h"

#include <cmath>
#include <cstring>
#include <iostream>
#include <vector>
#include <algorithm>

using namespace mfp;

namespace {

template <typename T>
void compare_results_t(uint32_t n, const T *a, const T *b)
{
  // measurement_class.h:60-89: relative squared error in T, printed only.
  T num = 0, den = 0;
  for (uint32_t i = 0; i < n; i++) {
    const T d = a[i] - b[i];
    const T d2 = d * d;
    num = num + d2;
    const T b2 = b[i] * b[i];
    den = den + b2;
  }
  const T err = num / den;
  if (err > 1e-9) {
    std::cerr << "ERROR: " << err << std::endl;
    for (uint32_t i = 0; i < n; i++) {
      if (std::abs(a[i] - b[i]) > 1e-8) {
        std::cerr << "a[" << i << "] = " << a[i] << "; b[" << i << "] = " << b[i] << std::endl;
        break;
      }
    }
  }
  std::cerr.flush();
}

void compare_results(int dtype, uint32_t n, const void *a, const void *b)
{
  if (dtype == MFP_F32)
    compare_results_t<float>(n, (const float *)a, (const float *)b);
  else
    compare_results_t<double>(n, (const double *)a, (const double *)b);
}

void fill_measurement(mfp_measurement *out, const char *label, double elapsed, double bytes,
                      double ops)
{
  std::memset(out, 0, sizeof(*out));
  std::strncpy(out->format, label, sizeof(out->format) - 1);
  out->elapsed = elapsed;
  out->load_store_bytes = bytes;
  out->operations_count = ops;
  const double giga = 1e9; // measurement_class.cpp:7-16
  out->effective_bandwidth = bytes / (elapsed * giga);
  out->computational_throughput = ops / (elapsed * giga);
}

// gpu_matrix_multiplier.cu:120-126 (CSR family byte model)
double csr_bytes(size_t s, uint32_t rows, uint64_t nnz)
{
  return (double)(nnz * s + nnz * s + nnz * 4 + 2ull * rows * 4 + (uint64_t)rows * s);
}

struct timed {
  hipEvent_t a = nullptr, b = nullptr;
  ~timed()
  {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int create()
  {
    MFP_HIP(hipEventCreate(&a));
    MFP_HIP(hipEventCreate(&b));
    return MFP_OK;
  }
};

int check_dtype(int dtype)
{
  MFP_ARG(dtype == MFP_F32 || dtype == MFP_F64, "the reference entry points take float or double");
  return ensure_init();
}

// Common tail: event-timed launch, y download, compare.
template <typename F>
int run_timed(F &&launch, double *seconds)
{
  timed ev;
  MFP_CHECK(ev.create());
  MFP_HIP(hipDeviceSynchronize());
  MFP_HIP(hipEventRecord(ev.a, nullptr));
  MFP_CHECK(launch());
  MFP_HIP(hipEventRecord(ev.b, nullptr));
  MFP_HIP(hipEventSynchronize(ev.b));
  float ms = 0;
  MFP_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
  *seconds = ms / 1000.0;
  return MFP_OK;
}

int finish(int dtype, uint32_t rows, const void *y_dev, void *reusable_vector,
           const void *reference_y)
{
  MFP_ARG(reusable_vector || rows == 0, "reusable_vector is null");
  MFP_CHECK(mfp_memcpy_d2h(reusable_vector, y_dev, (size_t)rows * dtype_size(dtype)));
  if (reference_y) compare_results(dtype, rows, reusable_vector, reference_y);
  return MFP_OK;
}

int upload_csr(int dtype, const mfp_host_csr *A, mfp_buffer Ad, mfp_buffer cd, mfp_buffer rd,
               mfp_buffer xd, mfp_buffer yd, double y_init)
{
  MFP_ARG(A && Ad && cd && rd && xd && yd, "null pointer");
  const size_t s = dtype_size(dtype);
  MFP_CHECK(mfp_buffer_resize(Ad, A->nnz * s));
  MFP_CHECK(mfp_buffer_resize(cd, A->nnz * 4));
  MFP_CHECK(mfp_buffer_resize(rd, ((size_t)A->rows + 1) * 4));
  MFP_CHECK(mfp_buffer_resize(xd, (size_t)A->cols * s));
  MFP_CHECK(mfp_buffer_resize(yd, (size_t)A->rows * s));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(Ad), A->data, A->nnz * s));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(cd), A->columns, A->nnz * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(rd), A->row_ptr, ((size_t)A->rows + 1) * 4));
  MFP_CHECK(mfp_fill(dtype, mfp_buffer_ptr(xd), A->cols, 1.0, nullptr));
  MFP_CHECK(mfp_fill(dtype, mfp_buffer_ptr(yd), A->rows, y_init, nullptr));
  return MFP_OK;
}

int upload_ell(int dtype, const mfp_host_ell *E, mfp_buffer Ad, mfp_buffer cd)
{
  const size_t s = dtype_size(dtype);
  const uint64_t slots = (uint64_t)E->rows * E->elements_in_rows;
  MFP_CHECK(mfp_buffer_resize(Ad, slots * s));
  MFP_CHECK(mfp_buffer_resize(cd, slots * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(Ad), E->data, slots * s));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(cd), E->columns, slots * 4));
  return MFP_OK;
}

int upload_coo(int dtype, const mfp_host_coo *C, mfp_buffer Ad, mfp_buffer cd, mfp_buffer rd)
{
  const size_t s = dtype_size(dtype);
  MFP_CHECK(mfp_buffer_resize(Ad, C->elements * s));
  MFP_CHECK(mfp_buffer_resize(cd, C->elements * 4));
  MFP_CHECK(mfp_buffer_resize(rd, C->elements * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(Ad), C->data, C->elements * s));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(cd), C->cols_, C->elements * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(rd), C->rows_, C->elements * 4));
  return MFP_OK;
}

int xy_buffers(int dtype, uint32_t rows, uint32_t cols, mfp_buffer xd, mfp_buffer yd, double y_init)
{
  const size_t s = dtype_size(dtype);
  MFP_CHECK(mfp_buffer_resize(xd, (size_t)cols * s));
  MFP_CHECK(mfp_buffer_resize(yd, (size_t)rows * s));
  MFP_CHECK(mfp_fill(dtype, mfp_buffer_ptr(xd), cols, 1.0, nullptr));
  MFP_CHECK(mfp_fill(dtype, mfp_buffer_ptr(yd), rows, y_init, nullptr));
  return MFP_OK;
}

// One 32-lane group per row, as the reference's warp (gpu_matrix_multiplier.cu:331-362,
// launched with warp_size = 32 at :412-417)
constexpr uint32_t REFERENCE_VECTOR_LANES = 32;

// The hybrid's COO tail on the host, one entry at a time in stored order
// (gpu_matrix_multiplier.cu:1163-1168); the product is rounded before the add.
template <typename T>
void host_coo_tail(const mfp_host_coo *C, const T *x, T *y)
{
#pragma clang fp contract(off)
  const T *a = (const T *)C->data;
  for (uint64_t e = 0; e < C->elements; e++) {
    const T p = a[e] * x[C->cols_[e]];
    y[C->rows_[e]] = y[C->rows_[e]] + p;
  }
}

struct two_streams {
  hipStream_t ell = nullptr, copy = nullptr;
  hipEvent_t copied = nullptr;
  ~two_streams()
  {
    if (copied) (void)hipEventDestroy(copied);
    if (copy) (void)hipStreamDestroy(copy);
    if (ell) (void)hipStreamDestroy(ell);
  }
  int create()
  {
    MFP_HIP(hipStreamCreateWithFlags(&ell, hipStreamNonBlocking));
    MFP_HIP(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
    MFP_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
    return MFP_OK;
  }
};

} // namespace

extern "C" {

int mfp_compare_results(int dtype, uint32_t n, const void *a, const void *b)
{
  MFP_ARG(dtype == MFP_F32 || dtype == MFP_F64, "dtype must be f32 or f64");
  MFP_ARG((a && b) || n == 0, "null pointer");
  compare_results(dtype, n, a, b);
  return MFP_OK;
}

int mfp_gpu_csr_spmv(int dtype, const mfp_host_csr *A, mfp_buffer Ad, mfp_buffer cd,
                     mfp_buffer rd, mfp_buffer xd, mfp_buffer yd, void *reusable_vector,
                     const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(out, "out is null");
  MFP_CHECK(upload_csr(dtype, A, Ad, cd, rd, xd, yd, 0.0));
  double sec = 0;
  MFP_CHECK(run_timed([&] {
    return launch_csr_scalar(dtype, A->rows, (const uint32_t *)mfp_buffer_ptr(rd),
                             (const uint32_t *)mfp_buffer_ptr(cd), mfp_buffer_ptr(Ad),
                             mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, &sec));
  MFP_CHECK(finish(dtype, A->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  fill_measurement(out, "GPU CSR", sec, csr_bytes(dtype_size(dtype), A->rows, A->nnz), 2.0 * (double)A->nnz);
  return MFP_OK;
}

int mfp_gpu_csr_vector_spmv(int dtype, const mfp_host_csr *A, mfp_buffer Ad, mfp_buffer cd,
                            mfp_buffer rd, mfp_buffer xd, mfp_buffer yd, void *reusable_vector,
                            const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(out, "out is null");
  // y pre-filled with 24 so an unwritten row shows (gpu_matrix_multiplier.cu:403)
  MFP_CHECK(upload_csr(dtype, A, Ad, cd, rd, xd, yd, 24.0));
  double sec = 0;
  MFP_CHECK(run_timed([&] {
    return launch_csr_vector(dtype, REFERENCE_VECTOR_LANES, A->rows, (const uint32_t *)mfp_buffer_ptr(rd),
                             (const uint32_t *)mfp_buffer_ptr(cd), mfp_buffer_ptr(Ad),
                             mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, &sec));
  MFP_CHECK(finish(dtype, A->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  fill_measurement(out, "GPU CSR (vector)", sec, csr_bytes(dtype_size(dtype), A->rows, A->nnz),
                   2.0 * (double)A->nnz);
  return MFP_OK;
}

int mfp_gpu_csr_adaptive_spmv(int dtype, const mfp_host_csr *A, mfp_buffer Ad, mfp_buffer cd,
                              mfp_buffer rd, mfp_buffer xd, mfp_buffer yd, void *reusable_vector,
                              const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(out, "out is null");
  MFP_CHECK(upload_csr(dtype, A, Ad, cd, rd, xd, yd, 0.0));
  const uint32_t W = default_nnz_per_wg(dtype), chunk = DEFAULT_CHUNK, long_row = DEFAULT_LONG_ROW;
  adaptive_layout lay;
  MFP_CHECK(build_adaptive_layout(dtype, A->rows, (const uint32_t *)mfp_buffer_ptr(rd), W,
                                  long_row, chunk, &lay, nullptr));
  adaptive_args g;
  g.rows = A->rows;
  g.rp = (const uint32_t *)mfp_buffer_ptr(rd);
  g.ci = (const uint32_t *)mfp_buffer_ptr(cd);
  g.blocks = lay.blocks;
  g.nblocks = lay.nblocks;
  g.nnz_per_wg = W;
  g.long_row = long_row;
  g.chunks = lay.chunks;
  g.nchunks = lay.nchunks;
  g.chunk = chunk;
  g.partials = lay.partials;
  double sec = 0;
  int st = run_timed([&] {
    return launch_csr_adaptive(dtype, g, mfp_buffer_ptr(Ad), mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, &sec);
  free_adaptive_layout(&lay);
  MFP_CHECK(st);
  MFP_CHECK(finish(dtype, A->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  // csr_adaptive_spmv.cu:320 records no bytes / flops
  fill_measurement(out, "GPU CSR-Adaptive", sec, 0.0, 0.0);
  return MFP_OK;
}

int mfp_gpu_ell_spmv(int dtype, const mfp_host_ell *E, mfp_buffer Ad, mfp_buffer cd,
                     mfp_buffer xd, mfp_buffer yd, void *reusable_vector, const void *reference_y,
                     mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(E && out, "null pointer");
  MFP_CHECK(upload_ell(dtype, E, Ad, cd));
  MFP_CHECK(xy_buffers(dtype, E->rows, E->cols, xd, yd, 0.0));
  double sec = 0;
  MFP_CHECK(run_timed([&] {
    return launch_ell(dtype, E->rows, E->elements_in_rows, (const uint32_t *)mfp_buffer_ptr(cd),
                      mfp_buffer_ptr(Ad), mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, &sec));
  MFP_CHECK(finish(dtype, E->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  // gpu_matrix_multiplier.cu:537-541: n_elements is an unsigned int product
  const size_t s = dtype_size(dtype);
  const uint32_t n = E->elements_in_rows * E->rows;
  const double bytes = (double)((uint64_t)n * s + (uint64_t)n * s + (uint64_t)n * 4 + (uint64_t)E->rows * s);
  fill_measurement(out, "GPU ELL", sec, bytes, 2.0 * (double)n);
  return MFP_OK;
}

int mfp_gpu_coo_spmv(int dtype, const mfp_host_coo *C, mfp_buffer Ad, mfp_buffer cd,
                     mfp_buffer rd, mfp_buffer xd, mfp_buffer yd, void *reusable_vector,
                     const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(C && out, "null pointer");
  MFP_CHECK(upload_coo(dtype, C, Ad, cd, rd));
  MFP_CHECK(xy_buffers(dtype, C->rows, C->cols, xd, yd, 0.0));
  double sec = 0;
  MFP_CHECK(run_timed([&] {
    return launch_coo_add(dtype, C->elements, (const uint32_t *)mfp_buffer_ptr(rd),
                          (const uint32_t *)mfp_buffer_ptr(cd), mfp_buffer_ptr(Ad),
                          mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, &sec));
  MFP_CHECK(finish(dtype, C->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  const size_t s = dtype_size(dtype);
  const double bytes = (double)(C->nnz * s + C->nnz * s + C->nnz * 4 + C->nnz * 4 + C->nnz * s);
  fill_measurement(out, "GPU COO", sec, bytes, 2.0 * (double)C->nnz);
  return MFP_OK;
}

int mfp_gpu_scoo_spmv(int dtype, int print_diff, const mfp_host_scoo *S, mfp_buffer Ad,
                      mfp_buffer r_index, mfp_buffer c_index, mfp_buffer xd, mfp_buffer yd,
                      void *reusable_vector, const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(S && out, "null pointer");
  MFP_ARG(S->slice_size > 0 && (uint64_t)S->slices_count * S->slice_size >= S->rows,
          "SCOO slices must cover every row");
  const size_t s = dtype_size(dtype);
  MFP_CHECK(mfp_buffer_resize(Ad, S->nnz * s));
  MFP_CHECK(mfp_buffer_resize(c_index, S->nnz * 4));
  MFP_CHECK(mfp_buffer_resize(r_index, S->nnz * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(Ad), S->values, S->nnz * s));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(c_index), S->c_index, S->nnz * 4));
  MFP_CHECK(mfp_memcpy_h2d(mfp_buffer_ptr(r_index), S->r_index, S->nnz * 4));
  MFP_CHECK(xy_buffers(dtype, S->rows, S->cols, xd, yd, 0.0));
  // slices_ptr: validated on the host before it reaches the kernel
  const uint32_t real = S->rows ? (S->rows + S->slice_size - 1) / S->slice_size : 0;
  for (uint32_t i = 0; i < real; i++)
    MFP_ARG(S->slices_ptr[i] <= S->slices_ptr[i + 1] && S->slices_ptr[i + 1] <= S->nnz,
            "slices_ptr must be non-decreasing and within nnz");
  uint32_t *sp = nullptr;
  MFP_CHECK(dmalloc((void **)&sp, ((size_t)S->slices_count + 1) * 4));
  int st = mfp_memcpy_h2d(sp, S->slices_ptr, ((size_t)S->slices_count + 1) * 4);
  double sec = 0;
  if (st == MFP_OK)
    st = run_timed([&] {
      return launch_scoo(dtype, S->rows, S->slice_size, S->lane_size, 1024,
                         (const uint32_t *)mfp_buffer_ptr(r_index),
                         (const uint32_t *)mfp_buffer_ptr(c_index), mfp_buffer_ptr(Ad), sp,
                         mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
    }, &sec);
  dfree(sp);
  MFP_CHECK(st);
  MFP_CHECK(finish(dtype, S->rows, mfp_buffer_ptr(yd), reusable_vector,
                   print_diff ? reference_y : nullptr));
  fill_measurement(out, "GPU SCOO", sec, 0.0, 0.0); // scoo_spmv.cu:221
  return MFP_OK;
}

static int hybrid_common(int dtype, const mfp_host_ell *E, const mfp_host_coo *C, mfp_buffer Ae,
                         mfp_buffer Ac, mfp_buffer ec, mfp_buffer cc, mfp_buffer cr, mfp_buffer xd,
                         mfp_buffer yd, bool fused, double *sec)
{
  MFP_ARG(E && C, "null pointer");
  MFP_CHECK(upload_ell(dtype, E, Ae, ec));
  MFP_CHECK(upload_coo(dtype, C, Ac, cc, cr));
  MFP_CHECK(xy_buffers(dtype, E->rows, E->cols, xd, yd, 0.0));
  return run_timed([&] {
    if (fused)
      return launch_hyb_atomic(dtype, E->rows, E->elements_in_rows,
                               (const uint32_t *)mfp_buffer_ptr(ec), mfp_buffer_ptr(Ae), C->elements,
                               (const uint32_t *)mfp_buffer_ptr(cr), (const uint32_t *)mfp_buffer_ptr(cc),
                               mfp_buffer_ptr(Ac), mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
    int st = launch_ell(dtype, E->rows, E->elements_in_rows, (const uint32_t *)mfp_buffer_ptr(ec),
                        mfp_buffer_ptr(Ae), mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
    if (st != MFP_OK) return st;
    return launch_coo_add(dtype, C->elements, (const uint32_t *)mfp_buffer_ptr(cr),
                          (const uint32_t *)mfp_buffer_ptr(cc), mfp_buffer_ptr(Ac),
                          mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), nullptr);
  }, sec);
}

int mfp_gpu_hybrid_spmv(int dtype, const mfp_host_ell *E, const mfp_host_coo *C, mfp_buffer Ae,
                        mfp_buffer Ac, mfp_buffer ec, mfp_buffer cc, mfp_buffer cr, mfp_buffer xd,
                        mfp_buffer yd, void *reusable_vector, const void *reference_y,
                        mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(out, "out is null");
  double sec = 0;
  MFP_CHECK(hybrid_common(dtype, E, C, Ae, Ac, ec, cc, cr, xd, yd, false, &sec));
  MFP_CHECK(finish(dtype, E->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  fill_measurement(out, "GPU Hybrid (TODO)", sec, 1.0, 1.0); // gpu_matrix_multiplier.cu:962-966
  return MFP_OK;
}

int mfp_gpu_hybrid_atomic_spmv(int dtype, const mfp_host_ell *E, const mfp_host_coo *C,
                               mfp_buffer Ae, mfp_buffer Ac, mfp_buffer ec, mfp_buffer cc,
                               mfp_buffer cr, mfp_buffer xd, mfp_buffer yd, void *reusable_vector,
                               const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(out, "out is null");
  double sec = 0;
  MFP_CHECK(hybrid_common(dtype, E, C, Ae, Ac, ec, cc, cr, xd, yd, true, &sec));
  MFP_CHECK(finish(dtype, E->rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  fill_measurement(out, "GPU Hybrid (atomic)", sec, 0.0, 0.0); // :1091-1095
  return MFP_OK;
}

int mfp_gpu_hybrid_cpu_coo_spmv(int dtype, const mfp_host_ell *E, const mfp_host_coo *C,
                                mfp_buffer Ae, mfp_buffer ec, mfp_buffer xd, mfp_buffer yd,
                                mfp_buffer tmp, void *cpu_y, void *reusable_vector,
                                const void *reference_y, mfp_measurement *out)
{
  MFP_CHECK(check_dtype(dtype));
  MFP_ARG(E && C && out && tmp, "null pointer");
  MFP_ARG(cpu_y || E->rows == 0, "cpu_y is null");
  MFP_ARG(C->rows == E->rows && C->cols == E->cols, "the ELL and COO parts differ in shape");
  for (uint64_t e = 0; e < C->elements; e++)
    MFP_ARG(C->rows_[e] < E->rows && C->cols_[e] < E->cols, "COO index out of range");
  const uint32_t rows = E->rows;
  const size_t s = dtype_size(dtype);
  MFP_CHECK(upload_ell(dtype, E, Ae, ec));
  MFP_CHECK(xy_buffers(dtype, rows, E->cols, xd, yd, 0.0));
  MFP_CHECK(mfp_buffer_resize(tmp, (size_t)rows * s));
  // host x = 1 over cols, cpu_y = 0 over rows (gpu_matrix_multiplier.cu:1131-1133)
  std::vector<double> x64;
  std::vector<float> x32;
  if (dtype == MFP_F32) {
    x32.assign(E->cols, 1.0f);
    std::fill_n((float *)cpu_y, rows, 0.0f);
  } else {
    x64.assign(E->cols, 1.0);
    std::fill_n((double *)cpu_y, rows, 0.0);
  }
  two_streams st;
  MFP_CHECK(st.create());
  timed ev;
  MFP_CHECK(ev.create());
  MFP_HIP(hipDeviceSynchronize());
  MFP_HIP(hipEventRecord(ev.a, st.ell));
  // the ELL part runs on the GPU while the host sums the COO tail; the tail
  // is copied on a second stream and added once both are done
  MFP_CHECK(launch_ell(dtype, rows, E->elements_in_rows, (const uint32_t *)mfp_buffer_ptr(ec),
                       mfp_buffer_ptr(Ae), mfp_buffer_ptr(xd), mfp_buffer_ptr(yd), st.ell));
  if (dtype == MFP_F32)
    host_coo_tail<float>(C, x32.data(), (float *)cpu_y);
  else
    host_coo_tail<double>(C, x64.data(), (double *)cpu_y);
  MFP_HIP(hipMemcpyAsync(mfp_buffer_ptr(tmp), cpu_y, (size_t)rows * s, hipMemcpyHostToDevice, st.copy));
  MFP_HIP(hipEventRecord(st.copied, st.copy));
  MFP_HIP(hipStreamWaitEvent(st.ell, st.copied, 0));
  MFP_CHECK(launch_vec_add(dtype, rows, mfp_buffer_ptr(tmp), mfp_buffer_ptr(yd), st.ell));
  MFP_HIP(hipEventRecord(ev.b, st.ell));
  MFP_HIP(hipEventSynchronize(ev.b));
  float ms = 0;
  MFP_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
  MFP_CHECK(finish(dtype, rows, mfp_buffer_ptr(yd), reusable_vector, reference_y));
  fill_measurement(out, "GPU Hybrid (CPU COO TODO)", ms / 1000.0, 1.0, 1.0); // :1209-1213
  return MFP_OK;
}

} // extern "C"
