// Internal declarations shared by the libmfp translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mfp.h"

namespace mfp {

// Per-thread last error text, returned by mfp_last_error().
void set_error(const std::string &msg);
const std::string &last_error();

// Return from the calling C-ABI function with MFP_ERR_HIP when a HIP call fails.
#define MFP_HIP(expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::mfp::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) +      \
                       " (" __FILE__ ":" + std::to_string(__LINE__) + ")");     \
      return MFP_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

#define MFP_CHECK(expr)                                                         \
  do {                                                                          \
    int s_ = (expr);                                                            \
    if (s_ != MFP_OK)                                                           \
      return s_;                                                                \
  } while (0)

#define MFP_ARG(cond, msg)                                                      \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::mfp::set_error(std::string("invalid argument: ") + (msg));              \
      return MFP_ERR_ARG;                                                       \
    }                                                                           \
  } while (0)

// Kernel marks for mfp_plan_kernel_times: while one is installed on the
// calling thread, the check that follows every kernel launch records an
// event on the marks' stream, so that consecutive events bracket consecutive
// kernels of one SpMV.
struct kernel_marks {
  static constexpr int kMax = 8;
  hipStream_t s = nullptr;
  hipEvent_t ev[kMax + 1] = {};
  hipEvent_t start[kMax] = {}; // kernel i's own start, when exact[i]
  bool exact[kMax] = {};
  int n = 0; // events recorded after ev[0]
};
extern thread_local kernel_marks *g_kernel_marks;

inline void mark_kernel_end()
{
  kernel_marks *k = g_kernel_marks;
  if (k && k->n < kernel_marks::kMax) (void)hipEventRecord(k->ev[1 + k->n++], k->s);
}

// A launcher that launches through hipExtLaunchKernelGGL asks here for the
// start / stop events of its kernel: true, with both, when marks are
// installed (the kernel's own start and end times are then recorded).
inline bool mark_kernel_events(hipEvent_t *start, hipEvent_t *stop)
{
  kernel_marks *k = g_kernel_marks;
  if (!k || k->n >= kernel_marks::kMax) return false;
  *start = k->start[k->n];
  *stop = k->ev[1 + k->n];
  k->exact[k->n++] = true;
  return true;
}

inline size_t dtype_size(int dtype)
{
  switch (dtype) {
  case MFP_F32: return 4;
  case MFP_F64: return 8;
  case MFP_BF16: return 2;
  case MFP_F16: return 2;
  default: return 0;
  }
}

inline hipStream_t as_stream(mfp_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t ceil_div_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Device properties cached by mfp_init.
struct device_props {
  int device = -1;
  int cu_count = 0;
  int xcd_count = 8;
  size_t lds_per_block = 0;
  size_t free_mem = 0, total_mem = 0;
  char name[256] = {0};
  char arch[64] = {0};
};
const device_props &props();
int ensure_init();

// Stream-ordered device allocation helpers (plain hipMalloc; plans own them).
int dmalloc(void **p, size_t bytes);
void dfree(void *p);

// ---- converter entry points used by plans (convert.hip) -------------------
int row_stats_device(uint32_t rows, const uint32_t *row_ptr_d, mfp_row_stats *out,
                     hipStream_t s);
int ell_build_device(int dtype, uint32_t rows, const uint32_t *row_ptr, const uint32_t *col,
                     const void *val, uint32_t width, void *ell_val, uint32_t *ell_col,
                     hipStream_t s);
int coo_tail_build_device(int dtype, uint32_t rows, uint64_t nnz, const uint32_t *row_ptr,
                          const uint32_t *col, const void *val, uint32_t element_start,
                          void **coo_val, uint32_t **coo_col, uint32_t **coo_row,
                          uint64_t *count, hipStream_t s);
int scoo_slices_device(uint32_t rows, uint64_t nnz, const uint32_t *row_ptr,
                       uint32_t slices_count, uint32_t slice_size, uint32_t *slices_ptr,
                       hipStream_t s);
int row_blocks_device(uint32_t rows, const uint32_t *row_ptr, uint32_t nnz_per_wg,
                      uint32_t **blocks_out, uint32_t *count_out, hipStream_t s);
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s,
                       uint32_t *total_out /* host, may be null */);
// Column-sorted sliced COO layout (plans with MFP_TUNE_SCOO_COLSORT).
// Slices 0..S_reg-1 hold R consecutive rows each; slices S_reg..S-1 hold the
// long rows, R per slice in row order (long_rows[] gives their rows,
// long_mask[] marks them).  A work item {slice, first entry, end entry, mode}
// is one workgroup; mode is a partial slot, or CS_DIRECT when the item is its
// slice's only one and writes the slice's rows of y.  fin[] lists finish
// blocks {slice, first slot, slots, first row}: the first NFR cover
// 4 CS_FIN_ROWS rows of an ordinary slice (a thread per four rows), the rest
// CS_FIN_LONG rows of a long-row slice (a wave per row, lanes over slots).
constexpr uint32_t CS_DIRECT = 0xFFFFFFFEu, CS_FIN_ROWS = 256, CS_FIN_LONG = CS_FIN_ROWS / 64;
struct scoo_cs_layout {
  uint32_t R = 0, rbits = 0, S_reg = 0, S_long = 0, S = 0, NL = 0;
  uint32_t NI = 0, NF = 0, NFR = 0, slots = 0;
  uint32_t kmax = 1;             // widest SpMM (columns of X) the partial slots hold
  uint32_t *sp = nullptr;        // S + 1 entry offsets of the slices
  uint32_t *long_rows = nullptr; // NL
  uint32_t *long_mask = nullptr; // ceil(rows / 32) bit words, when NL > 0
  uint4 *items = nullptr;        // NI
  uint4 *fin = nullptr;          // NF
  uint32_t *base = nullptr;      // NI first column of each item
  void *partial = nullptr;       // slots * R * kmax
  void *ev = nullptr;            // values in slice-major, column-sorted order
  uint32_t *ec = nullptr;        // (column - item base) << rbits | local row
  uint64_t device_bytes = 0;
};
int scoo_cs_build_device(int dtype, uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t *rp,
                         const uint32_t *ci, const void *val, uint32_t R, uint32_t long_threshold,
                         uint64_t item_entries, scoo_cs_layout *out, hipStream_t s,
                         bool own_partials = true);
void scoo_cs_free(scoo_cs_layout *l);
// Binned CSR over 16-bit values (plans of format MFP_CSR_ADAPTIVE): row lists
// per lane count G = 2 << b (list[bin_off[b] .. bin_off[b+1])), chunks
// {row, first entry, end entry, slot} of long rows, and per long row
// {row, first slot, slots}.
struct lp_binned_layout {
  uint32_t bin_off[7] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t nchunks = 0, nlong = 0;
  uint32_t *list = nullptr;
  uint4 *chunks = nullptr, *longs = nullptr;
  float2 *partial = nullptr;
  uint64_t device_bytes = 0;
};
int lp_binned_build(uint32_t rows, const uint32_t *rp_dev, lp_binned_layout *out, hipStream_t s,
                    bool own_partials = true);
void lp_binned_free(lp_binned_layout *l);
int launch_lp_binned(int dt, const lp_binned_layout &l, const uint32_t *rp, const uint32_t *ci, const void *a,
                     const void *x, void *y, hipStream_t s);
int fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
// looks up the current device's address of mfp_plan_spmv_noalloc's static
// scratch (plan.hip), once per device
int static_scratch_setup();
int fill_bytes_zero(void *p, size_t bytes, hipStream_t s);
int convert_values(int dtype_in, const void *in, int dtype_out, void *out, uint64_t n,
                   hipStream_t s);

} // namespace mfp
