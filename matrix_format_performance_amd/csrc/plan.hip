// Device-resident matrices and per-format plans.
#include "kernels.h"
#include "mfp_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace mfp {

// ---- CSR-Adaptive layout: row blocks + long-row chunk items ----------------
__global__ void k_chunk_counts(const uint32_t *__restrict__ blocks, uint32_t nb,
                               const uint32_t *__restrict__ rp, uint32_t long_row,
                               uint32_t chunk, uint32_t *__restrict__ cnt)
{
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const uint32_t r0 = blocks[b], r1 = blocks[b + 1];
  uint32_t c = 0;
  if (r1 == r0 + 1) {
    const uint32_t len = rp[r0 + 1] - rp[r0];
    if (len > long_row) c = (len + chunk - 1) / chunk;
  }
  cnt[b] = c;
}

__global__ void k_chunk_fill(const uint32_t *__restrict__ blocks, uint32_t nb,
                             const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ base,
                             uint4 *__restrict__ chunks)
{
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const uint32_t n = cnt[b];
  for (uint32_t c = 0; c < n; c++) chunks[base[b] + c] = make_uint4(blocks[b], c, base[b], n);
}

void free_adaptive_layout(adaptive_layout *l)
{
  dfree(l->blocks);
  dfree(l->chunks);
  dfree(l->partials);
  *l = adaptive_layout{};
}

int build_adaptive_layout(int dtype, uint32_t rows, const uint32_t *rp, uint32_t W,
                          uint32_t long_row, uint32_t chunk, adaptive_layout *out, hipStream_t s,
                          bool own_partials)
{
  MFP_ARG(W >= 1 && adaptive_vectors_per_lane(dtype, W) <= 16, "nnz_per_wg too large");
  MFP_ARG(chunk >= 256, "chunk must be >= 256");
  MFP_ARG(long_row >= chunk, "long_row must be >= chunk");
  adaptive_layout l;
  MFP_CHECK(row_blocks_device(rows, rp, W, &l.blocks, &l.nblocks, s));
  uint32_t *cnt = nullptr, *base = nullptr;
  int st = dmalloc((void **)&cnt, (uint64_t)l.nblocks * 4 + 4);
  if (st == MFP_OK) st = dmalloc((void **)&base, (uint64_t)l.nblocks * 4 + 4);
  if (st == MFP_OK && l.nblocks) {
    k_chunk_counts<<<ceil_div_u32(l.nblocks, 256), 256, 0, s>>>(l.blocks, l.nblocks, rp, long_row,
                                                                chunk, cnt);
    st = exclusive_scan_u32(cnt, base, l.nblocks, s, &l.nchunks);
  }
  if (st == MFP_OK) st = dmalloc((void **)&l.chunks, (uint64_t)l.nchunks * sizeof(uint4));
  if (st == MFP_OK && own_partials) st = dmalloc(&l.partials, (uint64_t)l.nchunks * dtype_size(dtype));
  if (st == MFP_OK && l.nchunks)
    k_chunk_fill<<<ceil_div_u32(l.nblocks, 256), 256, 0, s>>>(l.blocks, l.nblocks, cnt, base, l.chunks);
  if (st == MFP_OK && hipStreamSynchronize(s) != hipSuccess) st = MFP_ERR_HIP;
  dfree(cnt);
  dfree(base);
  if (st != MFP_OK) {
    free_adaptive_layout(&l);
    return st;
  }
  *out = l;
  return MFP_OK;
}

} // namespace mfp

using namespace mfp;

#include "objects.h"

namespace mfp {
int launch_csr_vector_lp(int dt, uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                         const void *a, const void *x, void *y, hipStream_t s);
int launch_ell_lp(int dt, bool off16, bool block, uint32_t rows, uint32_t w, const void *cols, const uint32_t *base,
                  const void *ev, const void *x, void *y, hipStream_t s);
int block_base_build(uint32_t rows, const uint32_t *rp, const uint32_t *ci, uint32_t *base, uint32_t *span,
                     hipStream_t s);
int row_span_max(uint32_t rows, const uint32_t *rp, const uint32_t *ci, uint32_t *out, hipStream_t s);
uint32_t ell_lp_stride(uint32_t rows);
int ell_lp_build(uint32_t rows, const uint32_t *rp, const uint32_t *ci, const void *a, uint32_t w,
                 bool off16, bool block, void *ev, void *cols, uint32_t *base, hipStream_t s);
int to_lp_values(const void *in, int din, void *out, int dout, uint64_t n, hipStream_t s);
} // namespace mfp

static bool is_lp(int dtype) { return dtype == MFP_BF16 || dtype == MFP_F16; }

static int matrix_stats(mfp_matrix m, mfp_row_stats *out, hipStream_t s)
{
  if (!m->stats_valid) {
    MFP_CHECK(row_stats_device(m->rows, m->row_ptr, &m->stats, s));
    m->stats_valid = true;
  }
  *out = m->stats;
  return MFP_OK;
}

static uint32_t next_pow2_u32(uint32_t v)
{
  uint32_t p = 1;
  while (p < v && p < 0x80000000u) p <<= 1;
  return p;
}

static uint32_t prev_pow2_u32(uint32_t v)
{
  if (v == 0) return 0;
  uint32_t p = 1;
  while ((p << 1) <= v && (p << 1) != 0) p <<= 1;
  return p;
}

// MI355X SCOO sizing: a slice's accumulator acc[slice_size][lane_size] takes
// 32 KB of the 160 KB LDS (4 resident workgroups per CU) in double, the type
// the kernel accumulates FP32 values in too, with at least 4 slices per CU
// when the row count allows.
static void scoo_mi355x_geometry(int dtype, uint32_t rows, uint32_t *slices, uint32_t *ss,
                                 uint32_t *L)
{
  (void)dtype;
  const uint32_t cus = (uint32_t)std::max(1, props().cu_count);
  const uint32_t acc_elems = (uint32_t)((32u * 1024u) / sizeof(double));
  uint32_t s = prev_pow2_u32(std::max<uint32_t>(1, rows / (4 * cus)));
  s = std::min<uint32_t>(std::max<uint32_t>(s, 64), acc_elems / 2);
  uint32_t lanes = std::min<uint32_t>(32, prev_pow2_u32(acc_elems / s));
  *ss = s;
  *L = std::max<uint32_t>(1, lanes);
  *slices = rows ? ceil_div_u32(rows, s) : 1;
}

static int plan_free_arrays(mfp_plan p)
{
  dfree(p->ell_val);
  dfree(p->ell_col);
  dfree(p->coo_val);
  dfree(p->coo_col);
  dfree(p->coo_row);
  dfree(p->sp);
  dfree(p->lp_cols);
  dfree(p->lp_base);
  dfree(p->pk_val);
  dfree(p->pk_off);
  dfree(p->pk_base);
  scoo_cs_free(&p->cs);
  lp_binned_free(&p->lpb);
  free_adaptive_layout(&p->ad);
  dfree(p->stream_meta);
  return MFP_OK;
}

// Plans over 16-bit values: CSR-vector, binned CSR and ELL with
// Dot2-compensated FP32 accumulation, and column-sorted SCOO with exact
// products added in double.
static int plan_create_lp(mfp_plan p, mfp_matrix_s *m, const mfp_tuning &t, hipStream_t s);

static uint64_t budget_of(const mfp_tuning &t)
{
  if (t.budget_bytes) return t.budget_bytes;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  return (uint64_t)(0.9 * (double)fr);
}

static int matrix_stats(mfp_matrix m, mfp_row_stats *out, hipStream_t s);

static int plan_create_scoo_cs(mfp_plan p, mfp_matrix_s *m, const mfp_tuning &t, hipStream_t s);

static int plan_create_lp(mfp_plan p, mfp_matrix_s *m, const mfp_tuning &t, hipStream_t s)
{
  if (p->format == MFP_SCOO && (t.flags & MFP_TUNE_SCOO_COLSORT)) return plan_create_scoo_cs(p, m, t, s);
  if (p->format == MFP_CSR_VECTOR) {
    if (t.lanes < 2 || t.lanes > 64 || (t.lanes & (t.lanes - 1))) {
      set_error("lanes must be a power of two in [2, 64]");
      return MFP_ERR_ARG;
    }
    return MFP_OK;
  }
  if (p->format == MFP_CSR_ADAPTIVE) {
    MFP_CHECK(lp_binned_build(m->rows, m->row_ptr, &p->lpb, s, !(t.flags & MFP_TUNE_STATIC_PARTIALS)));
    p->device_bytes = p->lpb.device_bytes;
    return MFP_OK;
  }
  if (p->format != MFP_ELL) {
    set_error("16-bit values are available in the CSR-vector, CSR-Adaptive (binned), ELL and column-sorted SCOO "
              "formats");
    return MFP_ERR_ARG;
  }
  mfp_row_stats rs;
  MFP_CHECK(matrix_stats(m, &rs, s));
  const uint32_t w = rs.max_elements_in_rows;
  uint32_t span = 0;
  MFP_CHECK(row_span_max(m->rows, m->row_ptr, m->columns, &span, s));
  const uint32_t stride = ell_lp_stride(m->rows); // rows rounded up to a multiple of 64
  // offsets from the first column of each 64-row block when every block's
  // columns span less than 65536 (one base per block), else per row
  uint32_t *bbase = nullptr;
  uint32_t bspan = 0xFFFFFFFFu;
  if (!(t.flags & MFP_TUNE_NO_OFFSET16) && m->rows) {
    MFP_CHECK(dmalloc((void **)&bbase, (uint64_t)stride / 64 * 4));
    const int bst = block_base_build(m->rows, m->row_ptr, m->columns, bbase, &bspan, s);
    if (bst != MFP_OK) {
      dfree(bbase);
      return bst;
    }
  }
  p->lp_block = bspan < 65536u;
  if (!p->lp_block) {
    dfree(bbase);
    bbase = nullptr;
  }
  p->off16 = p->lp_block || (span < 65536u && !(t.flags & MFP_TUNE_NO_OFFSET16));
  // with 16-bit offsets the slots are packed 4 per 16-byte word (w rounded up to 4)
  const uint64_t slots = (uint64_t)stride * (p->off16 ? (w + 3) / 4 * 4 : w);
  const uint64_t base_bytes = p->lp_block ? (uint64_t)stride / 64 * 4 : (p->off16 ? (uint64_t)stride * 4 : 0);
  const uint64_t need = slots * (p->off16 ? 4 : 6) + base_bytes;
  if (need > budget_of(t)) {
    dfree(bbase);
    set_error("ELL part needs " + std::to_string(need) + " bytes, above the budget");
    return MFP_ERR_BUDGET;
  }
  p->lp_base = bbase; // the plan owns it from here
  p->ell_w = w;
  // with 16-bit offsets, value and offset share one u32 slot in lp_cols
  if (!p->off16) MFP_CHECK(dmalloc(&p->ell_val, slots * 2));
  MFP_CHECK(dmalloc(&p->lp_cols, slots * 4));
  if (p->off16 && !p->lp_block) MFP_CHECK(dmalloc((void **)&p->lp_base, (uint64_t)stride * 4));
  MFP_CHECK(ell_lp_build(m->rows, m->row_ptr, m->columns, m->values, w, p->off16, p->lp_block, p->ell_val,
                         p->lp_cols, p->lp_base, s));
  MFP_HIP(hipStreamSynchronize(s));
  p->device_bytes = need;
  return MFP_OK;
}

// ---- the allocation-free SpMV ------------------------------------------------
// Partial results of mfp_plan_spmv_noalloc live in one static array per
// device (each device that runs it holds MFP_NOALLOC_SCRATCH_BYTES until the
// process ends).  In host code `scratch` names a host-side stand-in, not a
// device address: the address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address reaches kernels.
static __device__ __align__(256) unsigned char scratch[MFP_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[64] = {};

namespace mfp {
// set-up: called by the library's init for its device, and by the entry
// point for a device that init never saw
int static_scratch_setup()
{
  int dev = -1;
  MFP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device index beyond the static scratch table");
    return MFP_ERR_STATE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    MFP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  return MFP_OK;
}
} // namespace mfp

extern "C" {

int mfp_matrix_from_host_csr(mfp_matrix *out, int dtype, uint32_t rows, uint32_t cols,
                             uint64_t nnz, const uint32_t *row_ptr, const uint32_t *columns,
                             const void *values)
{
  MFP_ARG(out && row_ptr, "null pointer");
  MFP_ARG(dtype_size(dtype) != 0, "dtype");
  MFP_ARG(nnz < (1ull << 32), "nnz must fit in 32 bits");
  MFP_ARG(row_ptr[0] == 0 && row_ptr[rows] == nnz, "row_ptr must start at 0 and end at nnz");
  MFP_CHECK(ensure_init());
  auto *m = new mfp_matrix_s;
  m->dtype = dtype;
  m->rows = rows;
  m->cols = cols;
  m->nnz = nnz;
  const size_t vs = dtype_size(dtype);
  int st = dmalloc((void **)&m->row_ptr, ((uint64_t)rows + 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->columns, nnz * 4);
  if (st == MFP_OK) st = dmalloc(&m->values, nnz * vs);
  if (st == MFP_OK) st = mfp_memcpy_h2d(m->row_ptr, row_ptr, ((uint64_t)rows + 1) * 4);
  if (st == MFP_OK) st = mfp_memcpy_h2d(m->columns, columns, nnz * 4);
  if (st == MFP_OK) st = mfp_memcpy_h2d(m->values, values, nnz * vs);
  if (st != MFP_OK) {
    mfp_matrix_destroy(m);
    return st;
  }
  *out = m;
  return MFP_OK;
}

int mfp_matrix_from_triplets(mfp_matrix *out, int dtype, uint32_t rows, uint32_t cols,
                             uint64_t nnz, const uint32_t *tr_rows, const uint32_t *tr_cols,
                             const double *tr_vals, int inputs_on_device, mfp_stream sv)
{
  MFP_ARG(out, "null out");
  MFP_ARG(dtype == MFP_F32 || dtype == MFP_F64, "dtype must be f32 or f64");
  MFP_ARG(nnz < (1ull << 32), "nnz must fit in 32 bits");
  MFP_CHECK(ensure_init());
  hipStream_t s = as_stream(sv);
  const uint32_t *dr = tr_rows, *dc = tr_cols;
  const double *dv = tr_vals;
  void *tmp[3] = {nullptr, nullptr, nullptr};
  auto release = [&]() {
    for (void *p : tmp) dfree(p);
  };
  if (!inputs_on_device && nnz) {
    int st = dmalloc(&tmp[0], nnz * 4);
    if (st == MFP_OK) st = dmalloc(&tmp[1], nnz * 4);
    if (st == MFP_OK) st = dmalloc(&tmp[2], nnz * 8);
    if (st == MFP_OK) st = mfp_memcpy_h2d(tmp[0], tr_rows, nnz * 4);
    if (st == MFP_OK) st = mfp_memcpy_h2d(tmp[1], tr_cols, nnz * 4);
    if (st == MFP_OK) st = mfp_memcpy_h2d(tmp[2], tr_vals, nnz * 8);
    if (st != MFP_OK) {
      release();
      return st;
    }
    dr = (const uint32_t *)tmp[0];
    dc = (const uint32_t *)tmp[1];
    dv = (const double *)tmp[2];
  }
  auto *m = new mfp_matrix_s;
  m->dtype = dtype;
  m->rows = rows;
  m->cols = cols;
  m->nnz = nnz;
  int st = dmalloc((void **)&m->row_ptr, ((uint64_t)rows + 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->columns, nnz * 4);
  if (st == MFP_OK) st = dmalloc(&m->values, nnz * dtype_size(dtype));
  if (st == MFP_OK)
    st = csr_from_triplets_device(dtype, rows, cols, nnz, dr, dc, dv, m->row_ptr, m->columns,
                                  m->values, s);
  if (st == MFP_OK && hipStreamSynchronize(s) != hipSuccess) st = MFP_ERR_HIP;
  release();
  if (st != MFP_OK) {
    mfp_matrix_destroy(m);
    return st;
  }
  *out = m;
  return MFP_OK;
}

int mfp_matrix_view_get(mfp_matrix m, mfp_matrix_view *out)
{
  MFP_ARG(m && out, "null pointer");
  out->dtype = m->dtype;
  out->rows = m->rows;
  out->cols = m->cols;
  out->nnz = m->nnz;
  out->row_ptr = m->row_ptr;
  out->columns = m->columns;
  out->values = m->values;
  return MFP_OK;
}

int mfp_matrix_download(mfp_matrix m, uint32_t *row_ptr, uint32_t *columns, void *values)
{
  MFP_ARG(m, "null matrix");
  if (row_ptr) MFP_CHECK(mfp_memcpy_d2h(row_ptr, m->row_ptr, ((uint64_t)m->rows + 1) * 4));
  if (columns) MFP_CHECK(mfp_memcpy_d2h(columns, m->columns, m->nnz * 4));
  if (values) MFP_CHECK(mfp_memcpy_d2h(values, m->values, m->nnz * dtype_size(m->dtype)));
  return MFP_OK;
}

int mfp_matrix_convert(mfp_matrix src, int dtype, mfp_matrix *out)
{
  MFP_ARG(src && out, "null pointer");
  MFP_ARG(dtype_size(dtype) != 0, "dtype");
  MFP_ARG(src->dtype == MFP_F32 || src->dtype == MFP_F64, "source values must be f32 or f64");
  MFP_CHECK(ensure_init());
  auto *m = new mfp_matrix_s;
  m->dtype = dtype;
  m->rows = src->rows;
  m->cols = src->cols;
  m->nnz = src->nnz;
  int st = dmalloc((void **)&m->row_ptr, ((uint64_t)m->rows + 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->columns, m->nnz * 4);
  if (st == MFP_OK) st = dmalloc(&m->values, m->nnz * dtype_size(dtype));
  if (st == MFP_OK) st = mfp_memcpy_d2d(m->row_ptr, src->row_ptr, ((uint64_t)m->rows + 1) * 4, nullptr);
  if (st == MFP_OK) st = mfp_memcpy_d2d(m->columns, src->columns, m->nnz * 4, nullptr);
  if (st == MFP_OK)
    st = is_lp(dtype) ? to_lp_values(src->values, src->dtype, m->values, dtype, m->nnz, nullptr)
                      : convert_values(src->dtype, src->values, dtype, m->values, m->nnz, nullptr);
  if (st == MFP_OK && hipDeviceSynchronize() != hipSuccess) st = MFP_ERR_HIP;
  if (st != MFP_OK) {
    mfp_matrix_destroy(m);
    return st;
  }
  *out = m;
  return MFP_OK;
}

int mfp_matrix_destroy(mfp_matrix m)
{
  if (!m) return MFP_OK;
  dfree(m->row_ptr);
  dfree(m->columns);
  dfree(m->values);
  delete m;
  return MFP_OK;
}

int mfp_matrix_row_stats(mfp_matrix m, mfp_row_stats *out)
{
  MFP_ARG(m && out, "null pointer");
  MFP_CHECK(ensure_init());
  return matrix_stats(m, out, nullptr);
}

int mfp_tuning_default(mfp_matrix m, int format, mfp_tuning *t)
{
  MFP_ARG(m && t, "null pointer");
  MFP_ARG(format >= 0 && format < MFP_FORMAT_COUNT, "format");
  MFP_CHECK(ensure_init());
  std::memset(t, 0, sizeof(*t));
  const uint32_t avg = m->rows ? (uint32_t)((m->nnz + m->rows - 1) / m->rows) : 1;
  t->lanes = std::min<uint32_t>(64, std::max<uint32_t>(2, next_pow2_u32(std::max<uint32_t>(1, avg))));
  t->nnz_per_wg = default_nnz_per_wg(m->dtype);
  t->chunk = DEFAULT_CHUNK;
  t->long_row = DEFAULT_LONG_ROW;
  t->ell_width = 0;
  t->hyb_percent = 0.2; // main.cpp:240
  t->hyb_width = 0;
  scoo_mi355x_geometry(m->dtype, m->rows, &t->scoo_slices, &t->scoo_slice_size, &t->scoo_lanes);
  t->scoo_threads = 512;
  t->budget_bytes = 0;
  if (is_lp(m->dtype) && format == MFP_SCOO) {
    // 16-bit values have only the column-sorted SCOO
    t->flags = MFP_TUNE_SCOO_COLSORT;
    t->scoo_slice_size = 8192;
  }
  return MFP_OK;
}

// FP32 / FP64 ELL with MFP_TUNE_ELL_PACKED (spmv.hip k_ell_pk): 64-row slices,
// each with a base column (its rows' smallest first column) that every column
// of the slice lies within 65,536 of; w rounded up to words of 4 slots.
static int plan_create_ell_packed(mfp_plan p, mfp_matrix_s *m, uint32_t w, const mfp_tuning &t, hipStream_t s)
{
  const uint32_t stride = ell_lp_stride(m->rows);
  const uint32_t w4 = (w + 3) / 4;
  const uint64_t slots = (uint64_t)stride * 4u * w4;
  const uint64_t vs = dtype_size(m->dtype);
  const uint64_t need = slots * (vs + 2) + (uint64_t)stride / 64 * 4;
  if (need > budget_of(t)) {
    set_error("packed ELL needs " + std::to_string(need) + " bytes, above the budget");
    return MFP_ERR_BUDGET;
  }
  uint32_t span = 0;
  MFP_CHECK(dmalloc((void **)&p->pk_base, std::max<uint64_t>((uint64_t)stride / 64, 1) * 4));
  if (m->rows) MFP_CHECK(block_base_build(m->rows, m->row_ptr, m->columns, p->pk_base, &span, s));
  else MFP_CHECK(fill_u32(p->pk_base, 0, 1, s));
  if (span >= 65536u) {
    set_error("packed ELL: a 64-row slice spans " + std::to_string(span) + " columns, more than 16-bit offsets reach");
    return MFP_ERR_ARG;
  }
  MFP_CHECK(dmalloc(&p->pk_val, std::max<uint64_t>(slots, 1) * vs));
  MFP_CHECK(dmalloc((void **)&p->pk_off, std::max<uint64_t>(slots, 1) * 2));
  MFP_CHECK(ell_pk_build(m->dtype, m->rows, stride, m->row_ptr, m->columns, m->values, w4, p->pk_base, p->pk_val,
                         p->pk_off, s));
  MFP_HIP(hipStreamSynchronize(s));
  p->ell_packed = true;
  p->ell_w = w;
  p->pk_w4 = w4;
  p->device_bytes = need;
  return MFP_OK;
}

// Column-sorted SCOO: slice_size rows per slice (a power of two; a slice
// accumulates in LDS, 8 bytes per row for either value type).  Rows longer than max(256, mean entries per slice /
// 128) get their own slices: in an ordinary slice such a row would put many
// lanes of every wave on one LDS address.  Work items take about
// nnz / (2 CUs) entries each (at least 2048), or nnz / (slices * pieces)
// when scoo_pieces is set.  MFP_TUNE_SCOO_WHOLE_SLICES: no long-row slices
// and one item per slice.
static int plan_create_scoo_cs(mfp_plan p, mfp_matrix_s *m, const mfp_tuning &t, hipStream_t s)
{
  const uint32_t R = t.scoo_slice_size;
  if (R == 0 || (R & (R - 1)) || R > 65536 || (uint64_t)R * sizeof(double) > 163840) {
    set_error("column-sorted SCOO slice_size must be a power of two whose accumulator fits in 160 KB of LDS");
    return MFP_ERR_ARG;
  }
  const uint64_t S = m->rows ? ((uint64_t)m->rows + R - 1) / R : 0u;
  const uint64_t per_slice = S ? m->nnz / S : 0;
  const bool whole = (t.flags & MFP_TUNE_SCOO_WHOLE_SLICES) != 0;
  const uint32_t T = whole ? 0xFFFFFFFFu : (uint32_t)std::max<uint64_t>(256, per_slice / 128);
  uint64_t item;
  if (whole)
    item = ~0ull; // one item per band of a slice
  else if (t.scoo_pieces)
    item = std::max<uint64_t>(1, per_slice / t.scoo_pieces);
  else
    item = std::max<uint64_t>(2048, m->nnz / (2u * (uint64_t)std::max(1, props().cu_count)));
  p->colsort = true;
  p->scoo_threads = t.scoo_threads ? t.scoo_threads : 512;
  MFP_CHECK(scoo_cs_build_device(m->dtype, m->rows, m->cols, m->nnz, m->row_ptr, m->columns, m->values, R, T, item,
                                 &p->cs, s, !(t.flags & MFP_TUNE_STATIC_PARTIALS)));
  p->ss = R;
  p->slices = p->cs.S;
  p->L = p->cs.NI;
  p->device_bytes = p->cs.device_bytes;
  return MFP_OK;
}

int mfp_plan_create(mfp_plan *out, mfp_matrix m, int format, const mfp_tuning *tin, mfp_stream sv)
{
  MFP_ARG(out && m, "null pointer");
  MFP_ARG(format >= 0 && format < MFP_FORMAT_COUNT, "format");
  MFP_CHECK(ensure_init());
  hipStream_t s = as_stream(sv);
  mfp_tuning t;
  if (tin)
    t = *tin;
  else
    MFP_CHECK(mfp_tuning_default(m, format, &t));
  auto *p = new mfp_plan_s;
  p->format = format;
  p->dtype = m->dtype;
  p->m = m;
  p->t = t;
  p->static_partials = (t.flags & MFP_TUNE_STATIC_PARTIALS) != 0;
  const size_t vs = dtype_size(m->dtype);
  int st = MFP_OK;
  auto fail = [&](int code) {
    plan_free_arrays(p);
    delete p;
    return code;
  };
  if (is_lp(m->dtype)) {
    if (t.flags & MFP_TUNE_ELL_PACKED) { // the 16-bit-value ELL has 16-bit offsets of its own
      set_error("MFP_TUNE_ELL_PACKED is for ELL plans over f32 or f64 values");
      return fail(MFP_ERR_ARG);
    }
    st = plan_create_lp(p, m, t, s);
    if (st != MFP_OK) return fail(st);
    *out = p;
    return MFP_OK;
  }
  switch (format) {
  case MFP_CSR_SCALAR: break;
  case MFP_CSR_VECTOR:
    if (t.lanes < 2 || t.lanes > 64 || (t.lanes & (t.lanes - 1))) {
      set_error("lanes must be a power of two in [2, 64]");
      return fail(MFP_ERR_ARG);
    }
    break;
  case MFP_CSR_ADAPTIVE:
    st = build_adaptive_layout(m->dtype, m->rows, m->row_ptr, t.nnz_per_wg, t.long_row, t.chunk,
                               &p->ad, s, !(t.flags & MFP_TUNE_STATIC_PARTIALS));
    if (st != MFP_OK) return fail(st);
    st = stream_meta_device(m->rows, m->row_ptr, &p->stream_meta, &p->nstream, s);
    if (st != MFP_OK) return fail(st);
    p->device_bytes = ((uint64_t)p->ad.nblocks + 1) * 4 + ((uint64_t)p->nstream + 1) * 8 +
                      (uint64_t)p->ad.nchunks * (16 + (p->ad.partials ? vs : 0));
    break;
  case MFP_ELL:
  case MFP_HYB:
  case MFP_HYB_ATOMIC: {
    mfp_row_stats rs;
    st = matrix_stats(m, &rs, s);
    if (st != MFP_OK) return fail(st);
    uint32_t w;
    if (format == MFP_ELL) {
      w = t.ell_width ? t.ell_width : rs.max_elements_in_rows;
      if (t.ell_width && t.ell_width < rs.max_elements_in_rows) {
        set_error("ELL width below the longest row would drop entries; use the hybrid format");
        return fail(MFP_ERR_ARG);
      }
    } else if (t.hyb_width) {
      w = t.hyb_width;
    } else {
      // hybrid_matrix_class::allocate (common/matrix_converter.cpp:296-302)
      const double wd = (double)rs.avg_elements_in_rows +
                        (double)(rs.max_elements_in_rows - rs.avg_elements_in_rows) * t.hyb_percent;
      w = (uint32_t)wd;
    }
    if (t.flags & MFP_TUNE_ELL_PACKED) {
      if (format != MFP_ELL || (m->dtype != MFP_F32 && m->dtype != MFP_F64)) {
        set_error("MFP_TUNE_ELL_PACKED is for ELL plans over f32 or f64 values");
        return fail(MFP_ERR_ARG);
      }
      st = plan_create_ell_packed(p, m, w, t, s);
      if (st != MFP_OK) return fail(st);
      break;
    }
    const uint64_t slots = (uint64_t)m->rows * w;
    const uint64_t need = slots * (vs + 4);
    if (need > budget_of(t)) {
      set_error("ELL part needs " + std::to_string(need) + " bytes, above the budget");
      return fail(MFP_ERR_BUDGET);
    }
    p->ell_w = w;
    st = dmalloc(&p->ell_val, slots * vs);
    if (st == MFP_OK) st = dmalloc((void **)&p->ell_col, slots * 4);
    if (st == MFP_OK)
      st = ell_build_device(m->dtype, m->rows, m->row_ptr, m->columns, m->values, w, p->ell_val,
                            p->ell_col, s);
    if (st != MFP_OK) return fail(st);
    p->device_bytes = need;
    if (format != MFP_ELL) {
      st = coo_tail_build_device(m->dtype, m->rows, m->nnz, m->row_ptr, m->columns, m->values, w,
                                 &p->coo_val, &p->coo_col, &p->coo_row, &p->ncoo, s);
      if (st != MFP_OK) return fail(st);
      p->device_bytes += p->ncoo * (vs + 8);
    }
    break;
  }
  case MFP_SCOO:
    if (t.flags & MFP_TUNE_SCOO_COLSORT) {
      st = plan_create_scoo_cs(p, m, t, s);
      if (st != MFP_OK) return fail(st);
      break;
    }
    [[fallthrough]];
  case MFP_COO: {
    st = coo_tail_build_device(m->dtype, m->rows, m->nnz, m->row_ptr, m->columns, m->values, 0,
                               &p->coo_val, &p->coo_col, &p->coo_row, &p->ncoo, s);
    if (st != MFP_OK) return fail(st);
    p->device_bytes = p->ncoo * (vs + 8);
    if (format == MFP_SCOO) {
      p->ss = t.scoo_slice_size;
      p->L = t.scoo_lanes;
      p->slices = t.scoo_slices;
      p->scoo_threads = t.scoo_threads ? t.scoo_threads : 512;
      if (p->ss == 0 || p->L == 0 || (p->L & (p->L - 1)) ||
          (m->rows && (uint64_t)p->slices * p->ss < m->rows)) {
        set_error("SCOO geometry must cover every row with a power-of-two lane size");
        return fail(MFP_ERR_ARG);
      }
      st = dmalloc((void **)&p->sp, ((uint64_t)p->slices + 1) * 4);
      if (st == MFP_OK) st = scoo_slices_device(m->rows, m->nnz, m->row_ptr, p->slices, p->ss, p->sp, s);
      if (st != MFP_OK) return fail(st);
      p->device_bytes += ((uint64_t)p->slices + 1) * 4;
    }
    break;
  }
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("plan construction failed");
    return fail(MFP_ERR_HIP);
  }
  *out = p;
  return MFP_OK;
}

// Partial results a plan's SpMV writes and reads back: column-sorted SCOO
// slots, CSR-Adaptive long-row chunk sums, chunk sums of the binned 16-bit
// CSR.  Bytes for one SpMV (0 when the plan has none).
static uint64_t plan_partial_bytes(const mfp_plan_s *p)
{
  if (p->colsort) return (uint64_t)p->cs.slots * p->cs.R * (p->dtype == MFP_F64 ? 8 : 4);
  if (p->format == MFP_CSR_ADAPTIVE && is_lp(p->dtype)) return (uint64_t)p->lpb.nchunks * sizeof(float2);
  if (p->format == MFP_CSR_ADAPTIVE) return (uint64_t)p->ad.nchunks * dtype_size(p->dtype);
  return 0;
}

// y = A x; partials, when not null, replaces the plan's buffer of partial
// results (plan_partial_bytes of it)
static int plan_spmv(mfp_plan p, const void *x, void *y, hipStream_t s, void *partials)
{
  const mfp_matrix m = p->m;
  const int dt = p->dtype;
  if (is_lp(dt)) {
    if (p->format == MFP_CSR_VECTOR)
      return launch_csr_vector_lp(dt, p->t.lanes, m->rows, m->row_ptr, m->columns, m->values, x, y, s);
    if (p->format == MFP_CSR_ADAPTIVE) {
      lp_binned_layout l = p->lpb;
      if (partials) l.partial = static_cast<float2 *>(partials);
      return launch_lp_binned(dt, l, m->row_ptr, m->columns, m->values, x, y, s);
    }
    if (p->format == MFP_SCOO) {
      scoo_cs_layout l = p->cs;
      if (partials) l.partial = partials;
      return launch_scoo_cs(dt, m->rows, l, p->scoo_threads, x, y, s);
    }
    return launch_ell_lp(dt, p->off16, p->lp_block, m->rows, p->ell_w, p->lp_cols, p->lp_base, p->ell_val, x, y, s);
  }
  switch (p->format) {
  case MFP_CSR_SCALAR:
    return launch_csr_scalar(dt, m->rows, m->row_ptr, m->columns, m->values, x, y, s);
  case MFP_CSR_VECTOR:
    return launch_csr_vector(dt, p->t.lanes, m->rows, m->row_ptr, m->columns, m->values, x, y, s);
  case MFP_CSR_ADAPTIVE: {
    adaptive_args g;
    g.rows = m->rows;
    g.rp = m->row_ptr;
    g.ci = m->columns;
    g.blocks = p->ad.blocks;
    g.nblocks = p->ad.nblocks;
    g.nnz_per_wg = p->t.nnz_per_wg;
    g.long_row = p->t.long_row;
    g.chunks = p->ad.chunks;
    g.nchunks = p->ad.nchunks;
    g.chunk = p->t.chunk;
    g.partials = partials ? partials : p->ad.partials;
    return launch_csr_adaptive(dt, g, m->values, x, y, s);
  }
  case MFP_ELL:
    if (p->ell_packed)
      return launch_ell_pk(dt, m->rows, p->ell_w, p->pk_w4, p->pk_base, p->pk_val, p->pk_off, x, y, s);
    return launch_ell(dt, m->rows, p->ell_w, p->ell_col, p->ell_val, x, y, s);
  case MFP_COO:
    MFP_CHECK(fill_bytes_zero(y, (size_t)m->rows * dtype_size(dt), s));
    mark_kernel_end(); // the zero fill of y is its own interval under mfp_plan_kernel_times
    return launch_coo_add(dt, p->ncoo, p->coo_row, p->coo_col, p->coo_val, x, y, s);
  case MFP_SCOO:
    if (p->colsort) {
      scoo_cs_layout l = p->cs;
      if (partials) l.partial = partials;
      return launch_scoo_cs(dt, m->rows, l, p->scoo_threads, x, y, s);
    }
    return launch_scoo(dt, m->rows, p->ss, p->L, p->scoo_threads, p->coo_row, p->coo_col,
                       p->coo_val, p->sp, x, y, s);
  case MFP_HYB:
    MFP_CHECK(launch_ell(dt, m->rows, p->ell_w, p->ell_col, p->ell_val, x, y, s));
    return launch_coo_add(dt, p->ncoo, p->coo_row, p->coo_col, p->coo_val, x, y, s);
  case MFP_HYB_ATOMIC:
    MFP_CHECK(fill_bytes_zero(y, (size_t)m->rows * dtype_size(dt), s));
    mark_kernel_end();
    return launch_hyb_atomic(dt, m->rows, p->ell_w, p->ell_col, p->ell_val, p->ncoo, p->coo_row,
                             p->coo_col, p->coo_val, x, y, s);
  }
  set_error("unknown format");
  return MFP_ERR_ARG;
}

int mfp_plan_spmv(mfp_plan p, const void *x, void *y, mfp_stream sv)
{
  MFP_ARG(p && x && y, "null pointer");
  if (p->static_partials && plan_partial_bytes(p)) {
    set_error("plan created with MFP_TUNE_STATIC_PARTIALS has no buffer of partial results: "
              "run it with mfp_plan_spmv_noalloc");
    return MFP_ERR_STATE;
  }
  return plan_spmv(p, x, y, as_stream(sv), nullptr);
}

int mfp_plan_spmv_noalloc(mfp_plan p, const void *x, void *y, mfp_stream sv)
{
  MFP_ARG(p && x && y, "null pointer");
  MFP_CHECK(ensure_init());
  int dev = -1;
  MFP_HIP(hipGetDevice(&dev));
  MFP_ARG(dev >= 0 && dev < 64, "device index beyond the static scratch table");
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    ptr = g_scratch_ptr[dev];
  }
  if (!ptr) {
    MFP_CHECK(static_scratch_setup());
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    ptr = g_scratch_ptr[dev];
  }
  const uint64_t need = plan_partial_bytes(p);
  if (need > MFP_NOALLOC_SCRATCH_BYTES) {
    set_error("mfp_plan_spmv_noalloc: the plan's partial results take " + std::to_string(need) +
              " bytes, more than the static scratch of " + std::to_string(MFP_NOALLOC_SCRATCH_BYTES) +
              " bytes; use mfp_plan_spmv");
    return MFP_ERR_BUDGET;
  }
  return plan_spmv(p, x, y, as_stream(sv), ptr);
}

int mfp_plan_kernel_times(mfp_plan p, const void *x, void *y, mfp_stream sv, float *ms, int max_kernels,
                          int *n_kernels)
{
  MFP_ARG(p && x && y && ms && n_kernels, "null pointer");
  MFP_ARG(max_kernels > 0, "max_kernels must be positive");
  MFP_ARG(!g_kernel_marks, "mfp_plan_kernel_times does not nest");
  kernel_marks km;
  km.s = as_stream(sv);
  int made = 0, st = MFP_OK;
  for (; made <= kernel_marks::kMax; made++)
    if (hipEventCreate(&km.ev[made]) != hipSuccess ||
        (made < kernel_marks::kMax && hipEventCreate(&km.start[made]) != hipSuccess)) {
      set_error("hipEventCreate failed");
      st = MFP_ERR_HIP;
      break;
    }
  if (st == MFP_OK && hipEventRecord(km.ev[0], km.s) != hipSuccess) {
    set_error("hipEventRecord failed");
    st = MFP_ERR_HIP;
  }
  if (st == MFP_OK) {
    g_kernel_marks = &km;
    st = mfp_plan_spmv(p, x, y, sv);
    g_kernel_marks = nullptr;
  }
  if (st == MFP_OK && km.n == 0) { // nothing launched (no rows): one empty interval
    if (hipEventRecord(km.ev[1], km.s) != hipSuccess) st = MFP_ERR_HIP;
    km.n = 1;
  }
  if (st == MFP_OK && hipEventSynchronize(km.ev[km.n]) != hipSuccess) {
    set_error("hipEventSynchronize failed");
    st = MFP_ERR_HIP;
  }
  if (st == MFP_OK) {
    const int n = std::min(km.n, max_kernels);
    for (int i = 0; i < n && st == MFP_OK; i++)
      if (hipEventElapsedTime(&ms[i], km.exact[i] ? km.start[i] : km.ev[i], km.ev[i + 1]) != hipSuccess) {
        set_error("hipEventElapsedTime failed");
        st = MFP_ERR_HIP;
      }
    *n_kernels = km.n;
  }
  for (int i = 0; i <= kernel_marks::kMax; i++) {
    if (km.ev[i]) (void)hipEventDestroy(km.ev[i]);
    if (i < kernel_marks::kMax && km.start[i]) (void)hipEventDestroy(km.start[i]);
  }
  return st;
}

int mfp_plan_get_info(mfp_plan p, mfp_plan_info *o)
{
  MFP_ARG(p && o, "null pointer");
  std::memset(o, 0, sizeof(*o));
  o->format = p->format;
  o->dtype = p->dtype;
  o->rows = p->m->rows;
  o->cols = p->m->cols;
  o->nnz = p->m->nnz;
  o->device_bytes = p->device_bytes;
  o->ell_width = p->ell_w;
  o->coo_elements = p->ncoo;
  o->stored_elements = (uint64_t)p->m->rows * p->ell_w + p->ncoo;
  o->blocks = p->ad.nblocks;
  o->chunks = p->ad.nchunks;
  o->slices_count = p->slices;
  o->slice_size = p->ss;
  o->lane_size = p->L;
  o->lanes = p->t.lanes;
  o->nnz_per_wg = p->t.nnz_per_wg;
  o->column_bytes = p->off16 || p->ell_packed ? 2u : 4u;
  o->scoo_colsort = p->colsort ? 1u : 0u;
  o->scoo_packed = p->colsort ? 1u : 0u;
  o->scoo_long_rows = p->colsort ? p->cs.NL : 0u;
  if (p->colsort) o->stored_elements = p->m->nnz;
  return MFP_OK;
}

int mfp_plan_array(mfp_plan p, int which, const void **ptr, uint64_t *count)
{
  MFP_ARG(p && ptr && count, "null pointer");
  const uint64_t slots = (uint64_t)p->m->rows * p->ell_w;
  MFP_ARG(!((is_lp(p->dtype) || p->ell_packed) && (which == MFP_ARR_ELL_VALUES || which == MFP_ARR_ELL_COLS)),
          "packed ELL plans (16-bit values or MFP_TUNE_ELL_PACKED) keep their own layout");
  switch (which) {
  case MFP_ARR_ELL_VALUES: *ptr = p->ell_val; *count = slots; break;
  case MFP_ARR_ELL_COLS: *ptr = p->ell_col; *count = slots; break;
  case MFP_ARR_COO_VALUES: *ptr = p->coo_val; *count = p->ncoo; break;
  case MFP_ARR_COO_COLS: *ptr = p->coo_col; *count = p->ncoo; break;
  case MFP_ARR_COO_ROWS: *ptr = p->coo_row; *count = p->ncoo; break;
  case MFP_ARR_SCOO_SLICES_PTR: *ptr = p->sp; *count = p->sp ? (uint64_t)p->slices + 1 : 0; break;
  case MFP_ARR_ROW_BLOCKS: *ptr = p->ad.blocks; *count = p->ad.blocks ? (uint64_t)p->ad.nblocks + 1 : 0; break;
  default: MFP_ARG(false, "array id");
  }
  return MFP_OK;
}

int mfp_plan_destroy(mfp_plan p)
{
  if (!p) return MFP_OK;
  plan_free_arrays(p);
  delete p;
  return MFP_OK;
}

int mfp_fill_row_blocks(uint32_t rows, const uint32_t *row_ptr, uint32_t W, uint32_t *out,
                        uint32_t *count)
{
  MFP_ARG(row_ptr && count, "null pointer");
  MFP_CHECK(ensure_init());
  uint32_t *drp = nullptr, *blocks = nullptr;
  MFP_CHECK(dmalloc((void **)&drp, ((uint64_t)rows + 1) * 4));
  int st = mfp_memcpy_h2d(drp, row_ptr, ((uint64_t)rows + 1) * 4);
  uint32_t n = 0;
  if (st == MFP_OK) st = row_blocks_device(rows, drp, W, &blocks, &n, nullptr);
  if (st == MFP_OK && out) st = mfp_memcpy_d2h(out, blocks, ((uint64_t)n + 1) * 4);
  dfree(drp);
  dfree(blocks);
  if (st == MFP_OK) *count = n;
  return st;
}

// scoo_matrix_class sizing (common/scoo_matrix_class.cpp:13-62), restated.
int mfp_scoo_geometry(int dtype, uint32_t rows, uint32_t sm_count, uint64_t shm,
                      uint32_t *slices_count, uint32_t *slice_size, uint32_t *lane_size)
{
  MFP_ARG(slices_count && slice_size && lane_size, "null pointer");
  MFP_ARG(sm_count > 0, "sm_count must be positive");
  MFP_ARG(dtype == MFP_F32 || dtype == MFP_F64, "dtype");
  auto next_pow = [](uint32_t x) {
    --x;
    x |= x >> 1u;
    x |= x >> 2u;
    x |= x >> 4u;
    x |= x >> 8u;
    x |= x >> 16u;
    return ++x;
  };
  auto prev_pow = [&](uint32_t x) {
    const uint32_t n = next_pow(x);
    return n > x ? n / 2 : n;
  };
  const uint64_t vs = dtype_size(dtype);
  const uint32_t max_rows = (uint32_t)std::min<uint64_t>(512, shm / vs);
  uint32_t mult = 1;
  while (max_rows < rows / (sm_count * mult)) mult++;
  const uint32_t slices = sm_count * mult;
  const uint32_t ss = (rows + slices - 1) / slices;
  uint32_t lane = 0;
  if (ss) {
    const uint32_t inner = prev_pow((uint32_t)(shm / ((uint64_t)ss * vs)));
    lane = prev_pow(std::min(slices, inner));
    while (ss * lane > 1024) lane /= 2;
  }
  *slices_count = slices;
  *slice_size = ss;
  *lane_size = lane ? lane : 1;
  return MFP_OK;
}

} // extern "C"
