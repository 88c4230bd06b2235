/*
 * CPU oracle for the SpMV hot path.  TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker / CPU baseline, never by the product library.
 *
 * Plain C restatement of the reference's host algorithms; see
 * mfp_oracle.h for the file:line each function follows.
 */
#define _GNU_SOURCE
#include "mfp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---------------------------------------------------------------------- */
/* Triplets -> CSR (common/matrix_converter.cpp:15-87)                      */
/* ---------------------------------------------------------------------- */

/* Stable merge sort of idx[0..n) by key[idx[.]] (ties keep order). */
static void stable_sort_by_key(uint32_t *idx, uint32_t *tmp, uint32_t n,
                               const uint32_t *key)
{
  if (n < 2)
    return;
  if (n <= 16) {
    for (uint32_t i = 1; i < n; i++) {
      uint32_t v = idx[i];
      uint32_t j = i;
      while (j > 0 && key[idx[j - 1]] > key[v]) {
        idx[j] = idx[j - 1];
        j--;
      }
      idx[j] = v;
    }
    return;
  }
  uint32_t h = n / 2;
  stable_sort_by_key(idx, tmp, h, key);
  stable_sort_by_key(idx + h, tmp, n - h, key);
  uint32_t i = 0, j = h, o = 0;
  while (i < h && j < n)
    tmp[o++] = (key[idx[j]] < key[idx[i]]) ? idx[j++] : idx[i++];
  while (i < h)
    tmp[o++] = idx[i++];
  while (j < n)
    tmp[o++] = idx[j++];
  memcpy(idx, tmp, sizeof(uint32_t) * n);
}

int orc_csr_from_triplets(uint32_t rows, uint32_t cols, uint64_t nnz,
                          const uint32_t *tr_rows, const uint32_t *tr_cols,
                          const double *tr_vals, int dtype,
                          uint32_t *row_ptr, uint32_t *columns, void *data)
{
  for (uint64_t i = 0; i < nnz; i++)
    if (tr_rows[i] >= rows || tr_cols[i] >= cols)
      return -1;

  /* count per row (:35-36) */
  memset(row_ptr, 0, sizeof(uint32_t) * ((size_t)rows + 1));
  for (uint64_t i = 0; i < nnz; i++)
    row_ptr[tr_rows[i]]++;

  /* exclusive scan (:38-44) */
  uint32_t ptr = 0;
  for (uint64_t r = 0; r < (uint64_t)rows + 1; r++) {
    uint32_t t = row_ptr[r];
    row_ptr[r] = ptr;
    ptr += t;
  }

  /* scatter in input order (:46-57) */
  uint32_t *fill = calloc(rows ? rows : 1, sizeof(uint32_t));
  uint32_t *src = malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  for (uint64_t i = 0; i < nnz; i++) {
    uint32_t r = tr_rows[i];
    src[row_ptr[r] + fill[r]++] = (uint32_t)i;
  }
  free(fill);

  /* per-row sort by column (:59-85), stable for ties */
  uint32_t *tmp = malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  for (uint32_t r = 0; r < rows; r++)
    stable_sort_by_key(src + row_ptr[r], tmp, row_ptr[r + 1] - row_ptr[r],
                       tr_cols);
  free(tmp);

  for (uint64_t e = 0; e < nnz; e++) {
    uint32_t i = src[e];
    columns[e] = tr_cols[i];
    if (dtype == ORC_F32)
      ((float *)data)[e] = (float)tr_vals[i];
    else
      ((double *)data)[e] = tr_vals[i];
  }
  free(src);
  return 0;
}

void orc_rows_statistics(uint32_t rows, const uint32_t *row_ptr,
                         uint32_t out3[3], double *out_std)
{
  uint32_t mn = 0xFFFFFFFFu - 1u, mx = 0, sum = 0; /* :98, :100 (32-bit) */
  for (uint32_t r = 0; r < rows; r++) {
    uint32_t len = row_ptr[r + 1] - row_ptr[r];
    if (len > mx)
      mx = len;
    if (len < mn)
      mn = len;
    sum += len;
  }
  uint32_t avg = rows ? sum / rows : 0; /* :114 integer division */
  double acc = 0.0;
  for (uint32_t r = 0; r < rows; r++) {
    double d = (double)(row_ptr[r + 1] - row_ptr[r]) - (double)avg;
    acc += d * d; /* std::pow(d, 2) is exact for these integers */
  }
  out3[0] = mn;
  out3[1] = mx;
  out3[2] = avg;
  *out_std = rows ? sqrt(acc / (double)rows) : 0.0;
}

/* ---------------------------------------------------------------------- */
/* ELL / COO / hybrid                                                     */
/* ---------------------------------------------------------------------- */

void orc_ell_from_csr(uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data, int dtype,
                      uint32_t width, void *ell_data, uint32_t *ell_cols)
{
  size_t slots = (size_t)rows * width;
  size_t vs = dtype == ORC_F32 ? 4 : 8;
  memset(ell_data, 0, slots * vs);
  memset(ell_cols, 0, slots * sizeof(uint32_t));
  for (uint32_t r = 0; r < rows; r++) {
    uint32_t s = row_ptr[r], e = row_ptr[r + 1];
    uint32_t lim = e - s < width ? e - s : width;
    for (uint32_t k = 0; k < lim; k++) {
      size_t o = (size_t)r + (size_t)k * rows;
      ell_cols[o] = columns[s + k];
      if (dtype == ORC_F32)
        ((float *)ell_data)[o] = ((const float *)data)[s + k];
      else
        ((double *)ell_data)[o] = ((const double *)data)[s + k];
    }
  }
}

uint64_t orc_coo_count(uint32_t rows, const uint32_t *row_ptr,
                       uint32_t element_start)
{
  uint64_t n = 0;
  for (uint32_t r = 0; r < rows; r++) {
    uint32_t len = row_ptr[r + 1] - row_ptr[r];
    if (len > element_start)
      n += len - element_start;
  }
  return n;
}

void orc_coo_from_csr(uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data, int dtype,
                      uint32_t element_start, void *coo_data,
                      uint32_t *coo_cols, uint32_t *coo_rows)
{
  uint64_t id = 0;
  for (uint32_t r = 0; r < rows; r++) {
    for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; e++) {
      if (e - row_ptr[r] < element_start)
        continue;
      coo_cols[id] = columns[e];
      coo_rows[id] = r;
      if (dtype == ORC_F32)
        ((float *)coo_data)[id] = ((const float *)data)[e];
      else
        ((double *)coo_data)[id] = ((const double *)data)[e];
      id++;
    }
  }
}

uint32_t orc_hybrid_width(uint32_t rows, const uint32_t *row_ptr,
                          double percent)
{
  uint32_t s[3];
  double sd;
  orc_rows_statistics(rows, row_ptr, s, &sd);
  double w = (double)s[2] + (double)(s[1] - s[2]) * percent;
  return (uint32_t)w;
}

/* ---------------------------------------------------------------------- */
/* SCOO sizing and layout (common/scoo_matrix_class.cpp)                   */
/* ---------------------------------------------------------------------- */

uint32_t orc_next_pow_2(uint32_t x)
{
  --x;
  x |= x >> 1u;
  x |= x >> 2u;
  x |= x >> 4u;
  x |= x >> 8u;
  x |= x >> 16u;
  return ++x;
}

uint32_t orc_prev_pow_2(uint32_t x)
{
  uint32_t n = orc_next_pow_2(x);
  if (n > x)
    return n / 2;
  return n;
}

uint32_t orc_scoo_slices_count(uint32_t rows, uint32_t sm_count,
                               uint64_t shared_mem_size, int dtype)
{
  uint64_t vs = dtype == ORC_F32 ? 4 : 8;
  uint64_t q = shared_mem_size / vs;
  uint32_t max_rows_per_block = (uint32_t)(q < 512 ? q : 512);
  uint32_t m = 1;
  while (max_rows_per_block < rows / (sm_count * m))
    m++;
  return sm_count * m;
}

uint32_t orc_scoo_lane_size(uint32_t slices_count, uint32_t slice_size,
                            uint64_t shared_mem_size, int dtype)
{
  uint64_t vs = dtype == ORC_F32 ? 4 : 8;
  uint32_t inner = orc_prev_pow_2(
      (uint32_t)(shared_mem_size / ((uint64_t)slice_size * vs)));
  uint32_t lane = orc_prev_pow_2(slices_count < inner ? slices_count : inner);
  while (slice_size * lane > 1024)
    lane /= 2;
  return lane ? lane : 1;
}

void orc_scoo_build(uint32_t rows, const uint32_t *row_ptr,
                    const uint32_t *columns, const void *data, int dtype,
                    uint32_t slices_count, uint32_t slice_size,
                    void *values, uint32_t *c_index, uint32_t *r_index,
                    uint32_t *slices_ptr)
{
  memset(slices_ptr, 0, sizeof(uint32_t) * ((size_t)slices_count + 1));
  uint32_t offset = 0, slice_id = 0, last_slice_row = 0;
  for (uint32_t r = 0; r < rows; r++) {
    if (r - last_slice_row >= slice_size || slice_id == 0) {
      slices_ptr[slice_id++] = offset;
      last_slice_row = r;
    }
    for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; e++) {
      if (dtype == ORC_F32)
        ((float *)values)[offset] = ((const float *)data)[e];
      else
        ((double *)values)[offset] = ((const double *)data)[e];
      c_index[offset] = columns[e];
      r_index[offset] = r;
      offset++;
    }
  }
  slices_ptr[slice_id] = offset;
}

/* ---------------------------------------------------------------------- */
/* fill_row_blocks (gpu/csr_adaptive_spmv.cu:180-240)                       */
/* ---------------------------------------------------------------------- */

uint32_t orc_fill_row_blocks(int fill, uint32_t rows, const uint32_t *row_ptr,
                             uint32_t *row_blocks, uint32_t nnz_per_wg)
{
  if (fill)
    row_blocks[0] = 0;
  int64_t last_i = 0;
  uint32_t current_wg = 1;
  uint32_t nnz_sum = 0;
  for (int64_t i = 1; i <= (int64_t)rows; i++) {
    nnz_sum += row_ptr[i] - row_ptr[i - 1];
    if (nnz_sum == nnz_per_wg) {
      last_i = i;
      if (fill)
        row_blocks[current_wg] = (uint32_t)i;
      current_wg++;
      nnz_sum = 0;
    } else if (nnz_sum > nnz_per_wg) {
      if (i - last_i > 1) {
        if (fill)
          row_blocks[current_wg] = (uint32_t)(i - 1);
        current_wg++;
        i--;
      } else {
        if (fill)
          row_blocks[current_wg] = (uint32_t)i;
        current_wg++;
      }
      last_i = i;
      nnz_sum = 0;
    } else if (i - last_i > (int64_t)nnz_per_wg) {
      last_i = i;
      if (fill)
        row_blocks[current_wg] = (uint32_t)i;
      current_wg++;
      nnz_sum = 0;
    }
  }
  if (fill)
    row_blocks[current_wg] = rows;
  return current_wg;
}

/* ---------------------------------------------------------------------- */
/* SpMV                                                                   */
/* ---------------------------------------------------------------------- */

#define DEFINE_SEQ(T, FMA, SUF)                                               \
  static void spmv_seq_##SUF(uint32_t rows, const uint32_t *rp,             \
                             const uint32_t *ci, const T *a, const T *x,      \
                             T *y, uint32_t r0, uint32_t r1)                  \
  {                                                                           \
    (void)rows;                                                               \
    for (uint32_t r = r0; r < r1; r++) {                                      \
      T dot = 0;                                                              \
      for (uint32_t e = rp[r]; e < rp[r + 1]; e++)                            \
        dot = FMA(a[e], x[ci[e]], dot);                                       \
      y[r] = dot;                                                             \
    }                                                                         \
  }
DEFINE_SEQ(float, fmaf, f32)
DEFINE_SEQ(double, fma, f64)

void orc_spmv_csr_seq(int dtype, uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data,
                      const void *x, void *y)
{
  if (dtype == ORC_F32)
    spmv_seq_f32(rows, row_ptr, columns, data, x, y, 0, rows);
  else
    spmv_seq_f64(rows, row_ptr, columns, data, x, y, 0, rows);
}

/* Keep the multiply and the add as two rounded operations. */
static float mul_f32(float a, float b)
{
  volatile float p = a * b;
  return p;
}
static double mul_f64(double a, double b)
{
  volatile double p = a * b;
  return p;
}

void orc_spmv_csr_seq_nofma(int dtype, uint32_t rows, const uint32_t *row_ptr,
                            const uint32_t *columns, const void *data,
                            const void *x, void *y)
{
  for (uint32_t r = 0; r < rows; r++) {
    if (dtype == ORC_F32) {
      const float *a = data, *xx = x;
      float dot = 0;
      for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; e++)
        dot = dot + mul_f32(a[e], xx[columns[e]]);
      ((float *)y)[r] = dot;
    } else {
      const double *a = data, *xx = x;
      double dot = 0;
      for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; e++)
        dot = dot + mul_f64(a[e], xx[columns[e]]);
      ((double *)y)[r] = dot;
    }
  }
}

#define DEFINE_LANES(T, FMA, SUF)                                             \
  static void spmv_lanes_##SUF(uint32_t lanes, uint32_t rows,               \
                               const uint32_t *rp, const uint32_t *ci,        \
                               const T *a, const T *x, T *y)                  \
  {                                                                           \
    T dots[64];                                                               \
    for (uint32_t r = 0; r < rows; r++) {                                     \
      for (uint32_t l = 0; l < lanes; l++) {                                  \
        T d = 0;                                                              \
        for (uint32_t e = rp[r] + l; e < rp[r + 1]; e += lanes)               \
          d = FMA(a[e], x[ci[e]], d);                                         \
        dots[l] = d;                                                          \
      }                                                                       \
      for (uint32_t off = lanes / 2; off > 0; off /= 2)                       \
        for (uint32_t l = 0; l < off; l++)                                    \
          dots[l] = dots[l] + dots[l + off];                                  \
      y[r] = dots[0];                                                         \
    }                                                                         \
  }
DEFINE_LANES(float, fmaf, f32)
DEFINE_LANES(double, fma, f64)

void orc_spmv_csr_lanes(int dtype, uint32_t lanes, uint32_t rows,
                        const uint32_t *row_ptr, const uint32_t *columns,
                        const void *data, const void *x, void *y)
{
  if (dtype == ORC_F32)
    spmv_lanes_f32(lanes, rows, row_ptr, columns, data, x, y);
  else
    spmv_lanes_f64(lanes, rows, row_ptr, columns, data, x, y);
}

/* gpu/gpu_matrix_multiplier.cu:1163-1168: the host COO loop of the
 * host-COO hybrid, y[row[e]] += data[e] * x[col[e]] in stored order, the
 * product rounded before the add.  y is accumulated into, not cleared. */
void orc_spmv_coo_seq(int dtype, uint64_t n, const uint32_t *coo_rows,
                      const uint32_t *coo_cols, const void *coo_data,
                      const void *x, void *y)
{
  for (uint64_t e = 0; e < n; e++) {
    if (dtype == ORC_F32) {
      float *yy = y;
      yy[coo_rows[e]] = yy[coo_rows[e]] +
                        mul_f32(((const float *)coo_data)[e], ((const float *)x)[coo_cols[e]]);
    } else {
      double *yy = y;
      yy[coo_rows[e]] = yy[coo_rows[e]] +
                        mul_f64(((const double *)coo_data)[e], ((const double *)x)[coo_cols[e]]);
    }
  }
}

/* The summation order of this library's CSR-stream SpMM (one column of X;
 * every column is summed alike).  Row blocks from orc_fill_row_blocks(W).  A
 * block with more than W entries is one row, taken in chunks of W entries:
 * group g (of `groups`) adds entries c W + g (W / groups) + j, j < W / groups,
 * of chunk c = 0, 1, ... in turn with fma, then a halving tree
 * p[g] += p[g + h], h = groups/2 .. 1, gives p[0].  Otherwise entry e of the
 * block belongs to group e / (W / groups); a row's segment in each group is
 * an fma sum from 0 in entry order; a row inside one group is that sum, and a
 * row across groups ga..gb is seg[gb] + seg[gb-1] + ... + seg[ga], added in
 * that order.  An empty row is 0. */
#define STREAM_BODY(T, FMA)                                                        \
  do {                                                                             \
    const T *a = data, *xx = x;                                                    \
    T *yy = y;                                                                     \
    for (uint32_t b = 0; b < nb; b++) {                                            \
      const uint32_t r0 = blocks[b], r1 = blocks[b + 1];                           \
      const uint32_t e0 = row_ptr[r0], n = row_ptr[r1] - e0;                       \
      if (n > W) {                                                                 \
        for (uint32_t g = 0; g < groups; g++) {                                    \
          T p = 0;                                                                 \
          for (uint64_t cb = 0; cb < n; cb += W)                                   \
            for (uint32_t j = 0; j < epg; j++) {                                   \
              const uint64_t e = cb + (uint64_t)g * epg + j;                       \
              if (e < n) p = FMA(a[e0 + e], xx[columns[e0 + e]], p);               \
            }                                                                      \
          part[g] = (double)p;                                                     \
        }                                                                          \
        for (uint32_t h = groups / 2; h >= 1; h /= 2)                              \
          for (uint32_t g = 0; g < h; g++) part[g] = (double)((T)part[g] + (T)part[g + h]); \
        yy[r0] = (T)part[0];                                                       \
        continue;                                                                  \
      }                                                                            \
      for (uint32_t r = r0; r < r1; r++) {                                         \
        const uint32_t lo = row_ptr[r] - e0, hi = row_ptr[r + 1] - e0;            \
        if (lo == hi) {                                                            \
          yy[r] = 0;                                                               \
          continue;                                                                \
        }                                                                          \
        const uint32_t ga = lo / epg, gb = (hi - 1) / epg;                         \
        T sum = 0;                                                                 \
        for (uint32_t g = gb + 1; g-- > ga;) {                                     \
          const uint32_t s0 = g * epg > lo ? g * epg : lo;                         \
          const uint32_t s1 = (g + 1) * epg < hi ? (g + 1) * epg : hi;             \
          T seg = 0;                                                               \
          for (uint32_t e = s0; e < s1; e++)                                       \
            seg = FMA(a[e0 + e], xx[columns[e0 + e]], seg);                        \
          sum = g == gb ? seg : sum + seg;                                         \
        }                                                                          \
        yy[r] = sum;                                                               \
      }                                                                            \
    }                                                                              \
  } while (0)

int orc_spmm_stream_col(int dtype, uint32_t W, uint32_t groups, uint32_t rows,
                        const uint32_t *row_ptr, const uint32_t *columns,
                        const void *data, const void *x, void *y)
{
  if (W == 0 || groups == 0 || W % groups != 0 || (groups & (groups - 1)) != 0)
    return -1;
  const uint32_t epg = W / groups;
  const uint32_t nb = orc_fill_row_blocks(0, rows, row_ptr, NULL, W);
  uint32_t *blocks = malloc(((size_t)nb + 1) * sizeof(uint32_t));
  double *part = malloc((size_t)groups * sizeof(double));
  if (!blocks || !part) {
    free(blocks);
    free(part);
    return -1;
  }
  orc_fill_row_blocks(1, rows, row_ptr, blocks, W);
  if (dtype == ORC_F32)
    STREAM_BODY(float, fmaf);
  else
    STREAM_BODY(double, fma);
  free(blocks);
  free(part);
  return 0;
}
#undef STREAM_BODY

void orc_spmv_ell_seq(int dtype, uint32_t rows, uint32_t width,
                      const uint32_t *ell_cols, const void *ell_data,
                      const void *x, void *y)
{
  for (uint32_t r = 0; r < rows; r++) {
    if (dtype == ORC_F32) {
      const float *a = ell_data, *xx = x;
      float dot = 0;
      for (uint32_t k = 0; k < width; k++) {
        size_t o = (size_t)r + (size_t)k * rows;
        dot = fmaf(a[o], xx[ell_cols[o]], dot);
      }
      ((float *)y)[r] = dot;
    } else {
      const double *a = ell_data, *xx = x;
      double dot = 0;
      for (uint32_t k = 0; k < width; k++) {
        size_t o = (size_t)r + (size_t)k * rows;
        dot = fma(a[o], xx[ell_cols[o]], dot);
      }
      ((double *)y)[r] = dot;
    }
  }
}

/* Single-row reduction of elements [b, e) as done by one workgroup of
 * `threads` lanes: fma partials strided by `threads`, a 64-lane halving tree
 * per wave, then a halving tree over the waves. */
#define DEFINE_WG_ROW(T, FMA, SUF)                                            \
  static T wg_row_##SUF(uint32_t threads, uint32_t b, uint32_t e,            \
                        const uint32_t *ci, const T *a, const T *x)           \
  {                                                                           \
    T part[1024];                                                             \
    for (uint32_t t = 0; t < threads; t++) {                                  \
      T d = 0;                                                                \
      for (uint32_t i = b + t; i < e; i += threads)                           \
        d = FMA(a[i], x[ci[i]], d);                                           \
      part[t] = d;                                                            \
    }                                                                         \
    uint32_t waves = threads / 64;                                            \
    T wsum[16];                                                               \
    for (uint32_t w = 0; w < waves; w++) {                                    \
      T *v = part + 64 * w;                                                   \
      for (uint32_t off = 32; off > 0; off /= 2)                              \
        for (uint32_t l = 0; l < off; l++)                                    \
          v[l] = v[l] + v[l + off];                                           \
      wsum[w] = v[0];                                                         \
    }                                                                         \
    for (uint32_t off = waves / 2; off > 0; off /= 2)                         \
      for (uint32_t l = 0; l < off; l++)                                      \
        wsum[l] = wsum[l] + wsum[l + off];                                    \
    return wsum[0];                                                           \
  }
DEFINE_WG_ROW(float, fmaf, f32)
DEFINE_WG_ROW(double, fma, f64)

/* Segmented reduction of one multi-row block (CSR-Stream).  Lane i of B
 * sums the rounded products cache[i*E, i*E+E) (E = ceil(W/B)) left to right,
 * restarting at every row boundary.  A row that starts and ends inside one
 * lane is finished there.  The partial a lane carries into the next lane
 * (its last row continues) enters a Hillis-Steele segmented inclusive scan
 * over the B lanes: for off = 1, 2, 4, ...: a lane whose chain does not start
 * in it adds the value `off` lanes to its left (s = s[i-off] + s) and takes
 * that lane's flag.  A row ending in lane j that began earlier gets
 * scan[j-1] + (lane j's partial). */
#define DEFINE_STREAM(T, SUF)                                                 \
  static void stream_block_##SUF(uint32_t r0, uint32_t r1, const uint32_t *rp, \
                                 const T *cache, uint32_t W, uint32_t B,      \
                                 T *y, T *sv, T *sv2, unsigned char *sf,      \
                                 unsigned char *sf2, uint32_t *frow,          \
                                 T *fval)                                     \
  {                                                                           \
    const uint32_t base = rp[r0], n = rp[r1] - base, nr = r1 - r0;            \
    const uint32_t E = (W + B - 1) / B;                                       \
    for (uint32_t lr = 0; lr < nr; lr++)                                      \
      if (rp[r0 + lr] == rp[r0 + lr + 1])                                     \
        y[r0 + lr] = 0;                                                       \
    for (uint32_t i = 0; i < B; i++) {                                        \
      sv[i] = 0;                                                              \
      sf[i] = 1;                                                              \
      frow[i] = 0xFFFFFFFFu;                                                  \
      const uint32_t e0 = i * E;                                              \
      const uint32_t e1 = e0 + E < n ? e0 + E : n;                            \
      if (e0 >= e1)                                                           \
        continue;                                                             \
      uint32_t cur = 0;                                                       \
      while (rp[r0 + cur + 1] - base <= e0)                                   \
        cur++;                                                                \
      T d = 0;                                                                \
      uint32_t seg = e0;                                                      \
      for (uint32_t e = e0; e < e1; e++) {                                    \
        while (rp[r0 + cur + 1] - base <= e) {                                \
          if (seg < e) {                                                      \
            if (seg == rp[r0 + cur] - base)                                   \
              y[r0 + cur] = d;                                                \
            else {                                                            \
              frow[i] = cur;                                                  \
              fval[i] = d;                                                    \
            }                                                                 \
          }                                                                   \
          d = 0;                                                              \
          seg = e;                                                            \
          cur++;                                                              \
        }                                                                     \
        d = d + cache[e];                                                     \
      }                                                                       \
      const int starts = seg == rp[r0 + cur] - base;                          \
      const int ends = e1 == rp[r0 + cur + 1] - base;                         \
      if (ends) {                                                             \
        if (starts)                                                           \
          y[r0 + cur] = d;                                                    \
        else {                                                                \
          frow[i] = cur;                                                      \
          fval[i] = d;                                                        \
        }                                                                     \
      } else {                                                                \
        sv[i] = d;                                                            \
        sf[i] = (unsigned char)starts;                                        \
      }                                                                       \
    }                                                                         \
    for (uint32_t off = 1; off < B; off <<= 1) {                              \
      for (uint32_t i = 0; i < B; i++) {                                      \
        sv2[i] = sv[i];                                                       \
        sf2[i] = sf[i];                                                       \
        if (!sf[i] && i >= off) {                                             \
          sv2[i] = sv[i - off] + sv[i];                                       \
          sf2[i] = sf[i - off];                                               \
        }                                                                     \
      }                                                                       \
      memcpy(sv, sv2, sizeof(T) * B);                                         \
      memcpy(sf, sf2, B);                                                     \
    }                                                                         \
    for (uint32_t i = 1; i < B; i++)                                          \
      if (frow[i] != 0xFFFFFFFFu)                                             \
        y[r0 + frow[i]] = sv[i - 1] + fval[i];                                \
    (void)nr;                                                                 \
  }
DEFINE_STREAM(float, f32)
DEFINE_STREAM(double, f64)

#define DEFINE_ADAPTIVE(T, SUF)                                               \
  static void adaptive_##SUF(uint32_t rows, const uint32_t *rp,              \
                             const uint32_t *ci, const T *a, const T *x,      \
                             T *y, uint32_t W, uint32_t B, uint32_t long_row, \
                             uint32_t chunk)                                  \
  {                                                                           \
    uint32_t nb = orc_fill_row_blocks(0, rows, rp, NULL, W);                  \
    uint32_t *blk = malloc(sizeof(uint32_t) * (nb + 1));                      \
    orc_fill_row_blocks(1, rows, rp, blk, W);                                 \
    T *cache = malloc(sizeof(T) * (W + 1));                                   \
    T *sv = malloc(sizeof(T) * B), *sv2 = malloc(sizeof(T) * B);              \
    T *fval = malloc(sizeof(T) * B);                                          \
    unsigned char *sf = malloc(B), *sf2 = malloc(B);                          \
    uint32_t *frow = malloc(sizeof(uint32_t) * B);                            \
    for (uint32_t b = 0; b < nb; b++) {                                       \
      uint32_t r0 = blk[b], r1 = blk[b + 1];                                  \
      if (r1 <= r0)                                                           \
        continue;                                                             \
      if (r1 - r0 > 1) {                                                      \
        uint32_t base = rp[r0], n = rp[r1] - base;                            \
        for (uint32_t i = 0; i < n; i++)                                      \
          cache[i] = a[base + i] * x[ci[base + i]];                           \
        stream_block_##SUF(r0, r1, rp, cache, W, B, y, sv, sv2, sf, sf2,      \
                           frow, fval);                                       \
      } else {                                                                \
        uint32_t s = rp[r0], e = rp[r0 + 1];                                  \
        if (e - s > long_row) {                                               \
          T acc = 0;                                                          \
          int first = 1;                                                      \
          for (uint32_t c = s; c < e; c += chunk) {                           \
            uint32_t ce = c + chunk < e ? c + chunk : e;                      \
            T p = wg_row_##SUF(B, c, ce, ci, a, x);                           \
            acc = first ? p : acc + p;                                        \
            first = 0;                                                        \
          }                                                                   \
          y[r0] = acc;                                                        \
        } else {                                                              \
          y[r0] = wg_row_##SUF(B, s, e, ci, a, x);                            \
        }                                                                     \
      }                                                                       \
    }                                                                         \
    free(cache);                                                              \
    free(blk);                                                                \
    free(sv);                                                                 \
    free(sv2);                                                                \
    free(fval);                                                               \
    free(sf);                                                                 \
    free(sf2);                                                                \
    free(frow);                                                               \
  }
DEFINE_ADAPTIVE(float, f32)
DEFINE_ADAPTIVE(double, f64)

void orc_spmv_csr_adaptive(int dtype, uint32_t rows, const uint32_t *row_ptr,
                           const uint32_t *columns, const void *data,
                           const void *x, void *y, uint32_t nnz_per_wg,
                           uint32_t block_threads, uint32_t long_row,
                           uint32_t chunk)
{
  if (dtype == ORC_F32)
    adaptive_f32(rows, row_ptr, columns, data, x, y, nnz_per_wg,
                 block_threads, long_row, chunk);
  else
    adaptive_f64(rows, row_ptr, columns, data, x, y, nnz_per_wg,
                 block_threads, long_row, chunk);
}

void orc_spmv_csr_f64(int dtype, uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data,
                      const void *x, double *y64, double *absdot)
{
  for (uint32_t r = 0; r < rows; r++) {
    double d = 0.0, ad = 0.0;
    for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; e++) {
      double av, xv;
      if (dtype == ORC_F32) {
        av = ((const float *)data)[e];
        xv = ((const float *)x)[columns[e]];
      } else {
        av = ((const double *)data)[e];
        xv = ((const double *)x)[columns[e]];
      }
      d += av * xv;
      ad += fabs(av * xv);
    }
    y64[r] = d;
    if (absdot)
      absdot[r] = ad;
  }
}

/* ---------------------------------------------------------------------- */
/* Multi-thread CPU baseline (cpu/cpu_matrix_multiplier.cpp:229-308)       */
/* ---------------------------------------------------------------------- */

typedef struct {
  int dtype;
  uint32_t rows, r0, r1, threads;
  const uint32_t *rp, *ci;
  const void *a, *x;
  void *y;
  atomic_uint *started;
  double elapsed;
} mt_arg;

static void *mt_worker(void *p)
{
  mt_arg *m = p;
  atomic_fetch_add_explicit(m->started, 1u, memory_order_relaxed);
  while (atomic_load_explicit(m->started, memory_order_relaxed) < m->threads)
    __builtin_ia32_pause();
  double t0 = now_s();
  if (m->dtype == ORC_F32)
    spmv_seq_f32(m->rows, m->rp, m->ci, m->a, m->x, m->y, m->r0, m->r1);
  else
    spmv_seq_f64(m->rows, m->rp, m->ci, m->a, m->x, m->y, m->r0, m->r1);
  m->elapsed = now_s() - t0;
  return NULL;
}

double orc_spmv_csr_multi_thread(int dtype, uint32_t threads, uint32_t rows,
                                 const uint32_t *row_ptr,
                                 const uint32_t *columns, const void *data,
                                 const void *x, void *y)
{
  if (threads == 0)
    threads = 1;
  pthread_t *th = malloc(sizeof(pthread_t) * threads);
  mt_arg *args = malloc(sizeof(mt_arg) * threads);
  atomic_uint started;
  atomic_init(&started, 0u);
  uint32_t per = rows / threads;
  for (uint32_t t = 0; t < threads; t++) {
    mt_arg *m = &args[t];
    m->dtype = dtype;
    m->rows = rows;
    m->r0 = per * t;
    m->r1 = (t == threads - 1) ? rows : (t + 1) * per;
    m->threads = threads;
    m->rp = row_ptr;
    m->ci = columns;
    m->a = data;
    m->x = x;
    m->y = y;
    m->started = &started;
    m->elapsed = 0;
    pthread_create(&th[t], NULL, mt_worker, m);
  }
  double mx = 0.0;
  for (uint32_t t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    if (args[t].elapsed > mx)
      mx = args[t].elapsed;
  }
  free(th);
  free(args);
  return mx;
}

double orc_spmv_csr_seq_timed(int dtype, uint32_t reps, uint32_t rows,
                              const uint32_t *row_ptr, const uint32_t *columns,
                              const void *data, const void *x, void *y)
{
  if (reps == 0)
    reps = 1;
  double t0 = now_s();
  for (uint32_t i = 0; i < reps; i++)
    orc_spmv_csr_seq(dtype, rows, row_ptr, columns, data, x, y);
  return (now_s() - t0) / reps;
}

/* ---------------------------------------------------------------------- */
/* Measurement arithmetic                                                 */
/* ---------------------------------------------------------------------- */

double orc_compare_error(int dtype, uint32_t n, const void *a, const void *b)
{
  if (dtype == ORC_F32) {
    const float *pa = a, *pb = b;
    float num = 0, den = 0;
    for (uint32_t i = 0; i < n; i++) {
      float d = pa[i] - pb[i];
      num = num + mul_f32(d, d);
      den = den + mul_f32(pb[i], pb[i]);
    }
    return (double)(num / den);
  } else {
    const double *pa = a, *pb = b;
    double num = 0, den = 0;
    for (uint32_t i = 0; i < n; i++) {
      double d = pa[i] - pb[i];
      num = num + mul_f64(d, d);
      den = den + mul_f64(pb[i], pb[i]);
    }
    return num / den;
  }
}

void orc_measurement(double elapsed, double bytes, double ops, double out2[2])
{
  const double giga = 1e9;
  out2[0] = bytes / (elapsed * giga);
  out2[1] = ops / (elapsed * giga);
}

double orc_reference_bytes(int kind, int dtype, uint32_t rows, uint64_t nnz,
                           uint32_t ell_width)
{
  uint64_t s = dtype == ORC_F32 ? 4 : 8;
  if (kind == 0)
    return (double)(nnz * s + nnz * s + nnz * 4 + 2ull * rows * 4 + rows * s);
  if (kind == 1) {
    uint32_t n = ell_width * rows; /* unsigned int product, wraps mod 2^32 */
    return (double)((uint64_t)n * s + (uint64_t)n * s + (uint64_t)n * 4 +
                    (uint64_t)rows * s);
  }
  return (double)(nnz * s + nnz * s + nnz * 4 + nnz * 4 + nnz * s);
}
