// 64-d structure features of a device CSR matrix, the format selector's input.
// Two passes accumulate integer counters (so the result does not depend on
// atomic order); the host turns them into the feature vector in double.
//
//  0 log2(rows+1)/32          1 log2(cols+1)/32         2 log2(nnz+1)/40
//  3 log2(1+mean)/16          4 log2(1+max)/16          5 min(cv, 32)/32
//  6 empty-row fraction       7 log2(1+std)/16
//  8..31  row fraction per row-length bin: 0 for empty rows, 1 + floor(log2 len)
//         capped at 23
// 32..47  nnz fraction per bin floor(log2 len) capped at 15
// 48..55  nnz fraction per relative distance d = |c rows - r cols| / (rows cols)
//         from the scaled diagonal: d < 1e-4 4^k for k = 0..6, then d >= 0.4096
// 56 mean d (d quantised down to 2^-24)   57 fraction of entries with c == r
// 58 mean over non-empty rows of (last column - first column + 1) / cols
// 59 log2(rows max / nnz)/16 (ELL padding factor, 0 when nnz = 0)
// 60 nnz fraction in rows longer than 4096   61 nnz fraction in rows of <= 4
// 62 fraction of rows longer than 64         63 constant 1
// mean, std and cv are of the row length (population std).
#include "mfp_internal.h"
#include "objects.h"

#include <cmath>

namespace mfp {

struct feat_counters {
  unsigned long long len_hist[24];
  unsigned long long nnz_hist[16];
  unsigned long long dist_hist[8];
  unsigned long long sumsq, empty, nnz_big, nnz_small, rows_gt64, span_sum, nonempty, max_len;
  unsigned long long dist_sum, diag;
};

__device__ __forceinline__ int ilog2(uint32_t v) { return 31 - __clz(v); }

__global__ __launch_bounds__(256) void k_feat_rows(uint32_t rows, const uint32_t *__restrict__ rp,
                                                   const uint32_t *__restrict__ ci, feat_counters *out)
{
  __shared__ unsigned long long s_len[24], s_nnz[16], s_misc[7];
  __shared__ unsigned s_max;
  const int t = threadIdx.x;
  if (t < 24) s_len[t] = 0;
  if (t < 16) s_nnz[t] = 0;
  if (t < 7) s_misc[t] = 0;
  if (t == 0) s_max = 0;
  __syncthreads();
  unsigned long long sumsq = 0, empty = 0, big = 0, small = 0, gt64 = 0, span = 0, nonempty = 0;
  unsigned mx = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + t; r < rows; r += gridDim.x * blockDim.x) {
    const uint32_t b = rp[r], len = rp[r + 1] - b;
    sumsq += (unsigned long long)len * len;
    mx = max(mx, len);
    if (len == 0) {
      empty++;
      atomicAdd(&s_len[0], 1ull);
      continue;
    }
    const int lb = ilog2(len);
    atomicAdd(&s_len[min(23, lb + 1)], 1ull);
    atomicAdd(&s_nnz[min(15, lb)], (unsigned long long)len);
    if (len > 4096) big += len;
    if (len <= 4) small += len;
    if (len > 64) gt64++;
    const uint32_t c0 = ci[b], c1 = ci[b + len - 1];
    span += (unsigned long long)(c1 >= c0 ? c1 - c0 : c0 - c1) + 1;
    nonempty++;
  }
  atomicAdd(&s_misc[0], sumsq);
  atomicAdd(&s_misc[1], empty);
  atomicAdd(&s_misc[2], big);
  atomicAdd(&s_misc[3], small);
  atomicAdd(&s_misc[4], gt64);
  atomicAdd(&s_misc[5], span);
  atomicAdd(&s_misc[6], nonempty);
  atomicMax(&s_max, mx);
  __syncthreads();
  if (t < 24 && s_len[t]) atomicAdd(&out->len_hist[t], s_len[t]);
  if (t < 16 && s_nnz[t]) atomicAdd(&out->nnz_hist[t], s_nnz[t]);
  if (t == 0) {
    atomicAdd(&out->sumsq, s_misc[0]);
    atomicAdd(&out->empty, s_misc[1]);
    atomicAdd(&out->nnz_big, s_misc[2]);
    atomicAdd(&out->nnz_small, s_misc[3]);
    atomicAdd(&out->rows_gt64, s_misc[4]);
    atomicAdd(&out->span_sum, s_misc[5]);
    atomicAdd(&out->nonempty, s_misc[6]);
    atomicMax(&out->max_len, (unsigned long long)s_max);
  }
}

// One thread per run of ITEMS consecutive entries; the row of the first is
// found by binary search, later ones by advancing.
constexpr int FEAT_ITEMS = 16;

__global__ __launch_bounds__(256) void k_feat_entries(uint32_t rows, uint32_t cols, uint64_t nnz,
                                                      const uint32_t *__restrict__ rp,
                                                      const uint32_t *__restrict__ ci,
                                                      const double *__restrict__ thresholds,
                                                      feat_counters *out)
{
  __shared__ unsigned long long s_d[8], s_sum, s_diag;
  const int t = threadIdx.x;
  if (t < 8) s_d[t] = 0;
  if (t == 0) s_sum = s_diag = 0;
  __syncthreads();
  const uint64_t e0 = ((uint64_t)blockIdx.x * blockDim.x + t) * FEAT_ITEMS;
  unsigned long long dsum = 0, diag = 0;
  if (e0 < nnz) {
    uint32_t lo = 0, hi = rows; // last row with rp[row] <= e0
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (rp[mid] <= e0) lo = mid;
      else hi = mid;
    }
    uint32_t r = lo;
    const double denom = (double)rows * (double)cols;
    const uint64_t e1 = min(nnz, e0 + FEAT_ITEMS);
    for (uint64_t e = e0; e < e1; e++) {
      while (rp[r + 1] <= e) r++;
      const uint32_t c = ci[e];
      const uint64_t a = (uint64_t)c * rows, b = (uint64_t)r * cols;
      const double d = (double)(a > b ? a - b : b - a) / denom;
      int k = 0;
      while (k < 7 && d >= thresholds[k]) k++;
      atomicAdd(&s_d[k], 1ull);
      dsum += (unsigned long long)(d * 16777216.0);
      diag += c == r;
    }
  }
  atomicAdd(&s_sum, dsum);
  atomicAdd(&s_diag, diag);
  __syncthreads();
  if (t < 8 && s_d[t]) atomicAdd(&out->dist_hist[t], s_d[t]);
  if (t == 0) {
    atomicAdd(&out->dist_sum, s_sum);
    atomicAdd(&out->diag, s_diag);
  }
}

} // namespace mfp

using namespace mfp;

extern "C" int mfp_matrix_features(mfp_matrix m, float *out)
{
  MFP_ARG(m && out, "null pointer");
  MFP_ARG(m->rows > 0 && m->cols > 0, "empty matrix");
  feat_counters *dc = nullptr;
  double *dth = nullptr;
  MFP_CHECK(dmalloc((void **)&dc, sizeof(feat_counters)));
  int st = dmalloc((void **)&dth, 8 * sizeof(double));
  double th[8];
  for (int k = 0; k < 7; k++) th[k] = 1e-4 * std::pow(4.0, k);
  th[7] = 0;
  feat_counters h{};
  if (st == MFP_OK) st = mfp_memcpy_h2d(dth, th, sizeof th);
  if (st == MFP_OK) st = fill_bytes_zero(dc, sizeof(feat_counters), nullptr);
  if (st == MFP_OK) {
    const uint32_t g1 = std::min<uint32_t>(4096, ceil_div_u32(m->rows, 256));
    k_feat_rows<<<g1, 256>>>(m->rows, m->row_ptr, m->columns, dc);
    if (m->nnz) {
      const uint64_t g2 = (m->nnz + 256ull * FEAT_ITEMS - 1) / (256ull * FEAT_ITEMS);
      k_feat_entries<<<(uint32_t)g2, 256>>>(m->rows, m->cols, m->nnz, m->row_ptr, m->columns, dth, dc);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("feature kernels: ") + hipGetErrorString(e));
      st = MFP_ERR_HIP;
    }
  }
  if (st == MFP_OK) st = mfp_memcpy_d2h(&h, dc, sizeof h);
  dfree(dc);
  dfree(dth);
  MFP_CHECK(st);

  const double rows = m->rows, cols = m->cols, nnz = (double)m->nnz;
  const double mean = nnz / rows;
  const double var = std::max(0.0, (double)h.sumsq / rows - mean * mean);
  const double sd = std::sqrt(var);
  const double cv = mean > 0 ? sd / mean : 0.0;
  const double nz = nnz > 0 ? nnz : 1.0;
  double f[64] = {0};
  f[0] = std::log2(rows + 1) / 32;
  f[1] = std::log2(cols + 1) / 32;
  f[2] = std::log2(nnz + 1) / 40;
  f[3] = std::log2(1 + mean) / 16;
  f[4] = std::log2(1 + (double)h.max_len) / 16;
  f[5] = std::min(cv, 32.0) / 32;
  f[6] = (double)h.empty / rows;
  f[7] = std::log2(1 + sd) / 16;
  for (int b = 0; b < 24; b++) f[8 + b] = (double)h.len_hist[b] / rows;
  for (int b = 0; b < 16; b++) f[32 + b] = (double)h.nnz_hist[b] / nz;
  for (int b = 0; b < 8; b++) f[48 + b] = (double)h.dist_hist[b] / nz;
  f[56] = (double)h.dist_sum / 16777216.0 / nz;
  f[57] = (double)h.diag / nz;
  f[58] = h.nonempty ? (double)h.span_sum / (double)h.nonempty / cols : 0.0;
  f[59] = nnz > 0 ? std::log2(rows * (double)h.max_len / nnz) / 16 : 0.0;
  f[60] = (double)h.nnz_big / nz;
  f[61] = (double)h.nnz_small / nz;
  f[62] = (double)h.rows_gt64 / rows;
  f[63] = 1.0;
  for (int i = 0; i < 64; i++) out[i] = (float)f[i];
  return MFP_OK;
}
