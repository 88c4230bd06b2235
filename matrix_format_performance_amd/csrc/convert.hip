// Device-side matrix converters.
//
// The reference builds every layout on the host (common/matrix_converter.cpp,
// common/scoo_matrix_class.cpp, gpu/csr_adaptive_spmv.cu:180-240).  Here each
// layout is built by HIP kernels in HBM; the results are bit-identical to the
// reference's host algorithms (ties in triplets->CSR keep input order).
#include "mfp_internal.h"
#include "wave64.h"

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

namespace mfp {

// ---------------------------------------------------------------------------
// Exclusive scan of u32 (hierarchical: tile scan, scan of tile sums, add).
// ---------------------------------------------------------------------------
constexpr uint32_t SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *sh, uint32_t *total)
{
  // Hillis-Steele over 256 thread values in LDS.
  const uint32_t t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < SCAN_T; off <<= 1) {
    uint32_t add = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint32_t incl = sh[t];
  if (total) *total = sh[SCAN_T - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const uint32_t *__restrict__ in,
                                                       uint32_t *__restrict__ out, uint64_t n,
                                                       uint32_t *__restrict__ sums)
{
  __shared__ uint32_t tile[SCAN_TILE];
  __shared__ uint32_t sh[SCAN_T];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  const uint32_t t = threadIdx.x;
  for (uint32_t k = 0; k < SCAN_I; k++) {
    uint64_t i = base + k * SCAN_T + t;
    tile[k * SCAN_T + t] = i < n ? in[i] : 0u;
  }
  __syncthreads();
  uint32_t loc[SCAN_I];
  uint32_t acc = 0;
  for (uint32_t k = 0; k < SCAN_I; k++) {
    loc[k] = acc;
    acc += tile[t * SCAN_I + k];
  }
  uint32_t total = 0;
  uint32_t pre = block_exclusive_scan_256(acc, sh, &total);
  for (uint32_t k = 0; k < SCAN_I; k++) tile[t * SCAN_I + k] = loc[k] + pre;
  __syncthreads();
  for (uint32_t k = 0; k < SCAN_I; k++) {
    uint64_t i = base + k * SCAN_T + t;
    if (i < n) out[i] = tile[k * SCAN_T + t];
  }
  if (t == 0 && sums) sums[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t *__restrict__ out, uint64_t n, const uint32_t *__restrict__ sums)
{
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  const uint32_t add = sums[blockIdx.x];
  for (uint32_t k = 0; k < SCAN_I; k++) {
    uint64_t i = base + k * SCAN_T + threadIdx.x;
    if (i < n) out[i] += add;
  }
}

__global__ void k_scan_total(const uint32_t *in, const uint32_t *out, uint64_t n, uint32_t *total)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0u;
}

static int scan_rec(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  const uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint32_t *sums = nullptr;
  if (tiles > 1) MFP_CHECK(dmalloc((void **)&sums, tiles * 4));
  k_scan_tiles<<<(uint32_t)tiles, SCAN_T, 0, s>>>(in, out, n, sums);
  MFP_HIP(hipGetLastError());
  if (tiles > 1) {
    int st = scan_rec(sums, sums, tiles, s);
    if (st == MFP_OK) {
      k_scan_add<<<(uint32_t)tiles, SCAN_T, 0, s>>>(out, n, sums);
      st = hipGetLastError() == hipSuccess ? MFP_OK : MFP_ERR_HIP;
    }
    MFP_HIP(hipStreamSynchronize(s)); // sums is freed below
    dfree(sums);
    return st;
  }
  return MFP_OK;
}

// In-place (in == out) is allowed.  total_out (host) gets in[0..n) summed.
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s,
                       uint32_t *total_out)
{
  uint32_t last_in = 0;
  if (total_out && n) MFP_HIP(hipMemcpyAsync(&last_in, in + n - 1, 4, hipMemcpyDeviceToHost, s));
  if (total_out) MFP_HIP(hipStreamSynchronize(s));
  MFP_CHECK(scan_rec(in, out, n, s));
  if (total_out) {
    uint32_t last_out = 0;
    if (n) MFP_HIP(hipMemcpyAsync(&last_out, out + n - 1, 4, hipMemcpyDeviceToHost, s));
    MFP_HIP(hipStreamSynchronize(s));
    *total_out = n ? last_out + last_in : 0;
  }
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// Triplets -> CSR: counts + scan give row_ptr; a stable LSD radix sort of
// (row, col, input index) by col then row gives the element order.
// ---------------------------------------------------------------------------
__global__ void k_count_rows(const uint32_t *__restrict__ r, const uint32_t *__restrict__ c,
                             uint64_t n, uint32_t rows, uint32_t cols,
                             uint32_t *__restrict__ cnt, uint32_t *__restrict__ bad)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t rr = r[i], cc = c[i];
    if (rr >= rows || cc >= cols) {
      atomicOr(bad, 1u);
      continue;
    }
    atomicAdd(cnt + rr, 1u);
  }
}

__global__ void k_iota(uint32_t *__restrict__ p, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (uint32_t)i;
}

constexpr uint32_t RX_T = 256, RX_I = 8, RX_TILE = RX_T * RX_I;

__global__ __launch_bounds__(RX_T) void k_radix_hist(const uint32_t *__restrict__ key, uint64_t n,
                                                     uint32_t shift, uint32_t *__restrict__ hist,
                                                     uint32_t ntiles)
{
  __shared__ uint32_t cnt[256];
  const uint32_t t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RX_TILE;
  for (uint32_t k = 0; k < RX_I; k++) {
    uint64_t i = base + k * RX_T + t;
    if (i < n) atomicAdd(&cnt[(key[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)t * ntiles + blockIdx.x] = cnt[t];
}

// Stable scatter of three u32 arrays by the 8-bit digit of key.  Elements are
// ranked in tile order (round k = elements k*256 .. k*256+255, by thread).
__global__ __launch_bounds__(RX_T) void k_radix_scatter(
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ a0,
    const uint32_t *__restrict__ a1, const uint32_t *__restrict__ a2, uint32_t *__restrict__ o0,
    uint32_t *__restrict__ o1, uint32_t *__restrict__ o2, uint64_t n, uint32_t shift,
    const uint32_t *__restrict__ offs, uint32_t ntiles)
{
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[RX_T / 64][256];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  run[t] = offs[(uint64_t)t * ntiles + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * RX_TILE;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
  for (uint32_t k = 0; k < RX_I; k++) {
    for (uint32_t w = 0; w < RX_T / 64; w++) wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t i = base + k * RX_T + t;
    const bool valid = i < n;
    const uint32_t d = valid ? (key[i] >> shift) & 255u : 0u;
    uint64_t peers = __ballot(valid);
    for (uint32_t b = 0; b < 8; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (uint32_t w = 0; w < wave; w++) pre += wcnt[w][d];
      const uint32_t pos = run[d] + pre + rank;
      o0[pos] = a0[i];
      o1[pos] = a1[i];
      o2[pos] = a2[i];
    }
    __syncthreads();
    uint32_t tot = 0;
    for (uint32_t w = 0; w < RX_T / 64; w++) tot += wcnt[w][t];
    run[t] += tot;
  }
}

template <typename T>
__global__ void k_gather_values(const uint32_t *__restrict__ idx, const double *__restrict__ v,
                                T *__restrict__ out, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (T)v[idx[i]];
}

static uint32_t bit_length(uint32_t v)
{
  uint32_t b = 0;
  while (v) {
    b++;
    v >>= 1;
  }
  return b;
}

static uint32_t grid_for(uint64_t n, uint32_t bs = 256)
{
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + bs - 1) / bs, 1u << 16));
}

int csr_from_triplets_device(int dtype, uint32_t rows, uint32_t cols, uint64_t nnz,
                             const uint32_t *tr_rows, const uint32_t *tr_cols,
                             const double *tr_vals, uint32_t *row_ptr, uint32_t *columns,
                             void *values, hipStream_t s)
{
  MFP_ARG(nnz < (1ull << 32), "nnz must fit in 32 bits");
  uint32_t *bad = nullptr;
  MFP_CHECK(dmalloc((void **)&bad, 4));
  MFP_CHECK(fill_u32(bad, 0, 1, s));
  MFP_CHECK(fill_u32(row_ptr, 0, (uint64_t)rows + 1, s));
  if (nnz) {
    k_count_rows<<<grid_for(nnz), 256, 0, s>>>(tr_rows, tr_cols, nnz, rows, cols, row_ptr, bad);
    MFP_HIP(hipGetLastError());
  }
  uint32_t hbad = 0;
  MFP_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  MFP_HIP(hipStreamSynchronize(s));
  dfree(bad);
  if (hbad) {
    set_error("triplet row/col index out of range");
    return MFP_ERR_ARG;
  }
  MFP_CHECK(exclusive_scan_u32(row_ptr, row_ptr, (uint64_t)rows + 1, s, nullptr));
  if (nnz == 0) return MFP_OK;

  // ping-pong buffers: (row, col, idx)
  uint32_t *buf[6] = {};
  for (int i = 0; i < 6; i++) {
    int st = dmalloc((void **)&buf[i], nnz * 4);
    if (st != MFP_OK) {
      for (int j = 0; j < i; j++) dfree(buf[j]);
      return st;
    }
  }
  const uint32_t ntiles = (uint32_t)((nnz + RX_TILE - 1) / RX_TILE);
  uint32_t *hist = nullptr;
  int st = dmalloc((void **)&hist, (uint64_t)ntiles * 256 * 4);
  auto cleanup = [&]() {
    for (int i = 0; i < 6; i++) dfree(buf[i]);
    dfree(hist);
  };
  if (st != MFP_OK) {
    cleanup();
    return st;
  }
  // own copies of the keys, so the caller's arrays are never written
  MFP_HIP(hipMemcpyAsync(buf[0], tr_rows, nnz * 4, hipMemcpyDeviceToDevice, s));
  MFP_HIP(hipMemcpyAsync(buf[1], tr_cols, nnz * 4, hipMemcpyDeviceToDevice, s));
  k_iota<<<grid_for(nnz), 256, 0, s>>>(buf[2], nnz);
  uint32_t *cur[3] = {buf[0], buf[1], buf[2]};
  uint32_t *nxt[3] = {buf[3], buf[4], buf[5]};
  // passes: col digits (LSD), then row digits
  std::vector<std::pair<int, uint32_t>> passes;
  for (uint32_t sh = 0; sh < bit_length(cols ? cols - 1 : 0); sh += 8) passes.push_back({1, sh});
  for (uint32_t sh = 0; sh < bit_length(rows ? rows - 1 : 0); sh += 8) passes.push_back({0, sh});
  for (auto [which, sh] : passes) {
    const uint32_t *key = cur[which];
    k_radix_hist<<<ntiles, RX_T, 0, s>>>(key, nnz, sh, hist, ntiles);
    if (hipGetLastError() != hipSuccess) {
      cleanup();
      set_error("radix hist launch failed");
      return MFP_ERR_HIP;
    }
    st = exclusive_scan_u32(hist, hist, (uint64_t)ntiles * 256, s, nullptr);
    if (st != MFP_OK) {
      cleanup();
      return st;
    }
    k_radix_scatter<<<ntiles, RX_T, 0, s>>>(key, cur[0], cur[1], cur[2], nxt[0], nxt[1], nxt[2],
                                             nnz, sh, hist, ntiles);
    if (hipGetLastError() != hipSuccess) {
      cleanup();
      set_error("radix scatter launch failed");
      return MFP_ERR_HIP;
    }
    std::swap(cur, nxt);
  }
  // cur[1] = sorted columns, cur[2] = source index
  MFP_HIP(hipMemcpyAsync(columns, cur[1], nnz * 4, hipMemcpyDeviceToDevice, s));
  if (dtype == MFP_F32)
    k_gather_values<float><<<grid_for(nnz), 256, 0, s>>>(cur[2], tr_vals, (float *)values, nnz);
  else
    k_gather_values<double><<<grid_for(nnz), 256, 0, s>>>(cur[2], tr_vals, (double *)values, nnz);
  MFP_HIP(hipGetLastError());
  MFP_HIP(hipStreamSynchronize(s));
  cleanup();
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// Row statistics (get_rows_statistics, common/matrix_converter.cpp:95-127).
// min/max/sum exactly; sum of squared deviations from the integer average in
// u64 (exact; equals the reference's double sum whenever that sum is below
// 2^53, where every partial double sum is exact).
// ---------------------------------------------------------------------------
struct stats_acc {
  uint32_t mn, mx;
  unsigned long long sum, sq;
};

__global__ void k_row_minmax(const uint32_t *__restrict__ rp, uint32_t rows, stats_acc *acc)
{
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
  unsigned long long sum = 0;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < rows; i += stride) {
    uint32_t len = rp[i + 1] - rp[i];
    mn = min(mn, len);
    mx = max(mx, len);
    sum += len;
  }
  atomicMin(&acc->mn, mn);
  atomicMax(&acc->mx, mx);
  atomicAdd(&acc->sum, sum);
}

__global__ void k_row_sq(const uint32_t *__restrict__ rp, uint32_t rows, stats_acc *acc)
{
  const uint32_t avg = rows ? (uint32_t)acc->sum / rows : 0u; // 32-bit sum as the reference
  unsigned long long sq = 0;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < rows; i += stride) {
    long long d = (long long)(rp[i + 1] - rp[i]) - (long long)avg;
    sq += (unsigned long long)(d * d);
  }
  atomicAdd(&acc->sq, sq);
}

int row_stats_device(uint32_t rows, const uint32_t *rp, mfp_row_stats *out, hipStream_t s)
{
  stats_acc init{0xFFFFFFFFu - 1u, 0u, 0ull, 0ull}; // min starts at UINT_MAX-1 (:98)
  stats_acc *acc = nullptr;
  MFP_CHECK(dmalloc((void **)&acc, sizeof(stats_acc)));
  MFP_HIP(hipMemcpyAsync(acc, &init, sizeof(init), hipMemcpyHostToDevice, s));
  const uint32_t g = std::min<uint32_t>(grid_for(rows), 2048);
  if (rows) {
    k_row_minmax<<<g, 256, 0, s>>>(rp, rows, acc);
    k_row_sq<<<g, 256, 0, s>>>(rp, rows, acc);
    MFP_HIP(hipGetLastError());
  }
  stats_acc h{};
  MFP_HIP(hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  MFP_HIP(hipStreamSynchronize(s));
  dfree(acc);
  out->min_elements_in_rows = h.mn;
  out->max_elements_in_rows = h.mx;
  out->avg_elements_in_rows = rows ? (uint32_t)h.sum / rows : 0u;
  out->elements_in_rows_std_deviation = rows ? std::sqrt((double)h.sq / rows) : 0.0;
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// ELL (common/matrix_converter.cpp:139-203): data[row + k*rows], pad (0, 0),
// rows longer than the width truncated.  Every slot is written.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_ell_build(uint32_t rows, const uint32_t *__restrict__ rp,
                            const uint32_t *__restrict__ ci, const T *__restrict__ a,
                            uint32_t width, T *__restrict__ ev, uint32_t *__restrict__ ec)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const uint32_t s = rp[r], len = rp[r + 1] - s;
  for (uint32_t k = 0; k < width; k++) {
    const uint64_t o = (uint64_t)r + (uint64_t)k * rows;
    if (k < len) {
      ev[o] = a[s + k];
      ec[o] = ci[s + k];
    } else {
      ev[o] = (T)0;
      ec[o] = 0u;
    }
  }
}

int ell_build_device(int dtype, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                     const void *val, uint32_t width, void *ev, uint32_t *ec, hipStream_t s)
{
  if (rows == 0 || width == 0) return MFP_OK;
  const uint32_t g = ceil_div_u32(rows, 256);
  if (dtype == MFP_F32)
    k_ell_build<float><<<g, 256, 0, s>>>(rows, rp, ci, (const float *)val, width, (float *)ev, ec);
  else
    k_ell_build<double><<<g, 256, 0, s>>>(rows, rp, ci, (const double *)val, width, (double *)ev, ec);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// COO (common/matrix_converter.cpp:211-280): CSR order, entries whose in-row
// offset is >= element_start.  Element-parallel with a binary search for the
// row of each CSR element.
// ---------------------------------------------------------------------------
__global__ void k_tail_counts(uint32_t rows, const uint32_t *__restrict__ rp, uint32_t e0,
                              uint32_t *__restrict__ cnt)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const uint32_t len = rp[r + 1] - rp[r];
  cnt[r] = len > e0 ? len - e0 : 0u;
}

__device__ __forceinline__ uint32_t row_of(const uint32_t *__restrict__ rp, uint32_t rows, uint32_t e)
{
  // largest r with rp[r] <= e (and rp[r+1] > e)
  uint32_t lo = 0, hi = rows; // invariant: rp[lo] <= e < rp[hi]
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (rp[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename T>
__global__ void k_coo_build(uint32_t rows, uint64_t nnz, const uint32_t *__restrict__ rp,
                            const uint32_t *__restrict__ ci, const T *__restrict__ a,
                            uint32_t e0, const uint32_t *__restrict__ toff,
                            T *__restrict__ cv, uint32_t *__restrict__ cc,
                            uint32_t *__restrict__ cr)
{
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; e < nnz; e += stride) {
    const uint32_t r = row_of(rp, rows, (uint32_t)e);
    const uint32_t k = (uint32_t)e - rp[r];
    if (k < e0) continue;
    const uint32_t o = toff[r] + (k - e0);
    cv[o] = a[e];
    cc[o] = ci[e];
    cr[o] = r;
  }
}

int coo_tail_build_device(int dtype, uint32_t rows, uint64_t nnz, const uint32_t *rp,
                          const uint32_t *ci, const void *val, uint32_t e0, void **cv,
                          uint32_t **cc, uint32_t **cr, uint64_t *count, hipStream_t s)
{
  *cv = nullptr;
  *cc = nullptr;
  *cr = nullptr;
  uint32_t *toff = nullptr;
  MFP_CHECK(dmalloc((void **)&toff, ((uint64_t)rows + 1) * 4));
  uint32_t total = 0;
  if (rows) {
    k_tail_counts<<<ceil_div_u32(rows, 256), 256, 0, s>>>(rows, rp, e0, toff);
    MFP_HIP(hipGetLastError());
  }
  int st = exclusive_scan_u32(toff, toff, rows, s, &total);
  if (st != MFP_OK) {
    dfree(toff);
    return st;
  }
  *count = total;
  const size_t vs = dtype_size(dtype);
  st = dmalloc(cv, (uint64_t)total * vs);
  if (st == MFP_OK) st = dmalloc((void **)cc, (uint64_t)total * 4);
  if (st == MFP_OK) st = dmalloc((void **)cr, (uint64_t)total * 4);
  if (st != MFP_OK) {
    dfree(toff);
    return st;
  }
  if (total) {
    if (dtype == MFP_F32)
      k_coo_build<float><<<grid_for(nnz), 256, 0, s>>>(rows, nnz, rp, ci, (const float *)val, e0,
                                                       toff, (float *)*cv, *cc, *cr);
    else
      k_coo_build<double><<<grid_for(nnz), 256, 0, s>>>(rows, nnz, rp, ci, (const double *)val,
                                                        e0, toff, (double *)*cv, *cc, *cr);
    MFP_HIP(hipGetLastError());
  }
  MFP_HIP(hipStreamSynchronize(s));
  dfree(toff);
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// SCOO slice pointers (common/scoo_matrix_class.cpp:74-100): slice s starts
// at row s*slice_size; entry [real slices] = nnz; the rest stay 0.
// ---------------------------------------------------------------------------
__global__ void k_scoo_slices(uint32_t rows, uint64_t nnz, const uint32_t *__restrict__ rp,
                              uint32_t slices_count, uint32_t ss, uint32_t *__restrict__ sp)
{
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > slices_count) return;
  const uint32_t real = rows ? (uint32_t)(((uint64_t)rows + ss - 1) / ss) : 0u;
  sp[s] = s < real ? rp[(uint64_t)s * ss] : (s == real ? (uint32_t)nnz : 0u);
}

int scoo_slices_device(uint32_t rows, uint64_t nnz, const uint32_t *rp, uint32_t slices_count,
                       uint32_t ss, uint32_t *sp, hipStream_t s)
{
  MFP_ARG(ss > 0, "slice_size must be positive");
  k_scoo_slices<<<ceil_div_u32((uint64_t)slices_count + 1, 256), 256, 0, s>>>(rows, nnz, rp,
                                                                             slices_count, ss, sp);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

// ---------------------------------------------------------------------------
// Column-sorted sliced COO.  Rows longer than a threshold T ("long rows")
// are moved to their own slices, R of them per slice, in row order; every
// other row stays in the slice of R consecutive rows that holds it.  Inside a
// slice the entries are ordered by column (stable, so equal columns keep row
// order).  Each slice is cut into work items of about equal entry counts.
// When every item's column span fits in 32 - log2(R) bits, an entry stores
// one u32 (column - item base) << log2(R) | local row; otherwise a u32 column
// and a u16 local row.
// ---------------------------------------------------------------------------
__global__ void k_long_flags(uint32_t rows, const uint32_t *__restrict__ rp, uint32_t T,
                             uint32_t *__restrict__ flag)
{
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) flag[r] = rp[r + 1] - rp[r] > T ? 1u : 0u;
}

__global__ void k_long_list(uint32_t rows, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ rank,
                            uint32_t *__restrict__ list)
{
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows && flag[r]) list[rank[r]] = (uint32_t)r;
}

__global__ void k_long_mask(uint32_t rows, const uint32_t *__restrict__ flag, uint32_t *__restrict__ mask)
{
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if ((uint64_t)w * 32 >= rows) return;
  uint32_t m = 0;
  for (uint32_t b = 0; b < 32 && (uint64_t)w * 32 + b < rows; b++) m |= flag[(uint64_t)w * 32 + b] << b;
  mask[w] = m;
}

__global__ void k_entry_slice(uint32_t rows, uint64_t nnz, const uint32_t *__restrict__ rp, uint32_t R,
                              uint32_t rbits, uint32_t S_reg, const uint32_t *__restrict__ flag,
                              const uint32_t *__restrict__ rank, uint32_t *__restrict__ slice,
                              uint32_t *__restrict__ lrow)
{
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; e < nnz; e += stride) {
    const uint32_t r = row_of(rp, rows, (uint32_t)e);
    if (flag && flag[r]) {
      const uint32_t j = rank[r];
      slice[e] = S_reg + (j >> rbits);
      lrow[e] = j & (R - 1u);
    } else {
      slice[e] = r >> rbits;
      lrow[e] = r & (R - 1u);
    }
  }
}

// sp[s] = first sorted position whose slice is >= s
// bb[sl (nb + 1) + k] = first entry of slice sl whose column is >= k W
// (entries of a slice are sorted by column), for k = 0..nb
__global__ void k_band_bounds(uint32_t S, uint32_t nb, uint64_t W, const uint32_t *__restrict__ sp,
                              const uint32_t *__restrict__ col_sorted, uint32_t *__restrict__ bb)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)S * (nb + 1)) return;
  const uint32_t sl = (uint32_t)(t / (nb + 1)), k = (uint32_t)(t % (nb + 1));
  uint32_t lo = sp[sl], hi = sp[sl + 1];
  const uint64_t c0 = (uint64_t)k * W;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)col_sorted[mid] < c0)
      lo = mid + 1;
    else
      hi = mid;
  }
  bb[t] = lo;
}

__global__ void k_slice_bounds(uint32_t S, uint64_t nnz, const uint32_t *__restrict__ slice_sorted,
                               uint32_t *__restrict__ sp)
{
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > S) return;
  uint64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (slice_sorted[mid] < s)
      lo = mid + 1;
    else
      hi = mid;
  }
  sp[s] = (uint32_t)lo;
}

// Column span of each slice: last column - first column of its sorted entries.
__global__ void k_slice_span(uint32_t S, const uint32_t *__restrict__ sp, const uint32_t *__restrict__ col_sorted,
                             uint32_t *__restrict__ span)
{
  const uint32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= S) return;
  const uint32_t lo = sp[sl], hi = sp[sl + 1];
  span[sl] = hi > lo ? col_sorted[hi - 1] - col_sorted[lo] : 0u;
}

__global__ void k_item_base(uint32_t NI, const uint4 *__restrict__ items, const uint32_t *__restrict__ col,
                            uint32_t *__restrict__ base, uint32_t *__restrict__ max_span)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NI) return;
  const uint32_t e0 = items[i].y, e1 = items[i].z;
  if (e1 > e0) {
    base[i] = col[e0];
    atomicMax(max_span, col[e1 - 1] - col[e0]);
  } else {
    base[i] = 0;
  }
}

template <typename T>
__global__ void k_cs_scatter(uint64_t nnz, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ col,
                             const uint32_t *__restrict__ lrow, const T *__restrict__ a, uint32_t NI,
                             const uint4 *__restrict__ items, const uint32_t *__restrict__ base, uint32_t rbits,
                             T *__restrict__ ev, uint32_t *__restrict__ ec)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < nnz; i += stride) {
    const uint32_t src = perm[i];
    const uint32_t c = col[i];
    // the item holding sorted position i: last item with e0 <= i (items
    // are in position order and cover every position)
    uint32_t lo = 0, hi = NI;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (items[mid].y <= i)
        lo = mid;
      else
        hi = mid;
    }
    // inside an aligned block of 256 positions that lies wholly in one
    // item, position 64 j + l goes to word l, slot j: lane l's 16-byte word
    // then holds the block's positions l, 64 + l, 128 + l and 192 + l, so
    // each of the wave's four gathers reads 64 consecutive columns
    const uint64_t b0 = i & ~(uint64_t)255;
    const bool in_block = b0 >= items[lo].y && b0 + 256 <= items[lo].z;
    const uint64_t d = in_block ? b0 + 4 * (i & 63) + ((i >> 6) & 3) : i;
    // FP64 values of such a block: slots 0 and 1 of its 64 words first, as
    // 64 consecutive 16-byte pairs, then slots 2 and 3, so that each 16-byte
    // value load of a wave reads 1 KiB contiguous
    const uint64_t dv = sizeof(T) == 8 && in_block ? b0 + (((i >> 7) & 1) << 7) + 2 * (i & 63) + ((i >> 6) & 1) : d;
    ev[dv] = a[src];
    ec[d] = ((c - base[lo]) << rbits) | lrow[src];
  }
}

template <typename T>
static int cs_scatter_t(uint64_t nnz, const uint32_t *perm, const uint32_t *col, const uint32_t *lrow,
                        const void *a, uint32_t NI, const uint4 *items, const uint32_t *base, uint32_t rbits,
                        void *ev, uint32_t *ec, hipStream_t s)
{
  k_cs_scatter<T><<<grid_for(nnz), 256, 0, s>>>(nnz, perm, col, lrow, (const T *)a, NI, items, base, rbits,
                                                (T *)ev, ec);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

int scoo_cs_build_device(int dtype, uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t *rp,
                         const uint32_t *ci, const void *val, uint32_t R, uint32_t long_threshold,
                         uint64_t item_entries, scoo_cs_layout *out, hipStream_t s,
                         bool own_partials)
{
  MFP_ARG(R >= 1 && R <= 65536 && (R & (R - 1)) == 0,
          "column-sorted SCOO slice size must be a power of two <= 65536");
  MFP_ARG(nnz < (1ull << 32), "nnz must fit in 32 bits");
  MFP_ARG(item_entries >= 1, "entries per work item must be positive");
  *out = scoo_cs_layout{};
  uint32_t rbits = 0;
  while ((1u << rbits) < R) rbits++;
  out->R = R;
  out->rbits = rbits;
  const uint32_t S_reg = rows ? (uint32_t)(((uint64_t)rows + R - 1) / R) : 0u;
  out->S_reg = S_reg;
  const size_t vs = dtype_size(dtype);
  std::vector<void *> tmp;
  auto scratch = [&](void **p, uint64_t bytes) {
    const int st = dmalloc(p, std::max<uint64_t>(bytes, 4));
    if (st == MFP_OK) tmp.push_back(*p);
    return st;
  };
  auto done = [&](int st) {
    for (void *p : tmp) dfree(p);
    if (st != MFP_OK) scoo_cs_free(out);
    return st;
  };
  int st = MFP_OK;
  // long rows: flag, rank (exclusive scan) and list
  uint32_t *flag = nullptr, *rank = nullptr;
  uint32_t NL = 0;
  if (rows && long_threshold != 0xFFFFFFFFu) {
    st = scratch((void **)&flag, (uint64_t)rows * 4);
    if (st == MFP_OK) st = scratch((void **)&rank, (uint64_t)rows * 4);
    if (st != MFP_OK) return done(st);
    k_long_flags<<<ceil_div_u32(rows, 256), 256, 0, s>>>(rows, rp, long_threshold, flag);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
    st = exclusive_scan_u32(flag, rank, rows, s, &NL);
    if (st != MFP_OK) return done(st);
  }
  out->NL = NL;
  out->S_long = (uint32_t)(((uint64_t)NL + R - 1) / R);
  const uint32_t S = S_reg + out->S_long;
  out->S = S;
  st = dmalloc((void **)&out->long_rows, std::max<uint64_t>(NL, 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&out->sp, ((uint64_t)S + 1) * 4);
  if (st == MFP_OK && NL) st = dmalloc((void **)&out->long_mask, ((uint64_t)rows + 31) / 32 * 4);
  if (st != MFP_OK) return done(st);
  if (NL) {
    k_long_list<<<ceil_div_u32(rows, 256), 256, 0, s>>>(rows, flag, rank, out->long_rows);
    k_long_mask<<<ceil_div_u32(((uint64_t)rows + 31) / 32, 256), 256, 0, s>>>(rows, flag, out->long_mask);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
  }
  // (slice, column, source entry) sorted by column, then stably by slice
  uint32_t *buf[7] = {};
  for (int i = 0; i < 7 && st == MFP_OK; i++) st = scratch((void **)&buf[i], nnz * 4);
  const uint32_t ntiles = (uint32_t)((nnz + RX_TILE - 1) / RX_TILE);
  uint32_t *hist = nullptr, *span = nullptr;
  if (st == MFP_OK) st = scratch((void **)&hist, (uint64_t)std::max<uint32_t>(ntiles, 1) * 256 * 4);
  if (st == MFP_OK) st = scratch((void **)&span, 4);
  if (st != MFP_OK) return done(st);
  uint32_t *lrow = buf[6];
  uint32_t *cur[3] = {buf[0], buf[1], buf[2]};
  uint32_t *nxt[3] = {buf[3], buf[4], buf[5]};
  if (nnz) {
    k_entry_slice<<<grid_for(nnz), 256, 0, s>>>(rows, nnz, rp, R, rbits, S_reg, NL ? flag : nullptr, rank, buf[0],
                                                 lrow);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
    if (hipMemcpyAsync(buf[1], ci, nnz * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return done(MFP_ERR_HIP);
    k_iota<<<grid_for(nnz), 256, 0, s>>>(buf[2], nnz);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
    std::vector<std::pair<int, uint32_t>> passes;
    for (uint32_t sh = 0; sh < bit_length(cols ? cols - 1 : 0); sh += 8) passes.push_back({1, sh});
    for (uint32_t sh = 0; sh < bit_length(S ? S - 1 : 0); sh += 8) passes.push_back({0, sh});
    for (auto [which, sh] : passes) {
      k_radix_hist<<<ntiles, RX_T, 0, s>>>(cur[which], nnz, sh, hist, ntiles);
      if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
      st = exclusive_scan_u32(hist, hist, (uint64_t)ntiles * 256, s, nullptr);
      if (st != MFP_OK) return done(st);
      k_radix_scatter<<<ntiles, RX_T, 0, s>>>(cur[which], cur[0], cur[1], cur[2], nxt[0], nxt[1], nxt[2], nnz, sh,
                                               hist, ntiles);
      if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
      std::swap(cur, nxt);
    }
  }
  // slice boundaries, then work items on the host
  k_slice_bounds<<<ceil_div_u32((uint64_t)S + 1, 256), 256, 0, s>>>(S, nnz, cur[0], out->sp);
  if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
  std::vector<uint32_t> hsp((size_t)S + 1);
  if (hipMemcpyAsync(hsp.data(), out->sp, hsp.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(MFP_ERR_HIP);
  // Column bands of width W = 2^(32 - rbits): an item whose columns span
  // less than W packs (column - item base) and the local row into one u32.
  // A slice whose columns span less than W is one band; any other slice is
  // cut at the multiples of W.  Each non-empty band of a slice is cut into
  // about E_band / item_entries equal pieces.
  const uint64_t W = rbits ? (1ull << (32 - rbits)) : (1ull << 32);
  const uint32_t nb = (uint32_t)std::max<uint64_t>(1, ((uint64_t)cols + W - 1) / W);
  std::vector<uint32_t> hbb((size_t)S * (nb + 1));
  std::vector<uint32_t> slice_span(S, 0u);
  if (S && nb > 1) {
    uint32_t *bb = nullptr, *sspan = nullptr;
    st = scratch((void **)&bb, hbb.size() * 4);
    if (st == MFP_OK) st = scratch((void **)&sspan, (uint64_t)S * 4);
    if (st != MFP_OK) return done(st);
    k_band_bounds<<<ceil_div_u32(hbb.size(), 256), 256, 0, s>>>(S, nb, W, out->sp, cur[1], bb);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
    k_slice_span<<<ceil_div_u32(S, 256), 256, 0, s>>>(S, out->sp, cur[1], sspan);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hbb.data(), bb, hbb.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(slice_span.data(), sspan, (size_t)S * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(MFP_ERR_HIP);
    for (uint32_t sl = 0; sl < S; sl++) {
      if (slice_span[sl] >= W) continue;
      uint32_t *b = &hbb[(size_t)sl * (nb + 1)];
      b[0] = hsp[sl]; // the whole slice in band 0, the other bands empty
      for (uint32_t k = 1; k <= nb; k++) b[k] = hsp[sl + 1];
    }
  } else {
    for (uint32_t sl = 0; sl < S; sl++) {
      hbb[(size_t)sl * 2] = hsp[sl];
      hbb[(size_t)sl * 2 + 1] = hsp[sl + 1];
    }
  }
  std::vector<uint4> items;
  std::vector<uint4> fin, finl; // finish blocks {slice, first partial slot, slots, first row in the slice}
  uint32_t slots = 0;
  std::vector<std::pair<uint32_t, uint32_t>> cuts; // [e0, e1) of the slice's items
  for (uint32_t sl = 0; sl < S; sl++) {
    cuts.clear();
    for (uint32_t k = 0; k < nb; k++) {
      const uint32_t b0 = hbb[(size_t)sl * (nb + 1) + k], b1 = hbb[(size_t)sl * (nb + 1) + k + 1];
      const uint64_t Eb = b1 - b0;
      if (Eb == 0) continue;
      const uint64_t nk = std::max<uint64_t>(1, (Eb + item_entries / 2) / item_entries);
      for (uint64_t q = 0; q < nk; q++)
        cuts.push_back({b0 + (uint32_t)(Eb * q / nk), b0 + (uint32_t)(Eb * (q + 1) / nk)});
    }
    if (cuts.empty()) cuts.push_back({hsp[sl], hsp[sl]}); // an empty slice still writes its zeros
    const uint32_t n = (uint32_t)cuts.size();
    // a slice with one item writes y; with several, each item fills a
    // partial slot and finish blocks add the slots in slot order
    const uint32_t mode = n == 1 ? CS_DIRECT : slots;
    if (n > 1 && sl < S_reg) {
      const uint32_t lim = (uint32_t)std::min<uint64_t>(R, (uint64_t)rows - (uint64_t)sl * R);
      for (uint32_t r0 = 0; r0 < lim; r0 += 4 * CS_FIN_ROWS) fin.push_back(make_uint4(sl, slots, n, r0));
    } else if (n > 1) {
      // a long-row slice can have many items: a wave per row adds its slots
      const uint32_t lim = (uint32_t)std::min<uint64_t>(R, (uint64_t)NL - (uint64_t)(sl - S_reg) * R);
      for (uint32_t r0 = 0; r0 < lim; r0 += CS_FIN_LONG) finl.push_back(make_uint4(sl, slots, n, r0));
    }
    for (uint32_t k = 0; k < n; k++)
      items.push_back(make_uint4(sl, cuts[k].first, cuts[k].second, mode < CS_DIRECT ? mode + k : mode));
    if (mode < CS_DIRECT) slots += n;
  }
  out->NFR = (uint32_t)fin.size();
  fin.insert(fin.end(), finl.begin(), finl.end());
  out->NI = (uint32_t)items.size();
  out->NF = (uint32_t)fin.size();
  out->slots = slots;
  st = dmalloc((void **)&out->items, std::max<size_t>(items.size(), 1) * sizeof(uint4));
  if (st == MFP_OK) st = dmalloc((void **)&out->fin, std::max<size_t>(fin.size(), 1) * sizeof(uint4));
  if (st == MFP_OK) st = dmalloc((void **)&out->base, std::max<size_t>(items.size(), 1) * 4);
  // partial slots hold the vector type: float for FP32 and 16-bit values;
  // FP32 / FP64 plans also serve SpMM with k columns while R k doubles fit
  // in LDS, so their slots hold kmax columns
  const size_t ps = dtype == MFP_F64 ? 8 : 4;
  out->kmax = 1;
  if (dtype == MFP_F32 || dtype == MFP_F64)
    while (out->kmax < 16 && (uint64_t)R * out->kmax * 2 * 8 <= 163840) out->kmax *= 2;
  if (st == MFP_OK && slots && own_partials) st = dmalloc(&out->partial, (uint64_t)slots * R * out->kmax * ps);
  if (st == MFP_OK) st = dmalloc(&out->ev, std::max<uint64_t>(nnz, 1) * vs);
  if (st == MFP_OK) st = dmalloc((void **)&out->ec, std::max<uint64_t>(nnz, 1) * 4);
  if (st != MFP_OK) return done(st);
  if (!items.empty() &&
      hipMemcpyAsync(out->items, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice, s) != hipSuccess)
    return done(MFP_ERR_HIP);
  if (!fin.empty() &&
      hipMemcpyAsync(out->fin, fin.data(), fin.size() * sizeof(uint4), hipMemcpyHostToDevice, s) != hipSuccess)
    return done(MFP_ERR_HIP);
  st = fill_u32(span, 0, 1, s);
  if (st != MFP_OK) return done(st);
  if (out->NI) {
    k_item_base<<<ceil_div_u32(out->NI, 256), 256, 0, s>>>(out->NI, out->items, cur[1], out->base, span);
    if (hipGetLastError() != hipSuccess) return done(MFP_ERR_HIP);
  }
  uint32_t hspan = 0;
  if (hipMemcpyAsync(&hspan, span, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return done(MFP_ERR_HIP);
  // every item lies inside one column band, so its offsets fit beside the row
  if (rbits && (uint64_t)hspan >= (1ull << (32 - rbits))) {
    set_error("column-sorted SCOO: an item spans more than its column band");
    return done(MFP_ERR_STATE);
  }
  if (nnz) {
    st = (dtype == MFP_F32 ? cs_scatter_t<float> : dtype == MFP_F64 ? cs_scatter_t<double> : cs_scatter_t<uint16_t>)(
        nnz, cur[2], cur[1], lrow, val, out->NI, out->items, out->base, rbits, out->ev, out->ec, s);
    if (st != MFP_OK) return done(st);
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("column-sorted SCOO build failed");
    return done(MFP_ERR_HIP);
  }
  // Launch order: piece k of n covers about columns [k/n, (k+1)/n) of its
  // slice, so ordering the items by the centre of that band (then by slice)
  // gives each XCD's contiguous range of items one band of x for its L2.
  // The scatter above needed the items in entry order; the kernel does not.
  if (out->NI > 1) {
    std::vector<uint32_t> hbase(out->NI);
    if (hipMemcpyAsync(hbase.data(), out->base, (size_t)out->NI * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(MFP_ERR_HIP);
    std::vector<uint32_t> pieces(out->NI), order(out->NI);
    for (uint32_t i = 0; i < out->NI;) { // items of one slice are adjacent
      uint32_t j = i;
      while (j < out->NI && items[j].x == items[i].x) j++;
      for (uint32_t k = i; k < j; k++) pieces[k] = j - i;
      i = j;
    }
    std::vector<uint32_t> first(out->NI);
    for (uint32_t i = 0; i < out->NI; i++) first[i] = (i && items[i - 1].x == items[i].x) ? first[i - 1] : i;
    for (uint32_t i = 0; i < out->NI; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      // band centre (2k + 1) / (2n), compared exactly as a cross product
      const uint64_t ka = a - first[a], kb = b - first[b];
      return (2 * ka + 1) * pieces[b] < (2 * kb + 1) * pieces[a];
    });
    // The kernel maps workgroups to items with xcd_block: XCD x runs one
    // contiguous range of this order, first those of its items that fill its
    // CUs, then the rest as CUs free up.  Items of ordinary slices take about
    // twice as long as long-row items, so each range gets an even share of
    // them, placed first, then long-row items (both in band-centre order, so
    // each XCD still works on one band of x).  Left to the band-centre order
    // alone, one XCD can receive more ordinary items than it has CUs and run
    // two of them back to back while others idle.
    if (out->NI >= 32) {
      std::vector<uint32_t> ro, lo;
      for (uint32_t i : order) (items[i].x < S_reg ? ro : lo).push_back(i);
      const uint32_t nr = (uint32_t)ro.size(), nl = (uint32_t)lo.size();
      const uint32_t q = out->NI / 8, rem = out->NI % 8;
      uint32_t ri = 0, li = 0, p0 = 0;
      for (uint32_t x = 0; x < 8; x++) {
        const uint32_t len = q + (x < rem ? 1u : 0u);
        const uint32_t share = (uint32_t)((uint64_t)nr * (x + 1) / 8 - (uint64_t)nr * x / 8);
        uint32_t rx = std::min({share, len, nr - ri});
        const uint32_t lx = std::min(len - rx, nl - li);
        rx += std::min(len - rx - lx, nr - ri - rx); // no long-row items left: more ordinary ones
        for (uint32_t k = 0; k < rx; k++) order[p0++] = ro[ri++];
        for (uint32_t k = 0; k < lx; k++) order[p0++] = lo[li++];
      }
      // every range is filled exactly: the ranges' lengths add up to NI
      if (p0 != out->NI || ri != nr || li != nl) {
        set_error("column-sorted SCOO: work item ordering failed");
        return done(MFP_ERR_STATE);
      }
    }
    std::vector<uint4> it2(out->NI);
    std::vector<uint32_t> b2(out->NI);
    for (uint32_t i = 0; i < out->NI; i++) {
      it2[i] = items[order[i]];
      b2[i] = hbase[order[i]];
    }
    if (hipMemcpyAsync(out->items, it2.data(), it2.size() * sizeof(uint4), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(out->base, b2.data(), b2.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(MFP_ERR_HIP);
  }
  out->device_bytes = nnz * (vs + 4) + ((uint64_t)S + 1) * 4 +
                      (uint64_t)out->NI * 20 + (uint64_t)out->NF * 16 + (uint64_t)NL * 4 +
                      (NL ? ((uint64_t)rows + 31) / 32 * 4 : 0) +
                      (own_partials ? (uint64_t)slots * R * out->kmax * ps : 0);
  return done(MFP_OK);
}

void scoo_cs_free(scoo_cs_layout *l)
{
  dfree(l->sp);
  dfree(l->long_rows);
  dfree(l->long_mask);
  dfree(l->items);
  dfree(l->fin);
  dfree(l->base);
  dfree(l->partial);
  dfree(l->ev);
  dfree(l->ec);
  *l = scoo_cs_layout{};
}

// ---------------------------------------------------------------------------
// Row blocks (fill_row_blocks, gpu/csr_adaptive_spmv.cu:180-240) in parallel.
// The greedy partition is a chain 0 -> close(0) -> ...: close(s) is the row
// where a block started at s closes (or NONE when the loop ends first), and it
// depends on s only.  Chain nodes are found by pointer jumping: mark[0] = 1;
// round k marks J_k(s) for every marked s with J_k = close^(2^k).  Marked
// nodes in [0, rows] in order, followed by rows, are the reference's output.
// ---------------------------------------------------------------------------
__global__ void k_close(uint32_t rows, const uint32_t *__restrict__ rp, uint32_t W,
                        uint32_t *__restrict__ J)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > (uint64_t)rows + 1) return;
  const uint32_t NONE = rows + 1;
  if (s >= rows) {
    J[s] = NONE;
    return;
  }
  uint32_t sum = 0;
  uint32_t res = NONE;
  for (uint64_t i = s + 1; i <= rows; i++) {
    sum += rp[i] - rp[i - 1];
    if (sum == W) {
      res = (uint32_t)i;
      break;
    }
    if (sum > W) {
      res = (i - s > 1) ? (uint32_t)(i - 1) : (uint32_t)i;
      break;
    }
    if (i - s > W) {
      res = (uint32_t)i;
      break;
    }
  }
  J[s] = res;
}

__global__ void k_mark(uint64_t n, const uint32_t *__restrict__ J, uint32_t *__restrict__ mark)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (mark[s]) mark[J[s]] = 1u;
}

__global__ void k_jump(uint64_t n, const uint32_t *__restrict__ J, uint32_t *__restrict__ J2)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  J2[s] = J[J[s]];
}

__global__ void k_compact(uint32_t rows, const uint32_t *__restrict__ mark,
                          const uint32_t *__restrict__ pos, uint32_t *__restrict__ out)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > rows) return;
  if (mark[s]) out[pos[s]] = (uint32_t)s;
}

__global__ void k_store1(uint32_t *p, uint32_t v) { *p = v; }

// Device allocations freed on every return path unless released.
struct device_temps {
  std::vector<void *> held;
  int alloc(void **p, size_t bytes)
  {
    const int st = dmalloc(p, bytes);
    if (st == MFP_OK) held.push_back(*p);
    return st;
  }
  void release(void *p)
  {
    for (void *&q : held)
      if (q == p) q = nullptr;
  }
  ~device_temps()
  {
    for (void *q : held) dfree(q);
  }
};

int row_blocks_device(uint32_t rows, const uint32_t *rp, uint32_t W, uint32_t **blocks_out,
                      uint32_t *count_out, hipStream_t s)
{
  MFP_ARG(W > 0, "nnz_per_wg must be positive");
  const uint64_t n = (uint64_t)rows + 2; // nodes 0..rows, sentinel rows+1
  device_temps tmp;
  uint32_t *J = nullptr, *J2 = nullptr, *mark = nullptr;
  MFP_CHECK(tmp.alloc((void **)&J, n * 4));
  MFP_CHECK(tmp.alloc((void **)&J2, n * 4));
  MFP_CHECK(tmp.alloc((void **)&mark, n * 4));
  const uint32_t g = ceil_div_u32(n, 256);
  k_close<<<g, 256, 0, s>>>(rows, rp, W, J);
  MFP_CHECK(fill_u32(mark, 0, n, s));
  MFP_CHECK(fill_u32(mark, 1, 1, s));
  for (uint64_t span = 1; span < 2 * n; span <<= 1) {
    k_mark<<<g, 256, 0, s>>>(n, J, mark);
    k_jump<<<g, 256, 0, s>>>(n, J, J2);
    std::swap(J, J2);
  }
  MFP_HIP(hipGetLastError());
  // count = marked nodes in [0, rows]; positions by exclusive scan
  uint32_t count = 0;
  MFP_CHECK(exclusive_scan_u32(mark, J2, (uint64_t)rows + 1, s, &count));
  uint32_t *out = nullptr;
  MFP_CHECK(tmp.alloc((void **)&out, ((uint64_t)count + 1) * 4));
  k_compact<<<ceil_div_u32((uint64_t)rows + 1, 256), 256, 0, s>>>(rows, mark, J2, out);
  k_store1<<<1, 1, 0, s>>>(out + count, rows);
  MFP_HIP(hipGetLastError());
  MFP_HIP(hipStreamSynchronize(s));
  tmp.release(out);
  *blocks_out = out;
  *count_out = count;
  return MFP_OK;
}

} // namespace mfp
