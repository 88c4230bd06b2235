// Multiple right-hand sides (SpMM) and batched matrices.
//
// SpMM: Y = A X with X (cols x k) and Y (rows x k) row-major.  G lanes per
// row (CSR-vector order, per column of X: lane-strided fma, halving tree),
// each lane reading the k entries of an X row with 16-byte loads.
//
// Batch: many matrices of uneven sizes become one block-diagonal CSR in HBM
// (row pointers offset by the preceding nnz, columns by the preceding
// columns); X / Y are the matrices' vectors stacked in the same order.  One
// launch (CSR-Adaptive for k = 1, or the column-sorted SCOO's two kernels
// when asked; SpMM otherwise: CSR-vector with one lane count for all rows or
// rows binned by length with matched lane counts, or the column-sorted SCOO
// SpMM) covers the whole batch, and the launch sequence can be captured once
// in a HIP graph and replayed.
#include "kernels.h"
#include "mfp_internal.h"
#include "wave64.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace mfp {

// Binned SpMM: six bins, G = 2 << b lanes for rows of at most 2 << b
// entries (one entry per lane), G = 64 above 32.
constexpr uint32_t SPMM_BINS = 6;
inline uint32_t spmm_bin_of(uint32_t len)
{
  uint32_t b = 0;
  while (b + 1 < SPMM_BINS && (2u << b) < len) b++;
  return b;
}

// One row of Y in the CSR-vector order with G lanes (row >= rows: the lanes
// only take part in the reduction).  Every lane of the wave must call it.
template <typename T, int G, int KV>
__device__ __forceinline__ void spmm_row(uint32_t row, uint32_t rows, uint32_t lane, const uint32_t *__restrict__ rp,
                                         const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                         const T *__restrict__ X, T *__restrict__ Y)
{
  T acc[KV];
#pragma unroll
  for (int j = 0; j < KV; j++) acc[j] = 0;
  if (row < rows) {
    const uint32_t b = rp[row], e = rp[row + 1];
    uint32_t i = b + lane;
    // four entries per lane in flight (matrix and X rows), then fma in entry order
    for (; i + 3u * G < e; i += 4u * G) {
      T av[4];
      const T *xr[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        av[q] = ldg_stream(a + i + q * G);
        xr[q] = X + (uint64_t)ldg_stream(ci + i + q * G) * KV;
      }
      T xv[4][KV];
#pragma unroll
      for (int q = 0; q < 4; q++) load_row<T, KV>(xr[q], xv[q]);
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int j = 0; j < KV; j++) acc[j] = fma_t(av[q], xv[q][j], acc[j]);
    }
    for (; i < e; i += G) {
      const T av = ldg_stream(a + i);
      T xv[KV];
      load_row<T, KV>(X + (uint64_t)ldg_stream(ci + i) * KV, xv);
#pragma unroll
      for (int j = 0; j < KV; j++) acc[j] = fma_t(av, xv[j], acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < KV; j++) acc[j] = group_reduce<G>(acc[j]);
  if (lane == 0 && row < rows) {
#pragma unroll
    for (int j = 0; j < KV; j++) Y[(uint64_t)row * KV + j] = acc[j];
  }
}

template <typename T, int G, int KV, bool PERSISTENT>
__global__ __launch_bounds__(256) void k_spmm_vector(uint32_t rows, const uint32_t *__restrict__ rp,
                                                     const uint32_t *__restrict__ ci,
                                                     const T *__restrict__ a, const T *__restrict__ X,
                                                     T *__restrict__ Y)
{
  const uint32_t lane = threadIdx.x % G;
  const uint32_t rows_per_block = 256u / G;
  const uint32_t nblk = (rows + rows_per_block - 1) / rows_per_block;
  for (uint32_t blk = blockIdx.x; blk < nblk; blk += (PERSISTENT ? gridDim.x : nblk)) {
    spmm_row<T, G, KV>(blk * rows_per_block + threadIdx.x / G, rows, lane, rp, ci, a, X, Y);
    if (!PERSISTENT) break;
  }
}

// Rows binned by length: bin b (G = 2 << b lanes per row) holds the rows
// list[off[b] .. off[b + 1]); a tile is 256 / G of them, one workgroup
// (tile_off[b] = first tile of bin b; the grid-stride loop also serves a
// smaller grid), so each row is summed in the CSR-vector order with its
// bin's G.
// Binning fills the lanes of short rows but gives up the matrix's streaming
// order (a wave's rows are no longer adjacent): it pays when the X rows are
// wide (k = 16 in the 512-matrix batch), not at k = 4.
struct spmm_bins_args {
  uint32_t off[SPMM_BINS + 1];
  uint32_t tile_off[SPMM_BINS + 1];
};

template <typename T, int G, int KV>
__device__ __forceinline__ void spmm_bin_tile(uint32_t t, uint32_t n, const uint32_t *__restrict__ list,
                                              uint32_t rows, const uint32_t *__restrict__ rp,
                                              const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                              const T *__restrict__ X, T *__restrict__ Y)
{
  const uint32_t idx = t * (256u / G) + threadIdx.x / G;
  const uint32_t row = idx < n ? list[idx] : rows; // rows: no row, the lanes still reduce
  spmm_row<T, G, KV>(row, rows, threadIdx.x % G, rp, ci, a, X, Y);
}

template <typename T, int KV>
__global__ __launch_bounds__(256) void k_spmm_binned(spmm_bins_args bins, const uint32_t *__restrict__ list,
                                                     uint32_t rows, const uint32_t *__restrict__ rp,
                                                     const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                                     const T *__restrict__ X, T *__restrict__ Y)
{
  const uint32_t total = bins.tile_off[SPMM_BINS];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    uint32_t b = 0;
    while (b + 1 < SPMM_BINS && t >= bins.tile_off[b + 1]) b++;
    const uint32_t tt = t - bins.tile_off[b], n = bins.off[b + 1] - bins.off[b];
    const uint32_t *l = list + bins.off[b];
    switch (b) {
    case 0: spmm_bin_tile<T, 2, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    case 1: spmm_bin_tile<T, 4, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    case 2: spmm_bin_tile<T, 8, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    case 3: spmm_bin_tile<T, 16, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    case 4: spmm_bin_tile<T, 32, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    default: spmm_bin_tile<T, 64, KV>(tt, n, l, rows, rp, ci, a, X, Y); break;
    }
  }
}

template <typename T, int KV, bool P>
static int spmm_g(uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci, const T *a,
                  const T *X, T *Y, uint32_t grid, hipStream_t s)
{
  switch (lanes) {
  case 2: k_spmm_vector<T, 2, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  case 4: k_spmm_vector<T, 4, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  case 8: k_spmm_vector<T, 8, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  case 16: k_spmm_vector<T, 16, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  case 32: k_spmm_vector<T, 32, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  case 64: k_spmm_vector<T, 64, KV, P><<<grid, 256, 0, s>>>(rows, rp, ci, a, X, Y); break;
  default: set_error("lanes must be a power of two in [2, 64]"); return MFP_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

template <typename T, bool P>
static int spmm_k(uint32_t k, uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                  const T *a, const T *X, T *Y, uint32_t grid, hipStream_t s)
{
  switch (k) {
  case 1: return spmm_g<T, 1, P>(lanes, rows, rp, ci, a, X, Y, grid, s);
  case 2: return spmm_g<T, 2, P>(lanes, rows, rp, ci, a, X, Y, grid, s);
  case 4: return spmm_g<T, 4, P>(lanes, rows, rp, ci, a, X, Y, grid, s);
  case 8: return spmm_g<T, 8, P>(lanes, rows, rp, ci, a, X, Y, grid, s);
  case 16: return spmm_g<T, 16, P>(lanes, rows, rp, ci, a, X, Y, grid, s);
  default: set_error("k must be 1, 2, 4, 8 or 16"); return MFP_ERR_ARG;
  }
}

int launch_spmm_vector(int dtype, uint32_t k, uint32_t lanes, uint32_t rows, const uint32_t *rp,
                       const uint32_t *ci, const void *a, const void *X, void *Y, bool persistent,
                       hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  if ((k * dtype_size(dtype)) % 16 == 0 && ((uintptr_t)X % 16) != 0) {
    set_error("X must be 16-byte aligned when its rows are a multiple of 16 bytes");
    return MFP_ERR_ARG;
  }
  const uint32_t nblk = ceil_div_u32((uint64_t)rows * lanes, 256);
  uint32_t grid = nblk;
  if (persistent) grid = std::min<uint32_t>(nblk, (uint32_t)std::max(1, props().cu_count) * 8u);
  if (dtype == MFP_F32)
    return persistent ? spmm_k<float, true>(k, lanes, rows, rp, ci, (const float *)a, (const float *)X, (float *)Y, grid, s)
                      : spmm_k<float, false>(k, lanes, rows, rp, ci, (const float *)a, (const float *)X, (float *)Y, grid, s);
  return persistent ? spmm_k<double, true>(k, lanes, rows, rp, ci, (const double *)a, (const double *)X, (double *)Y, grid, s)
                    : spmm_k<double, false>(k, lanes, rows, rp, ci, (const double *)a, (const double *)X, (double *)Y, grid, s);
}

template <typename T>
static int spmm_binned_t(uint32_t k, const spmm_bins_args &bins, const uint32_t *list, uint32_t rows,
                         const uint32_t *rp, const uint32_t *ci, const T *a, const T *X, T *Y, hipStream_t s)
{
  const uint32_t tiles = bins.tile_off[SPMM_BINS];
  if (tiles == 0) return MFP_OK;
  const uint32_t grid = tiles; // one tile per workgroup (the loop below then runs once)
  switch (k) {
  case 1: k_spmm_binned<T, 1><<<grid, 256, 0, s>>>(bins, list, rows, rp, ci, a, X, Y); break;
  case 2: k_spmm_binned<T, 2><<<grid, 256, 0, s>>>(bins, list, rows, rp, ci, a, X, Y); break;
  case 4: k_spmm_binned<T, 4><<<grid, 256, 0, s>>>(bins, list, rows, rp, ci, a, X, Y); break;
  case 8: k_spmm_binned<T, 8><<<grid, 256, 0, s>>>(bins, list, rows, rp, ci, a, X, Y); break;
  case 16: k_spmm_binned<T, 16><<<grid, 256, 0, s>>>(bins, list, rows, rp, ci, a, X, Y); break;
  default: set_error("k must be 1, 2, 4, 8 or 16"); return MFP_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

static int launch_spmm_binned(int dtype, uint32_t k, const spmm_bins_args &bins, const uint32_t *list,
                              uint32_t rows, const uint32_t *rp, const uint32_t *ci, const void *a,
                              const void *X, void *Y, hipStream_t s)
{
  if ((k * dtype_size(dtype)) % 16 == 0 && ((uintptr_t)X % 16) != 0) {
    set_error("X must be 16-byte aligned when its rows are a multiple of 16 bytes");
    return MFP_ERR_ARG;
  }
  if (dtype == MFP_F32)
    return spmm_binned_t<float>(k, bins, list, rows, rp, ci, (const float *)a, (const float *)X, (float *)Y, s);
  return spmm_binned_t<double>(k, bins, list, rows, rp, ci, (const double *)a, (const double *)X, (double *)Y, s);
}

// Row lists per bin (row order kept inside a bin) from the row pointers.
static int spmm_bins_build(uint32_t rows, const uint32_t *rp_dev, spmm_bins_args *bins, uint32_t **list_out,
                           hipStream_t s)
{
  std::vector<uint32_t> rp((size_t)rows + 1);
  MFP_HIP(hipMemcpyAsync(rp.data(), rp_dev, rp.size() * 4, hipMemcpyDeviceToHost, s));
  MFP_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> cnt(SPMM_BINS, 0);
  for (uint32_t r = 0; r < rows; r++) cnt[spmm_bin_of(rp[r + 1] - rp[r])]++;
  bins->off[0] = 0;
  bins->tile_off[0] = 0;
  for (uint32_t b = 0; b < SPMM_BINS; b++) {
    bins->off[b + 1] = bins->off[b] + cnt[b];
    const uint32_t per_tile = 256u / (2u << b);
    bins->tile_off[b + 1] = bins->tile_off[b] + (cnt[b] + per_tile - 1) / per_tile;
  }
  std::vector<uint32_t> list((size_t)std::max<uint32_t>(rows, 1)), pos(bins->off, bins->off + SPMM_BINS);
  for (uint32_t r = 0; r < rows; r++) list[pos[spmm_bin_of(rp[r + 1] - rp[r])]++] = r;
  MFP_CHECK(dmalloc((void **)list_out, list.size() * 4));
  MFP_HIP(hipMemcpyAsync(*list_out, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
  MFP_HIP(hipStreamSynchronize(s));
  return MFP_OK;
}

// ---- CSR-stream SpMM ----------------------------------------------------------------
// A workgroup of STREAM_T lanes takes one row block of fill_row_blocks with
// STREAM_W entries (whole rows: at most STREAM_W entries and STREAM_W + 1
// rows, or one longer row alone).  Each block is described by a pair
// (first row, first entry), so the block's bounds arrive in one load.  The
// block's columns, values and row offsets are staged in LDS, all loads
// issued before the first wait; each entry's row is found by writing the
// row number at each row's first entry and taking a running maximum over
// the entries (a wave scan plus the wave totals).  Then groups of LPE lanes
// (the 16-byte pieces of an X row: KV s / 16 of them) take EPG consecutive
// entries each, read their columns from LDS as whole vectors, issue all
// their X-row gathers together, and sum the row segments in registers in
// entry order with fma.  A row inside one group is written by that group; a
// row that crosses groups is finished by the group holding its last entry:
// its own segment, then the segments of the earlier groups from the nearest
// back, added through LDS.  A longer row goes through LDS in chunks of
// STREAM_W entries, group g summing its EPG entries of each chunk in turn,
// then a halving tree over the groups.  Every lane is
// busy whatever the row lengths, and a group's gathers do not wait on one
// another, where CSR-vector leaves lanes idle on short rows and each row is
// one chain of dependent loads.  Blocks are remapped so that each XCD works
// on one contiguous range of them.
constexpr uint32_t STREAM_T = 256; // STREAM_W (kernels.h): entries per row block
static_assert(STREAM_W == 2 * STREAM_T, "the row scan gives each lane two entries");

typedef float sv_f4 __attribute__((ext_vector_type(4)));
typedef double sv_d2 __attribute__((ext_vector_type(2)));
template <typename T> struct vec16b;
template <> struct vec16b<float> {
  using V = sv_f4;
  static constexpr uint32_t N = 4;
};
template <> struct vec16b<double> {
  using V = sv_d2;
  static constexpr uint32_t N = 2;
};

// CNT consecutive LDS elements starting at a multiple of CNT, read as 16- or
// 8-byte vectors.
template <uint32_t CNT, typename E>
__device__ __forceinline__ void lds_run(const E *p, E (&out)[CNT])
{
  constexpr uint32_t B = CNT * sizeof(E);
  if constexpr (B % 16 == 0) {
#pragma unroll
    for (uint32_t i = 0; i < B / 16; i++) {
      const uint4 w = reinterpret_cast<const uint4 *>(p)[i];
      __builtin_memcpy(&out[i * (16 / sizeof(E))], &w, 16);
    }
  } else if constexpr (B % 8 == 0) {
#pragma unroll
    for (uint32_t i = 0; i < B / 8; i++) {
      const uint2 w = reinterpret_cast<const uint2 *>(p)[i];
      __builtin_memcpy(&out[i * (8 / sizeof(E))], &w, 8);
    }
  } else {
#pragma unroll
    for (uint32_t i = 0; i < CNT; i++) out[i] = p[i];
  }
}

// Piece l (16 bytes) of X row c; OFF32: X is below 4 GiB, so the address is
// X plus a 32-bit offset.
template <typename V, uint32_t LPE, bool OFF32, typename T>
__device__ __forceinline__ V stream_gather(const T *__restrict__ X, uint32_t c, uint32_t l)
{
  if constexpr (OFF32)
    return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(X) + (c * (LPE * 16u) + l * 16u));
  else
    return reinterpret_cast<const V *>(X)[(uint64_t)c * LPE + l];
}

// Workgroups per CU that the register budget is set for: 8 (the wave limit)
// for every f32 form and the narrow f64 ones; wider f64 X rows hold more
// registers per gather.
template <typename T, int KV>
constexpr int stream_min_wg()
{
  return sizeof(T) == 4 || KV <= 4 ? 8 : (KV == 8 ? 5 : 4);
}

template <typename T, int KV, bool OFF32>
__global__ __launch_bounds__(STREAM_T, (stream_min_wg<T, KV>())) void k_spmm_stream(
    const uint2 *__restrict__ meta, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
    const T *__restrict__ a, const T *__restrict__ X, T *__restrict__ Y)
{
  using V = typename vec16b<T>::V;
  constexpr uint32_t N = vec16b<T>::N, LPE = KV / N, NG = STREAM_T / LPE, EPG = STREAM_W / NG;
  static_assert(KV % N == 0 && LPE >= 1 && STREAM_W % NG == 0, "X rows must be whole 16-byte pieces");
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  __shared__ __align__(16) uint32_t s_col[STREAM_W];
  __shared__ __align__(16) T s_val[STREAM_W];
  __shared__ __align__(16) uint32_t s_row[STREAM_W];
  __shared__ uint32_t s_rp[STREAM_W + 2];
  __shared__ V s_seg[STREAM_T];
  __shared__ uint32_t s_wmax[STREAM_T / 64];
  const uint32_t bi = xcd_block(blockIdx.x, gridDim.x);
  const uint2 m0 = meta[bi], m1 = meta[bi + 1];
  const uint32_t r0 = m0.x, r1 = m1.x, e0 = m0.y, e1 = m1.y;
  const uint32_t tid = threadIdx.x, g = tid / LPE, l = tid % LPE, lane = tid & 63, wv = tid >> 6;
  const uint32_t n = e1 - e0, nr = r1 - r0;
  V *__restrict__ Yv = reinterpret_cast<V *>(Y);
  const V zero = (V)0;
  if (n > STREAM_W) {
    // one row (the partition puts a longer row alone), in chunks of
    // STREAM_W entries staged in LDS with the next chunk's loads in flight
    // during the current chunk's gathers.  Group g sums entries
    // g EPG .. g EPG + EPG - 1 of every chunk, chunk after chunk, with fma;
    // then a halving tree over the groups.
    V acc = zero;
    const uint32_t i1 = tid + STREAM_T, base = g * EPG;
    uint32_t c0 = 0, c1 = 0;
    T v0 = 0, v1 = 0;
    if (tid < n) {
      c0 = ldg_stream(ci + e0 + tid);
      v0 = ldg_stream(a + e0 + tid);
    }
    if (i1 < n) {
      c1 = ldg_stream(ci + e0 + i1);
      v1 = ldg_stream(a + e0 + i1);
    }
    for (uint32_t cb = 0; cb < n; cb += STREAM_W) {
      s_col[tid] = c0;
      s_col[i1] = c1;
      s_val[tid] = v0;
      s_val[i1] = v1;
      __syncthreads();
      const uint32_t rem = n - cb; // entries from this chunk's start to the row's end
      if (rem > STREAM_W) {
        const uint32_t nx = e0 + cb + STREAM_W, left = rem - STREAM_W;
        if (tid < left) {
          c0 = ldg_stream(ci + nx + tid);
          v0 = ldg_stream(a + nx + tid);
        }
        if (i1 < left) {
          c1 = ldg_stream(ci + nx + i1);
          v1 = ldg_stream(a + nx + i1);
        }
      }
      V xv[EPG];
      {
        uint32_t cc[EPG];
        lds_run<EPG>(&s_col[base], cc);
#pragma unroll
        for (uint32_t j = 0; j < EPG; j++) xv[j] = stream_gather<V, LPE, OFF32>(X, base + j < rem ? cc[j] : 0u, l);
      }
      T vv[EPG];
      lds_run<EPG>(&s_val[base], vv);
#pragma unroll
      for (uint32_t j = 0; j < EPG; j++)
        if (base + j < rem) {
#pragma unroll
          for (uint32_t i = 0; i < N; i++) acc[i] = fma_t(vv[j], xv[j][i], acc[i]);
        }
      __syncthreads();
    }
    s_seg[tid] = acc;
    __syncthreads();
    for (uint32_t h = NG / 2; h >= 1; h >>= 1) {
      if (g < h) s_seg[tid] = s_seg[tid] + s_seg[tid + h * LPE];
      __syncthreads();
    }
    if (g == 0) Yv[(uint64_t)r0 * LPE + l] = s_seg[l];
    return;
  }
  {
    // staging: entries tid and tid + STREAM_T, row offsets tid, + STREAM_T and + 2 STREAM_T (nr <= STREAM_W + 1)
    const uint32_t i1 = tid + STREAM_T, i2 = tid + 2 * STREAM_T;
    uint32_t c0 = 0, c1 = 0, p0 = 0, p1 = 0, p2 = 0;
    T v0 = 0, v1 = 0;
    if (tid < n) {
      c0 = ldg_stream(ci + e0 + tid);
      v0 = ldg_stream(a + e0 + tid);
    }
    if (i1 < n) {
      c1 = ldg_stream(ci + e0 + i1);
      v1 = ldg_stream(a + e0 + i1);
    }
    if (tid <= nr) p0 = rp[r0 + tid];
    if (i1 <= nr) p1 = rp[r0 + i1];
    if (i2 <= nr) p2 = rp[r0 + i2];
    if (tid <= nr) s_rp[tid] = p0 - e0;
    if (i1 <= nr) s_rp[i1] = p1 - e0;
    if (i2 <= nr) s_rp[i2] = p2 - e0;
    s_col[tid] = c0;
    s_col[i1] = c1;
    s_val[tid] = v0;
    s_val[i1] = v1;
    s_row[tid] = 0;
    s_row[i1] = 0;
  }
  __syncthreads();
  for (uint32_t r = tid; r < nr; r += STREAM_T) {
    const uint32_t b = s_rp[r], z = s_rp[r + 1];
    if (b == z) {
      for (uint32_t k = 0; k < LPE; k++) Yv[(uint64_t)(r0 + r) * LPE + k] = zero; // an empty row
    } else {
      s_row[b] = r; // the row's first entry
    }
  }
  __syncthreads();
  {
    // entry -> row: running maximum over the entries; lane tid holds 2 tid and 2 tid + 1
    const uint2 p = *reinterpret_cast<const uint2 *>(&s_row[2 * tid]);
    const uint32_t hi = max(p.x, p.y);
    uint32_t x = hi;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x = max(x, y);
    }
    if (lane == 63) s_wmax[wv] = x;
    __syncthreads();
    uint32_t ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0;
    for (uint32_t i = 0; i < wv; i++) ex = max(ex, s_wmax[i]);
    *reinterpret_cast<uint2 *>(&s_row[2 * tid]) = make_uint2(max(ex, p.x), max(ex, hi));
  }
  __syncthreads();
  const uint32_t base = g * EPG;
  V xv[EPG];
  {
    uint32_t cc[EPG];
    lds_run<EPG>(&s_col[base], cc);
#pragma unroll
    for (uint32_t j = 0; j < EPG; j++) // past the block: X row 0, never summed into a row
      xv[j] = stream_gather<V, LPE, OFF32>(X, base + j < n ? cc[j] : 0u, l);
  }
  uint32_t rr[EPG];
  T v[EPG];
  lds_run<EPG>(&s_row[base], rr);
  lds_run<EPG>(&s_val[base], v);
  V acc = zero, head = zero;
  uint32_t head_row = NONE, cur = base < n ? rr[0] : NONE;
#pragma unroll
  for (uint32_t j = 0; j < EPG; j++) {
    const bool ok = base + j < n;
    const uint32_t rj = ok ? rr[j] : NONE;
    const T vj = ok ? v[j] : (T)0;
    if (rj != cur) {
      // row cur ends inside this group
      if (s_rp[cur] < base) {
        head = acc;
        head_row = cur;
      } else {
        Yv[(uint64_t)(r0 + cur) * LPE + l] = acc;
      }
      acc = zero;
      cur = rj;
    }
#pragma unroll
    for (uint32_t i = 0; i < N; i++) acc[i] = fma_t(vj, xv[j][i], acc[i]);
  }
  if (cur != NONE) {
    if (s_rp[cur + 1] > base + EPG) {
      s_seg[tid] = acc; // row cur goes on in the next group
    } else if (s_rp[cur] < base) {
      head = acc;
      head_row = cur;
    } else {
      Yv[(uint64_t)(r0 + cur) * LPE + l] = acc;
    }
  }
  __syncthreads();
  if (head_row != NONE) {
    const uint32_t ga = s_rp[head_row] / EPG;
    V sum = head;
    for (uint32_t q = g; q-- > ga;) sum = sum + s_seg[q * LPE + l];
    Yv[(uint64_t)(r0 + head_row) * LPE + l] = sum;
  }
}

template <typename T, int KV>
static void spmm_stream_launch(uint32_t nblocks, bool off32, const uint2 *meta, const uint32_t *rp,
                               const uint32_t *ci, const T *a, const T *X, T *Y, hipStream_t s)
{
  if (off32)
    k_spmm_stream<T, KV, true><<<nblocks, STREAM_T, 0, s>>>(meta, rp, ci, a, X, Y);
  else
    k_spmm_stream<T, KV, false><<<nblocks, STREAM_T, 0, s>>>(meta, rp, ci, a, X, Y);
}

template <typename T>
static int spmm_stream_t(uint32_t k, uint32_t cols, uint32_t nblocks, const uint2 *meta, const uint32_t *rp,
                         const uint32_t *ci, const T *a, const T *X, T *Y, hipStream_t s)
{
  if (nblocks == 0) return MFP_OK;
  const bool off32 = (uint64_t)cols * k * sizeof(T) <= 0xFFFFFFFFull;
  switch (k) {
  case 2:
    if constexpr (sizeof(T) == 8) {
      spmm_stream_launch<T, 2>(nblocks, off32, meta, rp, ci, a, X, Y, s);
      break;
    }
    [[fallthrough]];
  case 1: set_error("the CSR-stream SpMM needs k values to fill whole 16-byte pieces (k >= 4 for f32, >= 2 for f64)");
    return MFP_ERR_ARG;
  case 4: spmm_stream_launch<T, 4>(nblocks, off32, meta, rp, ci, a, X, Y, s); break;
  case 8: spmm_stream_launch<T, 8>(nblocks, off32, meta, rp, ci, a, X, Y, s); break;
  case 16: spmm_stream_launch<T, 16>(nblocks, off32, meta, rp, ci, a, X, Y, s); break;
  default: set_error("k must be 1, 2, 4, 8 or 16"); return MFP_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

int launch_spmm_stream(int dtype, uint32_t k, uint32_t cols, uint32_t nblocks, const uint32_t *meta,
                       const uint32_t *rp, const uint32_t *ci, const void *a, const void *X, void *Y, hipStream_t s)
{
  if (((uintptr_t)X % 16) != 0 || ((uintptr_t)Y % 16) != 0) {
    set_error("the CSR-stream SpMM needs 16-byte aligned X and Y");
    return MFP_ERR_ARG;
  }
  const uint2 *m = reinterpret_cast<const uint2 *>(meta);
  if (dtype == MFP_F32)
    return spmm_stream_t<float>(k, cols, nblocks, m, rp, ci, (const float *)a, (const float *)X, (float *)Y, s);
  if (dtype == MFP_F64)
    return spmm_stream_t<double>(k, cols, nblocks, m, rp, ci, (const double *)a, (const double *)X, (double *)Y, s);
  set_error("SpMM values must be f32 or f64");
  return MFP_ERR_ARG;
}

// Row blocks of STREAM_W entries as (first row, first entry) pairs: n + 1 of
// them, the last one (rows, nnz).
__global__ void k_stream_meta(uint32_t n, const uint32_t *__restrict__ blocks, const uint32_t *__restrict__ rp,
                              uint32_t *__restrict__ meta)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const uint32_t r = blocks[i];
  meta[2 * i] = r;
  meta[2 * i + 1] = rp[r];
}

int stream_meta_device(uint32_t rows, const uint32_t *rp, uint32_t **meta, uint32_t *nblocks, hipStream_t s)
{
  *meta = nullptr;
  *nblocks = 0;
  uint32_t *blocks = nullptr, n = 0;
  int st = row_blocks_device(rows, rp, STREAM_W, &blocks, &n, s);
  if (st != MFP_OK) return st;
  uint32_t *m = nullptr;
  st = dmalloc((void **)&m, ((uint64_t)n + 1) * 8);
  if (st == MFP_OK) {
    k_stream_meta<<<(n + 1 + 255) / 256, 256, 0, s>>>(n, blocks, rp, m);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      set_error("row block pairs: kernel failed");
      st = MFP_ERR_HIP;
    }
  }
  dfree(blocks);
  if (st != MFP_OK) {
    dfree(m);
    return st;
  }
  *meta = m;
  *nblocks = n;
  return MFP_OK;
}

// ---- batch assembly ---------------------------------------------------------------
__global__ void k_offset_u32(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t n,
                             uint32_t add)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = in[i] + add;
}

int offset_copy_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t add, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  k_offset_u32<<<std::min<uint32_t>(ceil_div_u32(n, 256), 65536), 256, 0, s>>>(in, out, n, add);
  return hipGetLastError() == hipSuccess ? MFP_OK : MFP_ERR_HIP;
}

} // namespace mfp

using namespace mfp;

#include "objects.h"

struct mfp_batch_s {
  uint32_t count = 0;
  int dtype = MFP_F32;
  mfp_matrix big = nullptr; // block-diagonal CSR
  std::vector<uint32_t> row_off, col_off;
  mfp_plan k1 = nullptr;       // k = 1: CSR-Adaptive, or column-sorted SCOO
  uint32_t lanes = 8;          // k > 1; 0: lanes per row from its length (binned)
  spmm_bins_args bins{};       // lanes == 0
  uint32_t *bin_list = nullptr;
  uint32_t *stream_meta = nullptr; // lanes == MFP_BATCH_STREAM: row blocks of STREAM_W entries, (row, entry) pairs
  uint32_t nstream = 0;
  mfp_plan cs[5] = {};         // lanes == 1: column-sorted SCOO plan for k = 2 << (i - 1)
};

static uint32_t log2_k(uint32_t k)
{
  uint32_t i = 0;
  while ((1u << i) < k) i++;
  return i;
}

// lanes == 1: the column-sorted plan for k columns, built on first use (never
// inside a graph capture): the largest slice (a power of two, up to 16384
// rows) whose R k double accumulators fill at most half of the LDS, so that
// two workgroups share a CU, but never below 1024 rows; 1024 threads.
static int batch_cs_plan(mfp_batch b, uint32_t k, mfp_stream sv)
{
  const uint32_t i = log2_k(k);
  if (i == 0 || i > 4 || (1u << i) != k) {
    set_error("k must be 1, 2, 4, 8 or 16");
    return MFP_ERR_ARG;
  }
  if (b->cs[i]) return MFP_OK;
  mfp_tuning t;
  MFP_CHECK(mfp_tuning_default(b->big, MFP_SCOO, &t));
  t.flags = MFP_TUNE_SCOO_COLSORT;
  // On the 512-matrix batch the half-LDS slice beats the full-LDS one at
  // k = 2 (920 against 1,027 us), 4 (1,538 against 1,664) and 8 (3,077
  // against 3,238); at k = 16, 1024 rows (128 KB) beat 512 (6,526 against
  // 6,787 us).
  t.scoo_slice_size = 16384;
  while (t.scoo_slice_size > 1024 && (uint64_t)t.scoo_slice_size * k * 8 > 81920) t.scoo_slice_size /= 2;
  // one piece per slice: a slice of the block-diagonal batch seldom spans
  // more than one column band, so most slices write Y directly and skip the
  // partial round trip of R k values per slot (k = 4: 1.70 against 1.80 ms
  // with two pieces; the slices still number thousands for 256 CUs)
  t.scoo_pieces = 1;
  // k <= 4: no long-row slices and one item per slice, so no partial slots
  // and one kernel (k = 4: 1,484 against 1,523 us with long-row slices and
  // the finish; k = 2 even; k = 8 and 16 were 1-2% slower)
  if (k <= 4) t.flags |= MFP_TUNE_SCOO_WHOLE_SLICES;
  t.scoo_threads = 1024;
  return mfp_plan_create(&b->cs[i], b->big, MFP_SCOO, &t, sv);
}

struct mfp_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_launch(mfp_batch b, uint32_t k, const void *X, void *Y, hipStream_t s)
{
  const mfp_matrix m = b->big;
  if (k == 1) {
    const mfp_plan p = b->k1;
    if (p->format == MFP_SCOO) return launch_scoo_cs(m->dtype, m->rows, p->cs, p->scoo_threads, X, Y, s);
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
    g.partials = p->ad.partials;
    g.persistent = true;
    return launch_csr_adaptive(m->dtype, g, m->values, X, Y, s);
  }
  if (b->lanes == 0)
    return launch_spmm_binned(m->dtype, k, b->bins, b->bin_list, m->rows, m->row_ptr, m->columns, m->values, X,
                              Y, s);
  if (b->lanes == MFP_BATCH_STREAM)
    return launch_spmm_stream(m->dtype, k, m->cols, b->nstream, b->stream_meta, m->row_ptr, m->columns, m->values, X, Y,
                              s);
  if (b->lanes == 1) {
    const uint32_t i = log2_k(k);
    if (i > 4 || !b->cs[i]) {
      set_error("column-sorted batch SpMM: plan for this k missing");
      return MFP_ERR_STATE;
    }
    return launch_scoo_cs_spmm(m->dtype, k, m->rows, b->cs[i]->cs, b->cs[i]->scoo_threads, X, Y, s);
  }
  return launch_spmm_vector(m->dtype, k, b->lanes, m->rows, m->row_ptr, m->columns, m->values, X, Y,
                            false, s);
}

extern "C" {

int mfp_plan_spmm(mfp_plan p, uint32_t k, const void *X, void *Y, mfp_stream sv)
{
  MFP_ARG(p && X && Y, "null pointer");
  if (k == 1) return mfp_plan_spmv(p, X, Y, sv);
  if (p->format == MFP_SCOO && p->colsort) {
    if (p->cs.slots && !p->cs.partial) {
      set_error("plan created with MFP_TUNE_STATIC_PARTIALS has no partial slots for SpMM");
      return MFP_ERR_STATE;
    }
    return launch_scoo_cs_spmm(p->dtype, k, p->m->rows, p->cs, p->scoo_threads, X, Y, as_stream(sv));
  }
  if (p->format == MFP_CSR_ADAPTIVE && p->stream_meta) {
    const mfp_matrix m = p->m;
    return launch_spmm_stream(p->dtype, k, m->cols, p->nstream, p->stream_meta, m->row_ptr, m->columns, m->values, X,
                              Y, as_stream(sv));
  }
  MFP_ARG(p->format == MFP_CSR_VECTOR,
          "SpMM with k > 1 runs on CSR-vector, CSR-Adaptive (f32 / f64) and column-sorted SCOO plans");
  MFP_ARG(p->dtype == MFP_F32 || p->dtype == MFP_F64, "SpMM values must be f32 or f64");
  const mfp_matrix m = p->m;
  return launch_spmm_vector(p->dtype, k, p->t.lanes, m->rows, m->row_ptr, m->columns, m->values, X, Y,
                            false, as_stream(sv));
}

int mfp_batch_create(mfp_batch *out, const mfp_matrix *ms, uint32_t count, uint32_t lanes,
                     const mfp_tuning *k1_tuning, mfp_stream sv)
{
  MFP_ARG(out && ms && count > 0, "null pointer or empty batch");
  MFP_ARG(lanes <= 1 || lanes == MFP_BATCH_STREAM || (lanes <= 64 && (lanes & (lanes - 1)) == 0), "lanes");
  MFP_CHECK(ensure_init());
  hipStream_t s = as_stream(sv);
  const int dt = ms[0]->dtype;
  MFP_ARG(dt == MFP_F32 || dt == MFP_F64, "batched values must be f32 or f64");
  uint64_t rows = 0, cols = 0, nnz = 0;
  auto *b = new mfp_batch_s;
  b->count = count;
  b->dtype = dt;
  b->lanes = lanes;
  for (uint32_t i = 0; i < count; i++) {
    if (!ms[i] || ms[i]->dtype != dt) {
      delete b;
      set_error("every batched matrix needs the same value type");
      return MFP_ERR_ARG;
    }
    b->row_off.push_back((uint32_t)rows);
    b->col_off.push_back((uint32_t)cols);
    rows += ms[i]->rows;
    cols += ms[i]->cols;
    nnz += ms[i]->nnz;
  }
  if (rows >= (1ull << 32) || cols >= (1ull << 32) || nnz >= (1ull << 32)) {
    delete b;
    set_error("batch exceeds 32-bit rows / cols / nnz");
    return MFP_ERR_ARG;
  }
  b->row_off.push_back((uint32_t)rows);
  b->col_off.push_back((uint32_t)cols);
  auto *m = new mfp_matrix_s;
  m->dtype = dt;
  m->rows = (uint32_t)rows;
  m->cols = (uint32_t)cols;
  m->nnz = nnz;
  b->big = m;
  const size_t vs = dtype_size(dt);
  int st = dmalloc((void **)&m->row_ptr, (rows + 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->columns, nnz * 4);
  if (st == MFP_OK) st = dmalloc(&m->values, nnz * vs);
  uint64_t e0 = 0;
  for (uint32_t i = 0; i < count && st == MFP_OK; i++) {
    const mfp_matrix a = ms[i];
    // rows of matrix i: row_ptr + nnz offset (its last entry is the next one's first)
    st = offset_copy_u32(a->row_ptr, m->row_ptr + b->row_off[i], (uint64_t)a->rows + 1, (uint32_t)e0, s);
    if (st == MFP_OK) st = offset_copy_u32(a->columns, m->columns + e0, a->nnz, b->col_off[i], s);
    if (st == MFP_OK && a->nnz)
      st = hipMemcpyAsync((char *)m->values + e0 * vs, a->values, a->nnz * vs, hipMemcpyDeviceToDevice, s) ==
                   hipSuccess ? MFP_OK : MFP_ERR_HIP;
    e0 += a->nnz;
  }
  if (st == MFP_OK && hipStreamSynchronize(s) != hipSuccess) st = MFP_ERR_HIP;
  if (st == MFP_OK) {
    const int f = k1_tuning && (k1_tuning->flags & MFP_TUNE_SCOO_COLSORT) ? MFP_SCOO : MFP_CSR_ADAPTIVE;
    st = mfp_plan_create(&b->k1, m, f, k1_tuning, sv);
  }
  if (st == MFP_OK && lanes == 0) st = spmm_bins_build(m->rows, m->row_ptr, &b->bins, &b->bin_list, s);
  if (st == MFP_OK && lanes == MFP_BATCH_STREAM)
    st = stream_meta_device(m->rows, m->row_ptr, &b->stream_meta, &b->nstream, s);
  if (st != MFP_OK) {
    mfp_batch_destroy(b);
    return st;
  }
  *out = b;
  return MFP_OK;
}

int mfp_batch_info(mfp_batch b, uint32_t *rows, uint32_t *cols, uint64_t *nnz, uint32_t *row_offsets,
                   uint32_t *col_offsets)
{
  MFP_ARG(b, "null batch");
  if (rows) *rows = b->big->rows;
  if (cols) *cols = b->big->cols;
  if (nnz) *nnz = b->big->nnz;
  if (row_offsets) std::memcpy(row_offsets, b->row_off.data(), b->row_off.size() * 4);
  if (col_offsets) std::memcpy(col_offsets, b->col_off.data(), b->col_off.size() * 4);
  return MFP_OK;
}

int mfp_batch_matrix(mfp_batch b, mfp_matrix *out)
{
  MFP_ARG(b && out, "null pointer");
  *out = b->big;
  return MFP_OK;
}

int mfp_batch_spmm(mfp_batch b, uint32_t k, const void *X, void *Y, mfp_stream sv)
{
  MFP_ARG(b && X && Y, "null pointer");
  if (b->lanes == 1 && k > 1) MFP_CHECK(batch_cs_plan(b, k, sv));
  return batch_launch(b, k, X, Y, as_stream(sv));
}

int mfp_batch_graph_create(mfp_batch b, uint32_t k, const void *X, void *Y, mfp_graph *out)
{
  MFP_ARG(b && X && Y && out, "null pointer");
  MFP_CHECK(ensure_init());
  if (b->lanes == 1 && k > 1) MFP_CHECK(batch_cs_plan(b, k, nullptr)); // before the capture
  hipStream_t cs = nullptr;
  MFP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  auto *g = new mfp_graph_s;
  int st = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess ? MFP_OK : MFP_ERR_HIP;
  if (st == MFP_OK) {
    st = batch_launch(b, k, X, Y, cs);
    const hipError_t e = hipStreamEndCapture(cs, &g->graph);
    if (st == MFP_OK && e != hipSuccess) st = MFP_ERR_HIP;
  }
  if (st == MFP_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    st = MFP_ERR_HIP;
  (void)hipStreamDestroy(cs);
  if (st != MFP_OK) {
    set_error("graph capture of the batched launch failed");
    mfp_graph_destroy(g);
    return st;
  }
  *out = g;
  return MFP_OK;
}

int mfp_plan_graph_create(mfp_plan p, const void *x, void *y, mfp_graph *out)
{
  MFP_ARG(p && x && y && out, "null pointer");
  MFP_CHECK(ensure_init());
  hipStream_t cs = nullptr;
  MFP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  auto *g = new mfp_graph_s;
  int st = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess ? MFP_OK : MFP_ERR_HIP;
  if (st == MFP_OK) {
    st = mfp_plan_spmv(p, x, y, cs);
    const hipError_t e = hipStreamEndCapture(cs, &g->graph);
    if (st == MFP_OK && e != hipSuccess) st = MFP_ERR_HIP;
  }
  if (st == MFP_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    st = MFP_ERR_HIP;
  (void)hipStreamDestroy(cs);
  if (st != MFP_OK) {
    if (st == MFP_ERR_HIP) set_error("graph capture of the plan's SpMV failed");
    mfp_graph_destroy(g);
    return st;
  }
  *out = g;
  return MFP_OK;
}

int mfp_graph_launch(mfp_graph g, mfp_stream sv)
{
  MFP_ARG(g && g->exec, "null graph");
  MFP_HIP(hipGraphLaunch(g->exec, as_stream(sv)));
  return MFP_OK;
}

int mfp_graph_destroy(mfp_graph g)
{
  if (!g) return MFP_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return MFP_OK;
}

int mfp_batch_destroy(mfp_batch b)
{
  if (!b) return MFP_OK;
  mfp_plan_destroy(b->k1);
  for (mfp_plan p : b->cs) mfp_plan_destroy(p);
  mfp_matrix_destroy(b->big);
  dfree(b->bin_list);
  dfree(b->stream_meta);
  delete b;
  return MFP_OK;
}

} // extern "C"
