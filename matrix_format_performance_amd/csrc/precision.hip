// Precision paths.
//
// * FP64 oracle SpMV on the GPU (y64 = A x with the stored values widened to
//   double) and the residual report of an SpMV result against it.
// * Low-precision value storage: BF16 / FP16 values, FP32 x and y, FP32
//   accumulation compensated with the Dot2 scheme of Ogita, Rump and Oishi,
//   "Accurate Sum and Dot Product" (SIAM J. Sci. Comput. 26(6), 2005):
//   TwoProduct by fma (p = a*x, e = fma(a, x, -p)) and Knuth's TwoSum
//   (s' = s + p, z = s' - s, t = (s - (s' - z)) + (p - z)); the error terms
//   are summed in a second accumulator and added at the end.  Lane partials
//   (s, c) are combined with TwoSum as well.
// * ELL with 16-bit column offsets: when every row's column span is below
//   65536, columns are stored as (row base: u32, offset: u16), which with
//   16-bit values makes 4 bytes per stored entry instead of 8.
#include "kernels.h"
#include "mfp_internal.h"
#include "values16.h"
#include "wave64.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

#include <algorithm>
#include <cmath>

namespace mfp {


template <typename T>
__device__ __forceinline__ double widen(const T *p, uint64_t i, int dt)
{
  return (double)p[i];
}

__device__ __forceinline__ double value_f64(const void *vals, uint64_t i, int dt)
{
  switch (dt) {
  case MFP_F32: return (double)((const float *)vals)[i];
  case MFP_F64: return ((const double *)vals)[i];
  case MFP_BF16: return (double)lp_to_f32_bf16(((const uint16_t *)vals)[i]);
  default: return (double)lp_to_f32_f16(((const uint16_t *)vals)[i]);
  }
}

// ---- compensated accumulation (Dot2) ---------------------------------------------
struct comp {
  float s, c;
};

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &t)
{
  s = a + b;
  const float z = s - a;
  t = (a - (s - z)) + (b - z);
}

__device__ __forceinline__ void dot2_add(comp &acc, float a, float x)
{
  const float p = a * x;
  const float e = __builtin_fmaf(a, x, -p);
  float s, t;
  two_sum(acc.s, p, s, t);
  acc.s = s;
  acc.c = acc.c + (t + e);
}

__device__ __forceinline__ comp comp_merge(comp u, comp v)
{
  float s, t;
  two_sum(u.s, v.s, s, t);
  return comp{s, (u.c + v.c) + t};
}

template <int G>
__device__ __forceinline__ comp comp_group_reduce(comp v)
{
  if constexpr (G >= 64) v = comp_merge(v, comp{lane_down<32>(v.s), lane_down<32>(v.c)});
  if constexpr (G >= 32) v = comp_merge(v, comp{lane_down<16>(v.s), lane_down<16>(v.c)});
  if constexpr (G >= 16) v = comp_merge(v, comp{lane_down<8>(v.s), lane_down<8>(v.c)});
  if constexpr (G >= 8) v = comp_merge(v, comp{lane_down<4>(v.s), lane_down<4>(v.c)});
  if constexpr (G >= 4) v = comp_merge(v, comp{lane_down<2>(v.s), lane_down<2>(v.c)});
  if constexpr (G >= 2) v = comp_merge(v, comp{lane_down<1>(v.s), lane_down<1>(v.c)});
  return v;
}

// ---- FP64 oracle SpMV and residual ------------------------------------------------
// One wave per row: 64 lanes strided, double fma, halving tree.
__global__ __launch_bounds__(256) void k_spmv_f64(uint32_t rows, const uint32_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci, const void *vals,
                                                  int vdt, const void *x, int xdt,
                                                  double *__restrict__ y64, double *__restrict__ absdot)
{
  const uint32_t row = blockIdx.x * 4u + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  double d = 0.0, ad = 0.0;
  if (row < rows) {
    for (uint32_t i = rp[row] + lane; i < rp[row + 1]; i += 64) {
      const double p = value_f64(vals, i, vdt) * value_f64(x, ci[i], xdt);
      d += p;
      ad += fabs(p);
    }
  }
  d = group_reduce<64>(d);
  ad = group_reduce<64>(ad);
  if (lane == 0 && row < rows) {
    y64[row] = d;
    if (absdot) absdot[row] = ad;
  }
}

struct resid_acc {
  double num, den;
  unsigned long long max_norm_bits, max_abs_bits;
  unsigned long long worst; // (bits of max_norm << 32) is not representable; row kept separately
};

__global__ void k_residual(uint32_t n, const void *y, int ydt, const double *__restrict__ y64,
                           const double *__restrict__ absdot, resid_acc *acc)
{
  double num = 0, den = 0, mx = 0, mabs = 0;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const double v = value_f64(y, i, ydt);
    const double d = v - y64[i];
    num += d * d;
    den += y64[i] * y64[i];
    const double ad = fabs(d);
    mabs = fmax(mabs, ad);
    const double sc = absdot ? absdot[i] : fabs(y64[i]);
    const double r = sc > 0 ? ad / sc : (ad > 0 ? INFINITY : 0.0);
    mx = fmax(mx, r);
  }
  num = group_reduce<64>(num);
  den = group_reduce<64>(den);
  // max over the wave through the integer order of non-negative doubles
  unsigned long long bx = (unsigned long long)__double_as_longlong(mx);
  unsigned long long ba = (unsigned long long)__double_as_longlong(mabs);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(&acc->num, num);
    atomicAdd(&acc->den, den);
  }
  atomicMax(&acc->max_norm_bits, bx);
  atomicMax(&acc->max_abs_bits, ba);
}

// ---- value conversion for low-precision storage --------------------------------------
__global__ void k_to_lp(const void *in, int din, uint16_t *__restrict__ out, int dout, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    // values are rounded to FP32 first, then to the 16-bit type (nearest-even)
    const float f = din == MFP_F32 ? ((const float *)in)[i] : (float)((const double *)in)[i];
    uint16_t b;
    if (dout == MFP_BF16) {
      __hip_bfloat16 h = __float2bfloat16(f);
      b = *reinterpret_cast<uint16_t *>(&h);
    } else {
      __half h = __float2half(f);
      b = __half_as_ushort(h);
    }
    out[i] = b;
  }
}

// ---- low-precision CSR-vector (G lanes per row, Dot2 per lane) --------------------------
// G lanes per row, Dot2 per lane, compensated group reduction.  Block blk's
// group g takes row g (or list[g] with a row list); n rows in all.
template <int DT, int G, bool LIST>
__device__ __forceinline__ void lp_rows(uint32_t blk, uint32_t n, const uint32_t *__restrict__ rp,
                                        const uint32_t *__restrict__ ci, const uint16_t *__restrict__ a,
                                        const float *__restrict__ x, float *__restrict__ y,
                                        const uint32_t *__restrict__ list)
{
  const uint32_t g = (blk * 256u + threadIdx.x) / G;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t row = LIST ? (g < n ? list[g] : 0xFFFFFFFFu) : g;
  comp acc{0.f, 0.f};
  if (g < n) {
    const uint32_t b = rp[row], e = rp[row + 1];
    uint32_t i = b + lane;
    for (; i + 3u * G < e; i += 4u * G) {
      const uint16_t a0 = ldg_stream(a + i), a1 = ldg_stream(a + i + G);
      const uint16_t a2 = ldg_stream(a + i + 2 * G), a3 = ldg_stream(a + i + 3 * G);
      const uint32_t c0 = ldg_stream(ci + i), c1 = ldg_stream(ci + i + G);
      const uint32_t c2 = ldg_stream(ci + i + 2 * G), c3 = ldg_stream(ci + i + 3 * G);
      const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
      dot2_add(acc, lp_val<DT>(a0), x0);
      dot2_add(acc, lp_val<DT>(a1), x1);
      dot2_add(acc, lp_val<DT>(a2), x2);
      dot2_add(acc, lp_val<DT>(a3), x3);
    }
    for (; i < e; i += G) dot2_add(acc, lp_val<DT>(ldg_stream(a + i)), x[ldg_stream(ci + i)]);
  }
  acc = comp_group_reduce<G>(acc);
  if (lane == 0 && g < n) y[row] = acc.s + acc.c;
}

template <int DT, int G>
__global__ __launch_bounds__(256) void k_csr_vector_lp(uint32_t rows, const uint32_t *__restrict__ rp,
                                                       const uint32_t *__restrict__ ci,
                                                       const uint16_t *__restrict__ a,
                                                       const float *__restrict__ x, float *__restrict__ y)
{
  lp_rows<DT, G, false>(xcd_block(blockIdx.x, gridDim.x), rows, rp, ci, a, x, y, nullptr);
}

// Chunks of long rows: one wave per chunk {row, first entry, end entry,
// slot}; the wave's compensated sum goes to partial[slot].
template <int DT>
__device__ __forceinline__ void lp_chunks(uint32_t blk, uint32_t n, const uint4 *__restrict__ chunks,
                                          const uint32_t *__restrict__ ci, const uint16_t *__restrict__ a,
                                          const float *__restrict__ x, float2 *__restrict__ partial)
{
  const uint32_t c = blk * 4u + threadIdx.x / 64u;
  const uint32_t lane = threadIdx.x & 63u;
  comp acc{0.f, 0.f};
  uint4 ch = make_uint4(0, 0, 0, 0);
  if (c < n) {
    ch = chunks[c];
    uint32_t i = ch.y + lane;
    for (; i + 3u * 64u < ch.z; i += 4u * 64u) {
      const uint16_t a0 = ldg_stream(a + i), a1 = ldg_stream(a + i + 64), a2 = ldg_stream(a + i + 128),
                     a3 = ldg_stream(a + i + 192);
      const uint32_t c0 = ldg_stream(ci + i), c1 = ldg_stream(ci + i + 64), c2 = ldg_stream(ci + i + 128),
                     c3 = ldg_stream(ci + i + 192);
      const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
      dot2_add(acc, lp_val<DT>(a0), x0);
      dot2_add(acc, lp_val<DT>(a1), x1);
      dot2_add(acc, lp_val<DT>(a2), x2);
      dot2_add(acc, lp_val<DT>(a3), x3);
    }
    for (; i < ch.z; i += 64u) dot2_add(acc, lp_val<DT>(ldg_stream(a + i)), x[ldg_stream(ci + i)]);
  }
  acc = comp_group_reduce<64>(acc);
  if (lane == 0 && c < n) partial[ch.w] = make_float2(acc.s, acc.c);
}

template <int DT>
__global__ __launch_bounds__(256) void k_lp_chunks(uint32_t n, const uint4 *__restrict__ chunks,
                                                   const uint32_t *__restrict__ ci, const uint16_t *__restrict__ a,
                                                   const float *__restrict__ x, float2 *__restrict__ partial)
{
  lp_chunks<DT>(xcd_block(blockIdx.x, gridDim.x), n, chunks, ci, a, x, partial);
}

template <int DT, int G>
__global__ __launch_bounds__(256) void k_lp_rows(uint32_t n, const uint32_t *__restrict__ list,
                                                 const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                                 const uint16_t *__restrict__ a, const float *__restrict__ x,
                                                 float *__restrict__ y)
{
  lp_rows<DT, G, true>(xcd_block(blockIdx.x, gridDim.x), n, rp, ci, a, x, y, list);
}

// y[row] of each long row {row, first slot, slots}: its chunks merged in order
__global__ void k_lp_long_finish(uint32_t n, const uint4 *__restrict__ longs, const float2 *__restrict__ partial,
                                 float *__restrict__ y)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 l = longs[i];
  comp acc{partial[l.y].x, partial[l.y].y};
  for (uint32_t k = 1; k < l.z; k++) acc = comp_merge(acc, comp{partial[l.y + k].x, partial[l.y + k].y});
  y[l.x] = acc.s + acc.c;
}

// ---- low-precision ELL (Dot2, optional 16-bit column offsets) -------------------------------
// Sliced layout: the w slots of each block of 64 rows are contiguous, slot k
// of row r at (r / 64 * w + k) * 64 + r % 64, so a wave (one block) reads one
// contiguous run instead of w strided ones.  With 16-bit offsets a slot is
// one u32, offset << 16 | value bits; otherwise values (u16) and columns
// (u32) are separate arrays.  Thread per row; each row adds its slots in
// order.
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

template <int DT, bool OFF16>
struct ell_lp_row {
  const void *cols;
  const uint16_t *ev;
  const float *x;
  uint32_t r, w, b0;
  __device__ __forceinline__ uint64_t at(uint32_t k) const { return ((uint64_t)(r >> 6) * w + k) * 64u + (r & 63u); }
  // N slots from k: every load issued before the first use
  template <uint32_t N> __device__ __forceinline__ void group(uint32_t k, comp &acc) const
  {
    uint32_t a[N], c[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) {
      a[q] = ldg_stream(ev + at(k + q));
      c[q] = ldg_stream((const uint32_t *)cols + at(k + q));
    }
    float xv[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) xv[q] = x[c[q]];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) dot2_add(acc, lp_val<DT>((uint16_t)a[q]), xv[q]);
  }
};

// 16-bit offsets: packed slots, 4 consecutive slots of a row in one 16-byte
// word; word j of row r at ((r / 64) * w4 + j) * 64 + r % 64 (w4 = w / 4
// rounded up; padding slots hold value 0, offset 0).
template <int DT>
struct ell_lp_row16 {
  const u32x4v *slots;
  const float *x;
  uint32_t r, w4, b0;
  template <uint32_t N> __device__ __forceinline__ void group(uint32_t j, comp &acc) const
  {
    u32x4v v[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++)
      v[q] = __builtin_nontemporal_load(slots + ((uint64_t)(r >> 6) * w4 + j + q) * 64u + (r & 63u));
    float xv[N][4];
#pragma unroll
    for (uint32_t q = 0; q < N; q++)
#pragma unroll
      for (uint32_t e = 0; e < 4; e++) xv[q][e] = x[b0 + (v[q][e] >> 16)];
#pragma unroll
    for (uint32_t q = 0; q < N; q++)
#pragma unroll
      for (uint32_t e = 0; e < 4; e++) dot2_add(acc, lp_val<DT>((uint16_t)(v[q][e] & 0xFFFFu)), xv[q][e]);
  }
};

template <int DT, bool OFF16, bool BLOCK = false>
__global__ __launch_bounds__(256) void k_ell_lp(uint32_t rows, uint32_t stride, uint32_t w,
                                                const void *__restrict__ cols, const uint32_t *__restrict__ base,
                                                const uint16_t *__restrict__ ev, const float *__restrict__ x,
                                                float *__restrict__ y)
{
  const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (r >= rows) return;
  (void)stride;
  comp acc{0.f, 0.f};
  // groups of slots with no guarded loads, so each group's loads are in flight together
  if constexpr (OFF16) {
    const uint32_t w4 = (w + 3) / 4;
    // offsets from the row's first column, or from its 64-row block's
    const ell_lp_row16<DT> row{(const u32x4v *)cols, x, r, w4, BLOCK ? base[r >> 6] : base[r]};
    uint32_t j = 0;
    // a row of at most 8 words (32 slots) is one group: every load of the
    // row in flight before the first gather (3D 27-point stencil, 7 words:
    // 43.4 against 45.7 us in groups of 3, 3 and 1)
    if (w4 <= 8) {
      switch (w4) {
      case 8: row.template group<8>(0, acc); break;
      case 7: row.template group<7>(0, acc); break;
      case 6: row.template group<6>(0, acc); break;
      case 5: row.template group<5>(0, acc); break;
      case 4: row.template group<4>(0, acc); break;
      case 3: row.template group<3>(0, acc); break;
      case 2: row.template group<2>(0, acc); break;
      case 1: row.template group<1>(0, acc); break;
      default: break;
      }
      j = w4;
    }
    for (; j + 3 <= w4; j += 3) row.template group<3>(j, acc);
    if (j + 2 <= w4) {
      row.template group<2>(j, acc);
      j += 2;
    }
    if (j < w4) row.template group<1>(j, acc);
  } else {
    const ell_lp_row<DT, false> row{cols, ev, x, r, w, 0u};
    uint32_t k = 0;
    for (; k + 12 <= w; k += 12) row.template group<12>(k, acc);
    if (k + 8 <= w) {
      row.template group<8>(k, acc);
      k += 8;
    }
    if (k + 4 <= w) {
      row.template group<4>(k, acc);
      k += 4;
    }
    if (k + 2 <= w) {
      row.template group<2>(k, acc);
      k += 2;
    }
    if (k < w) row.template group<1>(k, acc);
  }
  y[r] = acc.s + acc.c;
}

// Row column span (last - first) maximum, for choosing 16-bit offsets.
__global__ void k_row_span(uint32_t rows, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                           uint32_t *mx)
{
  uint32_t m = 0;
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; r < rows; r += stride) {
    const uint32_t b = rp[r], e = rp[r + 1];
    if (e > b) m = max(m, ci[e - 1] - ci[b]);
  }
  atomicMax(mx, m);
}

// Per 64-row block: the smallest first column of its rows (0 for a block of
// empty rows) into base[], and the block's column span into *mx.
__global__ void k_block_base(uint32_t rows, uint32_t nblocks, const uint32_t *__restrict__ rp,
                             const uint32_t *__restrict__ ci, uint32_t *__restrict__ base, uint32_t *mx)
{
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  uint32_t lo = 0xFFFFFFFFu, hi = 0;
  for (uint32_t r = b * 64u; r < min(rows, b * 64u + 64u); r++) {
    const uint32_t s = rp[r], e = rp[r + 1];
    if (e > s) {
      lo = min(lo, ci[s]);
      hi = max(hi, ci[e - 1]);
    }
  }
  base[b] = lo == 0xFFFFFFFFu ? 0u : lo;
  if (lo != 0xFFFFFFFFu) atomicMax(mx, hi - lo);
}

int block_base_build(uint32_t rows, const uint32_t *rp, const uint32_t *ci, uint32_t *base, uint32_t *span,
                     hipStream_t s)
{
  uint32_t *d = nullptr;
  MFP_CHECK(dmalloc((void **)&d, 4));
  int st = fill_u32(d, 0, 1, s);
  const uint32_t nb = (uint32_t)(((uint64_t)rows + 63) / 64);
  if (st == MFP_OK && nb) {
    k_block_base<<<ceil_div_u32(nb, 256), 256, 0, s>>>(rows, nb, rp, ci, base, d);
    if (hipGetLastError() != hipSuccess) st = MFP_ERR_HIP;
  }
  if (st == MFP_OK && (hipMemcpyAsync(span, d, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                       hipStreamSynchronize(s) != hipSuccess))
    st = MFP_ERR_HIP;
  dfree(d);
  return st;
}

// ELL build for 16-bit values: column-major, padding (0, offset 0 / col 0).
__global__ void k_ell_lp_build(uint32_t rows, uint32_t stride, const uint32_t *__restrict__ rp,
                               const uint32_t *__restrict__ ci, const uint16_t *__restrict__ a, uint32_t w,
                               bool off16, bool block, uint16_t *__restrict__ ev, void *cols,
                               uint32_t *__restrict__ base)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= stride) return;
  const uint32_t s = r < rows ? rp[r] : 0u, len = r < rows ? rp[r + 1] - s : 0u;
  uint32_t b0 = len ? ci[s] : 0u;
  if (off16 && block)
    b0 = base[r >> 6];
  else if (off16)
    base[r] = b0;
  const uint32_t kend = off16 ? (w + 3) / 4 * 4 : w; // packed words also hold the padding slots
  for (uint32_t k = 0; k < kend; k++) {
    const bool in = k < len;
    const uint16_t v = in ? a[s + k] : (uint16_t)0;
    const uint32_t c = in ? ci[s + k] : 0u;
    if (off16) {
      const uint32_t w4 = (w + 3) / 4;
      const uint64_t o = (((uint64_t)(r >> 6) * w4 + k / 4) * 64u + (r & 63u)) * 4u + k % 4;
      ((uint32_t *)cols)[o] = (in ? (c - b0) << 16 : 0u) | v;
    } else {
      const uint64_t o = ((uint64_t)(r >> 6) * w + k) * 64u + (r & 63u);
      ev[o] = v;
      ((uint32_t *)cols)[o] = c;
    }
  }
}

static int launched_lp()
{
  mark_kernel_end();
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

int launch_spmv_f64(uint32_t rows, const uint32_t *rp, const uint32_t *ci, const void *vals, int vdt,
                    const void *x, int xdt, double *y64, double *absdot, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  k_spmv_f64<<<ceil_div_u32(rows, 4), 256, 0, s>>>(rows, rp, ci, vals, vdt, x, xdt, y64, absdot);
  return launched_lp();
}

int launch_csr_vector_lp(int dt, uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                         const void *a, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32((uint64_t)rows * lanes, 256);
  const uint16_t *av = (const uint16_t *)a;
  const float *xv = (const float *)x;
  float *yv = (float *)y;
#define LPV(D, G) k_csr_vector_lp<D, G><<<grid, 256, 0, s>>>(rows, rp, ci, av, xv, yv)
  if (dt == MFP_BF16) {
    switch (lanes) {
    case 2: LPV(MFP_BF16, 2); break;
    case 4: LPV(MFP_BF16, 4); break;
    case 8: LPV(MFP_BF16, 8); break;
    case 16: LPV(MFP_BF16, 16); break;
    case 32: LPV(MFP_BF16, 32); break;
    case 64: LPV(MFP_BF16, 64); break;
    default: set_error("lanes"); return MFP_ERR_ARG;
    }
  } else {
    switch (lanes) {
    case 2: LPV(MFP_F16, 2); break;
    case 4: LPV(MFP_F16, 4); break;
    case 8: LPV(MFP_F16, 8); break;
    case 16: LPV(MFP_F16, 16); break;
    case 32: LPV(MFP_F16, 32); break;
    case 64: LPV(MFP_F16, 64); break;
    default: set_error("lanes"); return MFP_ERR_ARG;
    }
  }
#undef LPV
  return launched_lp();
}

uint32_t ell_lp_stride(uint32_t rows) { return (uint32_t)(((uint64_t)rows + 63) / 64 * 64); }

int launch_ell_lp(int dt, bool off16, bool block, uint32_t rows, uint32_t w, const void *cols, const uint32_t *base,
                  const void *ev, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t stride = ell_lp_stride(rows);
  const uint32_t grid = ceil_div_u32(rows, 256);
  const uint16_t *e = (const uint16_t *)ev;
  const float *xv = (const float *)x;
  float *yv = (float *)y;
  if (dt == MFP_BF16) {
    if (off16 && block) k_ell_lp<MFP_BF16, true, true><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
    else if (off16) k_ell_lp<MFP_BF16, true><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
    else k_ell_lp<MFP_BF16, false><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
  } else {
    if (off16 && block) k_ell_lp<MFP_F16, true, true><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
    else if (off16) k_ell_lp<MFP_F16, true><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
    else k_ell_lp<MFP_F16, false><<<grid, 256, 0, s>>>(rows, stride, w, cols, base, e, xv, yv);
  }
  return launched_lp();
}

// Binned 16-bit CSR: rows grouped by length into lists for G = 2..64 lanes
// (G = the row length rounded up to a power of two), rows longer than
// LP_LONG split into chunks of LP_CHUNK entries, one wave each.
constexpr uint32_t LP_LONG = 4096, LP_CHUNK = 1024;

int lp_binned_build(uint32_t rows, const uint32_t *rp_dev, lp_binned_layout *out, hipStream_t s,
                    bool own_partials)
{
  *out = lp_binned_layout{};
  std::vector<uint32_t> rp((size_t)rows + 1);
  MFP_HIP(hipMemcpyAsync(rp.data(), rp_dev, rp.size() * 4, hipMemcpyDeviceToHost, s));
  MFP_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> bins[6];
  std::vector<uint4> chunks, longs;
  for (uint32_t r = 0; r < rows; r++) {
    const uint32_t len = rp[r + 1] - rp[r];
    if (len > LP_LONG) {
      const uint32_t n = (len + LP_CHUNK - 1) / LP_CHUNK;
      longs.push_back(make_uint4(r, (uint32_t)chunks.size(), n, 0));
      for (uint32_t k = 0; k < n; k++)
        chunks.push_back(make_uint4(r, rp[r] + k * LP_CHUNK, std::min(rp[r + 1], rp[r] + (k + 1) * LP_CHUNK),
                                    (uint32_t)chunks.size()));
      continue;
    }
    uint32_t b = 0;
    while (b < 5 && (2u << b) < len) b++;
    bins[b].push_back(r);
  }
  std::vector<uint32_t> list;
  for (int b = 0; b < 6; b++) {
    out->bin_off[b] = (uint32_t)list.size();
    list.insert(list.end(), bins[b].begin(), bins[b].end());
  }
  out->bin_off[6] = (uint32_t)list.size();
  out->nchunks = (uint32_t)chunks.size();
  out->nlong = (uint32_t)longs.size();
  int st = dmalloc((void **)&out->list, std::max<size_t>(list.size(), 1) * 4);
  if (st == MFP_OK) st = dmalloc((void **)&out->chunks, std::max<size_t>(chunks.size(), 1) * sizeof(uint4));
  if (st == MFP_OK) st = dmalloc((void **)&out->longs, std::max<size_t>(longs.size(), 1) * sizeof(uint4));
  if (st == MFP_OK && own_partials)
    st = dmalloc((void **)&out->partial, std::max<size_t>(chunks.size(), 1) * sizeof(float2));
  if (st != MFP_OK) {
    lp_binned_free(out);
    return st;
  }
  if (!list.empty()) MFP_HIP(hipMemcpyAsync(out->list, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
  if (!chunks.empty())
    MFP_HIP(hipMemcpyAsync(out->chunks, chunks.data(), chunks.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  if (!longs.empty())
    MFP_HIP(hipMemcpyAsync(out->longs, longs.data(), longs.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  MFP_HIP(hipStreamSynchronize(s));
  out->device_bytes = list.size() * 4 + (chunks.size() + longs.size()) * sizeof(uint4) +
                      (own_partials ? chunks.size() * 8 : 0);
  return MFP_OK;
}

void lp_binned_free(lp_binned_layout *l)
{
  dfree(l->list);
  dfree(l->chunks);
  dfree(l->longs);
  dfree(l->partial);
  *l = lp_binned_layout{};
}

// one launch per bin (a single launch over all bins measured slower: its
// registers are those of the widest variant), long-row chunks first
template <int DT>
static int lp_binned_t(const lp_binned_layout &l, const uint32_t *rp, const uint32_t *ci, const uint16_t *a,
                       const float *x, float *y, hipStream_t s)
{
  if (l.nchunks) {
    k_lp_chunks<DT><<<ceil_div_u32(l.nchunks, 4), 256, 0, s>>>(l.nchunks, l.chunks, ci, a, x, l.partial);
    MFP_CHECK(launched_lp());
  }
  for (int b = 0; b < 6; b++) {
    const uint32_t n = l.bin_off[b + 1] - l.bin_off[b];
    if (!n) continue;
    const uint32_t grid = ceil_div_u32((uint64_t)n * (2u << b), 256);
    const uint32_t *lst = l.list + l.bin_off[b];
    switch (b) {
    case 0: k_lp_rows<DT, 2><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    case 1: k_lp_rows<DT, 4><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    case 2: k_lp_rows<DT, 8><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    case 3: k_lp_rows<DT, 16><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    case 4: k_lp_rows<DT, 32><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    default: k_lp_rows<DT, 64><<<grid, 256, 0, s>>>(n, lst, rp, ci, a, x, y); break;
    }
    MFP_CHECK(launched_lp());
  }
  if (l.nchunks) {
    k_lp_long_finish<<<ceil_div_u32(l.nlong, 256), 256, 0, s>>>(l.nlong, l.longs, l.partial, y);
    MFP_CHECK(launched_lp());
  }
  return MFP_OK;
}

int launch_lp_binned(int dt, const lp_binned_layout &l, const uint32_t *rp, const uint32_t *ci, const void *a,
                     const void *x, void *y, hipStream_t s)
{
  if (dt == MFP_BF16)
    return lp_binned_t<MFP_BF16>(l, rp, ci, (const uint16_t *)a, (const float *)x, (float *)y, s);
  return lp_binned_t<MFP_F16>(l, rp, ci, (const uint16_t *)a, (const float *)x, (float *)y, s);
}

int row_span_max(uint32_t rows, const uint32_t *rp, const uint32_t *ci, uint32_t *out, hipStream_t s)
{
  uint32_t *d = nullptr;
  MFP_CHECK(dmalloc((void **)&d, 4));
  MFP_CHECK(fill_u32(d, 0, 1, s));
  if (rows) k_row_span<<<std::min<uint32_t>(ceil_div_u32(rows, 256), 4096), 256, 0, s>>>(rows, rp, ci, d);
  MFP_HIP(hipMemcpyAsync(out, d, 4, hipMemcpyDeviceToHost, s));
  MFP_HIP(hipStreamSynchronize(s));
  dfree(d);
  return MFP_OK;
}

int ell_lp_build(uint32_t rows, const uint32_t *rp, const uint32_t *ci, const void *a, uint32_t w,
                 bool off16, bool block, void *ev, void *cols, uint32_t *base, hipStream_t s)
{
  if (rows == 0 || w == 0) return MFP_OK;
  const uint32_t stride = ell_lp_stride(rows);
  k_ell_lp_build<<<ceil_div_u32(stride, 256), 256, 0, s>>>(rows, stride, rp, ci, (const uint16_t *)a, w, off16,
                                                           block, (uint16_t *)ev, cols, base);
  return launched_lp();
}

int to_lp_values(const void *in, int din, void *out, int dout, uint64_t n, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  k_to_lp<<<std::min<uint32_t>(ceil_div_u32(n, 256), 65536), 256, 0, s>>>(in, din, (uint16_t *)out, dout, n);
  return launched_lp();
}

} // namespace mfp

using namespace mfp;

extern "C" {

int mfp_spmv_f64_oracle(mfp_matrix m, int x_dtype, const void *x, double *y64, double *absdot,
                        mfp_stream sv)
{
  MFP_ARG(m && x && y64, "null pointer");
  MFP_ARG(x_dtype == MFP_F32 || x_dtype == MFP_F64, "x dtype must be f32 or f64");
  mfp_matrix_view v;
  MFP_CHECK(mfp_matrix_view_get(m, &v));
  return launch_spmv_f64(v.rows, v.row_ptr, v.columns, v.values, v.dtype, x, x_dtype, y64, absdot,
                         as_stream(sv));
}

int mfp_residual(uint32_t n, int y_dtype, const void *y, const double *y64, const double *absdot,
                 mfp_error_report *out)
{
  MFP_ARG(out && y && y64, "null pointer");
  MFP_ARG(y_dtype == MFP_F32 || y_dtype == MFP_F64, "y dtype must be f32 or f64");
  MFP_CHECK(ensure_init());
  resid_acc *acc = nullptr;
  MFP_CHECK(dmalloc((void **)&acc, sizeof(resid_acc)));
  MFP_CHECK(fill_bytes_zero(acc, sizeof(resid_acc), nullptr));
  if (n) {
    k_residual<<<std::min<uint32_t>(ceil_div_u32(n, 256), 2048), 256>>>(n, y, y_dtype, y64, absdot, acc);
    MFP_HIP(hipGetLastError());
  }
  resid_acc h{};
  MFP_HIP(hipMemcpy(&h, acc, sizeof(h), hipMemcpyDeviceToHost));
  dfree(acc);
  out->rel_l2 = h.den > 0 ? std::sqrt(h.num / h.den) : std::sqrt(h.num);
  out->max_row_err = __builtin_bit_cast(double, h.max_norm_bits);
  out->max_abs_err = __builtin_bit_cast(double, h.max_abs_bits);
  return MFP_OK;
}

} // extern "C"
