// SpMV kernels for gfx950: CSR-scalar, CSR-vector, CSR-Adaptive, ELL, COO,
// SCOO and the two hybrid ELL+COO forms.
//
// SpMV is bound by HBM bandwidth, so every kernel keeps several independent
// loads in flight per lane (loads are issued for a group of elements before
// any of them is used) and streams the matrix with nontemporal loads, which
// leaves L2 / Infinity Cache to the gathered x entries.
//
// Summation orders (bit-exact against the oracle's emulators where the
// kernel is deterministic):
//   CSR-scalar  sequential fma per row            (cpu_matrix_multiplier.cpp:20-61)
//   CSR-vector  G-lane strided fma + halving tree (cpu_matrix_multiplier.cpp:170-227, G=32)
//   ELL         sequential fma over the padded width
//   CSR-Adaptive: see k_csr_adaptive.
// COO, SCOO and the hybrids accumulate with atomics (order not fixed).
#include "kernels.h"
#include "mfp_internal.h"
#include "values16.h"
#include "wave64.h"

#include <hip/hip_ext.h>

#include <type_traits>

namespace mfp {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// 16-byte value vectors and the matching column-index vectors.
template <typename T> struct vec16;
template <> struct vec16<float> {
  using V = f32x4;
  using I = u32x4;
  static constexpr uint32_t N = 4;
};
template <> struct vec16<double> {
  using V = f64x2;
  using I = u32x2;
  static constexpr uint32_t N = 2;
};

// ---------------------------------------------------------------------------
// CSR-scalar: one lane per row, as gpu_matrix_multiplier.cu:22-43, but the
// wave stages its rows' contiguous nnz range through LDS with coalesced loads
// so each lane then walks only its own row.  Order: sequential fma.
// The staging of a chunk is split in two so that no load waits behind
// another: the values and column indices of chunk k + 1 are issued before
// chunk k is summed, and their x gathers right after; every load sits at a
// clamped in-range offset, so none is predicated.  A lane reads its row
// from LDS eight elements at a time, so a long row pays the LDS latency once
// per eight fma.
// ---------------------------------------------------------------------------
constexpr uint32_t SCALAR_PER_LANE = 8, SCALAR_CHUNK = 64 * SCALAR_PER_LANE;

template <typename T>
__device__ __forceinline__ void scalar_stage_loads(const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                                   uint32_t c0, uint32_t n, uint32_t lane, T (&av)[SCALAR_PER_LANE],
                                                   uint32_t (&cv)[SCALAR_PER_LANE])
{
#pragma unroll
  for (uint32_t k = 0; k < SCALAR_PER_LANE; k++) {
    const uint32_t j = min(lane + 64u * k, n - 1u); // n >= 1: a duplicate of the last element past the end
    av[k] = ldg_stream(a + c0 + j);
    cv[k] = ldg_stream(ci + c0 + j);
  }
}

template <typename T>
__global__ __launch_bounds__(64) void k_csr_scalar(uint32_t rows, const uint32_t *__restrict__ rp,
                                                   const uint32_t *__restrict__ ci,
                                                   const T *__restrict__ a, const T *__restrict__ x,
                                                   T *__restrict__ y)
{
  __shared__ T va[2][SCALAR_CHUNK];
  __shared__ T vx[2][SCALAR_CHUNK];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t r0 = blk * 64u;
  const uint32_t lane = threadIdx.x;
  const uint32_t r = r0 + lane;
  const uint32_t rend = min(r0 + 64u, rows);
  const uint32_t rs = rp[min(r, rows)];
  const uint32_t re = rp[min(r + 1u, rows)];
  const uint32_t b = rp[r0], e = rp[rend];
  T dot = 0;
  if (b < e) {
    T av[SCALAR_PER_LANE], xv[SCALAR_PER_LANE];
    uint32_t cv[SCALAR_PER_LANE];
    scalar_stage_loads(ci, a, b, min(SCALAR_CHUNK, e - b), lane, av, cv);
#pragma unroll
    for (uint32_t k = 0; k < SCALAR_PER_LANE; k++) xv[k] = x[cv[k]];
    uint32_t buf = 0;
    for (uint32_t c0 = b; c0 < e; c0 += SCALAR_CHUNK, buf ^= 1u) {
      const uint32_t n = min(SCALAR_CHUNK, e - c0);
      // slots past n hold duplicates and are never read
#pragma unroll
      for (uint32_t k = 0; k < SCALAR_PER_LANE; k++) {
        va[buf][lane + 64u * k] = av[k];
        vx[buf][lane + 64u * k] = xv[k];
      }
      __syncthreads();
      const uint32_t c1 = c0 + SCALAR_CHUNK;
      const bool next = c1 < e; // uniform over the wave
      if (next) scalar_stage_loads(ci, a, c1, min(SCALAR_CHUNK, e - c1), lane, av, cv);
      const uint32_t lo = max(rs, c0), hi = min(re, c0 + n);
      const T *pa = va[buf] - c0, *px = vx[buf] - c0;
      uint32_t i = lo;
      for (; i + 8u <= hi; i += 8u) {
        T qa[8], qx[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
          qa[q] = pa[i + q];
          qx[q] = px[i + q];
        }
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) dot = fma_t(qa[q], qx[q], dot);
      }
      for (; i < hi; i++) dot = fma_t(pa[i], px[i], dot);
      if (next) {
#pragma unroll
        for (uint32_t k = 0; k < SCALAR_PER_LANE; k++) xv[k] = x[cv[k]];
      }
      // the buffer written next is the one summed two chunks ago: the
      // barrier above already ordered those reads
    }
  }
  if (r < rows) y[r] = dot;
}

// ---------------------------------------------------------------------------
// CSR-vector: G lanes per row (G = 2..64; the reference uses a 32-lane warp,
// gpu_matrix_multiplier.cu:331-362).  Lane l accumulates elements l, l+G, ...
// with fma (four elements' loads issued together), then group_reduce<G>.
// ---------------------------------------------------------------------------
// Rows with at least eight elements per lane run a software pipeline: the
// values and column indices of the next eight are loaded (at clamped in-row
// offsets, so no load is predicated) before the x entries of the current
// eight are gathered and summed, so a long row waits for one memory latency
// per eight elements instead of two per four.
template <typename T, int G>
__global__ __launch_bounds__(256) void k_csr_vector(uint32_t rows, const uint32_t *__restrict__ rp,
                                                    const uint32_t *__restrict__ ci,
                                                    const T *__restrict__ a, const T *__restrict__ x,
                                                    T *__restrict__ y)
{
  constexpr uint32_t U = 8;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t row = (blk * 256u + threadIdx.x) / G;
  const uint32_t lane = threadIdx.x % G;
  T dot = 0;
  if (row < rows) {
    const uint32_t b = rp[row], e = rp[row + 1];
    uint32_t i = b + lane;
    if (i + (U - 1) * G < e) {
      T an[U];
      uint32_t cn[U];
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        an[u] = ldg_stream(a + i + u * G);
        cn[u] = ldg_stream(ci + i + u * G);
      }
      for (;;) {
        T ac[U];
        uint32_t cc[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
          ac[u] = an[u];
          cc[u] = cn[u];
        }
        const uint32_t inext = i + U * G;
        const bool more = inext + (U - 1) * G < e;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
          const uint32_t j = min(inext + u * G, e - 1u); // e > i: in range
          an[u] = ldg_stream(a + j);
          cn[u] = ldg_stream(ci + j);
        }
        T xv[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) xv[u] = x[cc[u]];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) dot = fma_t(ac[u], xv[u], dot);
        i = inext;
        if (!more) break;
      }
    }
    for (; i + 3u * G < e; i += 4u * G) {
      const T a0 = ldg_stream(a + i), a1 = ldg_stream(a + i + G);
      const T a2 = ldg_stream(a + i + 2 * G), a3 = ldg_stream(a + i + 3 * G);
      const uint32_t c0 = ldg_stream(ci + i), c1 = ldg_stream(ci + i + G);
      const uint32_t c2 = ldg_stream(ci + i + 2 * G), c3 = ldg_stream(ci + i + 3 * G);
      const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
      dot = fma_t(a0, x0, dot);
      dot = fma_t(a1, x1, dot);
      dot = fma_t(a2, x2, dot);
      dot = fma_t(a3, x3, dot);
    }
    for (; i < e; i += G) dot = fma_t(ldg_stream(a + i), x[ldg_stream(ci + i)], dot);
  }
  dot = group_reduce<G>(dot);
  if (lane == 0 && row < rows) y[row] = dot;
}

// ---------------------------------------------------------------------------
// CSR-Adaptive (csr_adaptive_spmv.cu:26-178 redesigned for 256-lane
// workgroups and LDS-sized row blocks).
//   multi-row block  (CSR-Stream): the block's nnz range is read with 16-byte
//     vector loads (K vectors per lane, all in flight together), rounded
//     products a*x staged in LDS (up to nnz_per_wg of them).  The reduction
//     is balanced whatever the row lengths: lane i sums the products
//     [i*E, i*E + E) (E = ceil(nnz_per_wg / 256)) left to right, restarting
//     at every row boundary; rows inside one lane finish there, and the
//     partial a lane carries into the next enters a 256-lane Hillis-Steele
//     segmented scan; a row ending in lane j that began earlier is
//     scan[j-1] + lane j's partial.
//   single-row block (CSR-Vector/VectorL): 256 lanes strided with fma, a
//     64-lane halving tree per wave, then a halving tree over the 4 waves.
//   rows longer than long_row: split into `chunk`-element pieces, one
//     workgroup each, reduced like a single-row block; k_adaptive_finish
//     (launched next on the same stream) adds a row's pieces left to right.
//     The pieces are dispatched first, round-robin over the XCDs.
// ---------------------------------------------------------------------------
constexpr uint32_t AD_T = 256;

template <typename T>
__device__ __forceinline__ T wg_row_sum(const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                        const T *__restrict__ x, uint32_t b, uint32_t e, T *wsum)
{
  T d = 0;
  uint32_t i = b + threadIdx.x;
  for (; i + 3 * AD_T < e; i += 4 * AD_T) {
    const T a0 = ldg_stream(a + i), a1 = ldg_stream(a + i + AD_T);
    const T a2 = ldg_stream(a + i + 2 * AD_T), a3 = ldg_stream(a + i + 3 * AD_T);
    const uint32_t c0 = ldg_stream(ci + i), c1 = ldg_stream(ci + i + AD_T);
    const uint32_t c2 = ldg_stream(ci + i + 2 * AD_T), c3 = ldg_stream(ci + i + 3 * AD_T);
    const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    d = fma_t(a0, x0, d);
    d = fma_t(a1, x1, d);
    d = fma_t(a2, x2, d);
    d = fma_t(a3, x3, d);
  }
  for (; i < e; i += AD_T) d = fma_t(ldg_stream(a + i), x[ldg_stream(ci + i)], d);
  d = group_reduce<64>(d);
  if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = d;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    const T w0 = wsum[0] + wsum[2];
    const T w1 = wsum[1] + wsum[3];
    r = w0 + w1;
  }
  return r;
}

template <typename T>
struct adaptive_smem {
  T *cache;     // staged products (nnz_per_wg)
  uint32_t *rs; // block-relative row starts (nnz_per_wg + 2)
  T *wsum, *sv;
  uint32_t *sf;
};

// One work item of the launch: hw is the item's dispatch index among grid
// items (pieces of long rows first, then row blocks in XCD order).
template <typename T, uint32_t K>
__device__ __forceinline__ void adaptive_item(
    uint32_t hw, uint32_t grid, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
    const T *__restrict__ a, const T *__restrict__ x, T *__restrict__ y,
    const uint32_t *__restrict__ blocks, uint32_t nnz_per_wg, uint32_t long_row,
    const uint4 *__restrict__ chunks, uint32_t nchunks, uint32_t nchunks_pad, uint32_t chunk,
    T *__restrict__ partials, const adaptive_smem<T> &sm)
{
  using VT = vec16<T>;
  constexpr uint32_t N = VT::N;
  T *cache = sm.cache;
  uint32_t *rs = sm.rs;
  T *wsum = sm.wsum, *sv = sm.sv;
  uint32_t *sf = sm.sf;
  const uint32_t tid = threadIdx.x;

  if (hw < nchunks_pad) {
    if (hw >= nchunks) return;
    const uint4 c4 = chunks[hw]; // {row, piece, first item, pieces}
    const uint32_t s = rp[c4.x], e = rp[c4.x + 1];
    const uint32_t cb = s + c4.y * chunk;
    const uint32_t ce = min(e, cb + chunk);
    const T p = wg_row_sum(ci, a, x, cb, ce, wsum);
    if (tid == 0) partials[c4.z + c4.y] = p;
    return;
  }

  const uint32_t item = xcd_block(hw - nchunks_pad, grid - nchunks_pad);
  const uint32_t r0 = blocks[item], r1 = blocks[item + 1];
  if (r1 <= r0) return;
  if (r1 - r0 == 1) {
    const uint32_t s = rp[r0], e = rp[r0 + 1];
    if (e - s > long_row) return; // handled by the chunk items
    const T d = wg_row_sum(ci, a, x, s, e, wsum);
    if (tid == 0) y[r0] = d;
    return;
  }

  // CSR-Stream: vectors [v0, v1) cover the block's elements [base, base + n)
  const uint32_t base = rp[r0];
  const uint32_t n = rp[r1] - base;
  const uint32_t v0 = base / N, v1 = (base + n + N - 1) / N;
  const typename VT::V *av = reinterpret_cast<const typename VT::V *>(a);
  const typename VT::I *cv = reinterpret_cast<const typename VT::I *>(ci);
  typename VT::V va[K];
  typename VT::I vc[K];
#pragma unroll
  for (uint32_t k = 0; k < K; k++) {
    const uint32_t v = v0 + tid + k * AD_T;
    if (v < v1) {
      va[k] = ldg_stream(av + v);
      vc[k] = ldg_stream(cv + v);
    }
  }
  T xv[K][N];
#pragma unroll
  for (uint32_t k = 0; k < K; k++) {
    const uint32_t v = v0 + tid + k * AD_T;
#pragma unroll
    for (uint32_t j = 0; j < N; j++) {
      const uint32_t e = v * N + j;
      xv[k][j] = (v < v1 && e >= base && e < base + n) ? x[vc[k][j]] : (T)0;
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < K; k++) {
    const uint32_t v = v0 + tid + k * AD_T;
#pragma unroll
    for (uint32_t j = 0; j < N; j++) {
      const uint32_t e = v * N + j;
      if (v < v1 && e >= base && e < base + n) cache[e - base] = va[k][j] * xv[k][j];
    }
  }
  // block-relative row starts rs[0..nr]
  const uint32_t nr = r1 - r0;
  for (uint32_t lr = tid; lr <= nr; lr += AD_T) rs[lr] = rp[r0 + lr] - base;
  __syncthreads();
  for (uint32_t lr = tid; lr < nr; lr += AD_T)
    if (rs[lr] == rs[lr + 1]) y[r0 + lr] = (T)0; // empty rows

  // lane's run [e0, e1): sums split at row boundaries
  const uint32_t E = (nnz_per_wg + AD_T - 1) / AD_T;
  const uint32_t e0 = tid * E;
  const uint32_t e1 = min(n, e0 + E);
  T carry_v = 0;       // partial carried into the next lane
  uint32_t carry_f = 1; // 1: no chain comes in from the left
  uint32_t frow = 0xFFFFFFFFu;
  T fval = 0;
  if (e0 < e1) {
    uint32_t lo = 0, hi = nr; // rs[lo] <= e0 < rs[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (rs[mid] <= e0)
        lo = mid;
      else
        hi = mid;
    }
    uint32_t cur = lo;
    uint32_t cur_start = rs[cur], cur_end = rs[cur + 1];
    uint32_t seg = e0;
    T d = 0;
    for (uint32_t e = e0; e < e1; e++) {
      while (cur_end <= e) {
        if (seg < e) {
          if (seg == cur_start) {
            y[r0 + cur] = d;
          } else {
            frow = cur;
            fval = d;
          }
        }
        d = 0;
        seg = e;
        cur++;
        cur_start = cur_end;
        cur_end = rs[cur + 1];
      }
      d = d + cache[e];
    }
    const bool starts = seg == cur_start;
    if (e1 == cur_end) {
      if (starts) {
        y[r0 + cur] = d;
      } else {
        frow = cur;
        fval = d;
      }
    } else {
      carry_v = d;
      carry_f = starts ? 1u : 0u;
    }
  }
  // segmented inclusive scan of the carried partials (Hillis-Steele)
  sv[tid] = carry_v;
  sf[tid] = carry_f;
  __syncthreads();
  for (uint32_t off = 1; off < AD_T; off <<= 1) {
    const bool take = sf[tid] == 0u && tid >= off;
    T pv = 0;
    uint32_t pf = 0;
    if (take) {
      pv = sv[tid - off];
      pf = sf[tid - off];
    }
    __syncthreads();
    if (take) {
      sv[tid] = pv + sv[tid];
      sf[tid] = pf;
    }
    __syncthreads();
  }
  if (frow != 0xFFFFFFFFu) y[r0 + frow] = sv[tid - 1] + fval;
}

#define ADAPTIVE_SMEM(T)                                                                       \
  extern __shared__ __align__(16) unsigned char smem_raw[];                                    \
  __shared__ T wsum_[4];                                                                       \
  __shared__ T sv_[AD_T];                                                                      \
  __shared__ uint32_t sf_[AD_T];                                                               \
  adaptive_smem<T> sm{reinterpret_cast<T *>(smem_raw),                                         \
                      reinterpret_cast<uint32_t *>(smem_raw + ((nnz_per_wg * sizeof(T) + 15u) & ~15u)), \
                      wsum_, sv_, sf_};

template <typename T, uint32_t K>
__global__ __launch_bounds__(AD_T) void k_csr_adaptive(
    const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci, const T *__restrict__ a,
    const T *__restrict__ x, T *__restrict__ y, const uint32_t *__restrict__ blocks,
    uint32_t nnz_per_wg, uint32_t long_row, const uint4 *__restrict__ chunks, uint32_t nchunks,
    uint32_t nchunks_pad, uint32_t chunk, T *__restrict__ partials)
{
  ADAPTIVE_SMEM(T)
  adaptive_item<T, K>(blockIdx.x, gridDim.x, rp, ci, a, x, y, blocks, nnz_per_wg, long_row, chunks,
                      nchunks, nchunks_pad, chunk, partials, sm);
}

// Persistent form: a fixed grid (a multiple of 8, so a workgroup keeps its
// XCD) walks the items; used for batched launches.
template <typename T, uint32_t K>
__global__ __launch_bounds__(AD_T) void k_csr_adaptive_persistent(
    uint32_t items, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
    const T *__restrict__ a, const T *__restrict__ x, T *__restrict__ y,
    const uint32_t *__restrict__ blocks, uint32_t nnz_per_wg, uint32_t long_row,
    const uint4 *__restrict__ chunks, uint32_t nchunks, uint32_t nchunks_pad, uint32_t chunk,
    T *__restrict__ partials)
{
  ADAPTIVE_SMEM(T)
  for (uint32_t hw = blockIdx.x; hw < items; hw += gridDim.x) {
    adaptive_item<T, K>(hw, items, rp, ci, a, x, y, blocks, nnz_per_wg, long_row, chunks, nchunks,
                        nchunks_pad, chunk, partials, sm);
    __syncthreads();
  }
}

// One lane per long row: add its pieces left to right.
// p[0] + p[stride] + ... + p[(n - 1) stride], added left to right (n >= 1).
// Loads go out eight at a time with clamped indices, so none waits on the
// sum before it is issued; the additions keep the order.
template <typename T>
__device__ __forceinline__ T sum_in_order(const T *__restrict__ p, uint32_t n, uint64_t stride)
{
  T d = p[0];
  for (uint32_t k = 1; k < n; k += 8) {
    T q[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) q[j] = p[(uint64_t)min(k + j, n - 1) * stride];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) d = (k + j < n) ? d + q[j] : d;
  }
  return d;
}

template <typename T>
__global__ void k_adaptive_finish(const uint4 *__restrict__ chunks, uint32_t nchunks,
                                  const T *__restrict__ partials, T *__restrict__ y)
{
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nchunks) return;
  const uint4 c4 = chunks[j];
  if (c4.y != 0) return;
  y[c4.x] = sum_in_order(partials + c4.z, c4.w, 1);
}

// ---------------------------------------------------------------------------
// ELL: one lane per row, data[row + k*rows] (coalesced across lanes by
// construction), sequential fma over the padded width
// (gpu_matrix_multiplier.cu:450-471).  ATOMIC adds instead of storing (the
// fused hybrid).
// ---------------------------------------------------------------------------
template <typename T, bool ATOMIC>
__device__ __forceinline__ void ell_row(uint32_t r, uint32_t rows, uint32_t w,
                                        const uint32_t *__restrict__ ec, const T *__restrict__ ev,
                                        const T *__restrict__ x, T *__restrict__ y)
{
  T d = 0;
  uint64_t o = r;
  // groups of 8 slots, then 4 / 2 / 1: each group's loads are all in flight
  // before the first fma, and the fma order stays slot order
  auto group = [&](auto n_const) {
    constexpr uint32_t N = decltype(n_const)::value;
    T a[N];
    uint32_t c[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) {
      a[q] = ldg_stream(ev + o + (uint64_t)q * rows);
      c[q] = ldg_stream(ec + o + (uint64_t)q * rows);
    }
    T xv[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) xv[q] = x[c[q]];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) d = fma_t(a[q], xv[q], d);
    o += (uint64_t)N * rows;
  };
  uint32_t k = 0;
  for (; k + 8 <= w; k += 8) group(std::integral_constant<uint32_t, 8>{});
  if (k + 4 <= w) {
    group(std::integral_constant<uint32_t, 4>{});
    k += 4;
  }
  if (k + 2 <= w) {
    group(std::integral_constant<uint32_t, 2>{});
    k += 2;
  }
  if (k < w) group(std::integral_constant<uint32_t, 1>{});
  if constexpr (ATOMIC)
    unsafeAtomicAdd(y + r, d);
  else
    y[r] = d;
}

template <typename T>
__global__ __launch_bounds__(256) void k_ell(uint32_t rows, uint32_t w, const uint32_t *__restrict__ ec,
                                             const T *__restrict__ ev, const T *__restrict__ x,
                                             T *__restrict__ y)
{
  const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (r < rows) ell_row<T, false>(r, rows, w, ec, ev, x, y);
}

// FP32 / FP64 ELL over a 64-row sliced layout with 16-bit column offsets
// (plans created with MFP_TUNE_ELL_PACKED).  Slots 4j .. 4j + 3 of row r are
// value word j (four values) and offset word j (four u16); with
// u = (r / 64) w4 + j, the offset word is at u 64 + r % 64, and the value
// word is a float4 at u 64 + r % 64 (FP32) or two double2 halves, slots
// 4j, 4j + 1 at 2u 64 + r % 64 and 4j + 2, 4j + 3 at (2u + 1) 64 + r % 64
// (FP64), so every 16-byte load of a wave reads 1 KiB contiguous.  A slot's
// column is base[r / 64] + its offset.  A slot takes 6 bytes instead of 8
// (FP64: 10 instead of 12).  The fma order is k_ell's, slots 0 .. w - 1 (the
// slots past w that fill the last word are skipped); a padding slot adds
// 0 x x[base].
typedef uint16_t pk_u16x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename T> struct pk_word;
template <> struct pk_word<float> {
  typedef f32x4 type;
  typedef f32x4 unit; // the array's element: one word
  static __device__ __forceinline__ type load(const unit *__restrict__ v, uint64_t u, uint32_t lr)
  {
    return __builtin_nontemporal_load(v + u * 64u + lr);
  }
};
template <> struct pk_word<double> {
  typedef f64x4 type;
  typedef f64x2 unit; // the array's element: half a word
  static __device__ __forceinline__ type load(const unit *__restrict__ v, uint64_t u, uint32_t lr)
  {
    const f64x2 lo = __builtin_nontemporal_load(v + u * 128u + lr);
    const f64x2 hi = __builtin_nontemporal_load(v + u * 128u + 64u + lr);
    return type{lo.x, lo.y, hi.x, hi.y};
  }
};

template <typename T>
__global__ __launch_bounds__(256) void k_ell_pk(uint32_t rows, uint32_t w, uint32_t w4,
                                                const uint32_t *__restrict__ base,
                                                const typename pk_word<T>::unit *__restrict__ val,
                                                const pk_u16x4 *__restrict__ off, const T *__restrict__ x,
                                                T *__restrict__ y)
{
  typedef typename pk_word<T>::type V;
  const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (r >= rows) return;
  const uint32_t blk = r >> 6, lr = r & 63u;
  const T *__restrict__ xb = x + base[blk];
  T d = 0;
  // N words of 4 slots, every load of the group in flight before the first fma
  auto group = [&](uint32_t j, auto n_const) {
    constexpr uint32_t N = decltype(n_const)::value;
    pk_u16x4 o[N];
    V v[N];
#pragma unroll
    for (uint32_t q = 0; q < N; q++) {
      const uint64_t u = (uint64_t)blk * w4 + j + q;
      o[q] = __builtin_nontemporal_load(off + u * 64u + lr);
      v[q] = pk_word<T>::load(val, u, lr);
    }
    T xv[N][4];
#pragma unroll
    for (uint32_t q = 0; q < N; q++)
#pragma unroll
      for (uint32_t e = 0; e < 4; e++) xv[q][e] = xb[o[q][e]];
#pragma unroll
    for (uint32_t q = 0; q < N; q++)
#pragma unroll
      for (uint32_t e = 0; e < 4; e++)
        if (4u * (j + q) + e < w) d = fma_t(v[q][e], xv[q][e], d);
  };
  uint32_t j = 0;
  // a row of 5 to 8 words (up to 32 slots) is one group, every load of the
  // row in flight before the first gather (3D 27-point stencil, 7 words:
  // FP32 62.2 against 64.0 us in groups of 4, 2 and 1, FP64 111.5 against
  // 113.5 us, though FP64 then holds 148 VGPRs, 3 waves per SIMD)
  if (w4 >= 5 && w4 <= 8) {
    switch (w4) {
    case 8: group(0, std::integral_constant<uint32_t, 8>{}); break;
    case 7: group(0, std::integral_constant<uint32_t, 7>{}); break;
    case 6: group(0, std::integral_constant<uint32_t, 6>{}); break;
    default: group(0, std::integral_constant<uint32_t, 5>{}); break;
    }
    j = w4;
  }
  for (; j + 4 <= w4; j += 4) group(j, std::integral_constant<uint32_t, 4>{});
  if (j + 2 <= w4) {
    group(j, std::integral_constant<uint32_t, 2>{});
    j += 2;
  }
  if (j < w4) group(j, std::integral_constant<uint32_t, 1>{});
  y[r] = d;
}

// The packed layout of k_ell_pk from CSR: a thread per row (rows rounded up
// to 64; the extra rows and the slots past a row's length hold value 0,
// offset 0).
template <typename T>
__global__ void k_ell_pk_build(uint32_t rows, uint32_t stride, const uint32_t *__restrict__ rp,
                               const uint32_t *__restrict__ ci, const T *__restrict__ a, uint32_t w4,
                               const uint32_t *__restrict__ base, T *__restrict__ val,
                               uint16_t *__restrict__ off)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= stride) return;
  const uint32_t s = r < rows ? rp[r] : 0u, len = r < rows ? rp[r + 1] - s : 0u;
  const uint32_t blk = r >> 6, lr = r & 63u, b0 = base[blk];
  for (uint32_t k = 0; k < 4u * w4; k++) {
    const bool in = k < len;
    const uint64_t u = (uint64_t)blk * w4 + k / 4u;
    off[(u * 64u + lr) * 4u + k % 4u] = in ? (uint16_t)(ci[s + k] - b0) : (uint16_t)0;
    // FP32: word u of row lr is one float4; FP64: two double2 halves
    const uint64_t at = sizeof(T) == 4 ? (u * 64u + lr) * 4u + k % 4u
                                       : ((u * 2u + (k % 4u) / 2u) * 64u + lr) * 2u + k % 2u;
    val[at] = in ? a[s + k] : (T)0;
  }
}

// y += t, element-wise (the host-COO hybrid's merge of the tail sums,
// gpu_matrix_multiplier.cu:1098-1108)
template <typename T>
__global__ __launch_bounds__(256) void k_vec_add(uint32_t n, const T *__restrict__ t, T *__restrict__ y)
{
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (i < n) y[i] = y[i] + t[i];
}

// ---------------------------------------------------------------------------
// COO (gpu_matrix_multiplier.cu:622-638): a wave takes 4 x 64 consecutive
// (row-sorted) entries, all loads in flight together.  Per 64-entry tile a
// segmented suffix sum over equal row keys leaves each run's total in its
// first lane, which issues one atomic add.
// ---------------------------------------------------------------------------
constexpr uint32_t COO_TILES = 4;

template <typename T>
__device__ __forceinline__ void coo_segment_add(uint32_t r, T p, bool valid, T *__restrict__ y)
{
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t r2 = __shfl_down(r, off, 64);
    const T p2 = __shfl_down(p, off, 64);
    if (lane + off < 64u && r2 == r) p = p + p2;
  }
  const uint32_t rprev = __shfl_up(r, 1, 64);
  if (valid && (lane == 0 || rprev != r)) unsafeAtomicAdd(y + r, p);
}

template <typename T>
__device__ __forceinline__ void coo_wave(uint64_t base, uint64_t n, const uint32_t *__restrict__ cr,
                                         const uint32_t *__restrict__ cc, const T *__restrict__ cv,
                                         const T *__restrict__ x, T *__restrict__ y)
{
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t r[COO_TILES], c[COO_TILES];
  T v[COO_TILES], xv[COO_TILES];
#pragma unroll
  for (uint32_t k = 0; k < COO_TILES; k++) {
    const uint64_t i = base + 64u * k + lane;
    const bool ok = i < n;
    r[k] = ok ? ldg_stream(cr + i) : 0xFFFFFFFFu;
    c[k] = ok ? ldg_stream(cc + i) : 0u;
    v[k] = ok ? ldg_stream(cv + i) : (T)0;
  }
#pragma unroll
  for (uint32_t k = 0; k < COO_TILES; k++) xv[k] = (base + 64u * k + lane < n) ? x[c[k]] : (T)0;
#pragma unroll
  for (uint32_t k = 0; k < COO_TILES; k++)
    coo_segment_add(r[k], v[k] * xv[k], base + 64u * k + lane < n, y);
}

template <typename T>
__global__ __launch_bounds__(256) void k_coo(uint64_t n, const uint32_t *__restrict__ cr,
                                             const uint32_t *__restrict__ cc,
                                             const T *__restrict__ cv, const T *__restrict__ x,
                                             T *__restrict__ y)
{
  const uint64_t span = 64u * COO_TILES;
  const uint64_t waves = (uint64_t)gridDim.x * 4u;
  uint64_t w = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
  for (; w * span < n; w += waves) coo_wave(w * span, n, cr, cc, cv, x, y);
}

// Fused hybrid (gpu_matrix_multiplier.cu:969-1002): workgroups [0, ell_blocks)
// run ELL rows with atomic adds, the rest run the COO tail.  y must be zero.
template <typename T>
__global__ __launch_bounds__(256) void k_hyb_atomic(uint32_t rows, uint32_t w,
                                                    const uint32_t *__restrict__ ec,
                                                    const T *__restrict__ ev, uint64_t ncoo,
                                                    const uint32_t *__restrict__ cr,
                                                    const uint32_t *__restrict__ cc,
                                                    const T *__restrict__ cv,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    uint32_t ell_blocks)
{
  const uint32_t b = blockIdx.x;
  if (b < ell_blocks) {
    const uint32_t r = b * 256u + threadIdx.x;
    if (r < rows) ell_row<T, true>(r, rows, w, ec, ev, x, y);
    return;
  }
  const uint64_t span = 64u * COO_TILES;
  const uint64_t waves = (uint64_t)(gridDim.x - ell_blocks) * 4u;
  uint64_t wv = (uint64_t)(b - ell_blocks) * 4u + (threadIdx.x >> 6);
  for (; wv * span < ncoo; wv += waves) coo_wave(wv * span, ncoo, cr, cc, cv, x, y);
}

// ---------------------------------------------------------------------------
// SCOO (scoo_spmv.cu:46-129): one workgroup per slice of slice_size rows.
// Products are added with LDS atomics (ds_add) into acc[row][lane], lane =
// thread % lane_size, then each row's lanes are summed in order.  FP32
// values accumulate in double (A) when acc[slice_size][lane_size] fits in
// LDS that way: the float product is exact in double, and gfx950's
// ds_add_f64 sustains several times the lane rate of ds_add_f32.
// ---------------------------------------------------------------------------
template <typename T, typename A>
__global__ void k_scoo(uint32_t rows, uint32_t ss, uint32_t L, const uint32_t *__restrict__ cr,
                       const uint32_t *__restrict__ cc, const T *__restrict__ cv,
                       const uint32_t *__restrict__ sp, const T *__restrict__ x,
                       T *__restrict__ y)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  A *acc = reinterpret_cast<A *>(smem_raw);
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t r0 = s * ss;
  const uint32_t lim = min(ss, rows - r0);
  const uint32_t bt = blockDim.x, tid = threadIdx.x;
  for (uint32_t i = tid; i < lim * L; i += bt) acc[i] = (A)0;
  __syncthreads();
  const uint32_t b = sp[s], e = sp[s + 1];
  const uint32_t lane = tid & (L - 1u);
  uint32_t i = b + tid;
  for (; i + 3 * bt < e; i += 4 * bt) {
    uint32_t lr[4], c[4];
    T v[4], xv[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      lr[k] = ldg_stream(cr + i + k * bt) - r0;
      c[k] = ldg_stream(cc + i + k * bt);
      v[k] = ldg_stream(cv + i + k * bt);
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) xv[k] = x[c[k]];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
      if (lr[k] < lim) atomicAdd(acc + lr[k] * L + lane, (A)v[k] * (A)xv[k]);
  }
  for (; i < e; i += bt) {
    const uint32_t lr = ldg_stream(cr + i) - r0;
    const A p = (A)ldg_stream(cv + i) * (A)x[ldg_stream(cc + i)];
    if (lr < lim) atomicAdd(acc + lr * L + lane, p);
  }
  __syncthreads();
  for (uint32_t lr = tid; lr < lim; lr += bt) {
    A d = 0;
    for (uint32_t l = 0; l < L; l++) d = d + acc[lr * L + l];
    y[r0 + lr] = (T)d;
  }
}

// ---------------------------------------------------------------------------
// Column-sorted SCOO: one workgroup per work item.  The item's slice rows
// accumulate in LDS in double for either value type: a product of two floats
// is exact in double, and on gfx950 ds_add_f64 sustains about 7x the lane
// rate of ds_add_f32 (2.3 against 0.33 lanes per clock per CU with random
// addresses), the rate that bounds this kernel.  Its entries come in column order, so the x
// reads sweep one column range and consecutive lanes share x lines.  Each
// aligned block of 256 positions inside an item is stored word-interleaved:
// a lane's 16-byte load brings positions l, 64 + l, 128 + l and 192 + l, so
// every gather instruction of a wave covers 64 consecutive sorted columns
// (a few x lines) instead of 256 (several times as many L1 lines).  Two
// words per lane are in flight; the unaligned head and tail go one entry
// per lane, in position order.  A slice's
// single item writes y; several write partial slots that k_scoo_cs_finish
// adds in slot order.  Long-row slices work the same way, their local rows
// mapped through long_rows; ordinary slices skip the long rows.  Each y entry
// is written once, with no global atomics (the order of the LDS additions
// still varies from run to run).
// ---------------------------------------------------------------------------
// entry word: (column - item base) << rbits | local row
__device__ __forceinline__ void cs_decode(uint32_t w, uint32_t cbase, uint32_t rbits, uint32_t &c, uint32_t &lr)
{
  c = cbase + (w >> rbits);
  lr = w & ((1u << rbits) - 1u);
}

__device__ __forceinline__ bool cs_is_long(const uint32_t *__restrict__ mask, uint32_t row)
{
  return mask && ((mask[row >> 5] >> (row & 31u)) & 1u);
}

// Stored values of a column-sorted plan: the vector type T itself (DT < 0)
// or 16-bit BF16 / FP16 (DT = MFP_BF16 / MFP_F16, with T = float).  Every
// product is formed in double, exactly: a float, BF16 or FP16 value times a
// float x needs at most 48 significant bits.
template <typename T, int DT>
struct cs_values {
  using V = typename std::conditional<(DT < 0), T, uint16_t>::type;
  __device__ __forceinline__ static double one(const V *__restrict__ ev, uint32_t i)
  {
    if constexpr (DT < 0)
      return (double)ldg_stream(ev + i);
    else
      return (double)lp_val<DT>(ldg_stream(ev + i));
  }
  // the four values of 16-byte entry word g
  __device__ __forceinline__ static void word(const V *__restrict__ ev, uint32_t g, double (&v)[4])
  {
    if constexpr (DT >= 0) {
      const u32x2 h = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(ev) + g);
      v[0] = (double)lp_val<DT>((uint16_t)(h[0] & 0xFFFFu));
      v[1] = (double)lp_val<DT>((uint16_t)(h[0] >> 16));
      v[2] = (double)lp_val<DT>((uint16_t)(h[1] & 0xFFFFu));
      v[3] = (double)lp_val<DT>((uint16_t)(h[1] >> 16));
    } else if constexpr (sizeof(T) == 4) {
      const f32x4 f = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(ev) + g);
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) v[j] = (double)f[j];
    } else {
      // FP64: slots 0, 1 of the block's 64 words, then slots 2, 3 (word g
      // is word g % 64 of block g / 64)
      const f64x2 *p = reinterpret_cast<const f64x2 *>(ev) + (g >> 6) * 128u + (g & 63u);
      const f64x2 d0 = __builtin_nontemporal_load(p);
      const f64x2 d1 = __builtin_nontemporal_load(p + 64);
      v[0] = d0[0];
      v[1] = d0[1];
      v[2] = d1[0];
      v[3] = d1[1];
    }
  }
};

template <typename T, int DT>
__global__ void k_scoo_cs(uint32_t rows, uint32_t R, uint32_t rbits, uint32_t S_reg, uint32_t NL,
                          const uint4 *__restrict__ items, const uint32_t *__restrict__ ibase,
                          const typename cs_values<T, DT>::V *__restrict__ ev, const uint32_t *__restrict__ ec,
                          const T *__restrict__ x, T *__restrict__ y,
                          T *__restrict__ partial, const uint32_t *__restrict__ long_rows,
                          const uint32_t *__restrict__ long_mask)
{
  using CV = cs_values<T, DT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *acc = reinterpret_cast<double *>(smem_raw);
  const uint4 it = items[xcd_block(blockIdx.x, gridDim.x)];
  const uint32_t sl = it.x, e0 = it.y, e1 = it.z, slot = it.w;
  const uint32_t lim = sl < S_reg ? min(R, rows - sl * R) : min(R, NL - (sl - S_reg) * R);
  const uint32_t bt = blockDim.x, tid = threadIdx.x;
  for (uint32_t i = tid; i < lim; i += bt) acc[i] = 0.0;
  __syncthreads();
  const uint32_t cbase = ibase[xcd_block(blockIdx.x, gridDim.x)];
  // head: up to the first multiple of 256 (positions stored in order)
  const uint32_t ea = (uint32_t)min<uint64_t>(e1, ((uint64_t)e0 + 255u) & ~(uint64_t)255);
  for (uint32_t i = e0 + tid; i < ea; i += bt) {
    uint32_t c, lr;
    cs_decode(ldg_stream(ec + i), cbase, rbits, c, lr);
    atomicAdd(acc + lr, CV::one(ev, i) * (double)x[c]);
  }
  // body: whole blocks of 256 positions, stored so that 16-byte word l of a
  // block holds positions l, 64 + l, 128 + l, 192 + l; two words per lane in
  // flight.  The decode does not depend on which positions a word holds.
  const uint32_t g0 = ea >> 2, g1 = max(g0, (e1 & ~255u) >> 2); // words [g0, g1) lie inside [ea, e1)
  uint32_t g = g0 + tid;
  for (; g + bt < g1; g += 2 * bt) {
    u32x4 w[2];
    double v[2][4];
#pragma unroll
    for (uint32_t k = 0; k < 2; k++) {
      w[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(ec) + g + k * bt);
      CV::word(ev, g + k * bt, v[k]);
    }
    uint32_t c[2][4], lr[2][4];
    T xv[2][4];
#pragma unroll
    for (uint32_t k = 0; k < 2; k++)
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) cs_decode(w[k][j], cbase, rbits, c[k][j], lr[k][j]);
#pragma unroll
    for (uint32_t k = 0; k < 2; k++)
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) xv[k][j] = x[c[k][j]];
#pragma unroll
    for (uint32_t k = 0; k < 2; k++)
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) atomicAdd(acc + lr[k][j], v[k][j] * (double)xv[k][j]);
  }
  for (; g < g1; g += bt) {
    const u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(ec) + g);
    double v[4];
    CV::word(ev, g, v);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      uint32_t c, lr;
      cs_decode(w[j], cbase, rbits, c, lr);
      atomicAdd(acc + lr, v[j] * (double)x[c]);
    }
  }
  // tail: entries after the last whole group
  for (uint32_t i = max(ea, g1 << 2) + tid; i < e1; i += bt) {
    uint32_t c, lr;
    cs_decode(ldg_stream(ec + i), cbase, rbits, c, lr);
    atomicAdd(acc + lr, CV::one(ev, i) * (double)x[c]);
  }
  __syncthreads();
  if (slot == CS_DIRECT && sl >= S_reg) {
    const uint32_t *lrows = long_rows + (uint64_t)(sl - S_reg) * R;
    for (uint32_t lr = tid; lr < lim; lr += bt) y[lrows[lr]] = (T)acc[lr];
  } else if (slot == CS_DIRECT) {
    const uint32_t r0 = sl * R;
    for (uint32_t lr = tid; lr < lim; lr += bt)
      if (!cs_is_long(long_mask, r0 + lr)) y[r0 + lr] = (T)acc[lr];
  } else {
    T *o = partial + (uint64_t)slot * R;
    for (uint32_t lr = tid; lr < lim; lr += bt) o[lr] = (T)acc[lr];
  }
}

// four consecutive values p[0..3] (16-byte aligned)
template <typename T>
__device__ __forceinline__ void load4(const T *__restrict__ p, T (&v)[4])
{
  if constexpr (sizeof(T) == 4) {
    const f32x4 a = *reinterpret_cast<const f32x4 *>(p);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) v[j] = a[j];
  } else {
    const f64x2 a = *reinterpret_cast<const f64x2 *>(p);
    const f64x2 b = *reinterpret_cast<const f64x2 *>(p + 2);
    v[0] = a[0];
    v[1] = a[1];
    v[2] = b[0];
    v[3] = b[1];
  }
}

// sum_in_order for four adjacent rows at once: d[j] = p[j] + p[stride + j]
// + ... in slot order, with 16-byte loads (p 16-byte aligned, stride a
// multiple of 4)
template <typename T>
__device__ __forceinline__ void sum4_in_order(const T *__restrict__ p, uint32_t n, uint64_t stride, T (&d)[4])
{
  load4(p, d);
  for (uint32_t k = 1; k < n; k += 4) {
    T q[4][4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) load4(p + (uint64_t)min(k + i, n - 1) * stride, q[i]);
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) d[j] = (k + i < n) ? d[j] + q[i][j] : d[j];
  }
}

// One workgroup per finish block.  Blocks below NFR: 4 CS_FIN_ROWS rows of
// an ordinary slice, four consecutive rows per thread, y[row] = the slice's
// slots added in slot order (long rows skipped).  The others: CS_FIN_LONG rows of a long-row
// slice, a wave per row; lane l adds slots l, l + 64, ... in order and the
// wave's halving tree adds the lanes, into y[long_rows[(slice - S_reg) R + lr]].
template <typename T>
__global__ __launch_bounds__(CS_FIN_ROWS) void k_scoo_cs_finish(uint32_t rows, uint32_t R, uint32_t S_reg,
                                                                uint32_t NL, uint32_t NFR,
                                                                const uint4 *__restrict__ fin,
                                                                const uint32_t *__restrict__ long_rows,
                                                                const uint32_t *__restrict__ long_mask,
                                                                const T *__restrict__ partial, T *__restrict__ y)
{
  const uint4 f = fin[blockIdx.x];
  const uint32_t sl = f.x, slot0 = f.y, n = f.z;
  if (blockIdx.x >= NFR) {
    const uint32_t lr = f.w + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (lr >= min(R, NL - (sl - S_reg) * R)) return; // uniform across the wave
    // the destination row and up to 8 of the lane's slots are loaded together
    const uint32_t dst = long_rows[(uint64_t)(sl - S_reg) * R + lr];
    const uint32_t cnt = lane < n ? (n - lane + 63u) >> 6 : 0u;
    T d = cnt ? sum_in_order(partial + (uint64_t)(slot0 + lane) * R + lr, cnt, (uint64_t)64 * R) : (T)0;
    d = group_reduce<64>(d);
    if (lane == 0) y[dst] = d;
    return;
  }
  const uint32_t lim = min(R, rows - sl * R);
  const uint32_t lr0 = f.w + 4 * threadIdx.x;
  if (lr0 >= lim || threadIdx.x >= CS_FIN_ROWS) return;
  const T *p = partial + (uint64_t)slot0 * R + lr0;
  const uint32_t row0 = sl * R + lr0;
  T d[4];
  uint32_t lm = 0; // bit j: row0 + j is a long row (its y comes from a long-row slice)
  if ((R & 3u) == 0 && lr0 + 4 <= lim) {
    // row0 is a multiple of 4, so its four mask bits share one word
    if (long_mask) lm = (long_mask[row0 >> 5] >> (row0 & 31u)) & 15u;
    sum4_in_order(p, n, R, d);
  } else {
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      d[j] = lr0 + j < lim ? sum_in_order(p + j, n, R) : (T)0;
      lm |= (lr0 + j < lim && cs_is_long(long_mask, row0 + j)) ? 1u << j : 0u;
    }
  }
  if (lm == 0 && lr0 + 4 <= lim && (reinterpret_cast<uintptr_t>(y + row0) & 15u) == 0) {
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<f32x4 *>(y + row0) = f32x4{d[0], d[1], d[2], d[3]};
    } else {
      reinterpret_cast<f64x2 *>(y + row0)[0] = f64x2{d[0], d[1]};
      reinterpret_cast<f64x2 *>(y + row0)[1] = f64x2{d[2], d[3]};
    }
    return;
  }
#pragma unroll
  for (uint32_t j = 0; j < 4; j++)
    if (lr0 + j < lim && !((lm >> j) & 1u)) y[row0 + j] = d[j];
}

// ---------------------------------------------------------------------------
// Column-sorted SCOO SpMM, Y = A X for KV columns (X cols x KV, Y rows x KV,
// row-major): the SpMV kernel above with KV accumulators per slice row
// (acc[column][row], doubles, R KV 8 bytes of LDS: lanes adding one column
// for random rows then spread over the banks, where acc[row][column] would
// put them KV doubles apart) and one X row (KV values, 16-byte loads) per
// entry.  Partial slots hold R KV values, row-major like Y.
// ---------------------------------------------------------------------------
template <typename T, int KV>
__global__ void k_scoo_cs_k(uint32_t rows, uint32_t R, uint32_t rbits, uint32_t S_reg, uint32_t NL,
                            const uint4 *__restrict__ items, const uint32_t *__restrict__ ibase,
                            const T *__restrict__ ev, const uint32_t *__restrict__ ec, const T *__restrict__ X,
                            T *__restrict__ Y, T *__restrict__ partial, const uint32_t *__restrict__ long_rows,
                            const uint32_t *__restrict__ long_mask)
{
  using CV = cs_values<T, -1>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *acc = reinterpret_cast<double *>(smem_raw);
  const uint4 it = items[xcd_block(blockIdx.x, gridDim.x)];
  const uint32_t sl = it.x, e0 = it.y, e1 = it.z, slot = it.w;
  const uint32_t lim = sl < S_reg ? min(R, rows - sl * R) : min(R, NL - (sl - S_reg) * R);
  const uint32_t bt = blockDim.x, tid = threadIdx.x;
  for (uint32_t i = tid; i < R * KV; i += bt) acc[i] = 0.0;
  __syncthreads();
  const uint32_t cbase = ibase[xcd_block(blockIdx.x, gridDim.x)];
  auto add_row = [&](uint32_t c, uint32_t lr, double v) {
    T xr[KV];
    load_row<T, KV>(X + (uint64_t)c * KV, xr);
#pragma unroll
    for (int j = 0; j < KV; j++) atomicAdd(acc + j * R + lr, v * (double)xr[j]);
  };
  const uint32_t ea = (uint32_t)min<uint64_t>(e1, ((uint64_t)e0 + 255u) & ~(uint64_t)255);
  for (uint32_t i = e0 + tid; i < ea; i += bt) {
    uint32_t c, lr;
    cs_decode(ldg_stream(ec + i), cbase, rbits, c, lr);
    add_row(c, lr, CV::one(ev, i));
  }
  // body: one 16-byte entry word (four entries) per lane per step
  const uint32_t g0 = ea >> 2, g1 = max(g0, (e1 & ~255u) >> 2);
  for (uint32_t g = g0 + tid; g < g1; g += bt) {
    const u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(ec) + g);
    double v[4];
    CV::word(ev, g, v);
    uint32_t c[4], lr[4];
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) cs_decode(w[j], cbase, rbits, c[j], lr[j]);
    T xr[4][KV];
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) load_row<T, KV>(X + (uint64_t)c[j] * KV, xr[j]);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < KV; q++) atomicAdd(acc + q * R + lr[j], v[j] * (double)xr[j][q]);
  }
  for (uint32_t i = max(ea, g1 << 2) + tid; i < e1; i += bt) {
    uint32_t c, lr;
    cs_decode(ldg_stream(ec + i), cbase, rbits, c, lr);
    add_row(c, lr, CV::one(ev, i));
  }
  __syncthreads();
  if (slot == CS_DIRECT && sl >= S_reg) {
    const uint32_t *lrows = long_rows + (uint64_t)(sl - S_reg) * R;
    for (uint32_t i = tid; i < lim * KV; i += bt)
      Y[(uint64_t)lrows[i / KV] * KV + i % KV] = (T)acc[(i % KV) * R + i / KV];
  } else if (slot == CS_DIRECT) {
    const uint64_t r0 = (uint64_t)sl * R;
    for (uint32_t i = tid; i < lim * KV; i += bt)
      if (!cs_is_long(long_mask, (uint32_t)(r0 + i / KV))) Y[r0 * KV + i] = (T)acc[(i % KV) * R + i / KV];
  } else {
    T *o = partial + (uint64_t)slot * R * KV;
    for (uint32_t i = tid; i < lim * KV; i += bt) o[i] = (T)acc[(i % KV) * R + i / KV];
  }
}

// The SpMM finish: as k_scoo_cs_finish, over (row, column) pairs; each
// pair's slots are added in slot order.
template <typename T>
__global__ __launch_bounds__(CS_FIN_ROWS) void k_scoo_cs_finish_k(uint32_t rows, uint32_t R, uint32_t KV,
                                                                  uint32_t S_reg, uint32_t NL, uint32_t NFR,
                                                                  const uint4 *__restrict__ fin,
                                                                  const uint32_t *__restrict__ long_rows,
                                                                  const uint32_t *__restrict__ long_mask,
                                                                  const T *__restrict__ partial,
                                                                  T *__restrict__ Y)
{
  const uint4 f = fin[blockIdx.x];
  const uint32_t sl = f.x, slot0 = f.y, n = f.z;
  const uint64_t stride = (uint64_t)R * KV;
  const T *p = partial + (uint64_t)slot0 * stride;
  if (blockIdx.x >= NFR) {
    const uint32_t lim = min(R, NL - (sl - S_reg) * R);
    for (uint32_t i = threadIdx.x; i < CS_FIN_LONG * KV; i += CS_FIN_ROWS) {
      const uint32_t lr = f.w + i / KV, j = i % KV;
      if (lr < lim) Y[(uint64_t)long_rows[(uint64_t)(sl - S_reg) * R + lr] * KV + j] = sum_in_order(p + (uint64_t)lr * KV + j, n, stride);
    }
    return;
  }
  const uint32_t lim = min(R, rows - sl * R);
  for (uint32_t i = threadIdx.x; i < 4 * CS_FIN_ROWS * KV; i += CS_FIN_ROWS) {
    const uint32_t lr = f.w + i / KV, j = i % KV;
    const uint32_t row = sl * R + lr;
    if (lr < lim && !cs_is_long(long_mask, row))
      Y[(uint64_t)row * KV + j] = sum_in_order(p + (uint64_t)lr * KV + j, n, stride);
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int launched()
{
  mark_kernel_end();
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

// launch status of a kernel launched with its own start / stop events
static int launched_timed()
{
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP);
}

// Launch a kernel and check the launch.  Under mfp_plan_kernel_times the
// kernel records its own start and end (hipExtLaunchKernelGGL events);
// otherwise it is a plain launch followed by the usual mark.
template <typename... P, typename... A>
static int launch_k(void (*kernel)(P...), dim3 grid, dim3 block, uint32_t lds, hipStream_t s, A... args)
{
  hipEvent_t e0, e1;
  if (mark_kernel_events(&e0, &e1)) {
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0u, static_cast<P>(args)...);
    return launched_timed();
  }
  hipLaunchKernelGGL(kernel, grid, block, lds, s, static_cast<P>(args)...);
  return launched();
}

template <typename T>
static int csr_vector_t(uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                        const T *a, const T *x, T *y, hipStream_t s)
{
  const uint32_t grid = ceil_div_u32((uint64_t)rows * lanes, 256);
  const dim3 g(grid), b(256);
  switch (lanes) {
  case 2: return launch_k(k_csr_vector<T, 2>, g, b, 0, s, rows, rp, ci, a, x, y);
  case 4: return launch_k(k_csr_vector<T, 4>, g, b, 0, s, rows, rp, ci, a, x, y);
  case 8: return launch_k(k_csr_vector<T, 8>, g, b, 0, s, rows, rp, ci, a, x, y);
  case 16: return launch_k(k_csr_vector<T, 16>, g, b, 0, s, rows, rp, ci, a, x, y);
  case 32: return launch_k(k_csr_vector<T, 32>, g, b, 0, s, rows, rp, ci, a, x, y);
  case 64: return launch_k(k_csr_vector<T, 64>, g, b, 0, s, rows, rp, ci, a, x, y);
  default: set_error("lanes must be a power of two in [2, 64]"); return MFP_ERR_ARG;
  }
}

int launch_csr_scalar(int dtype, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                      const void *a, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32(rows, 64);
  if (dtype == MFP_F32)
    return launch_k(k_csr_scalar<float>, dim3(grid), dim3(64), 0, s, rows, rp, ci, (const float *)a,
                    (const float *)x, (float *)y);
  return launch_k(k_csr_scalar<double>, dim3(grid), dim3(64), 0, s, rows, rp, ci, (const double *)a,
                  (const double *)x, (double *)y);
}

int launch_csr_vector(int dtype, uint32_t lanes, uint32_t rows, const uint32_t *rp,
                      const uint32_t *ci, const void *a, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  if (dtype == MFP_F32)
    return csr_vector_t<float>(lanes, rows, rp, ci, (const float *)a, (const float *)x, (float *)y, s);
  return csr_vector_t<double>(lanes, rows, rp, ci, (const double *)a, (const double *)x, (double *)y, s);
}

// Vectors a lane may need for a block of up to W elements at any alignment.
uint32_t adaptive_vectors_per_lane(int dtype, uint32_t W)
{
  const uint32_t N = dtype == MFP_F32 ? 4 : 2;
  const uint32_t maxvec = (W + 2 * N - 2) / N;
  uint32_t k = (maxvec + AD_T - 1) / AD_T;
  uint32_t p = 1;
  while (p < k) p <<= 1;
  return p;
}

template <typename T, uint32_t K>
static int adaptive_launch_k(const adaptive_args &g, const T *a, const T *x, T *y, hipStream_t s,
                             size_t lds, uint32_t items, uint32_t nchunks_pad, uint32_t persistent_grid)
{
  const void *fn = persistent_grid ? (const void *)k_csr_adaptive_persistent<T, K>
                                   : (const void *)k_csr_adaptive<T, K>;
  if (lds > 65536) MFP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (persistent_grid)
    return launch_k(k_csr_adaptive_persistent<T, K>, dim3(persistent_grid), dim3(AD_T), (uint32_t)lds, s, items,
                    g.rp, g.ci, a, x, y, g.blocks, g.nnz_per_wg, g.long_row, g.chunks, g.nchunks, nchunks_pad,
                    g.chunk, (T *)g.partials);
  return launch_k(k_csr_adaptive<T, K>, dim3(items), dim3(AD_T), (uint32_t)lds, s, g.rp, g.ci, a, x, y, g.blocks,
                  g.nnz_per_wg, g.long_row, g.chunks, g.nchunks, nchunks_pad, g.chunk, (T *)g.partials);
}

// Dynamic LDS of k_csr_adaptive: the staged products, then the row starts.
size_t adaptive_lds_bytes(int dtype, uint32_t W)
{
  return ((W * dtype_size(dtype) + 15u) & ~size_t(15)) + ((size_t)W + 2) * 4;
}

template <typename T>
static int adaptive_t(const adaptive_args &g, const T *a, const T *x, T *y, hipStream_t s)
{
  // chunk items first; padded to a multiple of 8 so block ids keep their XCD
  const uint32_t nchunks_pad = (g.nchunks + 7u) & ~7u;
  const uint32_t items = g.nblocks ? g.nblocks + nchunks_pad : nchunks_pad;
  const size_t lds = adaptive_lds_bytes(sizeof(T) == 4 ? MFP_F32 : MFP_F64, g.nnz_per_wg);
  const uint32_t K = adaptive_vectors_per_lane(sizeof(T) == 4 ? MFP_F32 : MFP_F64, g.nnz_per_wg);
  // persistent: 8 workgroups per CU (the LDS/wave occupancy), a multiple of 8
  uint32_t pgrid = 0;
  if (g.persistent) {
    pgrid = std::max<uint32_t>(8, (uint32_t)std::max(1, props().cu_count) * 8u);
    pgrid = std::min<uint32_t>(pgrid, (items + 7u) & ~7u);
    pgrid = std::max<uint32_t>(8, pgrid & ~7u);
  }
  if (items) {
    int st;
    switch (K) {
    case 1: st = adaptive_launch_k<T, 1>(g, a, x, y, s, lds, items, nchunks_pad, pgrid); break;
    case 2: st = adaptive_launch_k<T, 2>(g, a, x, y, s, lds, items, nchunks_pad, pgrid); break;
    case 4: st = adaptive_launch_k<T, 4>(g, a, x, y, s, lds, items, nchunks_pad, pgrid); break;
    case 8: st = adaptive_launch_k<T, 8>(g, a, x, y, s, lds, items, nchunks_pad, pgrid); break;
    case 16: st = adaptive_launch_k<T, 16>(g, a, x, y, s, lds, items, nchunks_pad, pgrid); break;
    default: set_error("nnz_per_wg too large"); return MFP_ERR_ARG;
    }
    MFP_CHECK(st);
  }
  if (g.nchunks)
    MFP_CHECK(launch_k(k_adaptive_finish<T>, dim3(ceil_div_u32(g.nchunks, 256)), dim3(256), 0, s, g.chunks,
                       g.nchunks, (const T *)g.partials, y));
  return MFP_OK;
}

int launch_csr_adaptive(int dtype, const adaptive_args &g, const void *a, const void *x, void *y,
                        hipStream_t s)
{
  if (dtype == MFP_F32) return adaptive_t<float>(g, (const float *)a, (const float *)x, (float *)y, s);
  return adaptive_t<double>(g, (const double *)a, (const double *)x, (double *)y, s);
}

int launch_ell_pk(int dtype, uint32_t rows, uint32_t w, uint32_t w4, const uint32_t *base, const void *val,
                  const uint16_t *off, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const dim3 grid(ceil_div_u32(rows, 256));
  if (dtype == MFP_F32)
    return launch_k(k_ell_pk<float>, grid, dim3(256), 0, s, rows, w, w4, base, (const f32x4 *)val,
                    (const pk_u16x4 *)off, (const float *)x, (float *)y);
  return launch_k(k_ell_pk<double>, grid, dim3(256), 0, s, rows, w, w4, base, (const f64x2 *)val,
                  (const pk_u16x4 *)off, (const double *)x, (double *)y);
}

int ell_pk_build(int dtype, uint32_t rows, uint32_t stride, const uint32_t *rp, const uint32_t *ci, const void *a,
                 uint32_t w4, const uint32_t *base, void *val, uint16_t *off, hipStream_t s)
{
  if (stride == 0) return MFP_OK;
  const uint32_t g = ceil_div_u32(stride, 256);
  if (dtype == MFP_F32)
    k_ell_pk_build<float><<<g, 256, 0, s>>>(rows, stride, rp, ci, (const float *)a, w4, base, (float *)val, off);
  else
    k_ell_pk_build<double><<<g, 256, 0, s>>>(rows, stride, rp, ci, (const double *)a, w4, base, (double *)val, off);
  return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("packed ELL build failed"), MFP_ERR_HIP);
}

int launch_ell(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
               const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32(rows, 256);
  if (dtype == MFP_F32)
    return launch_k(k_ell<float>, dim3(grid), dim3(256), 0, s, rows, w, ec, (const float *)ev, (const float *)x,
                    (float *)y);
  return launch_k(k_ell<double>, dim3(grid), dim3(256), 0, s, rows, w, ec, (const double *)ev, (const double *)x,
                  (double *)y);
}

int launch_vec_add(int dtype, uint32_t n, const void *t, void *y, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32(n, 256);
  if (dtype == MFP_F32)
    return launch_k(k_vec_add<float>, dim3(grid), dim3(256), 0, s, n, (const float *)t, (float *)y);
  return launch_k(k_vec_add<double>, dim3(grid), dim3(256), 0, s, n, (const double *)t, (double *)y);
}

static uint32_t coo_grid(uint64_t n)
{
  const uint64_t per_block = 4u * 64u * COO_TILES; // 4 waves x COO_TILES tiles per pass
  const uint64_t tiles = (n + per_block - 1) / per_block;
  const uint64_t cap = (uint64_t)std::max(1, props().cu_count) * 16u;
  return (uint32_t)std::max<uint64_t>(1, std::min(tiles, cap));
}

int launch_coo_add(int dtype, uint64_t n, const uint32_t *cr, const uint32_t *cc, const void *cv,
                   const void *x, void *y, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  const uint32_t grid = coo_grid(n);
  if (dtype == MFP_F32)
    return launch_k(k_coo<float>, dim3(grid), dim3(256), 0, s, n, cr, cc, (const float *)cv, (const float *)x,
                    (float *)y);
  return launch_k(k_coo<double>, dim3(grid), dim3(256), 0, s, n, cr, cc, (const double *)cv, (const double *)x,
                  (double *)y);
}

int launch_hyb_atomic(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
                      uint64_t ncoo, const uint32_t *cr, const uint32_t *cc, const void *cv,
                      const void *x, void *y, hipStream_t s)
{
  const uint32_t eb = w ? ceil_div_u32(rows, 256) : 0u;
  const uint32_t cb = ncoo ? coo_grid(ncoo) : 0u;
  if (eb + cb == 0) return MFP_OK;
  if (dtype == MFP_F32)
    return launch_k(k_hyb_atomic<float>, dim3(eb + cb), dim3(256), 0, s, rows, w, ec, (const float *)ev, ncoo, cr,
                    cc, (const float *)cv, (const float *)x, (float *)y, eb);
  return launch_k(k_hyb_atomic<double>, dim3(eb + cb), dim3(256), 0, s, rows, w, ec, (const double *)ev, ncoo, cr,
                  cc, (const double *)cv, (const double *)x, (double *)y, eb);
}

int launch_scoo(int dtype, uint32_t rows, uint32_t ss, uint32_t L, uint32_t threads,
                const uint32_t *cr, const uint32_t *cc, const void *cv, const uint32_t *sp,
                const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  if (ss == 0 || L == 0 || (L & (L - 1)) != 0) {
    set_error("SCOO geometry: slice_size > 0 and a power-of-two lane_size are required");
    return MFP_ERR_ARG;
  }
  if (threads == 0 || threads > 1024 || threads % 64 != 0) {
    set_error("SCOO workgroup size must be a multiple of 64 up to 1024");
    return MFP_ERR_ARG;
  }
  const size_t lds = (size_t)ss * L * dtype_size(dtype);
  if (lds > 160u * 1024u) {
    set_error("SCOO slice_size * lane_size exceeds the 160 KB LDS of a workgroup");
    return MFP_ERR_ARG;
  }
  const uint32_t real = ceil_div_u32(rows, ss);
  auto go = [&](auto kern, size_t bytes, auto tv) -> int {
    using TV = decltype(tv);
    if (bytes > 65536)
      MFP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return launch_k(kern, dim3(real), dim3(threads), (uint32_t)bytes, s, rows, ss, L, cr, cc, (const TV *)cv, sp,
                    (const TV *)x, (TV *)y);
  };
  if (dtype == MFP_F64)
    MFP_CHECK(go(k_scoo<double, double>, lds, 0.0));
  else if (2 * lds <= 160u * 1024u)
    MFP_CHECK(go(k_scoo<float, double>, 2 * lds, 0.0f));
  else
    MFP_CHECK(go(k_scoo<float, float>, lds, 0.0f));
  return MFP_OK;
}

template <typename T, int DT>
static int scoo_cs_t(uint32_t rows, const scoo_cs_layout &l, uint32_t threads, const void *x, void *y,
                     hipStream_t s)
{
  using V = typename cs_values<T, DT>::V;
  const size_t lds = (size_t)l.R * sizeof(double);
  if (lds > 65536)
    MFP_HIP(hipFuncSetAttribute((const void *)k_scoo_cs<T, DT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  MFP_CHECK(launch_k(k_scoo_cs<T, DT>, dim3(l.NI), dim3(threads), (uint32_t)lds, s, rows, l.R, l.rbits, l.S_reg,
                     l.NL, l.items, l.base, (const V *)l.ev, l.ec, (const T *)x, (T *)y, (T *)l.partial,
                     l.long_rows, l.long_mask));
  if (l.NF)
    MFP_CHECK(launch_k(k_scoo_cs_finish<T>, dim3(l.NF), dim3(CS_FIN_ROWS), 0, s, rows, l.R, l.S_reg, l.NL, l.NFR,
                       l.fin, l.long_rows, l.long_mask, (const T *)l.partial, (T *)y));
  return MFP_OK;
}

template <typename T, int KV>
static int scoo_cs_k_t(uint32_t rows, const scoo_cs_layout &l, uint32_t threads, const void *X, void *Y,
                       hipStream_t s)
{
  const size_t lds = (size_t)l.R * KV * sizeof(double);
  if (lds > 65536)
    MFP_HIP(hipFuncSetAttribute((const void *)k_scoo_cs_k<T, KV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  k_scoo_cs_k<T, KV><<<l.NI, threads, lds, s>>>(rows, l.R, l.rbits, l.S_reg, l.NL, l.items, l.base,
                                                (const T *)l.ev, l.ec, (const T *)X, (T *)Y, (T *)l.partial,
                                                l.long_rows, l.long_mask);
  MFP_CHECK(launched());
  if (l.NF) {
    k_scoo_cs_finish_k<T><<<l.NF, CS_FIN_ROWS, 0, s>>>(rows, l.R, KV, l.S_reg, l.NL, l.NFR, l.fin, l.long_rows,
                                                        l.long_mask, (const T *)l.partial, (T *)Y);
    MFP_CHECK(launched());
  }
  return MFP_OK;
}

int launch_scoo_cs_spmm(int dtype, uint32_t k, uint32_t rows, const scoo_cs_layout &l, uint32_t threads,
                        const void *X, void *Y, hipStream_t s)
{
  if (dtype != MFP_F32 && dtype != MFP_F64) {
    set_error("column-sorted SpMM values must be f32 or f64");
    return MFP_ERR_ARG;
  }
  if (k > l.kmax) {
    set_error("column-sorted SpMM: slice_size * k doubles exceed the 160 KB LDS (use smaller slices)");
    return MFP_ERR_ARG;
  }
  if ((k * dtype_size(dtype)) % 16 == 0 && ((uintptr_t)X % 16) != 0) {
    set_error("X must be 16-byte aligned when its rows are a multiple of 16 bytes");
    return MFP_ERR_ARG;
  }
  if (threads == 0 || threads > 1024 || threads % 64) {
    set_error("SCOO workgroup size must be a multiple of 64 up to 1024");
    return MFP_ERR_ARG;
  }
  if (rows == 0 || l.NI == 0) return MFP_OK;
  const bool f32 = dtype == MFP_F32;
  switch (k) {
  case 2: return f32 ? scoo_cs_k_t<float, 2>(rows, l, threads, X, Y, s) : scoo_cs_k_t<double, 2>(rows, l, threads, X, Y, s);
  case 4: return f32 ? scoo_cs_k_t<float, 4>(rows, l, threads, X, Y, s) : scoo_cs_k_t<double, 4>(rows, l, threads, X, Y, s);
  case 8: return f32 ? scoo_cs_k_t<float, 8>(rows, l, threads, X, Y, s) : scoo_cs_k_t<double, 8>(rows, l, threads, X, Y, s);
  case 16: return f32 ? scoo_cs_k_t<float, 16>(rows, l, threads, X, Y, s) : scoo_cs_k_t<double, 16>(rows, l, threads, X, Y, s);
  }
  set_error("k must be 1, 2, 4, 8 or 16");
  return MFP_ERR_ARG;
}

int launch_scoo_cs(int dtype, uint32_t rows, const scoo_cs_layout &l, uint32_t threads, const void *x, void *y,
                   hipStream_t s)
{
  if (rows == 0 || l.NI == 0) return MFP_OK;
  const size_t lds = (size_t)l.R * sizeof(double);
  if (lds > 163840) {
    set_error("column-sorted SCOO slice exceeds the LDS of a workgroup");
    return MFP_ERR_ARG;
  }
  if (threads == 0 || threads > 1024 || threads % 64) {
    set_error("SCOO workgroup size must be a multiple of 64 up to 1024");
    return MFP_ERR_ARG;
  }
  switch (dtype) {
  case MFP_F32: return scoo_cs_t<float, -1>(rows, l, threads, x, y, s);
  case MFP_F64: return scoo_cs_t<double, -1>(rows, l, threads, x, y, s);
  case MFP_BF16: return scoo_cs_t<float, MFP_BF16>(rows, l, threads, x, y, s);
  case MFP_F16: return scoo_cs_t<float, MFP_F16>(rows, l, threads, x, y, s);
  }
  set_error("column-sorted SCOO: unknown value type");
  return MFP_ERR_ARG;
}

} // namespace mfp
