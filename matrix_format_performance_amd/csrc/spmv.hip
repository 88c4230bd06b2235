// SpMV kernels for gfx950: CSR-scalar, CSR-vector, CSR-Adaptive, ELL, COO,
// SCOO and the two hybrid ELL+COO forms.
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
#include "wave64.h"

namespace mfp {

// ---------------------------------------------------------------------------
// CSR-scalar: one lane per row, as gpu_matrix_multiplier.cu:22-43, but the
// wave stages its rows' contiguous nnz range through LDS with coalesced loads
// so each lane then walks only its own row.  Order: sequential fma.
// ---------------------------------------------------------------------------
constexpr uint32_t SCALAR_CHUNK = 512;

template <typename T>
__global__ __launch_bounds__(64) void k_csr_scalar(uint32_t rows, const uint32_t *__restrict__ rp,
                                                   const uint32_t *__restrict__ ci,
                                                   const T *__restrict__ a, const T *__restrict__ x,
                                                   T *__restrict__ y)
{
  __shared__ T va[SCALAR_CHUNK];
  __shared__ T vx[SCALAR_CHUNK];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t r0 = blk * 64u;
  const uint32_t lane = threadIdx.x;
  const uint32_t r = r0 + lane;
  const uint32_t rend = min(r0 + 64u, rows);
  const uint32_t rs = rp[min(r, rows)];
  const uint32_t re = rp[min(r + 1u, rows)];
  const uint32_t b = rp[r0], e = rp[rend];
  T dot = 0;
  for (uint32_t c0 = b; c0 < e; c0 += SCALAR_CHUNK) {
    const uint32_t n = min(SCALAR_CHUNK, e - c0);
    for (uint32_t j = lane; j < n; j += 64u) {
      va[j] = ldg_stream(a + c0 + j);
      vx[j] = x[ldg_stream(ci + c0 + j)];
    }
    __syncthreads();
    const uint32_t lo = max(rs, c0), hi = min(re, c0 + n);
    for (uint32_t i = lo; i < hi; i++) dot = fma_t(va[i - c0], vx[i - c0], dot);
    __syncthreads();
  }
  if (r < rows) y[r] = dot;
}

// ---------------------------------------------------------------------------
// CSR-vector: G lanes per row (G = 2..64; the reference uses a 32-lane warp,
// gpu_matrix_multiplier.cu:331-362).  Lane l accumulates elements l, l+G, ...
// with fma, then group_reduce<G> (halving tree).
// ---------------------------------------------------------------------------
template <typename T, int G>
__global__ __launch_bounds__(256) void k_csr_vector(uint32_t rows, const uint32_t *__restrict__ rp,
                                                    const uint32_t *__restrict__ ci,
                                                    const T *__restrict__ a, const T *__restrict__ x,
                                                    T *__restrict__ y)
{
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t row = (blk * 256u + threadIdx.x) / G;
  const uint32_t lane = threadIdx.x % G;
  T dot = 0;
  if (row < rows) {
    const uint32_t b = rp[row], e = rp[row + 1];
    for (uint32_t i = b + lane; i < e; i += G) dot = fma_t(ldg_stream(a + i), x[ldg_stream(ci + i)], dot);
  }
  dot = group_reduce<G>(dot);
  if (lane == 0 && row < rows) y[row] = dot;
}

// ---------------------------------------------------------------------------
// CSR-Adaptive (csr_adaptive_spmv.cu:26-178 redesigned for 256-lane
// workgroups and LDS-sized row blocks).
//   multi-row block  (CSR-Stream): rounded products a*x staged in LDS (up to
//     nnz_per_wg of them), then each row summed by t lanes strided by t
//     (t = min(64, prev_pow2(256 / rows))) and a halving tree; t == 1 sums
//     each row sequentially.
//   single-row block (CSR-Vector/VectorL): 256 lanes strided with fma, a
//     64-lane halving tree per wave, then a halving tree over the 4 waves.
//   rows longer than long_row: split into `chunk`-element pieces, one
//     workgroup each; the last piece to finish adds the partials left to
//     right (ticket counter; it resets the counter for the next call).
// ---------------------------------------------------------------------------
constexpr uint32_t AD_T = 256;

template <typename T>
__device__ __forceinline__ T wg_row_sum(const uint32_t *__restrict__ ci, const T *__restrict__ a,
                                        const T *__restrict__ x, uint32_t b, uint32_t e, T *wsum)
{
  T d = 0;
  for (uint32_t i = b + threadIdx.x; i < e; i += AD_T) d = fma_t(ldg_stream(a + i), x[ldg_stream(ci + i)], d);
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
__global__ __launch_bounds__(AD_T) void k_csr_adaptive(
    uint32_t rows, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
    const T *__restrict__ a, const T *__restrict__ x, T *__restrict__ y,
    const uint32_t *__restrict__ blocks, uint32_t nblocks, uint32_t long_row,
    const uint4 *__restrict__ chunks, uint32_t chunk, T *__restrict__ partials,
    uint32_t *__restrict__ tickets)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *cache = reinterpret_cast<T *>(smem_raw);
  __shared__ T wsum[4];
  const uint32_t item = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t tid = threadIdx.x;

  if (item >= nblocks) {
    const uint4 c4 = chunks[item - nblocks]; // {row, piece, first item, pieces}
    const uint32_t s = rp[c4.x], e = rp[c4.x + 1];
    const uint32_t cb = s + c4.y * chunk;
    const uint32_t ce = min(e, cb + chunk);
    const T p = wg_row_sum(ci, a, x, cb, ce, wsum);
    if (tid == 0) {
      partials[c4.z + c4.y] = p;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const uint32_t old = __hip_atomic_fetch_add(tickets + c4.z, 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      if (old == c4.w - 1u) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        T acc = __hip_atomic_load(partials + c4.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (uint32_t i = 1; i < c4.w; i++)
          acc = acc + __hip_atomic_load(partials + c4.z + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        y[c4.x] = acc;
        __hip_atomic_store(tickets + c4.z, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }

  const uint32_t r0 = blocks[item], r1 = blocks[item + 1];
  if (r1 <= r0) return;
  if (r1 - r0 == 1) {
    const uint32_t s = rp[r0], e = rp[r0 + 1];
    if (e - s > long_row) return; // handled by the chunk items
    const T d = wg_row_sum(ci, a, x, s, e, wsum);
    if (tid == 0) y[r0] = d;
    return;
  }

  // CSR-Stream
  const uint32_t base = rp[r0];
  const uint32_t n = rp[r1] - base;
  for (uint32_t i = tid; i < n; i += AD_T) cache[i] = ldg_stream(a + base + i) * x[ldg_stream(ci + base + i)];
  __syncthreads();
  const uint32_t nr = r1 - r0;
  const uint32_t q = AD_T / nr;
  uint32_t t = 1;
  if (q >= 2) {
    t = 1u << (31 - __clz(q));
    if (t > 64) t = 64;
  }
  if (t == 1) {
    for (uint32_t lr = tid; lr < nr; lr += AD_T) {
      const uint32_t f = rp[r0 + lr] - base, l = rp[r0 + lr + 1] - base;
      T d = 0;
      for (uint32_t j = f; j < l; j++) d = d + cache[j];
      y[r0 + lr] = d;
    }
  } else {
    const uint32_t lr = tid / t, k = tid % t;
    T d = 0;
    if (lr < nr) {
      const uint32_t f = rp[r0 + lr] - base, l = rp[r0 + lr + 1] - base;
      for (uint32_t j = f + k; j < l; j += t) d = d + cache[j];
    }
    d = group_reduce_rt(d, t);
    if (k == 0 && lr < nr) y[r0 + lr] = d;
  }
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
  uint32_t k = 0;
  for (; k + 4 <= w; k += 4) {
    const T a0 = ldg_stream(ev + o), a1 = ldg_stream(ev + o + rows);
    const T a2 = ldg_stream(ev + o + 2ull * rows), a3 = ldg_stream(ev + o + 3ull * rows);
    const uint32_t c0 = ldg_stream(ec + o), c1 = ldg_stream(ec + o + rows);
    const uint32_t c2 = ldg_stream(ec + o + 2ull * rows), c3 = ldg_stream(ec + o + 3ull * rows);
    d = fma_t(a0, x[c0], d);
    d = fma_t(a1, x[c1], d);
    d = fma_t(a2, x[c2], d);
    d = fma_t(a3, x[c3], d);
    o += 4ull * rows;
  }
  for (; k < w; k++, o += rows) d = fma_t(ldg_stream(ev + o), x[ldg_stream(ec + o)], d);
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

// ---------------------------------------------------------------------------
// COO (gpu_matrix_multiplier.cu:622-638): a wave takes 64 consecutive
// entries (row-sorted); a segmented suffix sum over equal row keys leaves
// each run's total in its first lane, which issues one atomic add.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void coo_tile(uint64_t base, uint64_t n, const uint32_t *__restrict__ cr,
                                         const uint32_t *__restrict__ cc, const T *__restrict__ cv,
                                         const T *__restrict__ x, T *__restrict__ y)
{
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t i = base + lane;
  const bool valid = i < n;
  const uint32_t r = valid ? ldg_stream(cr + i) : 0xFFFFFFFFu;
  T p = valid ? ldg_stream(cv + i) * x[ldg_stream(cc + i)] : (T)0;
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
__global__ __launch_bounds__(256) void k_coo(uint64_t n, const uint32_t *__restrict__ cr,
                                             const uint32_t *__restrict__ cc,
                                             const T *__restrict__ cv, const T *__restrict__ x,
                                             T *__restrict__ y)
{
  const uint64_t waves = (uint64_t)gridDim.x * 4u;
  uint64_t w = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
  for (; w * 64u < n; w += waves) coo_tile(w * 64u, n, cr, cc, cv, x, y);
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
  const uint64_t waves = (uint64_t)(gridDim.x - ell_blocks) * 4u;
  uint64_t wv = (uint64_t)(b - ell_blocks) * 4u + (threadIdx.x >> 6);
  for (; wv * 64u < ncoo; wv += waves) coo_tile(wv * 64u, ncoo, cr, cc, cv, x, y);
}

// ---------------------------------------------------------------------------
// SCOO (scoo_spmv.cu:46-129): one workgroup per slice of slice_size rows.
// Products are added with LDS atomics (ds_add) into acc[row][lane], lane =
// thread % lane_size, then each row's lanes are summed in order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_scoo(uint32_t rows, uint32_t ss, uint32_t L, const uint32_t *__restrict__ cr,
                       const uint32_t *__restrict__ cc, const T *__restrict__ cv,
                       const uint32_t *__restrict__ sp, const T *__restrict__ x,
                       T *__restrict__ y)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *acc = reinterpret_cast<T *>(smem_raw);
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t r0 = s * ss;
  const uint32_t lim = min(ss, rows - r0);
  const uint32_t bt = blockDim.x, tid = threadIdx.x;
  for (uint32_t i = tid; i < lim * L; i += bt) acc[i] = (T)0;
  __syncthreads();
  const uint32_t b = sp[s], e = sp[s + 1];
  const uint32_t lane = tid & (L - 1u);
  for (uint32_t i = b + tid; i < e; i += bt) {
    const uint32_t lr = ldg_stream(cr + i) - r0;
    const T p = ldg_stream(cv + i) * x[ldg_stream(cc + i)];
    if (lr < lim) atomicAdd(acc + lr * L + lane, p);
  }
  __syncthreads();
  for (uint32_t lr = tid; lr < lim; lr += bt) {
    T d = 0;
    for (uint32_t l = 0; l < L; l++) d = d + acc[lr * L + l];
    y[r0 + lr] = d;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int launched() { return hipGetLastError() == hipSuccess ? MFP_OK : (set_error("kernel launch failed"), MFP_ERR_HIP); }

template <typename T>
static int csr_vector_t(uint32_t lanes, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                        const T *a, const T *x, T *y, hipStream_t s)
{
  const uint32_t grid = ceil_div_u32((uint64_t)rows * lanes, 256);
  switch (lanes) {
  case 2: k_csr_vector<T, 2><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  case 4: k_csr_vector<T, 4><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  case 8: k_csr_vector<T, 8><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  case 16: k_csr_vector<T, 16><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  case 32: k_csr_vector<T, 32><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  case 64: k_csr_vector<T, 64><<<grid, 256, 0, s>>>(rows, rp, ci, a, x, y); break;
  default: set_error("lanes must be a power of two in [2, 64]"); return MFP_ERR_ARG;
  }
  return launched();
}

int launch_csr_scalar(int dtype, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                      const void *a, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32(rows, 64);
  if (dtype == MFP_F32)
    k_csr_scalar<float><<<grid, 64, 0, s>>>(rows, rp, ci, (const float *)a, (const float *)x, (float *)y);
  else
    k_csr_scalar<double><<<grid, 64, 0, s>>>(rows, rp, ci, (const double *)a, (const double *)x, (double *)y);
  return launched();
}

int launch_csr_vector(int dtype, uint32_t lanes, uint32_t rows, const uint32_t *rp,
                      const uint32_t *ci, const void *a, const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  if (dtype == MFP_F32)
    return csr_vector_t<float>(lanes, rows, rp, ci, (const float *)a, (const float *)x, (float *)y, s);
  return csr_vector_t<double>(lanes, rows, rp, ci, (const double *)a, (const double *)x, (double *)y, s);
}

int launch_csr_adaptive(int dtype, const adaptive_args &g, const void *a, const void *x, void *y,
                        hipStream_t s)
{
  const uint32_t grid = g.nblocks + g.nchunks;
  if (grid == 0) return MFP_OK;
  const size_t lds = (size_t)g.nnz_per_wg * dtype_size(dtype);
  if (lds > 65536) {
    const void *fn = dtype == MFP_F32 ? (const void *)k_csr_adaptive<float> : (const void *)k_csr_adaptive<double>;
    MFP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  if (dtype == MFP_F32)
    k_csr_adaptive<float><<<grid, AD_T, lds, s>>>(g.rows, g.rp, g.ci, (const float *)a, (const float *)x,
                                                  (float *)y, g.blocks, g.nblocks, g.long_row,
                                                  g.chunks, g.chunk, (float *)g.partials, g.tickets);
  else
    k_csr_adaptive<double><<<grid, AD_T, lds, s>>>(g.rows, g.rp, g.ci, (const double *)a, (const double *)x,
                                                   (double *)y, g.blocks, g.nblocks, g.long_row,
                                                   g.chunks, g.chunk, (double *)g.partials, g.tickets);
  return launched();
}

int launch_ell(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
               const void *x, void *y, hipStream_t s)
{
  if (rows == 0) return MFP_OK;
  const uint32_t grid = ceil_div_u32(rows, 256);
  if (dtype == MFP_F32)
    k_ell<float><<<grid, 256, 0, s>>>(rows, w, ec, (const float *)ev, (const float *)x, (float *)y);
  else
    k_ell<double><<<grid, 256, 0, s>>>(rows, w, ec, (const double *)ev, (const double *)x, (double *)y);
  return launched();
}

static uint32_t coo_grid(uint64_t n)
{
  const uint64_t tiles = (n + 255) / 256; // 4 waves x 64 entries per block pass
  const uint64_t cap = (uint64_t)std::max(1, props().cu_count) * 16u;
  return (uint32_t)std::max<uint64_t>(1, std::min(tiles, cap));
}

int launch_coo_add(int dtype, uint64_t n, const uint32_t *cr, const uint32_t *cc, const void *cv,
                   const void *x, void *y, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  const uint32_t grid = coo_grid(n);
  if (dtype == MFP_F32)
    k_coo<float><<<grid, 256, 0, s>>>(n, cr, cc, (const float *)cv, (const float *)x, (float *)y);
  else
    k_coo<double><<<grid, 256, 0, s>>>(n, cr, cc, (const double *)cv, (const double *)x, (double *)y);
  return launched();
}

int launch_hyb_atomic(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
                      uint64_t ncoo, const uint32_t *cr, const uint32_t *cc, const void *cv,
                      const void *x, void *y, hipStream_t s)
{
  const uint32_t eb = w ? ceil_div_u32(rows, 256) : 0u;
  const uint32_t cb = ncoo ? coo_grid(ncoo) : 0u;
  if (eb + cb == 0) return MFP_OK;
  if (dtype == MFP_F32)
    k_hyb_atomic<float><<<eb + cb, 256, 0, s>>>(rows, w, ec, (const float *)ev, ncoo, cr, cc,
                                                (const float *)cv, (const float *)x, (float *)y, eb);
  else
    k_hyb_atomic<double><<<eb + cb, 256, 0, s>>>(rows, w, ec, (const double *)ev, ncoo, cr, cc,
                                                 (const double *)cv, (const double *)x, (double *)y, eb);
  return launched();
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
  if (dtype == MFP_F32) {
    if (lds > 65536) MFP_HIP(hipFuncSetAttribute((const void *)k_scoo<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_scoo<float><<<real, threads, lds, s>>>(rows, ss, L, cr, cc, (const float *)cv, sp, (const float *)x, (float *)y);
  } else {
    if (lds > 65536) MFP_HIP(hipFuncSetAttribute((const void *)k_scoo<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_scoo<double><<<real, threads, lds, s>>>(rows, ss, L, cr, cc, (const double *)cv, sp, (const double *)x, (double *)y);
  }
  return launched();
}

} // namespace mfp
