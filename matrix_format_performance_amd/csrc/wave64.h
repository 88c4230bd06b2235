// Wave64 cross-lane primitives for gfx950 (CDNA4).
//
// The reference reduces with a 32-lane __shfl_down_sync tree
// (gpu/reduce.cuh:6-20).  Here a wavefront has 64 lanes and the moves are
// DPP row shifts (offsets 1..8, inside a 16-lane row), ds_swizzle SWAP16
// (offset 16, inside 32 lanes) and v_permlane32_swap (offset 32).
//
// group_reduce<G>: lanes are split in aligned groups of G (power of two,
// <= 64).  On return the first lane of each group holds the halving-tree sum
//   v[i] += v[i + off]   for off = G/2, G/4, ..., 1
// which is the order of the reference's shfl-down tree for G == 32.  Every
// lane of the wave must call it (no divergence around the call).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfp {

// value of lane (i + OFF) for the lanes that need it (see group_reduce).
template <int OFF>
__device__ __forceinline__ uint32_t lane_down_u32(uint32_t v)
{
  static_assert(OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8 || OFF == 16 || OFF == 32,
                "offset");
  if constexpr (OFF <= 8) {
    // DPP row_shl:OFF -- lane i reads lane i+OFF within its 16-lane row.
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + OFF, 0xF, 0xF, true);
  } else if constexpr (OFF == 16) {
    // ds_swizzle bit mode, xor_mask 16 inside each 32-lane half.
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
  } else {
    // v_permlane32_swap: the low half of the second result holds lanes 32..63.
    auto pr = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return pr[1];
  }
}

template <int OFF>
__device__ __forceinline__ float lane_down(float v)
{
  return __uint_as_float(lane_down_u32<OFF>(__float_as_uint(v)));
}

template <int OFF>
__device__ __forceinline__ double lane_down(double v)
{
  uint64_t b = (uint64_t)__double_as_longlong(v);
  uint32_t lo = lane_down_u32<OFF>((uint32_t)b);
  uint32_t hi = lane_down_u32<OFF>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int OFF>
__device__ __forceinline__ uint32_t lane_down(uint32_t v)
{
  return lane_down_u32<OFF>(v);
}

template <int G, typename T>
__device__ __forceinline__ T group_reduce(T v)
{
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group size");
  if constexpr (G >= 64) v = v + lane_down<32>(v);
  if constexpr (G >= 32) v = v + lane_down<16>(v);
  if constexpr (G >= 16) v = v + lane_down<8>(v);
  if constexpr (G >= 8) v = v + lane_down<4>(v);
  if constexpr (G >= 4) v = v + lane_down<2>(v);
  if constexpr (G >= 2) v = v + lane_down<1>(v);
  return v;
}

// Runtime group size (t in {1,2,4,...,64}); uniform across the wave.
template <typename T>
__device__ __forceinline__ T group_reduce_rt(T v, uint32_t t)
{
  switch (t) {
  case 64: return group_reduce<64>(v);
  case 32: return group_reduce<32>(v);
  case 16: return group_reduce<16>(v);
  case 8: return group_reduce<8>(v);
  case 4: return group_reduce<4>(v);
  case 2: return group_reduce<2>(v);
  default: return v;
  }
}

// Workgroup ids are dispatched round-robin over the XCDs (each with its own
// L2).  Remap so that XCD k works on one contiguous range of logical blocks:
// neighbouring rows (and the x entries they share) stay in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
  constexpr uint32_t X = 8;
  if (nb < X * 4) return b;
  const uint32_t xcd = b % X, idx = b / X, q = nb / X, rem = nb % X;
  return xcd < rem ? xcd * (q + 1) + idx : rem * (q + 1) + (xcd - rem) * q + idx;
}

// A row of X (KV values, aligned to KV * sizeof(T)) with 16-byte loads where
// the row is a multiple of 16 bytes.
template <typename T, int KV>
__device__ __forceinline__ void load_row(const T *__restrict__ p, T (&out)[KV])
{
  if constexpr ((KV * sizeof(T)) % 16 == 0) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    constexpr int N = KV * sizeof(T) / 16;
    constexpr int PER = 16 / sizeof(T);
#pragma unroll
    for (int v = 0; v < N; v++) {
      const u4 w = *reinterpret_cast<const u4 *>(p + v * PER);
      T tmp[PER];
      __builtin_memcpy(tmp, &w, 16);
#pragma unroll
      for (int j = 0; j < PER; j++) out[v * PER + j] = tmp[j];
    }
  } else {
#pragma unroll
    for (int j = 0; j < KV; j++) out[j] = p[j];
  }
}

template <typename T>
__device__ __forceinline__ T ldg_stream(const T *p)
{
  return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

} // namespace mfp
