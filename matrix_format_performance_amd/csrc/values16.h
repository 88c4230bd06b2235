// 16-bit value storage (BF16 / FP16) shared by the kernels that read it.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mfp.h"

namespace mfp {

__device__ __forceinline__ float lp_to_f32_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float lp_to_f32_f16(uint16_t h) { return __half2float(__ushort_as_half(h)); }

template <int DT>
__device__ __forceinline__ float lp_val(uint16_t v)
{
  if constexpr (DT == MFP_BF16)
    return lp_to_f32_bf16(v);
  else
    return lp_to_f32_f16(v);
}

} // namespace mfp
